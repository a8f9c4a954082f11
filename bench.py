#!/usr/bin/env python3
"""Throughput benchmark of the TWMC hot path on one MI355X.

Workload: 2D 64x64 periodic Bose-Hubbard lattice, uniform hopping through the
FFT (J=1, U=0.05, detuning omega=1, F=2, gamma=1, dt=0.002), 16384
trajectories, FP32 state with FP64 ensemble sums, vacuum Wigner start.  One
benchmark step is one stored frame of that run: 500 Euler-Maruyama steps of
all 16384 trajectories plus the frame's ensemble reduction (the full run is
100 such steps).  The state lives in HBM for the whole run; the timed region
holds only kernel work.

Prints one JSON line with the metric (trajectory.site-steps/s), the roofline
of the step kernel (HIP-event average launch duration, FP32 VALU bound), the
time to steady state of the plain (vacuum) start vs the learned warm start on
a few held-out (F, Delta) points of the same lattice (--ws-points; the full
64-point evaluation is --workload warmstart64), and a CPU baseline: the oracle
in the reference's own per-step structure (FP64, std::mt19937 +
normal_distribution), threaded over this host's cores, on a bounded sample of
the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "trajectory·site-steps/s (2D 64×64, 1 GPU); time-to-steady-state plain vs warm"
PEAK_FP32 = 157.3e12           # MI355X FP32 vector peak (FLOP/s)
PEAK_BF16 = 2.5e15             # MI355X dense BF16 MFMA peak (FLOP/s)
PEAK_FP64 = 78.6e12            # MI355X FP64 vector peak (FLOP/s)

# square64: the headline workload (default).  lieb384: the disordered Lieb
# chain with matrix hopping as the dense MFMA product (the banded kernel timed
# beside it), measured with --workload lieb384.
N_TRAJ = 16384
FRAME_STEPS = 500
FLOPS_PER_SITE_STEP = 150.0    # 2 FFTs x 5 N log2 N / N + 30 elementwise, N = 4096
SPEC = dict(nx=64, ny=64, dt=0.002, gamma=1.0, J=1.0, U=0.05, omega=1.0, F=2.0, beta_init=0.0,
            beta_sigma=0.5, frame_steps=FRAME_STEPS, seed=20260101, t_start=0.0)
LIEB_CELLS, LIEB_TRAJ, LIEB_FRAME = 128, 8192, 200


def lieb_spec():
    from pincosimulator_amd.lattices import lieb_coupling, onsite_disorder
    n = 3 * LIEB_CELLS
    C = lieb_coupling(LIEB_CELLS, 1.0, 1.0, pbc=False, sigma=0.1, seed=7)
    E = onsite_disorder(n, 0.2, 7)
    return dict(nx=1, ny=n, dt=0.005, gamma=1.0, J=0.0, U=0.1, F=1.5, omega=E, beta_init=0.0,
                beta_sigma=0.5, frame_steps=LIEB_FRAME, seed=20260102, t_start=0.0, lattice="matrix",
                coupling=C, detuning=0.0, onsite="detuning-E")


WORKLOAD_TEXT = {
    "square64": "2D 64x64 PBC uniform-hopping TWMC, J=1 U=0.05 omega=1 F=2 gamma=1 dt=0.002, "
                "16384 trajectories, vacuum Wigner start; one step = one 500-step frame "
                "incl. FP64 ensemble reduction",
    "lieb384": "1D Lieb chain 128 cells x 3 = 384 sites OBC, J=1 with bond disorder 0.1 and on-site "
               "disorder 0.2 (one realisation, seed 7), U=0.1 F=1.5 gamma=1 dt=0.005, 8192 trajectories, "
               "hopping as a dense 384x384 BF16x3 MFMA GEMM (banded_kernel: the same product as the half-width-2 "
               "band stencil); one step = one 200-step frame",
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(threads, spec=None, target_s=12.0):
    import oracle
    spec = SPEC if spec is None else spec
    if spec.get("lattice") == "matrix":
        p = dict(spec, lattice=oracle.LAT_DENSE, E=spec["omega"])
    else:
        p = dict(spec, lattice=oracle.LAT_FFT)
    n_sites = spec["nx"] * spec["ny"]
    # pilot to size the sample to ~target_s of wall time
    n_traj_p, steps_p = threads, 5
    t = oracle.time_batch(p, n_traj_p, threads, steps_p, faithful=True)
    rate = n_traj_p * steps_p / t  # trajectory-steps per second
    n_traj = 4 * threads
    steps = max(5, int(target_s * rate / n_traj))
    t = oracle.time_batch(p, n_traj, threads, steps, faithful=True)
    value = n_traj * n_sites * steps / t
    kind = ("sparse coupling product" if spec.get("lattice") == "matrix"
            else "FFT hopping with the oracle's own FFT, FFTW not being installed")
    cpu = "unknown CPU"
    topo = ""
    try:
        # model and topology as lscpu reports them: sockets (physical ids),
        # cores per socket ("cpu cores"), threads per core (siblings / cores)
        with open("/proc/cpuinfo") as f:
            lines = f.read().splitlines()
        field = lambda name: [l.split(":", 1)[1].strip() for l in lines if l.split(":", 1)[0].strip() == name]
        cpu = next(iter(field("model name")), cpu)
        sockets = len(set(field("physical id"))) or 1
        cores = int(next(iter(field("cpu cores")), "0") or 0)
        siblings = int(next(iter(field("siblings")), "0") or 0)
        usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        if cores:
            topo = (f"{sockets} socket(s) x {cores} cores x {max(1, siblings // cores)} threads per core; "
                    f"{usable} CPUs usable by this process; ")
    except (OSError, ValueError):
        pass
    return {
        "value": value, "unit": "trajectory·site-steps/s", "cores": threads, "kind": "port",
        "sample": f"{n_traj} trajectories x {steps} steps of the same {n_sites}-site system "
                  f"({t:.1f} s wall), FP64, reference per-step structure ({kind}) with std::mt19937 "
                  f"noise, one trajectory per task on {threads} std::threads ({cpu}; {topo}{os.cpu_count()} logical CPUs "
                  f"visible, threads = min(16, logical CPUs))",
    }


def physical_cores(threads):
    """Physical cores among `threads` logical CPUs, from the threads per core
    that /proc/cpuinfo reports (siblings / cpu cores)."""
    try:
        with open("/proc/cpuinfo") as f:
            lines = f.read().splitlines()
        field = lambda name: [l.split(":", 1)[1].strip() for l in lines if l.split(":", 1)[0].strip() == name]
        cores = int(next(iter(field("cpu cores")), "0") or 0)
        siblings = int(next(iter(field("siblings")), "0") or 0)
        per_core = max(1, siblings // cores) if cores else 1
    except (OSError, ValueError):
        per_core = 1
    return max(1, threads // per_core)


def sweep_points():
    from pincosimulator_amd.workloads import sweep512_points
    return sweep512_points()


def sweep_traffic():
    """HBM bytes per step launch of the sweep kernel, from its counter profile."""
    try:
        return json.load(open(os.path.join(ROOT, "profiles", "sweep_512points_pmc.json")))["hbm_bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def bench_sweep(args):
    """One step = one 500-step frame of all 512 points (the full run is 60)."""
    from pincosimulator_amd import Sweep
    pts = sweep_points()
    sw = Sweep(pts, 0.002, 500, 20260103)
    t_at = time.perf_counter()
    tb = sw.batching() if args.no_autotune else sw.autotune(steps=50)
    if not args.no_autotune:
        log(f"bench: autotuned the stacking in {time.perf_counter() - t_at:.1f} s (outside the timed region)")
    log(f"bench: sweep stacking per shape: " + ", ".join(f"{pts[k]['nx']}x{pts[k]['ny']}:{tb[k]}" for k in range(9)))
    if args.warmup:
        sw.run(args.warmup * 500 * 0.002, graph=True)
    t0 = time.perf_counter()
    out = sw.run(args.steps * 500 * 0.002, graph=True)
    t1 = time.perf_counter()
    site_steps = sum(p["n_traj"] * p["nx"] * p["ny"] for p in pts) * 500.0 * args.steps
    wall = t1 - t0
    value = site_steps / wall
    log(f"bench: sweep {args.steps} frames: wall {wall:.3f} s, graph kernel time {out['kernel_ms']:.1f} ms, "
        f"{value:.4e} site-steps/s")
    # algorithmic work of each point's lattice (the reference's FFT hopping,
    # 2 x 5 N log2 N / N + 30 per site-step) over the graph's kernel time
    import math
    flops = sum(p["n_traj"] * p["nx"] * p["ny"] * (10.0 * math.log2(p["nx"] * p["ny"]) + 30.0)
                for p in pts) * 500.0 * args.steps
    achieved = flops / (out["kernel_ms"] * 1e-3)
    res = {
        "metric": METRIC, "value": value, "unit": "trajectory·site-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "512-point sweep, F in [0.5,4] x detuning in [-2,3], shapes round-robin "
                               "{1D 32..512, 2D 16..64}, 128 trajectories each, U=0.05 gamma=1 dt=0.002, "
                               "one persistent launch per 500-step frame, hipGraph frame loop",
                   "stacking": {f"{pts[k]['nx']}x{pts[k]['ny']}": tb[k] for k in range(9)},
                   "autotuned": not args.no_autotune,
                   "kernel_ms": out["kernel_ms"]},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32, "traffic": sweep_traffic(), "kernel": "sweep_step_kernel<false>",
                     "avg_launch_ms": out["kernel_ms"] / args.steps,
                     "flops_basis": "algorithmic: 10 log2 N + 30 flops per site-step of each point's lattice "
                                    "(the reference's FFT hopping), over the graph's kernel time per frame "
                                    "(step launch plus the frame's capture launch)"},
    }
    sw.close()
    print(json.dumps(res), flush=True)


def warmstart_summary(n_points, n_traj):
    """Time to steady state, cold (vacuum) vs learned warm start, on the first
    n_points held-out (F, Delta) points of the 64x64 lattice."""
    import numpy as np
    from pincosimulator_amd import warmstart as ws
    net = ws.WarmStart.load()
    F, D = ws.held_out_points(64)
    rows = []
    pick = np.unique(np.round(np.linspace(0, 63, n_points)).astype(int)) if n_points < 64 else np.arange(64)
    for k in pick:
        t0 = time.perf_counter()
        r = ws.evaluate_point(float(F[k]), float(D[k]), net, n_traj=n_traj, log=log)
        r["seconds"] = time.perf_counter() - t0
        rows.append(r)
    net.close()
    ok = [r for r in rows if r["cold_steps"] and r["warm_steps"] is not None]
    acc = [r for r in ok if r["accepted"]]
    # a rejected warm start falls back to the cold start (ratio 1)
    ratio = np.array([r["cold_steps"] / max(r["warm_steps"], 100) if r["accepted"] else 1.0 for r in ok])
    return {
        "points": len(rows), "point_indices": [int(k) for k in pick], "trajectories": n_traj,
        "accepted": f"{len(acc)}/{len(rows)}",
        "cold_steps_median": float(np.median([r["cold_steps"] for r in ok])) if ok else None,
        "warm_steps_median": float(np.median([r["warm_steps"] for r in acc])) if acc else None,
        "speedup_median": float(np.median(ratio)) if ok else None,
        "speedup_geomean": float(np.exp(np.log(ratio).mean())) if ok else None,
        "speedup_min": float(ratio.min()) if ok else None,
        "criterion": "steps until <n> and g2(0) enter and hold (10 frames) the 3 sigma band around the cold "
                     "run's t>=25 mean (frames every 100 steps; warm and cold runs use independent "
                     "trajectories; a warm run counts only if its t in [15,20] mean agrees "
                     "with the cold value within 3 sigma; otherwise ratio 1; a warm run already steady at its "
                     "first frame counts as one frame, 100 steps)",
        "network": "MLP (F, Delta, U, J, L) -> 3x256 GELU -> mean field + 32 radial n(|k|) bins, FP32, "
                   "trained by the library's HIP kernels on 4096 sweep steady states (seed 99)",
        "held_out": "(F, Delta) uniform over [0.5,4] x [-2,3] from Philox seed 100, 64x64, U=0.05, J=1; "
                    "point_indices are evenly spread over the 64 held-out points",
        "per_point": [{"F": r["F"], "D": r["D"], "cold_steps": r["cold_steps"], "warm_steps": r.get("warm_steps"),
                       "accepted": r.get("accepted")} for r in rows],
    }


def bench_warmstart(args):
    """Full held-out evaluation: value = median cold/warm time-to-steady-state ratio."""
    s = warmstart_summary(args.ws_points, args.ws_traj)
    res = {
        "metric": "time-to-steady-state, plain (vacuum) vs learned warm start (2D 64x64, 1 GPU)",
        "value": s["speedup_median"], "unit": "x (cold steps / warm steps, median over points)",
        "n_gpus": 1, "steps": args.ws_points, "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"{args.ws_points} held-out (F, Delta) points, 64x64 PBC, "
                               f"{args.ws_traj} trajectories each, dt=0.002", **s},
    }
    print(json.dumps(res), flush=True)


def time_frames(eng, n_traj, W, K):
    """W untimed frames, then K frames timed on the wall clock and with HIP
    events around each launch; the rest of the run (its final frame) runs
    outside the timed region.  Returns (engine, info, stats, wall seconds)."""
    import numpy as np
    info = eng.begin(np.arange(n_traj, dtype=np.uint64))
    assert info["n_segments"] >= W + K, info
    log(f"bench: kernel={info['kernel']} segments={info['n_segments']} warmup={W} steps={K}")
    eng.advance(W)
    eng.sync()
    eng.reset_stats()
    eng.set_timing(True)
    t0 = time.perf_counter()
    eng.advance(K)
    eng.sync()
    t1 = time.perf_counter()
    st = eng.stats()
    eng.set_timing(False)
    eng.advance(info["n_segments"])
    return eng, info, st, t1 - t0


def measure_engine(base, n_traj, fsteps, kernel, precision, W, K):
    """W untimed and K timed frames of one Engine workload; returns
    (site-steps/s over the wall clock, average step-kernel launch in ms, info)."""
    import numpy as np
    from pincosimulator_amd import Engine, System
    spec = dict(base, t_end=(W + K) * fsteps * base["dt"])
    eng, info, st, wall = time_frames(Engine(System(**spec), precision=precision, kernel=kernel), n_traj, W, K)
    res = eng.read()
    eng.close()
    if not np.all(np.isfinite(res.sum_n)):
        sys.exit("non-finite ensemble sums")
    site_steps = float(n_traj) * base["nx"] * base["ny"] * fsteps * K
    return site_steps / wall, st["step_kernel_ms"] / max(st["step_launches"], 1), info


def profile_fields(name):
    """traffic and the limiter's counters from a committed rocprofv3 summary
    of the same kernel and workload (profiles/<name>), or nothing."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        prof = json.load(open(path))
    except (OSError, ValueError):
        return {}
    pmc = {k: prof[k] for k in ("valu_insts_per_site_step", "valu_issue_utilisation", "mfma_util") if k in prof}
    pmc["source"] = os.path.relpath(path, ROOT)
    return {"traffic": prof.get("hbm_bytes_per_launch"), "pmc": pmc}


def secondary_lines():
    """The other measured rows, in the default line: the 64x64 workload in FP64
    (the reference's precision), the Lieb chain on MFMA and on the banded
    kernel, and the 512-point sweep."""
    out = {}
    v, ms, info = measure_engine(SPEC, N_TRAJ, FRAME_STEPS, "generic", 64, 1, 2)
    a = float(N_TRAJ) * 4096 * FRAME_STEPS * FLOPS_PER_SITE_STEP / (ms * 1e-3)
    log(f"bench: FP64 64x64: {v:.4e} site-steps/s, step kernel avg {ms:.2f} ms")
    out["fp64"] = {
        "value": v, "unit": "trajectory·site-steps/s", "dtype": "f64", "kernel": info["kernel"] + "_step_kernel",
        "workload": "the 64x64 workload above (16384 trajectories, 500-step frames, 2 timed) with an FP64 state "
                    "and FP64 Box-Muller; the steps apply the exact nearest-neighbour form of the reference's "
                    "FFT hopping, the frame captures the FFT",
        "roofline": {"bound": "valu", "achieved": a / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": a / PEAK_FP64, "avg_launch_ms": ms,
                     "flops_basis": "algorithmic: 150 flops per site-step, as the line's roofline",
                     **profile_fields("generic_fp64_64x64_16384traj_pmc.json")}}
    lieb = lieb_spec()
    n = lieb["nx"] * lieb["ny"]
    v, ms, _ = measure_engine(lieb, LIEB_TRAJ, LIEB_FRAME, "mfma", 32, 1, 10)
    a = float(LIEB_TRAJ) * n * LIEB_FRAME * 3 * 4.0 * n / (ms * 1e-3)
    vb, msb, _ = measure_engine(lieb, LIEB_TRAJ, LIEB_FRAME, "band", 32, 1, 10)
    log(f"bench: lieb384: MFMA {v:.4e} site-steps/s ({ms:.3f} ms per launch), band {vb:.4e} ({msb:.3f} ms)")
    out["lieb384"] = {
        "workload": WORKLOAD_TEXT["lieb384"] + "; 10 timed frames",
        "mfma": {"value": v, "unit": "trajectory·site-steps/s", "kernel": "mfma16_step_kernel", "avg_launch_ms": ms,
                 "roofline": {"bound": "mfma", "achieved": a / 1e12, "peak": PEAK_BF16 / 1e12, "unit": "TFLOP/s",
                              "frac": a / PEAK_BF16,
                              "flops_basis": "executed BF16 MFMA flops: 3 products x 4 N per site-step",
                              **profile_fields("mfma_lieb384_8192traj_pmc.json")}},
        "band": {"value": vb, "unit": "trajectory·site-steps/s", "kernel": "band_step_kernel", "avg_launch_ms": msb}}
    import math
    from pincosimulator_amd import Sweep
    pts = sweep_points()
    sw = Sweep(pts, 0.002, 500, 20260103)
    sw.autotune(steps=50)
    sw.run(500 * 0.002, graph=True)
    K = 10
    t0 = time.perf_counter()
    r = sw.run(K * 500 * 0.002, graph=True)
    wall = time.perf_counter() - t0
    sw.close()
    ss = sum(p["n_traj"] * p["nx"] * p["ny"] for p in pts) * 500.0 * K
    fl = sum(p["n_traj"] * p["nx"] * p["ny"] * (10.0 * math.log2(p["nx"] * p["ny"]) + 30.0) for p in pts) * 500.0 * K
    a = fl / (r["kernel_ms"] * 1e-3)
    log(f"bench: sweep512: {ss / wall:.4e} site-steps/s over {K} frames")
    out["sweep512"] = {
        "value": ss / wall, "unit": "trajectory·site-steps/s", "kernel": "sweep_step_kernel",
        "workload": "512-point sweep (shapes round-robin, 128 trajectories each), stacking autotuned first, "
                    f"graph-captured frame loop, {K} timed 500-step frames",
        "roofline": {"bound": "valu", "achieved": a / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": a / PEAK_FP32, "traffic": sweep_traffic(),
                     "flops_basis": "algorithmic: 10 log2 N + 30 flops per site-step of each point's lattice"}}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", default="square64", choices=["square64", "lieb384", "sweep512", "warmstart64"])
    ap.add_argument("--ws-points", type=int, default=4,
                    help="held-out points of the warm-start leg (square64) or evaluation (warmstart64)")
    ap.add_argument("--ws-traj", type=int, default=16384)
    ap.add_argument("--no-warmstart", action="store_true", help="square64: skip the warm-start leg")
    ap.add_argument("--no-secondary", action="store_true",
                    help="square64: skip the FP64, Lieb-chain and sweep rows of the line")
    ap.add_argument("--precision", type=int, default=32, choices=[32, 64],
                    help="square64 / lieb384: state precision (64: the reference's, on the LDS kernel)")
    ap.add_argument("--kernel", default="auto", choices=["auto", "stencil64", "fused64", "generic", "mfma", "band"],
                    help="step kernel (auto: the fastest one that supports the workload)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="sweep512: keep the default stacking instead of autotuning it first")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one 256 MiB device-to-device copy (hipMemcpy) whose HBM counters "
                         "calibrate FETCH_SIZE / WRITE_SIZE in a --pmc profile")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs the hot path on a single MI355X; "
              f"--gpus {args.gpus} is not supported.", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        sys.exit("--steps must be >= 1 and --warmup >= 0")

    import numpy as np
    from pincosimulator_amd import Engine, System

    W, K = args.warmup, args.steps
    if args.copy_calibration:
        # through the HIP runtime libtwmc_hip.so uses (no second runtime in the process)
        from pincosimulator_amd.device import DeviceBuffer, synchronize
        a = DeviceBuffer.from_array(np.ones(64 << 20, dtype=np.float32))
        b = DeviceBuffer(a.nbytes)
        synchronize()
        b.copy_from(a)
        synchronize()
        log("bench: calibration copy of 256 MiB done")
        a.free()
        b.free()
    if args.workload == "sweep512":
        return bench_sweep(args)
    if args.workload == "warmstart64":
        return bench_warmstart(args)
    if args.workload == "lieb384":
        base, n_traj, fsteps = lieb_spec(), LIEB_TRAJ, LIEB_FRAME
    else:
        base, n_traj, fsteps = SPEC, N_TRAJ, FRAME_STEPS
    n_sites = base["nx"] * base["ny"]
    spec = dict(base, t_end=(W + K) * fsteps * base["dt"])
    # the Lieb configuration names the dense GEMM on MFMA: that is its line;
    # the banded kernel (the library's AUTO choice there) is timed beside it
    kernel = "mfma" if (args.workload == "lieb384" and args.kernel == "auto" and args.precision == 32) else args.kernel
    eng, info, st, wall = time_frames(Engine(System(**spec), precision=args.precision, kernel=kernel), n_traj, W, K)
    res = eng.read()
    if not np.all(np.isfinite(res.sum_n)):
        sys.exit("non-finite ensemble sums")

    site_steps = float(n_traj) * n_sites * fsteps * K
    value = site_steps / wall
    avg_ms = st["step_kernel_ms"] / max(st["step_launches"], 1)
    if args.workload == "lieb384" and info["kernel"] == "mfma":
        # executed MFMA work: 3 BF16 products of the real N x N hopping on Re and Im
        flops_site = 3 * 4.0 * n_sites
        peak, bound, kname = PEAK_BF16, "mfma", "mfma16_step_kernel"
        basis = "executed BF16 MFMA flops: 3 products (hi.hi, hi.lo, lo.hi) x 4 N per site-step"
    elif args.workload == "lieb384":
        # the banded kernel: executed FP32 arithmetic per site-step, the
        # (2W+1)-point complex hopping (W = 2 for the Lieb chain) and the
        # drift / noise / update; its limiter is VALU issue (roofline.pmc)
        flops_site = 4.0 * 5 + 30.0
        peak, bound, kname = PEAK_FP32, "valu", info["kernel"] + "_step_kernel"
        basis = ("executed FP32 flops: 4 (2W + 1) = 20 for the banded hopping (W = 2) + 30 for drift, noise "
                 "scaling and update per site-step")
    else:
        # the workload's algorithmic work per site-step (the reference's FFT
        # path: 2 x 5 N log2 N / N + 30); the nearest-neighbour kernel evaluates
        # the same linear map with fewer executed flops
        flops_site = FLOPS_PER_SITE_STEP
        peak, bound = (PEAK_FP64 if args.precision == 64 else PEAK_FP32), "valu"
        kname = info["kernel"] + "_step_kernel"
        basis = ("algorithmic: 150 flops per site-step of the reference's FFT hopping "
                 "(2 x 5 N log2 N / N + 30, N = 4096)")
    flops_per_launch = float(n_traj) * n_sites * fsteps * flops_site
    achieved = flops_per_launch / (avg_ms * 1e-3)
    log(f"bench: wall {wall:.3f} s, {value:.4e} site-steps/s, step kernel avg {avg_ms:.2f} ms "
        f"over {st['step_launches']} launches, reduce avg "
        f"{st['reduce_kernel_ms'] / max(st['reduce_launches'], 1):.3f} ms, grid {st['grid']}")

    traffic, counters = None, {}
    tag = info["kernel"] + ("_fp64" if args.precision == 64 else "")
    pmc = os.path.join(ROOT, "profiles", f"{tag}_64x64_16384traj_pmc.json" if args.workload == "square64"
                       else f"{tag}_lieb384_8192traj_pmc.json")
    if os.path.exists(pmc):
        try:
            prof = json.load(open(pmc))
            traffic = prof.get("hbm_bytes_per_launch")
            # the kernel's own limiter, measured with rocprofv3 --pmc on the same workload
            for key in ("valu_insts_per_site_step", "valu_issue_utilisation", "mfma_util"):
                if key in prof:
                    counters[key] = prof[key]
            if counters:
                counters["source"] = os.path.relpath(pmc, ROOT)
        except Exception:
            traffic = None

    out = {
        "metric": METRIC, "value": value, "unit": "trajectory·site-steps/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": wall / K * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == 64 else "fp32",
        "data": "synthetic",
        "config": {
            "workload": WORKLOAD_TEXT[args.workload],
            "lattice": "64x64" if args.workload == "square64" else "Lieb 128x3 OBC",
            "trajectories": n_traj, "euler_steps_per_step": fsteps,
            "state": ("complex128" if args.precision == 64 else "complex64") + " in registers/LDS, FP64 ensemble sums",
            "kernel": info["kernel"],
            "grid": st["grid"], "block": st["block"],
        },
        "roofline": {
            "bound": bound, "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": traffic,
            "kernel": kname, "avg_launch_ms": avg_ms,
            "flops_per_launch": flops_per_launch, "flops_basis": basis, "pmc": counters or None,
        },
    }
    eng.close()
    if args.workload == "lieb384" and args.kernel == "auto":
        e2, i2, s2, w2 = time_frames(Engine(System(**spec), precision=32, kernel="band"), n_traj, W, K)
        e2.close()
        b_ms = s2["step_kernel_ms"] / max(s2["step_launches"], 1)
        log(f"bench: banded kernel: wall {w2:.3f} s, {site_steps / w2:.4e} site-steps/s, step kernel avg "
            f"{b_ms:.2f} ms")
        out["banded_kernel"] = {
            "value": site_steps / w2, "ms_per_step": w2 / K * 1e3, "avg_launch_ms": b_ms,
            "note": "the same workload with the coupling applied as its (2W+1)-point band stencil (W = 2), "
                    "the library's AUTO choice for banded couplings"}
    if args.workload == "square64" and not args.no_secondary:
        out.update(secondary_lines())
    if args.workload == "square64" and not args.no_warmstart and args.ws_points > 0:
        log(f"bench: warm-start leg, {args.ws_points} held-out points")
        out["time_to_steady_state"] = warmstart_summary(args.ws_points, args.ws_traj)
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"bench: CPU baseline on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(threads, base)
        # the same sample on one thread per physical core, when SMT leaves
        # fewer physical cores than logical CPUs among those used
        phys = physical_cores(threads)
        if phys < threads:
            log(f"bench: CPU baseline on {phys} threads (one per physical core)")
            b = cpu_baseline(phys, base, target_s=8.0)
            out["cpu_baseline"]["per_physical_core_run"] = {"value": b["value"], "cores": phys,
                                                            "sample": b["sample"]}
        else:
            out["cpu_baseline"]["per_physical_core_run"] = {
                "value": out["cpu_baseline"]["value"], "cores": threads,
                "sample": "the run above: /proc/cpuinfo shows one thread per core, so its threads are physical cores"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
