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
of the step kernel (HIP-event average launch duration, FP32 VALU bound), and a
CPU baseline: the oracle in the reference's own per-step structure (FP64,
std::mt19937 + normal_distribution), threaded over this host's cores, on a
bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "trajectory·site-steps/s (2D 64×64, 1 GPU); time-to-steady-state plain vs warm"
N_TRAJ = 16384
FRAME_STEPS = 500
FLOPS_PER_SITE_STEP = 150.0    # 2 FFTs x 5 N log2 N / N + 30 elementwise, N = 4096
PEAK_FP32 = 157.3e12           # MI355X FP32 vector peak (FLOP/s)
SPEC = dict(nx=64, ny=64, dt=0.002, gamma=1.0, J=1.0, U=0.05, omega=1.0, F=2.0, beta_init=0.0,
            beta_sigma=0.5, frame_steps=FRAME_STEPS, seed=20260101, t_start=0.0)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(threads, target_s=12.0):
    import oracle
    p = dict(SPEC, lattice=oracle.LAT_FFT)
    # pilot to size the sample to ~target_s of wall time
    n_traj_p, steps_p = threads, 5
    t = oracle.time_batch(p, n_traj_p, threads, steps_p, faithful=True)
    rate = n_traj_p * steps_p / t  # trajectory-steps per second
    n_traj = 4 * threads
    steps = max(5, int(target_s * rate / n_traj))
    t = oracle.time_batch(p, n_traj, threads, steps, faithful=True)
    value = n_traj * 4096 * steps / t
    return {
        "value": value, "unit": "trajectory·site-steps/s", "cores": threads, "kind": "port",
        "sample": f"{n_traj} trajectories x {steps} steps of the same 64x64 system "
                  f"({t:.1f} s wall), FP64, reference per-step structure with std::mt19937 noise, "
                  f"one trajectory per task on {threads} std::threads",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one 256 MiB device-to-device copy (torch) whose HBM counters "
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
        import torch
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("bench: calibration copy of 256 MiB done")
        del a, b
    spec = dict(SPEC, t_end=(W + K) * FRAME_STEPS * SPEC["dt"])
    eng = Engine(System(**spec), precision=32)
    info = eng.begin(np.arange(N_TRAJ, dtype=np.uint64))
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
    eng.advance(info["n_segments"])  # finish the run (final frame) outside the timed region
    res = eng.read()
    if not np.all(np.isfinite(res.sum_n)):
        sys.exit("non-finite ensemble sums")

    wall = t1 - t0
    site_steps = float(N_TRAJ) * 4096 * FRAME_STEPS * K
    value = site_steps / wall
    avg_ms = st["step_kernel_ms"] / max(st["step_launches"], 1)
    flops_per_launch = float(N_TRAJ) * 4096 * FRAME_STEPS * FLOPS_PER_SITE_STEP
    achieved = flops_per_launch / (avg_ms * 1e-3)
    log(f"bench: wall {wall:.3f} s, {value:.4e} site-steps/s, step kernel avg {avg_ms:.2f} ms "
        f"over {st['step_launches']} launches, reduce avg "
        f"{st['reduce_kernel_ms'] / max(st['reduce_launches'], 1):.3f} ms, grid {st['grid']}")

    traffic = None
    pmc = os.path.join(ROOT, "profiles", "fused64_64x64_16384traj_pmc.json")
    if os.path.exists(pmc):
        try:
            traffic = json.load(open(pmc)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    out = {
        "metric": METRIC, "value": value, "unit": "trajectory·site-steps/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": wall / K * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {
            "workload": "2D 64x64 PBC uniform-hopping TWMC, J=1 U=0.05 omega=1 F=2 gamma=1 dt=0.002, "
                        "16384 trajectories, vacuum Wigner start; one step = one 500-step frame "
                        "incl. FP64 ensemble reduction",
            "lattice": "64x64", "trajectories": N_TRAJ, "euler_steps_per_step": FRAME_STEPS,
            "state": "complex64 in registers/LDS, FP64 ensemble sums", "kernel": info["kernel"],
            "grid": st["grid"], "block": st["block"],
        },
        "roofline": {
            "bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32, "traffic": traffic,
            "kernel": "fused64_step_kernel", "avg_launch_ms": avg_ms,
            "flops_per_launch": flops_per_launch,
        },
    }
    eng.close()
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"bench: CPU baseline on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
