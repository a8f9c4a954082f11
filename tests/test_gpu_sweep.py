"""Batched uneven sweep (one persistent launch per frame): every point equals
the same system run alone; the hipGraph-captured frame loop equals the eager
one on every replay; the stacking chosen by the autotuner changes nothing but
speed."""
import numpy as np
import pytest

from helpers import OBS_SIGMA_BAR
from pincosimulator_amd import Engine, Sweep, System

pytestmark = pytest.mark.gpu

DT, FS, SEED = 0.002, 50, 99


def points():
    return [
        dict(nx=1, ny=32, n_traj=10, traj_id0=0, J=1.0, gamma=1.0, omega=0.5, U=0.05, F=1.0),
        dict(nx=1, ny=64, n_traj=7, traj_id0=1000, J=1.0, gamma=1.0, omega=-1.0, U=0.05, F=2.5),
        dict(nx=16, ny=16, n_traj=3, traj_id0=2000, J=1.0, gamma=1.0, omega=1.0, U=0.05, F=2.0),
        dict(nx=48, ny=1, n_traj=5, traj_id0=3000, J=1.0, gamma=1.0, omega=0.0, U=0.05, F=0.7),
        dict(nx=12, ny=10, n_traj=4, traj_id0=4000, J=1.0, gamma=1.0, omega=2.0, U=0.05, F=3.0),
    ]


def single(p, t_end):
    s = System(nx=p["nx"], ny=p["ny"], dt=DT, gamma=p["gamma"], J=p["J"], omega=p["omega"], U=p["U"], F=p["F"],
               beta_init=0.0, beta_sigma=0.5, frame_steps=FS, seed=SEED, t_start=0.0, t_end=t_end)
    eng = Engine(s, precision=32, kernel="generic")
    r = eng.run(np.arange(p["n_traj"]) + p["traj_id0"])
    eng.close()
    return r


def test_sweep_points_equal_single_runs():
    t_end = 0.4  # 200 steps, 5 frames
    sw = Sweep(points(), DT, FS, SEED)
    out = sw.run(t_end, graph=False)
    assert out["n_frames"] == 5
    for k, p in enumerate(points()):
        r = single(p, t_end)
        d2 = np.max(np.abs(out["m2"][k] - r.traj_m2)) / np.max(np.abs(r.traj_m2))
        d4 = np.max(np.abs(out["m4"][k] - r.traj_m4)) / np.max(np.abs(r.traj_m4))
        print(f"point {k} ({p['nx']}x{p['ny']}): |d m2| {d2:.2e} |d m4| {d4:.2e}")
        assert d2 < 1e-5 and d4 < 1e-5
    sw.close()


def test_sweep_kspace_bins_and_mean_field():
    """Radial-bin sums of |DFT beta|^2 and the mean-field sums equal the same
    quantities binned from a single run's per-mode ensemble sums."""
    t_end = 0.2
    sw = Sweep(points(), DT, FS, SEED)
    out = sw.run(t_end, graph=False)
    for k, p in enumerate(points()):
        bins = sw.mode_bins(k).reshape(-1)
        assert bins[0] == 0 and np.count_nonzero(bins == 0) == 1 and bins.max() <= 31
        s = System(nx=p["nx"], ny=p["ny"], dt=DT, gamma=p["gamma"], J=p["J"], omega=p["omega"], U=p["U"],
                   F=p["F"], beta_init=0.0, beta_sigma=0.5, frame_steps=FS, seed=SEED, t_start=0.0, t_end=t_end)
        r = Engine(s, precision=32, kernel="generic").run(np.arange(p["n_traj"]) + p["traj_id0"])
        ref = np.zeros((r.n_frames, 32))
        for b in range(32):
            ref[:, b] = r.sum_nk[:, bins == b].sum(axis=1)
        got = out["ksum"][k, :, :32]
        assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref)), k
        mf = r.sum_beta.sum(axis=1)
        assert np.max(np.abs(out["ksum"][k, :, 32] + 1j * out["ksum"][k, :, 33] - mf)) <= 1e-5 * np.max(np.abs(mf)) + 1e-6
    sw.close()


def test_graph_replays_equal_eager():
    sw = Sweep(points(), DT, FS, SEED)
    eager = sw.run(0.3, graph=False)
    for rep in range(3):
        g = sw.run(0.3, graph=True)
        for k in range(len(points())):
            np.testing.assert_array_equal(g["m2"][k], eager["m2"][k])
            np.testing.assert_array_equal(g["m4"][k], eager["m4"][k])
        again = sw.run(0.3, graph=False)  # eager work between replays
        np.testing.assert_array_equal(again["m2"][0], eager["m2"][0])
    sw.close()


def test_autotune_changes_nothing_but_speed():
    sw = Sweep(points(), DT, FS, SEED)
    sw.set_batching([1] * len(points()))
    ref = sw.run(0.2, graph=False)
    tb = sw.autotune(steps=10)
    print("autotuned stacking per point:", tb)
    for k, p in enumerate(points()):
        assert 1 <= tb[k] <= max(1, 4096 // (p["nx"] * p["ny"]))
    got = sw.run(0.2, graph=True)
    for k in range(len(points())):
        np.testing.assert_array_equal(got["m2"][k], ref["m2"][k])
    sw.close()


def test_stacked_2d_items_equal_unstacked():
    # 2D points stacked two trajectories per step item: 3 and 5 trajectories
    # give a last item of one trajectory (the single-trajectory path) beside
    # stacked ones; every trajectory must equal the unstacked run bit for bit
    pts = [dict(nx=16, ny=16, n_traj=3, traj_id0=2000, J=1.0, gamma=1.0, omega=1.0, U=0.05, F=2.0),
           dict(nx=12, ny=10, n_traj=5, traj_id0=4000, J=1.0, gamma=1.0, omega=2.0, U=0.05, F=3.0),
           dict(nx=2, ny=8, n_traj=5, traj_id0=5000, J=0.7, gamma=1.0, omega=0.3, U=0.05, F=1.0)]
    sw = Sweep(pts, DT, FS, SEED)
    sw.set_batching([1, 1, 1])
    ref = sw.run(0.2, graph=False)
    sw.set_batching([2, 2, 4])
    assert sw.batching() == [2, 2, 4]
    got = sw.run(0.2, graph=True)
    for k in range(len(pts)):
        np.testing.assert_array_equal(got["m2"][k], ref["m2"][k])
        np.testing.assert_array_equal(got["m4"][k], ref["m4"][k])
        # k-space bins are FP64 atomic sums over items: equal up to summation order
        np.testing.assert_allclose(got["ksum"][k], ref["ksum"][k], rtol=1e-12, atol=1e-12)
    sw.close()


def test_sweep_full_length_error_vs_fp64():
    """The sweep configuration's length (30000 steps at dt 0.002, frames every
    500): three points of different shapes, 256 trajectories each, against
    FP64 single runs (the LDS kernel, the reference's FFT form) of the same
    trajectory ids and seed -- per-point <n> and g2(0) per frame in standard
    errors of the FP64 ensemble."""
    from pincosimulator_amd import observables
    pts = [dict(nx=1, ny=64, n_traj=256, traj_id0=0, J=1.0, gamma=1.0, omega=-1.0, U=0.05, F=2.5),
           dict(nx=16, ny=16, n_traj=256, traj_id0=1000, J=1.0, gamma=1.0, omega=1.0, U=0.05, F=2.0),
           dict(nx=32, ny=32, n_traj=256, traj_id0=5000, J=1.0, gamma=1.0, omega=0.5, U=0.05, F=3.0)]
    t_end = 60.0
    sw = Sweep(pts, 0.002, 500, SEED)
    sw.autotune(steps=20)
    out = sw.run(t_end, graph=True)
    sw.close()
    for k, p in enumerate(pts):
        s = System(nx=p["nx"], ny=p["ny"], dt=0.002, gamma=1.0, J=1.0, omega=p["omega"], U=0.05, F=p["F"],
                   beta_init=0.0, beta_sigma=0.5, frame_steps=500, seed=SEED, t_start=0.0, t_end=t_end)
        eng = Engine(s, precision=64, kernel="generic")
        r = eng.run(np.arange(256) + p["traj_id0"])
        eng.close()
        m2, m4 = out["m2"][k], out["m4"][k]
        assert m2.shape == r.traj_m2.shape == (256, 61)
        n64, se_n = observables.mean_occupation(r)
        g64, se_g = observables.g2_jackknife(r)
        n32 = m2.mean(axis=0) - 0.5
        g32 = (m4.mean(axis=0) - 2 * m2.mean(axis=0) + 0.5) / (m2.mean(axis=0) - 0.5) ** 2
        dn = np.max(np.abs(n32 - n64)[1:] / se_n[1:])
        dg = np.max(np.abs(g32 - g64)[1:] / se_g[1:])
        print(f"point {k} ({p['nx']}x{p['ny']}): <n> {dn:.2e}, g2(0) {dg:.2e} standard errors over 60 frames")
        assert dn < OBS_SIGMA_BAR and dg < OBS_SIGMA_BAR


def test_config_sweep_512_points_in_one_launch():
    """The sweep configuration's own 512-point list (32 x 16 grid of F and
    detuning, the nine shapes round-robin: chains of 32, 64, 128, 256 and 512
    sites, squares of side 16, 32, 48 and 64; 128 trajectories each) in one
    persistent launch per frame, stacking autotuned, frame loop graph-captured,
    2 frames of 500 steps.  One point of each shape, spread over the grid,
    against the same system run alone (per-trajectory site means), and its
    radial k-space bins and mean field against the single run's per-mode sums."""
    from pincosimulator_amd.workloads import SWEEP_SHAPES, sweep512_points
    pts = sweep512_points()
    t_end = 2.0
    sw = Sweep(pts, 0.002, 500, SEED)
    tb = sw.autotune(steps=20)
    print("autotuned stacking per shape:", {SWEEP_SHAPES[s]: tb[s] for s in range(9)})
    out = sw.run(t_end, graph=True)
    assert out["n_frames"] == 3
    for s in range(9):
        k = s + 9 * ((7 * s + 3) % 56)  # the s-th shape, at varying grid positions
        p = pts[k]
        assert (p["nx"], p["ny"]) == SWEEP_SHAPES[s]
        sysp = System(nx=p["nx"], ny=p["ny"], dt=0.002, gamma=1.0, J=1.0, omega=p["omega"], U=0.05, F=p["F"],
                      beta_init=0.0, beta_sigma=0.5, frame_steps=500, seed=SEED, t_start=0.0, t_end=t_end)
        eng = Engine(sysp, precision=32, kernel="generic")
        r = eng.run(np.arange(128) + p["traj_id0"])
        eng.close()
        d2 = np.max(np.abs(out["m2"][k] - r.traj_m2)) / np.max(np.abs(r.traj_m2))
        d4 = np.max(np.abs(out["m4"][k] - r.traj_m4)) / np.max(np.abs(r.traj_m4))
        bins = sw.mode_bins(k).reshape(-1)
        ref = np.stack([r.sum_nk[:, bins == b].sum(axis=1) for b in range(32)], axis=1)
        dk = np.max(np.abs(out["ksum"][k, :, :32] - ref)) / np.max(np.abs(ref))
        mf = r.sum_beta.sum(axis=1)
        dm = np.max(np.abs(out["ksum"][k, :, 32] + 1j * out["ksum"][k, :, 33] - mf)) / max(np.max(np.abs(mf)), 1.0)
        print(f"point {k} {p['nx']}x{p['ny']} F={p['F']:.2f} D={p['omega']:.2f}: m2 {d2:.1e} m4 {d4:.1e} "
              f"k-bins {dk:.1e} mean field {dm:.1e}")
        assert d2 < 1e-5 and d4 < 1e-5 and dk < 1e-5 and dm < 1e-5
    sw.close()
