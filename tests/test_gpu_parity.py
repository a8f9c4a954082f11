"""FP64 HIP kernels vs the CPU oracle, per trajectory (same Philox increments).

Bar: every stored frame of every trajectory within 1e-9 relative (max |diff| /
max |beta|) after 1000 steps; frame counts, step counts and t_final identical.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from helpers import c1_spec, oracle_params, oracle_traj, rel_err, system
from pincosimulator_amd import Engine, observables
from pincosimulator_amd.observables import from_trajectories

pytestmark = pytest.mark.gpu

TOL64 = 1e-9


def _compare(spec, ids, precision=64, kernel="generic", **kw):
    eng = Engine(system(spec), precision=precision, kernel=kernel)
    res = eng.run(ids, traj=True, final=True, **kw)
    ok_kw = {k: v for k, v in kw.items() if k in ("t_start", "t_end", "store_initial", "step0")}
    if "init" in kw:
        ok_kw["init_mode"] = {"settings": 0, "fixedpoint": 1, "previous": 2}[kw["init"]]
    if "beta_prev" in kw:
        ok_kw["beta_prev"] = None
    frames, finals, tfs = [], [], []
    for j, t in enumerate(ids):
        okw = dict(ok_kw)
        if "beta_prev" in kw:
            okw["beta_prev"] = kw["beta_prev"][j]
        fr, bf, tf = oracle.run_trajectory(
            {**spec, "lattice": oracle.LAT_DENSE if spec.get("lattice") == "matrix" else oracle.LAT_FFT,
             "E": spec.get("omega", 0.0)}, int(t), **okw)
        frames.append(fr)
        finals.append(bf)
        tfs.append(tf)
    frames = np.array(frames)
    eng.close()
    return res, frames, np.array(finals), tfs


def test_chain64_fp64_1000_steps():
    """1D chain N=64 (configuration 1 parameters), 1000 steps, 16 trajectories."""
    spec = c1_spec(t_end=5.0)  # 1000 steps of dt = 0.005, frames every 200
    ids = np.arange(16) + 7
    res, frames, finals, tfs = _compare(spec, ids)
    assert res.n_steps == 1000 and res.n_frames == 6
    assert res.traj.shape == frames.shape
    assert rel_err(res.traj, frames) < TOL64
    assert rel_err(res.beta_final, finals) < TOL64
    assert res.t_final == tfs[0]


def test_chain64_configuration_in_full_fp64():
    """The 1D chain configuration as specified -- N=64 PBC, 256 trajectories,
    20000 steps of dt = 0.005, 101 frames, complex128 -- on the FP64 GPU path
    against the FP64 oracle run on every trajectory (a thread pool of oracle
    calls): per-trajectory deviation over all frames, and the GPU's FP64
    ensemble sums and observables against the same estimators built from the
    oracle's frames."""
    spec = c1_spec()
    ids = np.arange(256)
    print("chain64 in full: FP64, 256 trajectories x 20000 steps", flush=True)
    eng = Engine(system(spec), precision=64)
    res = eng.run(ids, traj=True, final=True)
    eng.close()
    assert res.kernel == "generic" and res.n_steps == 20000 and res.n_frames == 101
    p = oracle_params(spec)
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        out = list(ex.map(lambda t: oracle.run_trajectory(p, int(t)), ids))
    frames = np.array([o[0] for o in out])
    dev = rel_err(res.traj, frames)
    dev_final = rel_err(res.beta_final, np.array([o[1] for o in out]))
    ref = from_trajectories(frames, (1, 64))
    dsum = {k: rel_err(getattr(res, k), ref[k]) for k in ("sum_n", "sum_n2", "sum_nk", "sum_beta")}
    n_gpu, _ = observables.mean_occupation(res)
    g_gpu, g_se = observables.g2_jackknife(res)
    a2 = np.abs(frames) ** 2
    n_ora = a2.mean(axis=(0, 2)) - 0.5
    print(f"  per trajectory, all frames: {dev:.2e}; final states {dev_final:.2e}")
    print("  ensemble sums: " + ", ".join(f"{k} {v:.2e}" for k, v in dsum.items()))
    print(f"  <n> per frame: max |GPU - oracle| = {np.max(np.abs(n_gpu - n_ora)):.2e}; "
          f"g2(0) at the last frame {g_gpu[-1]:.4f} +- {g_se[-1]:.4f}")
    assert dev < TOL64 and dev_final < TOL64
    assert all(v < TOL64 for v in dsum.values())
    assert np.max(np.abs(n_gpu - n_ora)) < TOL64


@pytest.mark.parametrize("nx,ny", [(8, 8), (12, 10), (48, 48), (16, 15), (2, 2)])
def test_square_fp64(nx, ny):
    spec = dict(nx=nx, ny=ny, dt=0.002, gamma=1.0, J=1.0, U=0.05, omega=1.0, F=2.0, beta_init=0.0,
                beta_sigma=0.5, frame_steps=250, seed=99, t_start=0.0, t_end=2.0)
    ids = np.arange(6) * 1000 + 3
    res, frames, finals, _ = _compare(spec, ids)
    assert res.n_steps == 1000
    assert rel_err(res.traj, frames) < TOL64


@pytest.mark.parametrize("nx,ny,n_traj", [(64, 64, 600), (8, 8, 5000), (48, 48, 600)])
def test_fp64_more_trajectories_than_workgroups(nx, ny, n_traj):
    """The persistent step launch: more trajectories than workgroups, so
    workgroups step a second and third trajectory after the first (the
    64x64 and 48x48 launches, 1024 threads and two LDS copies of the state per
    workgroup, one workgroup per CU; 8x8, 256 threads).
    Trajectories from the first, second and last round of the grid against the
    oracle."""
    spec = dict(nx=nx, ny=ny, dt=0.002, gamma=1.0, J=1.0, U=0.05, omega=1.0, F=2.0, beta_init=0.0,
                beta_sigma=0.5, frame_steps=20, seed=2025, t_start=0.0, t_end=0.08)
    ids = np.arange(n_traj, dtype=np.uint64) * 3 + 11
    print(f"fp64 {nx}x{ny}, {n_traj} trajectories", flush=True)
    eng = Engine(system(spec), precision=64, kernel="generic")
    res = eng.run(ids, traj=True)
    eng.close()
    assert res.traj.shape == (n_traj, 3, nx * ny)
    pick = [0, 1, 255, 256, 300, n_traj - 1]
    ref = np.array([oracle.run_trajectory(dict(spec, lattice=oracle.LAT_FFT), int(ids[j]))[0] for j in pick])
    err = rel_err(res.traj[pick], ref)
    print(f"  max relative deviation {err:.2e}")
    assert err < TOL64


@pytest.mark.parametrize("nx,ny,n_traj", [(64, 64, 600), (32, 48, 2000), (8, 8, 5000)])
def test_fp64_frame_sums_over_many_trajectories_per_workgroup(nx, ny, n_traj):
    """The capture launch with more trajectories than workgroups, each
    workgroup summing several trajectories per site in registers (64x64 and
    32x48: 1024 threads, two and one site pairs per thread; 8x8: 256 threads):
    the ensemble sums and per-trajectory means equal the same estimators over
    the stored frames, which the oracle pins trajectory by trajectory, to 1e-12
    (FP64 sums in another order)."""
    spec = dict(nx=nx, ny=ny, dt=0.002, gamma=1.0, J=1.0, U=0.05, omega=1.0, F=2.0, beta_init=0.0,
                beta_sigma=0.5, frame_steps=20, seed=31, t_start=0.0, t_end=0.08)
    ids = np.arange(n_traj, dtype=np.uint64) * 5 + 2
    print(f"fp64 frame sums {nx}x{ny}, {n_traj} trajectories", flush=True)
    eng = Engine(system(spec), precision=64, kernel="generic")
    res = eng.run(ids, traj=True)
    eng.close()
    ref = from_trajectories(res.traj, (nx, ny))
    errs = {k: rel_err(getattr(res, k), ref[k]) for k in ("sum_n", "sum_n2", "sum_nk", "sum_beta", "traj_m2", "traj_m4")}
    print("  " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()))
    assert all(v < 1e-12 for v in errs.values())
    pick = [0, n_traj // 2, n_traj - 1]
    ora = np.array([oracle.run_trajectory(dict(spec, lattice=oracle.LAT_FFT), int(ids[j]))[0] for j in pick])
    assert rel_err(res.traj[pick], ora) < TOL64


def test_fp64_64x64_subset():
    spec = dict(nx=64, ny=64, dt=0.002, gamma=1.0, J=1.0, U=0.05, omega=1.0, F=2.0, beta_init=0.0,
                beta_sigma=0.5, frame_steps=500, seed=2024, t_start=0.0, t_end=2.0)
    ids = np.array([0, 1, 16383])
    res, frames, _, _ = _compare(spec, ids)
    assert res.n_frames == 3
    assert rel_err(res.traj, frames) < TOL64


def test_noisy_parameters_and_disorder():
    """Per-trajectory Gaussian/uniform disorder of U, omega, F and beta_init."""
    rng = np.random.default_rng(5)
    spec = dict(nx=1, ny=32, dt=0.005, gamma=1.0, J=0.45, frame_steps=100, seed=77, t_start=0.0, t_end=5.0,
                U={"base": 0.05, "noise": ("gaussian", 0.02 + 0.01j)},
                omega={"base": rng.normal(size=32) + 0j, "noise": ("uniform", 0.3)},
                F={"base": 1.0, "noise": ("gaussian", 0.3)},
                beta_init={"base": 0.1 + 0.2j, "noise": ("gaussian", 0.1 + 0.1j)}, beta_sigma=0.5)
    ids = np.arange(5) + 100
    res, frames, _, _ = _compare(spec, ids)
    assert np.all(np.isfinite(frames))
    assert rel_err(res.traj, frames) < TOL64


def test_lieb_matrix_hopping():
    """Lieb chain through matrix hopping: h = -i beta.C (TWMCLiebSolver)."""
    ny = 24
    C = oracle.lieb_coupling(ny, 4.0, 4.0, False)
    spec = dict(nx=1, ny=3 * ny, dt=0.0005 * 10, gamma=1.0, J=0.0, U=0.1, omega=0.2 * np.arange(3 * ny) / 72,
                F=1.0, beta_init=0.0, beta_sigma=0.5, frame_steps=100, seed=7, t_start=0.0, t_end=5.0,
                lattice="matrix", coupling=C, detuning=2.0, onsite="detuning-E")
    ids = np.arange(4)
    res, frames, _, _ = _compare(spec, ids)
    assert res.n_steps == 1000
    assert rel_err(res.traj, frames) < TOL64


def test_init_modes_and_resume():
    """FixedPoint start; ReadFromPreviousData continuation with step0 resumes the
    same Philox stream, so split runs equal one uninterrupted run."""
    spec = c1_spec(t_end=2.0, frame_steps=100)
    ids = np.arange(3)
    res, frames, _, _ = _compare(spec, ids, init="fixedpoint")
    assert rel_err(res.traj, frames) < TOL64
    eng = Engine(system(spec), precision=64, kernel="generic")
    full = eng.run(ids, traj=True, final=True)
    first = eng.run(ids, t_end=1.0, traj=True, final=True)
    second = eng.run(ids, t_start=1.0, t_end=2.0, init="previous", beta_prev=first.beta_final,
                     store_initial=False, step0=first.n_steps, traj=True, final=True)
    assert first.n_steps + second.n_steps == full.n_steps
    np.testing.assert_array_equal(second.beta_final, full.beta_final)
    np.testing.assert_array_equal(second.traj[:, -1], full.traj[:, -1])
    eng.close()


def test_ensemble_sums_match_trajectories():
    """Engine FP64 ensemble sums equal sums over the stored trajectories."""
    from pincosimulator_amd.observables import from_trajectories
    spec = dict(nx=12, ny=10, dt=0.002, gamma=1.0, J=1.0, U=0.05, omega=1.0, F=2.0, beta_init=0.0,
                beta_sigma=0.5, frame_steps=100, seed=5, t_start=0.0, t_end=1.0)
    eng = Engine(system(spec), precision=64, kernel="generic")
    res = eng.run(np.arange(40), traj=True)
    ref = from_trajectories(res.traj, (12, 10))
    for k in ("sum_n", "sum_n2", "sum_nk", "sum_beta", "traj_m2", "traj_m4"):
        assert rel_err(getattr(res, k), ref[k]) < 1e-12, k
    eng.close()


def test_truncated_schedule_matches_reference_loop():
    """t_end not a multiple of dt_obs: frame count and t_final follow the
    reference's FP64 time loop exactly."""
    spec = c1_spec(t_end=1.2345, frame_steps=37, dt=0.0123)
    res, frames, _, tfs = _compare(spec, np.arange(2))
    assert res.n_frames == frames.shape[1]
    assert res.t_final == tfs[0]
    assert rel_err(res.traj, frames) < TOL64
