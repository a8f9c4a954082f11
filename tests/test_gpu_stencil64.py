"""64x64 FP32 kernel with the hopping in its exact nearest-neighbour form,
checked against the FP64 oracle (which runs the reference's FFT form) per
trajectory, against the FFT kernel, and through its ensemble sums (captured by
the register-FFT capture kernel and the site-major sums kernel)."""
import numpy as np
import pytest

import oracle
from helpers import FP32_BAR, OBS_SIGMA_BAR, c2_spec, rel_err, system
from pincosimulator_amd import Engine, observables
from pincosimulator_amd.observables import from_trajectories

pytestmark = pytest.mark.gpu


def test_stencil_vs_fp64_oracle_per_trajectory():
    """FP32 vs FP64 oracle, same Philox stream, 1000 steps: bounded at FP32_BAR of
    max |beta| (measured 1.3e-6), as the FFT kernel."""
    spec = c2_spec(t_end=2.0, frame_steps=500)
    ids = np.array([0, 5, 16383])
    print("stencil64 vs oracle", flush=True)
    eng = Engine(system(spec), precision=32, kernel="stencil64")
    res = eng.run(ids, traj=True, final=True)
    assert res.kernel == "stencil64" and res.n_steps == 1000
    for j, t in enumerate(ids):
        fr, bf, _ = oracle.run_trajectory(dict(spec, lattice=oracle.LAT_FFT), int(t))
        err = rel_err(res.traj[j], fr)
        print(f"traj {t}: FP32 stencil vs FP64 oracle max rel err over 3 frames = {err:.3e}")
        assert err < FP32_BAR
        assert rel_err(res.beta_final[j], bf) < FP32_BAR
    eng.close()


def test_stencil_vs_fft_kernel_fp32():
    spec = c2_spec(t_end=0.2, frame_steps=50)
    ids = np.arange(8)
    a = Engine(system(spec), precision=32, kernel="stencil64").run(ids, traj=True)
    b = Engine(system(spec), precision=32, kernel="fused64").run(ids, traj=True)
    assert rel_err(a.traj, b.traj) < 2e-5


def test_stencil_ensemble_sums_and_many_trajectories():
    """More trajectories than resident waves (the persistent loop), sums vs
    the stored frames."""
    spec = c2_spec(t_end=0.1, frame_steps=25, U=0.2, F=1.0)
    eng = Engine(system(spec), precision=32, kernel="stencil64")
    res = eng.run(np.arange(50), traj=True)
    ref = from_trajectories(res.traj, (64, 64))
    for k in ("sum_n", "sum_n2", "sum_beta", "traj_m2", "traj_m4"):
        assert rel_err(getattr(res, k), ref[k]) < 1e-9, k
    assert rel_err(res.sum_nk, ref["sum_nk"]) < 1e-5
    eng.close()
    big = Engine(system(c2_spec(t_end=0.02, frame_steps=10)), precision=32, kernel="stencil64")
    r = big.run(np.arange(5000))
    n, se = observables.mean_occupation(r)
    assert np.all(np.isfinite(r.sum_n)) and n.shape == (2,)
    big.close()


def test_stencil_observables_vs_fp64():
    spec = c2_spec(t_end=4.0, frame_steps=500)
    ids = np.arange(256)
    r32 = Engine(system(spec), precision=32, kernel="stencil64").run(ids)
    r64 = Engine(system(spec), precision=64, kernel="generic").run(ids)
    n32, _ = observables.mean_occupation(r32)
    n64, se64 = observables.mean_occupation(r64)
    g32, gse = observables.g2_jackknife(r32)
    g64, _ = observables.g2_jackknife(r64)
    assert np.all(np.abs(n32 - n64) < OBS_SIGMA_BAR * se64 + 1e-12)
    assert np.all(np.abs(g32[1:] - g64[1:]) < OBS_SIGMA_BAR * gse[1:])


def test_stencil_complex_interaction_vs_fp64_oracle():
    """Im U != 0 (two-body loss) takes the kernel variant that keeps the
    |beta|^2-dependent real part of the drift factor."""
    spec = c2_spec(t_end=0.8, frame_steps=200, U=0.05 + 0.02j)
    ids = np.array([2, 9])
    print("stencil64 complex U vs oracle", flush=True)
    eng = Engine(system(spec), precision=32, kernel="stencil64")
    res = eng.run(ids, traj=True)
    eng.close()
    for j, t in enumerate(ids):
        fr, _, _ = oracle.run_trajectory(dict(spec, lattice=oracle.LAT_FFT), int(t))
        err = rel_err(res.traj[j], fr)
        print(f"  traj {t}: {err:.3e}")
        assert err < FP32_BAR


def test_full_length_fp32_error_vs_fp64():
    """The 64x64 configuration at its full length (50000 steps, 101 frames),
    1024 trajectories: FP32 (stencil64) against FP64 (the LDS kernel, the
    reference's FFT form) on the same seed and trajectory ids.  States the
    FP32 error per observable in standard errors of the FP64 ensemble, and
    the per-trajectory deviation of the final state."""
    spec = c2_spec(t_end=100.0, frame_steps=500)
    ids = np.arange(1024)
    print("full-length FP32 vs FP64 (50000 steps)", flush=True)
    r32 = Engine(system(spec), precision=32, kernel="stencil64").run(ids, final=True)
    r64 = Engine(system(spec), precision=64, kernel="generic").run(ids, final=True)
    assert r32.n_steps == 50000 and r32.sum_n.shape[0] == 101
    n32, _ = observables.mean_occupation(r32)
    n64, se_n = observables.mean_occupation(r64)
    g32, _ = observables.g2_jackknife(r32)
    g64, se_g = observables.g2_jackknife(r64)
    k32 = observables.momentum_occupation(r32)[:, 0, 0]
    k64 = observables.momentum_occupation(r64)[:, 0, 0]
    # frame 0 is the shared initial state (no spread to divide by); n(k=0) has no
    # per-trajectory record, so its bound is relative
    dn = np.max(np.abs(n32 - n64)[1:] / se_n[1:])
    dg = np.max(np.abs(g32 - g64)[1:] / se_g[1:])
    dk = np.max(np.abs(k32 - k64) / np.abs(k64))
    dt = rel_err(r32.beta_final, r64.beta_final)
    print(f"  <n>: max |FP32 - FP64| = {dn:.2e} standard errors over 100 frames")
    print(f"  g2(0): max {dg:.2e} standard errors; n(k=0): max relative deviation {dk:.2e}")
    print(f"  final state, 1024 trajectories: max relative deviation {dt:.2e}")
    assert dn < OBS_SIGMA_BAR and dg < OBS_SIGMA_BAR and dk < 1e-5
