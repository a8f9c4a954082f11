"""Register-resident 64x64 FP32 kernel: checked against the FP64 oracle per
trajectory (FP32 tolerance stated in each test), against the FP32 LDS kernel,
and through its ensemble sums."""
import numpy as np
import pytest

import oracle
from helpers import FP32_BAR, OBS_SIGMA_BAR, c2_spec, rel_err, system
from pincosimulator_amd import Engine, observables
from pincosimulator_amd.observables import from_trajectories

pytestmark = pytest.mark.gpu


def test_kernel_selection_for_64x64_fp32():
    """AUTO takes the nearest-neighbour form; the register FFT is selectable."""
    eng = Engine(system(c2_spec(t_end=0.002 * 10, frame_steps=10)), precision=32)
    assert eng.run(np.arange(4)).kernel == "stencil64"
    eng.close()
    eng = Engine(system(c2_spec(t_end=0.002 * 10, frame_steps=10)), precision=32, kernel="fused64")
    assert eng.run(np.arange(4)).kernel == "fused64"
    eng.close()


def test_fused_vs_fp64_oracle_per_trajectory():
    """FP32 vs FP64 oracle, same Philox stream, 1000 steps: the FP32 path's
    per-trajectory error, bounded at FP32_BAR of max |beta| (measured 1.3e-6)."""
    spec = c2_spec(t_end=2.0, frame_steps=500)
    ids = np.array([0, 5, 16383])
    eng = Engine(system(spec), precision=32, kernel="fused64")
    res = eng.run(ids, traj=True, final=True)
    assert res.kernel == "fused64" and res.n_steps == 1000
    for j, t in enumerate(ids):
        fr, bf, _ = oracle.run_trajectory(dict(spec, lattice=oracle.LAT_FFT), int(t))
        err = rel_err(res.traj[j], fr)
        print(f"traj {t}: FP32 fused vs FP64 oracle max rel err over 3 frames = {err:.3e}")
        assert err < FP32_BAR
    eng.close()


def test_fused_vs_generic_fp32():
    spec = c2_spec(t_end=0.2, frame_steps=50)
    ids = np.arange(8)
    a = Engine(system(spec), precision=32, kernel="fused64").run(ids, traj=True)
    b = Engine(system(spec), precision=32, kernel="generic").run(ids, traj=True)
    assert rel_err(a.traj, b.traj) < 2e-5


def test_fused_ensemble_sums_match_dump():
    spec = c2_spec(t_end=0.1, frame_steps=25, U=0.2, F=1.0)
    eng = Engine(system(spec), precision=32, kernel="fused64")
    res = eng.run(np.arange(50), traj=True)
    ref = from_trajectories(res.traj, (64, 64))
    for k in ("sum_n", "sum_n2", "sum_beta", "traj_m2", "traj_m4"):
        assert rel_err(getattr(res, k), ref[k]) < 1e-9, k
    # n_k comes from the kernel's FP32 FFT; the check recomputes it in FP64
    assert rel_err(res.sum_nk, ref["sum_nk"]) < 1e-5
    eng.close()


def test_fused_observables_vs_fp64():
    """Ensemble observables FP32 (fused) vs FP64 (generic) on the same seed:
    agreement within OBS_SIGMA_BAR standard errors; the error is printed per observable."""
    spec = c2_spec(t_end=4.0, frame_steps=500)
    ids = np.arange(256)
    r32 = Engine(system(spec), precision=32, kernel="fused64").run(ids)
    r64 = Engine(system(spec), precision=64, kernel="generic").run(ids)
    n32, se32 = observables.mean_occupation(r32)
    n64, se64 = observables.mean_occupation(r64)
    g32, gse = observables.g2_jackknife(r32)
    g64, _ = observables.g2_jackknife(r64)
    nk32 = observables.momentum_occupation(r32)[-1]
    nk64 = observables.momentum_occupation(r64)[-1]
    print("frame <n>32 <n>64 |d|/se  g2_32 g2_64 |d|/se")
    for f in range(r32.n_frames):
        print(f, n32[f], n64[f], abs(n32[f] - n64[f]) / max(se64[f], 1e-30), g32[f], g64[f],
              abs(g32[f] - g64[f]) / max(gse[f], 1e-30))
    print("n_k max |d| =", np.max(np.abs(nk32 - nk64)), " n_k(0) =", nk64[0, 0])
    assert np.all(np.abs(n32 - n64) < OBS_SIGMA_BAR * se64 + 1e-12)
    assert np.all(np.abs(g32[1:] - g64[1:]) < OBS_SIGMA_BAR * gse[1:])
    assert np.max(np.abs(nk32 - nk64)) < 1e-2 * max(1.0, np.max(np.abs(nk64)))
