"""Matrix hopping with a banded coupling on the register-resident band kernel:
per trajectory against the FP64 oracle (Lieb chains open and periodic,
asymmetric band matrices -- C_ij != C_ji pins the direction of the lane
rotations --, sizes padded to 64 S sites, a time-dependent drive), against
the MFMA kernel, ensemble sums against the dump, and the kernel selection.
FP32 bar: FP32_BAR of max |beta| per trajectory (measured 3e-7 to 6e-7)."""
import numpy as np
import pytest

import oracle
from helpers import FP32_BAR, MFMA_BAR, OBS_SIGMA_BAR, rel_err
from pincosimulator_amd import Engine, System
from pincosimulator_amd._lib import TWMCError
from pincosimulator_amd.observables import from_trajectories
from test_gpu_mfma import c3_spec, lieb_disordered

pytestmark = pytest.mark.gpu


def matrix_spec(C, **kw):
    n = C.shape[0]
    rng = np.random.default_rng(n)
    s = dict(nx=1, ny=n, dt=0.005, gamma=1.0, J=0.0, U=0.1, F=1.2, omega=0.3 * rng.standard_normal(n),
             beta_init=0.0, beta_sigma=0.5, frame_steps=50, seed=4242, t_start=0.0, t_end=1.0, lattice="matrix",
             coupling=C, detuning=0.2, onsite="detuning-E")
    s.update(kw)
    return s


def band_matrix(n, w, cyclic, seed):
    """Random real band matrix, C_ij != C_ji, nonzeros at (cyclic) distance <= w."""
    rng = np.random.default_rng(seed)
    C = np.zeros((n, n))
    for i in range(n):
        for d in range(-w, w + 1):
            j = i + d
            if cyclic:
                j %= n
            elif not 0 <= j < n:
                continue
            if d != 0:
                C[i, j] = rng.uniform(-1.0, 1.0)
    return C


def run_vs_oracle(spec, ids, kernel="band"):
    eng = Engine(System(**spec), precision=32, kernel=kernel)
    res = eng.run(ids, traj=True)
    eng.close()
    assert res.kernel == kernel
    worst = 0.0
    for j, t in enumerate(ids):
        ref, _, _ = oracle.run_trajectory(dict(spec, lattice=oracle.LAT_DENSE, E=spec["omega"]), int(t))
        worst = max(worst, rel_err(res.traj[j], ref))
    return worst


@pytest.mark.parametrize("cells,pbc", [(8, False), (128, False), (128, True), (100, False)])
def test_lieb_vs_fp64_oracle(cells, pbc):
    C = lieb_disordered(cells, pbc=pbc)
    spec = c3_spec(cells=cells, t_end=1.0, frame_steps=100)
    spec["coupling"] = C
    print(f"band Lieb cells={cells} pbc={pbc}", flush=True)
    err = run_vs_oracle(spec, np.array([0, 5, 63, 64, 70]))
    print(f"  max relative deviation {err:.2e}")
    assert err < FP32_BAR


@pytest.mark.parametrize("n,w,cyclic", [(200, 3, False), (256, 2, True), (128, 1, True), (20, 2, False),
                                         (512, 3, True), (384, 2, True), (330, 2, False)])
def test_asymmetric_band_vs_fp64_oracle(n, w, cyclic):
    C = band_matrix(n, w, cyclic, seed=n + w)
    assert not np.allclose(C, C.T)
    spec = matrix_spec(C)
    print(f"band asymmetric n={n} w={w} cyclic={cyclic}", flush=True)
    err = run_vs_oracle(spec, np.array([1, 2, 77]))
    print(f"  max relative deviation {err:.2e}")
    assert err < FP32_BAR


@pytest.mark.parametrize("cells", [16, 128])
def test_time_dependent_drive(cells):
    """F(t) on the banded kernel; 128 cells (384 sites, two cells per lane)
    takes the variant that leaves out the Lieb chain's structural zeros."""
    rng = np.random.default_rng(5)
    times = np.array([0.1, 0.45, 0.8])
    vals = rng.normal(size=(3, 3 * cells)) + 1j * rng.normal(size=(3, 3 * cells))
    spec = dict(nx=1, ny=3 * cells, lattice="matrix", coupling=oracle.lieb_coupling(cells, 1.0, 1.0, False),
                onsite="detuning-E", detuning=0.5, omega=0.0, dt=0.005, gamma=1.0, J=0.0, U=0.1, F=0.0,
                beta_init=0.0, beta_sigma=0.5, frame_steps=40, seed=808, t_start=0.0, t_end=1.2, F_t=(times, vals))
    ids = np.array([3, 9])
    eng = Engine(System(**spec), precision=32, kernel="band")
    res = eng.run(ids, traj=True)
    eng.close()
    p = dict(spec, E=0.0, lattice=oracle.LAT_DENSE)
    worst = max(rel_err(res.traj[j], oracle.run_trajectory(p, int(t))[0]) for j, t in enumerate(ids))
    print(f"band F(t) vs oracle: {worst:.2e}")
    assert worst < FP32_BAR


def test_band_vs_mfma_c3():
    spec = c3_spec(cells=128, t_end=0.5, frame_steps=50)
    ids = np.arange(40)
    a = Engine(System(**spec), precision=32, kernel="band").run(ids, traj=True)
    b = Engine(System(**spec), precision=32, kernel="mfma").run(ids, traj=True)
    err = rel_err(a.traj, b.traj)
    print(f"band vs MFMA FP32 after 100 steps: {err:.3e}")
    assert err < MFMA_BAR


def test_band_ensemble_sums_match_dump():
    spec = c3_spec(cells=64, t_end=0.2, frame_steps=20)
    eng = Engine(System(**spec), precision=32, kernel="band")
    res = eng.run(np.arange(70), traj=True)
    eng.close()
    ref = from_trajectories(res.traj, (1, 192))
    for k in ("sum_n", "sum_n2", "sum_beta", "traj_m2", "traj_m4"):
        assert rel_err(getattr(res, k), ref[k]) < 1e-9, k
    assert np.all(res.sum_nk == 0)  # matrix hopping has no k-space sums


def test_kernel_selection():
    lieb = c3_spec(cells=32, t_end=0.1, frame_steps=10)
    eng = Engine(System(**lieb), precision=32)
    assert eng.run(np.arange(4)).kernel == "band"
    eng.close()
    dense = np.random.default_rng(2).uniform(-0.1, 0.1, (96, 96))
    spec = matrix_spec(dense, t_end=0.1, frame_steps=10)
    eng = Engine(System(**spec), precision=32)
    assert eng.run(np.arange(4)).kernel == "mfma"
    eng.close()
    with pytest.raises(TWMCError):
        Engine(System(**spec), precision=32, kernel="band")
    # periodic Lieb chain of 300 sites: the closing bond is not banded (300 != 64 S)
    pbc = c3_spec(cells=100, t_end=0.1, frame_steps=10)
    pbc["coupling"] = lieb_disordered(100, pbc=True)
    eng = Engine(System(**pbc), precision=32)
    assert eng.run(np.arange(4)).kernel == "mfma"
    eng.close()


@pytest.mark.parametrize("kernel", ["band", "mfma"])
def test_lieb_full_length_fp32_error_vs_fp64(kernel):
    """The Lieb configuration at its full length (20000 steps, frames every
    200), 512 trajectories: FP32 (band or the BF16x3 MFMA product) against the
    FP64 sparse product on the same seed and ids, per observable."""
    from pincosimulator_amd import observables
    spec = c3_spec(cells=128, t_end=100.0, frame_steps=200)
    ids = np.arange(512)
    print(f"Lieb full length, {kernel} vs FP64", flush=True)
    r32 = Engine(System(**spec), precision=32, kernel=kernel).run(ids, final=True)
    r64 = Engine(System(**spec), precision=64, kernel="generic").run(ids, final=True)
    assert r32.n_steps == 20000
    n32, _ = observables.mean_occupation(r32)
    n64, se_n = observables.mean_occupation(r64)
    g32, _ = observables.g2_jackknife(r32)
    g64, se_g = observables.g2_jackknife(r64)
    dn = np.max(np.abs(n32 - n64)[1:] / se_n[1:])
    dg = np.max(np.abs(g32 - g64)[1:] / se_g[1:])
    dt = rel_err(r32.beta_final, r64.beta_final)
    print(f"  <n>: {dn:.2e} standard errors, g2(0): {dg:.2e}; final states {dt:.2e} relative")
    assert dn < OBS_SIGMA_BAR and dg < OBS_SIGMA_BAR
