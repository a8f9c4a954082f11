"""Matrix hopping on MFMA (BF16 hi/lo x3, FP32 accumulate) for disordered Lieb
chains: the GEMM error against FP64, per-trajectory agreement with the FP64
oracle and with the FP32 sparse-product kernel, and ensemble sums."""
import numpy as np
import pytest

import oracle
from helpers import MFMA_BAR, rel_err
from pincosimulator_amd import Engine, System, noisy_matrix
from pincosimulator_amd.observables import from_trajectories

pytestmark = pytest.mark.gpu


def lieb_disordered(cells, J=1.0, sigma_J=0.1, seed=7, pbc=False):
    """Lieb coupling with Gaussian bond disorder, bonds in the reference's order
    (AB, BC, inter-cell C-B per cell, closing bond last); the same draws as the
    host solver's LiebCoupling (Philox seed, trajectory 0, tag 7)."""
    bonds = []
    for i in range(cells):
        bonds += [(3 * i, 3 * i + 1), (3 * i + 1, 3 * i + 2)]
        if i == cells - 1:
            if pbc:
                bonds.append((3 * i + 2, 1))
            break
        bonds.append((3 * i + 2, 3 * (i + 1) + 1))
    z = noisy_matrix({"base": 0.0, "noise": ("gaussian", sigma_J)}, len(bonds), seed, 0, 7)
    n = 3 * cells
    C = np.zeros((n, n))
    for b, (i, j) in enumerate(bonds):
        C[i, j] += J + z[b].real
        C[j, i] += J + z[b].real
    return C


def c3_spec(cells=128, **kw):
    n = 3 * cells
    C = lieb_disordered(cells)
    E = noisy_matrix({"base": 0.0, "noise": ("gaussian", 0.2)}, n, 7, 0, 3).real
    s = dict(nx=1, ny=n, dt=0.005, gamma=1.0, J=0.0, U=0.1, F=1.5, omega=E, beta_init=0.0, beta_sigma=0.5,
             frame_steps=200, seed=2025, t_start=0.0, t_end=1.0, lattice="matrix", coupling=C, detuning=0.0,
             onsite="detuning-E")
    s.update(kw)
    return s


def test_lieb_coupling_disorder_matches_oracle_structure():
    C = lieb_disordered(5)
    base = oracle.lieb_coupling(5, 1.0, 1.0, False)
    assert np.array_equal(C != 0, base != 0) and np.allclose(C, C.T)
    assert 0.02 < np.std(C[C != 0] - 1.0) < 0.3


def test_bf16x3_gemm_error_vs_fp64():
    """dt = 1, gamma = U = F = 0, detuning = E: one step gives h = -i C^T beta
    exactly up to the GEMM; its error against FP64 is the BF16x3 error."""
    cells = 128
    n = 3 * cells
    C = lieb_disordered(cells)
    rng = np.random.default_rng(1)
    B = 64
    b0 = rng.standard_normal((B, n)) + 1j * rng.standard_normal((B, n))
    spec = dict(nx=1, ny=n, dt=1.0, gamma=0.0, J=0.0, U=0.0, F=0.0, omega=0.0, beta_init=0.0, beta_sigma=0.0,
                frame_steps=1, seed=1, t_start=0.0, t_end=1.0, lattice="matrix", coupling=C, detuning=0.0,
                onsite="detuning-E")
    eng = Engine(System(**spec), precision=32, kernel="mfma")
    res = eng.run(np.arange(B), init="previous", beta_prev=b0, final=True)
    assert res.kernel == "mfma" and res.n_steps == 1
    b0f = b0.astype(np.complex64).astype(np.complex128)
    h = res.beta_final - b0f
    ref = -1j * (b0f @ C)
    err = np.max(np.abs(h - ref)) / np.max(np.abs(ref))
    print(f"BF16x3 MFMA hopping vs FP64 GEMM: max relative error {err:.3e}")
    assert err < 5e-5
    eng.close()


def test_mfma_vs_fp64_oracle_per_trajectory():
    spec = c3_spec(cells=32, t_end=1.0, frame_steps=100)
    ids = np.array([0, 3, 31, 32, 40])  # spans two 32-trajectory blocks, second partial
    eng = Engine(System(**spec), precision=32, kernel="mfma")
    res = eng.run(ids, traj=True)
    assert res.n_steps == 200
    worst = 0.0
    for j, t in enumerate(ids):
        fr, _, _ = oracle.run_trajectory(dict(spec, lattice=oracle.LAT_DENSE, E=spec["omega"]), int(t))
        worst = max(worst, rel_err(res.traj[j], fr))
    print(f"MFMA FP32 vs FP64 oracle, 200 steps: max relative error {worst:.3e}")
    assert worst < MFMA_BAR
    eng.close()


def test_mfma_vs_sparse_fp32():
    spec = c3_spec(cells=128, t_end=0.5, frame_steps=50)
    ids = np.arange(40)
    a = Engine(System(**spec), precision=32, kernel="mfma").run(ids, traj=True)
    b = Engine(System(**spec), precision=32, kernel="generic").run(ids, traj=True)
    err = rel_err(a.traj, b.traj)
    print(f"MFMA vs sparse-product FP32 after 100 steps: {err:.3e}")
    assert err < MFMA_BAR


def test_mfma_ensemble_sums_match_dump():
    spec = c3_spec(cells=64, t_end=0.2, frame_steps=20)
    eng = Engine(System(**spec), precision=32, kernel="mfma")
    res = eng.run(np.arange(70), traj=True)
    ref = from_trajectories(res.traj, (1, 192))
    for k in ("sum_n", "sum_n2", "sum_beta", "traj_m2", "traj_m4"):
        assert rel_err(getattr(res, k), ref[k]) < 1e-9, k
    eng.close()


@pytest.mark.parametrize("n", [128, 250, 380])
def test_overlapped_step_sizes_vs_fp64_oracle(n):
    """Dense random couplings at the sizes whose step runs the noise and drift
    of one site pair per k block between the GEMM's MFMAs (padded sizes 128,
    256 and 384: one, two and three row tiles per wave), 250 and 380 with
    padded sites."""
    rng = np.random.default_rng(n)
    C = rng.uniform(-0.05, 0.05, (n, n))
    C = C + C.T
    spec = dict(nx=1, ny=n, dt=0.005, gamma=1.0, J=0.0, U=0.1, F=1.2, omega=0.2 * rng.standard_normal(n),
                beta_init=0.0, beta_sigma=0.5, frame_steps=50, seed=91, t_start=0.0, t_end=0.5, lattice="matrix",
                coupling=C, detuning=0.3, onsite="detuning-E")
    ids = np.array([0, 7, 33])  # two 32-trajectory blocks
    print(f"mfma {n} sites", flush=True)
    eng = Engine(System(**spec), precision=32, kernel="mfma")
    res = eng.run(ids, traj=True)
    eng.close()
    err = max(rel_err(res.traj[j], oracle.run_trajectory(dict(spec, lattice=oracle.LAT_DENSE, E=spec["omega"]),
                                                         int(t))[0]) for j, t in enumerate(ids))
    print(f"  MFMA {n} sites vs FP64 oracle, 100 steps: {err:.2e}")
    assert err < MFMA_BAR
