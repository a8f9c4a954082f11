"""Edges of the supported range on the GPU: the largest lattices the
LDS-resident kernel holds (FP32 100x100, FP64 70x70 -- 70 = 2 5 7 also takes
the direct-sum radix-7 pass), the first size past it (a clear error), the
MFMA kernel at its 512-site maximum, and empty or malformed batches."""
import numpy as np
import pytest

import oracle
from helpers import FP32_BAR, MFMA_BAR, rel_err, system
from pincosimulator_amd import Engine, System
from pincosimulator_amd._lib import TWMCError

pytestmark = pytest.mark.gpu


def spec2d(nx, ny, **kw):
    s = dict(nx=nx, ny=ny, dt=0.005, gamma=1.0, J=1.0, omega=0.7, U=0.08, F=1.3, beta_init=0.1, beta_sigma=0.5,
             frame_steps=50, seed=31337, t_start=0.0, t_end=0.5)
    s.update(kw)
    return s


@pytest.mark.parametrize("nx,ny,precision,tol", [(100, 100, 32, FP32_BAR), (70, 70, 64, 1e-9)])
def test_largest_lds_lattice_matches_oracle(nx, ny, precision, tol):
    spec = spec2d(nx, ny)
    ids = np.array([4, 11])
    print(f"generic {nx}x{ny} FP{precision}", flush=True)
    eng = Engine(system(spec), precision=precision, kernel="generic")
    res = eng.run(ids, traj=True)
    eng.close()
    assert res.kernel == "generic" and res.n_steps == 100
    err = max(rel_err(res.traj[j], oracle.run_trajectory(dict(spec, lattice=oracle.LAT_FFT), int(t))[0])
              for j, t in enumerate(ids))
    print(f"  max relative deviation {err:.2e}")
    assert err < tol


def test_first_size_past_the_lds_is_rejected():
    with pytest.raises(TWMCError, match="too large"):
        Engine(system(spec2d(101, 101)), precision=32, kernel="generic")
    with pytest.raises(TWMCError, match="too large"):
        Engine(system(spec2d(72, 72)), precision=64, kernel="generic")


def test_mfma_at_512_sites_matches_oracle():
    n = 512
    rng = np.random.default_rng(9)
    C = rng.uniform(-0.05, 0.05, (n, n))
    C = C + C.T
    spec = dict(nx=1, ny=n, dt=0.005, gamma=1.0, J=0.0, U=0.1, F=1.0, omega=0.2 * rng.standard_normal(n),
                beta_init=0.0, beta_sigma=0.5, frame_steps=50, seed=77, t_start=0.0, t_end=0.5, lattice="matrix",
                coupling=C, detuning=0.3, onsite="detuning-E")
    ids = np.array([0, 31, 32, 39])  # two 32-trajectory blocks, the second partial
    print("mfma 512 sites", flush=True)
    eng = Engine(System(**spec), precision=32, kernel="mfma")
    res = eng.run(ids, traj=True)
    eng.close()
    err = max(rel_err(res.traj[j], oracle.run_trajectory(dict(spec, lattice=oracle.LAT_DENSE, E=spec["omega"]),
                                                         int(t))[0]) for j, t in enumerate(ids))
    print(f"  MFMA 512 sites vs oracle: {err:.2e}")
    assert err < MFMA_BAR


def test_empty_and_malformed_batches_are_errors():
    eng = Engine(system(spec2d(8, 8)), precision=32)
    with pytest.raises(TWMCError, match="empty"):
        eng.run(np.array([], dtype=np.uint64))
    eng.close()
    with pytest.raises((TWMCError, ValueError)):
        Engine(system(spec2d(0, 8)), precision=32)
    with pytest.raises((TWMCError, ValueError)):
        Engine(system(spec2d(8, 8, dt=-0.1)), precision=32)
