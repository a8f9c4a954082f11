"""Batches of one and three trajectories on every kernel (the grid, slot and
chunk arithmetic at its smallest): per trajectory against the oracle, and the
ensemble sums against the dump."""
import numpy as np
import pytest

import oracle
from helpers import FP32_BAR, MFMA_BAR, rel_err, system
from pincosimulator_amd import Engine, System
from pincosimulator_amd.observables import from_trajectories
from test_gpu_mfma import c3_spec

pytestmark = pytest.mark.gpu


def square(nx, ny):
    return dict(nx=nx, ny=ny, dt=0.004, gamma=1.0, J=1.0, omega=0.5, U=0.06, F=1.1, beta_init=0.0, beta_sigma=0.5,
                frame_steps=40, seed=2718, t_start=0.0, t_end=0.32)


CASES = [("stencil64", square(64, 64), 32), ("fused64", square(64, 64), 32), ("stack", square(16, 12), 32),
         ("generic", square(6, 10), 32), ("generic", square(6, 10), 64),
         ("band", c3_spec(cells=16, t_end=0.32, frame_steps=40), 32),
         ("mfma", c3_spec(cells=16, t_end=0.32, frame_steps=40), 32)]


@pytest.mark.parametrize("kernel,spec,prec", CASES, ids=[f"{c[0]}-fp{c[2]}" for c in CASES])
@pytest.mark.parametrize("ids", [[41], [7, 2, 5]], ids=["one", "three"])
def test_small_batch(kernel, spec, prec, ids):
    ids = np.array(ids, dtype=np.uint64)
    print(f"{kernel} FP{prec} batch of {ids.size}", flush=True)
    eng = Engine(System(**spec), precision=prec, kernel=kernel)
    res = eng.run(ids, traj=True)
    eng.close()
    assert res.kernel == kernel
    lat = oracle.LAT_DENSE if spec.get("lattice") == "matrix" else oracle.LAT_FFT
    p = dict(spec, lattice=lat, E=spec["omega"]) if lat == oracle.LAT_DENSE else dict(spec, lattice=lat)
    err = max(rel_err(res.traj[j], oracle.run_trajectory(p, int(t))[0]) for j, t in enumerate(ids))
    tol = {64: 1e-9}.get(prec, MFMA_BAR if kernel == "mfma" else FP32_BAR)
    assert err < tol, err
    shape = (spec["nx"], spec["ny"])
    ref = from_trajectories(res.traj, shape)
    for k in ("sum_n", "sum_n2", "sum_beta", "traj_m2", "traj_m4"):
        assert rel_err(getattr(res, k), ref[k]) < 1e-9, k
