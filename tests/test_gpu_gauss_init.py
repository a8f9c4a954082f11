"""Warm-start initial state (TWMC_INIT_GAUSSIAN) on the GPU vs the oracle's
Gaussian sampler: the FP64 state equals the oracle's sample per trajectory,
and the trajectory that starts from it equals the oracle's run from the same
state (same Philox increments); the FP32 64x64 kernel starts from the same
sample to FP32 rounding."""
import numpy as np
import pytest

import oracle
from helpers import c2_spec, rel_err, system
from pincosimulator_amd import Engine

pytestmark = pytest.mark.gpu


def nk_profile(nx, ny):
    kx = 2 * np.pi * np.fft.fftfreq(nx)[:, None]
    ky = 2 * np.pi * np.fft.fftfreq(ny)[None, :]
    nk = 0.3 * np.exp(-(kx ** 2 + ky ** 2)) + 0.01
    nk[0, 0] = 0.1
    return nk


@pytest.mark.parametrize("nx,ny", [(1, 64), (12, 10), (16, 16)])
def test_gaussian_start_fp64_matches_oracle(nx, ny):
    spec = dict(nx=nx, ny=ny, dt=0.005, gamma=1.0, J=1.0, U=0.1, omega=0.5, F=1.0, beta_init=0.0,
                beta_sigma=0.5, frame_steps=50, seed=31337, t_start=0.0, t_end=1.0)
    mean = 0.9 - 0.4j
    nk = nk_profile(nx, ny)
    ids = np.arange(5) + 900
    print(f"gaussian start fp64 {nx}x{ny}", flush=True)
    eng = Engine(system(spec), precision=64, kernel="generic")
    res = eng.run(ids, traj=True, init="gaussian", gauss_mean=mean, gauss_nk=nk)
    eng.close()
    p = dict(spec, lattice=oracle.LAT_FFT)
    for j, t in enumerate(ids):
        b0 = oracle.gaussian_state(nx, ny, spec["seed"], int(t), mean, nk).reshape(-1)
        assert rel_err(res.traj[j, 0], b0) < 1e-12
        fr, bf, tf = oracle.run_trajectory(p, int(t), init_mode=oracle.INIT_PREVIOUS, beta_prev=b0)
        assert fr.shape == res.traj[j].shape
        assert rel_err(res.traj[j], fr) < 1e-9


@pytest.mark.parametrize("kernel", ["stencil64", "fused64"])
def test_gaussian_start_fp32_64x64(kernel):
    spec = c2_spec(t_end=0.2, frame_steps=50)
    mean = 1.2 + 0.5j
    nk = nk_profile(64, 64)
    ids = np.arange(64) + 5000
    print(f"gaussian start fp32 64x64 {kernel}", flush=True)
    eng = Engine(system(spec), precision=32, kernel=kernel)
    res = eng.run(ids, traj=True, init="gaussian", gauss_mean=mean, gauss_nk=nk)
    eng.close()
    assert res.kernel == kernel
    for j in (0, 17, 63):
        b0 = oracle.gaussian_state(64, 64, spec["seed"], int(ids[j]), mean, nk).reshape(-1)
        assert rel_err(res.traj[j, 0], b0) < 2e-6
    # populations of the stored initial frame (k = 0 with the coherent part removed)
    n = 4096
    k = res.sum_nk[0].reshape(64, 64) / len(ids) / n - 0.5
    k[0, 0] -= n * abs(mean) ** 2
    sel = np.ones((64, 64), bool)
    sel[0, 0] = False
    # per mode the estimate has standard error (n_k + 1/2)/sqrt(B); averaged over 4095 modes
    err = (k[sel] - nk[sel]).mean()
    assert abs(err) < 4 * np.sqrt(((nk[sel] + 0.5) ** 2).mean() / len(ids) / sel.sum())
