"""The FP64 step kernel's Wiener increments, recovered exactly from single
Euler steps, equal the oracle's Philox + libm Box-Muller to 1e-13.

With J = U = F = 0 the update of site i is
    b' = b (1 + dt (-i omega_i - gamma/2)) + sqrt(gamma dt / 4) xi_i,
so xi_i = (b' - b (1 + dt (-i omega_i - gamma/2))) / sqrt(gamma dt / 4) from
two consecutive stored frames (one step apart).  The kernel draws xi from its
own range-specialised FP64 Box-Muller (ln, sin and cos within 2 ulp); the
oracle from glibc's log / sin / cos on the same Philox words.  Covered: the
site-uniform launch (the parameters read once) and site-varying omega on a
small lattice, and the 64x64 site-uniform launch (1024 threads, two LDS
copies of the state), there for the first step of every site.
"""
import numpy as np
import pytest

import oracle
from helpers import system
from pincosimulator_amd import Engine

pytestmark = pytest.mark.gpu

DT, GAMMA, SEED = 0.002, 1.0, 20261203


def _spec(omega, nx, ny):
    return dict(nx=nx, ny=ny, dt=DT, gamma=GAMMA, J=0.0, U=0.0, omega=omega, F=0.0, beta_init=0.3 + 0.1j,
                beta_sigma=0.0, frame_steps=1, seed=SEED, t_start=0.0, t_end=3 * DT)


# 8x6: the 64-thread launch with two trajectories per workgroup; 64x64 with
# site-uniform parameters: the 1024-thread double-buffered launch
@pytest.mark.parametrize("site_uniform,nx,ny", [(True, 8, 6), (False, 8, 6), (True, 64, 64)])
def test_fp64_increments_equal_oracle_normals(site_uniform, nx, ny):
    n = nx * ny
    omega = 0.7 if site_uniform else 0.5 + 0.25 * np.cos(np.arange(n))
    spec = _spec(omega, nx, ny)
    ids = np.array([0, 1, 7, 4097, (1 << 33) + 5], dtype=np.uint64)
    eng = Engine(system(spec), precision=64, kernel="generic")
    res = eng.run(ids, traj=True)
    eng.close()
    traj = res.traj  # (B, frames, n): frames at steps 0, 1, 2, 3
    assert traj.shape == (ids.size, 4, n)
    w = np.broadcast_to(np.asarray(omega, dtype=np.float64), (n,))
    g = 1.0 + DT * (-1j * w - GAMMA / 2.0)
    sig = np.sqrt(GAMMA * DT / 4.0)
    worst = 0.0
    for j, t in enumerate(ids):
        for step in range(3 if n <= 64 else 1):
            xi = (traj[j, step + 1] - traj[j, step] * g) / sig
            ref = np.array([oracle.normal(SEED, int(t), oracle.TAG_WIENER, step, i) for i in range(n)])
            worst = max(worst, float(np.max(np.abs(xi - ref))))
    print(f"max |xi - oracle| = {worst:.2e}")
    assert worst < 1e-13
