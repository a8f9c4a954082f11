"""Parameter sets of the benchmark workloads, shared by bench.py and the tests."""
from __future__ import annotations

import numpy as np

# lattice shapes of the uneven sweep: chains of 32..512 sites and squares of side 16..64
SWEEP_SHAPES = [(1, 32), (1, 64), (1, 128), (1, 256), (1, 512), (16, 16), (32, 32), (48, 48), (64, 64)]


def sweep512_points(n_traj=128):
    """512 points on a 32 x 16 grid of F in [0.5, 4] and detuning in [-2, 3],
    shapes assigned round-robin, n_traj trajectories each (ids 128 k + j),
    U = 0.05, gamma = 1, J = 1, vacuum start; dt = 0.002, frames every 500 steps."""
    Fs = np.linspace(0.5, 4.0, 32)
    Ds = np.linspace(-2.0, 3.0, 16)
    pts = []
    for k in range(512):
        nx, ny = SWEEP_SHAPES[k % len(SWEEP_SHAPES)]
        pts.append(dict(nx=nx, ny=ny, n_traj=n_traj, traj_id0=128 * k, J=1.0, gamma=1.0,
                        omega=float(Ds[k // 32]), U=0.05, F=float(Fs[k % 32]), beta_init=0.0, beta_sigma=0.5))
    return pts


# held-out points of the warm-start summary in the benchmark line: evenly spread over the 64
WARMSTART_POINT_INDICES = [0, 21, 42, 63]
