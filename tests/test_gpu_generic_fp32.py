"""The LDS kernel on awkward shapes -- dimensions of length 1, 2, odd and
prime lengths (the FFT's direct-sum passes for prime factors above 5) -- and
on power-of-two lattices (the neighbour indices as masks), against the oracle
that applies the reference's FFT form.  The step launches use the hopping's
nearest-neighbour form and the FFT only at captures.  FP32: 2e-5 of max
|beta| after 200 steps; FP64: 1e-9."""
import numpy as np
import pytest

import oracle
from helpers import rel_err, system
from pincosimulator_amd import Engine

pytestmark = pytest.mark.gpu


SHAPES = [(1, 2), (1, 3), (2, 2), (2, 5), (16, 15), (5, 7), (1, 60), (1, 127), (7, 11), (13, 14), (4, 32), (32, 1)]


@pytest.mark.parametrize("nx,ny", SHAPES)
def test_fp32_nearest_neighbour_matches_fft_oracle(nx, ny):
    spec = dict(nx=nx, ny=ny, dt=0.005, gamma=1.0, J=1.3, omega=0.4, U=0.1, F=1.1, beta_init=0.2,
                beta_sigma=0.5, frame_steps=100, seed=515, t_start=0.0, t_end=1.0)
    ids = np.arange(6) + 40
    print(f"fp32 generic {nx}x{ny}", flush=True)
    eng = Engine(system(spec), precision=32, kernel="generic")
    res = eng.run(ids, traj=True)
    eng.close()
    assert res.kernel == "generic"
    ref = np.array([oracle.run_trajectory(dict(spec, lattice=oracle.LAT_FFT), int(t))[0] for t in ids])
    assert res.traj.shape == ref.shape
    err = rel_err(res.traj, ref)
    print(f"  max relative deviation {err:.2e}")
    assert err < 2e-5
    # k-space sums still come from the FFT at the captures
    nk_ref = (np.abs(np.fft.fft2(ref.reshape(len(ids), -1, nx, ny), axes=(-2, -1))) ** 2).sum(axis=0)
    assert rel_err(res.sum_nk, nk_ref.reshape(res.sum_nk.shape)) < 2e-5


# power-of-two lattices (ny >= 2, nx = 1 or >= 2) take the split-layout step
# launch whose neighbour indices are masks; other even row lengths ((13, 14),
# (8, 6), (48, 48)) the split layout with the division form; odd row lengths
# and (32, 1) the site-order layout; 2049 to 4096 sites the 1024-thread launch,
# double-buffered in the split layouts: (4, 1024), (128, 32), (32, 128), (48, 48)
@pytest.mark.parametrize("nx,ny", [(1, 7), (5, 7), (1, 127), (7, 11), (13, 14),
                                   (1, 2), (1, 32), (2, 8), (4, 2), (16, 4), (8, 6), (32, 1),
                                   (4, 1024), (128, 32), (32, 128), (48, 48)])
def test_fp64_awkward_and_power_of_two_shapes_match_oracle(nx, ny):
    spec = dict(nx=nx, ny=ny, dt=0.005, gamma=1.0, J=1.3, omega=0.4, U=0.1, F=1.1, beta_init=0.2,
                beta_sigma=0.5, frame_steps=50, seed=516, t_start=0.0, t_end=0.5)
    ids = np.arange(4) + 7
    print(f"fp64 generic {nx}x{ny}", flush=True)
    eng = Engine(system(spec), precision=64, kernel="generic")
    res = eng.run(ids, traj=True)
    eng.close()
    ref = np.array([oracle.run_trajectory(dict(spec, lattice=oracle.LAT_FFT), int(t))[0] for t in ids])
    err = rel_err(res.traj, ref)
    print(f"  max relative deviation {err:.2e}")
    assert err < 1e-9


# site-varying parameters (a site-dependent detuning, per-trajectory U and F
# disorder): the split-layout step launch that reads the parameter tables per
# site, with mask indices on power-of-two lattices and the division form on
# other even row lengths ((6, 10), (12, 14), (40, 48))
@pytest.mark.parametrize("nx,ny,prec", [(16, 16, 32), (16, 16, 64), (1, 64, 32), (64, 64, 64), (8, 32, 32),
                                        (6, 10, 64), (12, 14, 32), (40, 48, 32)])
def test_site_varying_parameters_split_layout(nx, ny, prec):
    n = nx * ny
    spec = dict(nx=nx, ny=ny, dt=0.005, gamma=1.0, J=1.1, omega=0.4 + 0.3 * np.cos(0.7 * np.arange(n)),
                U={"base": 0.1, "noise": ("gaussian", 0.02 + 0.01j)}, F={"base": 1.2, "noise": ("uniform", 0.3)},
                beta_init=0.1, beta_sigma=0.5, frame_steps=50, seed=518, t_start=0.0, t_end=0.5)
    ids = np.arange(4) + 21
    print(f"fp{prec} generic {nx}x{ny}, site-varying parameters", flush=True)
    eng = Engine(system(spec), precision=prec, kernel="generic")
    res = eng.run(ids, traj=True)
    eng.close()
    ref = np.array([oracle.run_trajectory(dict(spec, lattice=oracle.LAT_FFT), int(t))[0] for t in ids])
    err = rel_err(res.traj, ref)
    print(f"  max relative deviation {err:.2e}")
    assert err < (1e-9 if prec == 64 else 2e-5)


# lattices of up to 64 sites: 64-thread step launches that pack 2 (up to 64
# sites) or 4 (up to 32 sites) trajectories per workgroup, each with its own
# LDS state; odd batch sizes leave groups of the last workgroup past the end
# of the batch (they must store nothing); per-trajectory disorder reads each
# group's own parameter tables
@pytest.mark.parametrize("nx,ny,prec,n_traj", [(1, 32, 64, 7), (4, 8, 32, 9), (8, 8, 64, 5), (1, 64, 32, 3),
                                               (2, 12, 64, 11), (6, 10, 32, 13)])
def test_packed_trajectories_small_lattices(nx, ny, prec, n_traj):
    n = nx * ny
    spec = dict(nx=nx, ny=ny, dt=0.005, gamma=1.0, J=0.9, omega=0.3 + 0.2 * np.sin(0.9 * np.arange(n)),
                U={"base": 0.1, "noise": ("gaussian", 0.02)}, F={"base": 1.1, "noise": ("uniform", 0.2)},
                beta_init=0.1, beta_sigma=0.5, frame_steps=40, seed=519, t_start=0.0, t_end=0.4)
    ids = np.arange(n_traj) * 5 + 3
    print(f"fp{prec} generic {nx}x{ny}, {n_traj} trajectories", flush=True)
    eng = Engine(system(spec), precision=prec, kernel="generic")
    res = eng.run(ids, traj=True)
    eng.close()
    ref = np.array([oracle.run_trajectory(dict(spec, lattice=oracle.LAT_FFT), int(t))[0] for t in ids])
    err = rel_err(res.traj, ref)
    print(f"  max relative deviation {err:.2e}")
    assert err < (1e-9 if prec == 64 else 2e-5)
    # site-uniform parameters too (the other instantiation)
    spec_u = dict(spec, omega=0.3, U=0.1, F=1.1)
    eng = Engine(system(spec_u), precision=prec, kernel="generic")
    res = eng.run(ids, traj=True)
    eng.close()
    ref = np.array([oracle.run_trajectory(dict(spec_u, lattice=oracle.LAT_FFT), int(t))[0] for t in ids])
    assert rel_err(res.traj, ref) < (1e-9 if prec == 64 else 2e-5)
