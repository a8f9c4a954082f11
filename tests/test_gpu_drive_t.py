"""Time-dependent drive F(t) (TWMCLiebSolver.cpp:252-256, piecewise-linear
table of NoisyMatrix.cpp:164-264) on the GPU vs the oracle: FP64 generic kernel
per trajectory within 1e-9 for Lieb (matrix) and FFT lattices, the FP32 MFMA
kernel within the BF16x3 bar, and the .ini path with a table file that has
gaps ('*') filled by interpolation."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from helpers import MFMA_BAR, rel_err, system
from pincosimulator_amd import Engine
from pincosimulator_amd.pinco_format import Archive

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIM = os.path.join(ROOT, "pincosimulator_amd", "twmc_sim")


def table(n, seed=3):
    rng = np.random.default_rng(seed)
    times = np.array([0.1, 0.45, 0.8])
    vals = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
    return times, vals


def lieb_spec(cells, **kw):
    C = oracle.lieb_coupling(cells, 1.0, 1.0, False)
    s = dict(nx=1, ny=3 * cells, lattice="matrix", coupling=C, onsite="detuning-E", detuning=0.5, omega=0.0,
             dt=0.005, gamma=1.0, J=0.0, U=0.1, F=0.0, beta_init=0.0, beta_sigma=0.5, frame_steps=40, seed=808,
             t_start=0.0, t_end=1.2, F_t=table(3 * cells))
    s.update(kw)
    return s


def oracle_frames(spec, ids):
    p = dict(spec, E=spec.get("omega", 0.0))
    p["lattice"] = oracle.LAT_DENSE if spec.get("lattice") == "matrix" else oracle.LAT_FFT
    return np.array([oracle.run_trajectory(p, int(t))[0] for t in ids])


@pytest.mark.parametrize("kind", ["lieb", "chain", "square"])
def test_drive_table_fp64_matches_oracle(kind):
    if kind == "lieb":
        spec = lieb_spec(8)
    else:
        nx, ny = (1, 32) if kind == "chain" else (6, 10)
        spec = dict(nx=nx, ny=ny, dt=0.005, gamma=1.0, J=1.0, omega=0.5, U=0.1, F=0.0, beta_init=0.0,
                    beta_sigma=0.5, frame_steps=40, seed=808, t_start=0.0, t_end=1.2, F_t=table(nx * ny))
    ids = np.arange(4) + 70
    print(f"F(t) fp64 {kind}", flush=True)
    eng = Engine(system(spec), precision=64, kernel="generic")
    res = eng.run(ids, traj=True)
    eng.close()
    ref = oracle_frames(spec, ids)
    assert res.traj.shape == ref.shape
    assert rel_err(res.traj, ref) < 1e-9
    # the drive matters: a static F = F(t_0) run differs
    static = oracle_frames(dict(spec, F_t=None, F=spec["F_t"][1][0]), ids[:1])
    assert rel_err(static[0], ref[0]) > 1e-3


def test_drive_table_mfma_matches_oracle():
    spec = lieb_spec(32, t_end=1.0, frame_steps=50)
    ids = np.arange(64) + 1000
    print("F(t) fp32 mfma", flush=True)
    eng = Engine(system(spec), precision=32, kernel="mfma")
    res = eng.run(ids, traj=True)
    eng.close()
    assert res.kernel == "mfma"
    ref = oracle_frames(spec, ids[:8])
    assert rel_err(res.traj[:8], ref) < MFMA_BAR


def test_drive_table_mfma_overlapped_step_matches_oracle():
    """384 sites: the MFMA step that reads the F(t) table during the GEMM."""
    spec = lieb_spec(128, t_end=0.6, frame_steps=40)
    ids = np.array([5, 17, 40])
    print("F(t) fp32 mfma, 384 sites", flush=True)
    eng = Engine(system(spec), precision=32, kernel="mfma")
    res = eng.run(ids, traj=True)
    eng.close()
    assert res.kernel == "mfma"
    assert rel_err(res.traj, oracle_frames(spec, ids)) < MFMA_BAR


def test_drive_table_is_not_sent_to_the_fused_kernel():
    spec = dict(nx=64, ny=64, dt=0.002, gamma=1.0, J=1.0, omega=1.0, U=0.05, F=0.0, beta_init=0.0,
                beta_sigma=0.5, frame_steps=10, seed=1, t_start=0.0, t_end=0.02,
                F_t=(np.array([0.0, 1.0]), np.ones((2, 4096)) * np.array([[1.0], [2.0]])))
    eng = Engine(system(spec), precision=32, kernel="auto")
    res = eng.run(np.arange(2))
    eng.close()
    assert res.kernel == "generic"


LIEB_INI = """Manager = "TWMC"
lattice = "lieb"
nx = 8
ny = 1
PBC = false
J_AB = 1.0
J_BC = 1.0
detuning = 0.5
E = 0.0
U = 0.1
F_t = "_Ft_lieb.dat"
gamma = 1
beta_init = 0
beta_init_sigma = 0.5
t_start = 0
t_end = 1.2
timestep = 0.005
frames_freq = 5
n_traj = 3
SEED = 11
precision = 64
"""


def test_ini_time_dependent_drive_with_gaps(tmp_path):
    n = 24
    times, vals = table(n, seed=5)
    vals = vals.real  # real table file: one value per site
    folder = str(tmp_path / "ft")
    os.makedirs(folder)
    with open(os.path.join(folder, "_Ft_lieb.dat"), "w") as f:
        for k, t in enumerate(times):
            row = [("*" if (k == 1 and i % 5 == 0) else f"{vals[k, i]:.17g}") for i in range(n)]
            f.write(f"{t}\n{' '.join(row)}\n\n")
    with open(os.path.join(folder, "run.ini"), "w") as f:
        f.write(LIEB_INI)
    p = subprocess.run([SIM, "-i", os.path.join(folder, "run.ini"), "-o", folder], capture_output=True, text=True,
                       timeout=300, cwd=folder)
    print(p.stdout[-1500:], p.stderr[-1500:])
    assert p.returncode == 0, p.stderr
    # gaps filled by linear interpolation between the neighbouring times
    filled = vals.copy()
    for i in range(0, n, 5):
        filled[1, i] = np.interp(times[1], times[[0, 2]], vals[[0, 2], i])
    arc = Archive(folder)
    spec = lieb_spec(8, seed=11, frame_steps=40, F_t=(times, filled.astype(complex)))
    for tid in arc.ids():
        d = arc.load(tid)
        ref = oracle_frames(spec, [tid])[0]
        assert rel_err(d["traj"].reshape(ref.shape), ref) < 1e-9
