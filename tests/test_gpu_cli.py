"""End to end on the GPU: .ini -> twmc_sim -> GPU solver plugin -> PincoFormat
archive, compared with the FP64 oracle trajectory by trajectory; resume of a
run folder continues the same Philox stream."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from helpers import FP32_BAR, rel_err
from pincosimulator_amd.pinco_format import Archive

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIM = os.path.join(ROOT, "pincosimulator_amd", "twmc_sim")

CHAIN = """Manager = "TWMC"
nx = 16
ny = 1
PBC = true
J = 1.0
omega = 0.5
U = 0.1
F = 1.0
gamma = 1
beta_init = 0
beta_init_sigma = 0.5
t_start = 0
t_end = {t_end}
timestep = 0.005
frames_freq = 5
n_traj = 6
SEED = 42
precision = 64
"""


def run_sim(folder, ini_text, *over):
    os.makedirs(folder, exist_ok=True)
    ini = os.path.join(folder, "run.ini")
    with open(ini, "w") as f:
        f.write(ini_text)
    p = subprocess.run([SIM, "-i", ini, "-o", folder, *over], capture_output=True, text=True, timeout=300)
    print(p.stdout[-2000:], p.stderr[-2000:])
    assert p.returncode == 0, p.stderr
    return Archive(folder)


def oracle_frames(tid, t_end):
    p = dict(nx=1, ny=16, lattice=oracle.LAT_FFT, dt=0.005, gamma=1.0, J=1.0, omega=0.5, U=0.1, F=1.0,
             beta_init=0.0, beta_sigma=0.5, frame_steps=40, seed=42, t_start=0.0, t_end=t_end)
    fr, _, tf = oracle.run_trajectory(p, tid)
    return fr, tf


def test_ini_to_archive_matches_oracle(tmp_path):
    arc = run_sim(str(tmp_path / "run"), CHAIN.format(t_end=1.0))
    assert arc.ids() == [1, 2, 3, 4, 5, 6]
    assert "traj" in arc.names and arc.formats[arc.names.index("traj")] == 22
    for tid in arc.ids():
        d = arc.load(tid)
        fr, tf = oracle_frames(tid, 1.0)
        assert d["traj"].shape == (6, 1, 16)
        assert rel_err(d["traj"].reshape(6, 16), fr) < 1e-9
        assert d["t"][1] == tf
    t = np.loadtxt(str(tmp_path / "run" / "_t.dat"), comments="%")
    assert len(t) == 6 and abs(t[-1] - 1.0) < 1e-12
    obs = np.loadtxt(str(tmp_path / "run" / "_observables.dat"), comments="%")
    assert obs.shape == (6, 3)


def test_resume_continues_the_same_stream(tmp_path):
    folder = str(tmp_path / "resume")
    run_sim(folder, CHAIN.format(t_end=1.0))
    arc = run_sim(folder, CHAIN.format(t_end=2.0))
    assert arc.ids() == [1, 2, 3, 4, 5, 6]
    assert len(arc.entries) == 12  # one continuation entry per trajectory
    for tid in (1, 6):
        d = arc.load(tid)
        fr, _ = oracle_frames(tid, 2.0)
        assert d["traj"].shape[0] == 11
        assert rel_err(d["traj"].reshape(11, 16), fr) < 1e-9
        last = arc.load(tid, last_frame_only=True)
        assert np.array_equal(last["traj"][0], d["traj"][-1])


SQUARE32 = """Manager = "TWMC"
nx = {n}
ny = {n}
PBC = true
J = 1.0
omega = 1.0
U = 0.05
F = 2.0
gamma = 1
beta_init = 0
beta_init_sigma = 0.5
t_start = 0
t_end = 0.4
timestep = 0.002
frames_freq = 5
n_traj = 4
SEED = 7
precision = 32
"""


@pytest.mark.parametrize("n", [16, 64])
def test_fp32_ini_runs_on_the_fast_kernels(tmp_path, n):
    """precision = 32 in the .ini: the plugin's AUTO choice is the stacked
    stencil (16x16) or the register-resident 64x64 kernel; the archive's
    trajectories match the FP64 oracle to the FP32 bar."""
    arc = run_sim(str(tmp_path / f"sq{n}"), SQUARE32.format(n=n))
    assert arc.ids() == [1, 2, 3, 4]
    p = dict(nx=n, ny=n, lattice=oracle.LAT_FFT, dt=0.002, gamma=1.0, J=1.0, omega=1.0, U=0.05, F=2.0,
             beta_init=0.0, beta_sigma=0.5, frame_steps=100, seed=7, t_start=0.0, t_end=0.4)
    for tid in arc.ids():
        d = arc.load(tid)
        fr, _, _ = oracle.run_trajectory(p, tid)
        assert d["traj"].shape == (3, n, n)
        err = rel_err(d["traj"].reshape(3, n * n), fr)
        assert 1e-10 < err < FP32_BAR  # FP32 rounding is visible: the FP32 kernels ran
