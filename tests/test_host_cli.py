"""C++ host layer on the CPU: .ini parsing, matrix input forms, solver
selection and the frame schedule, through `twmc_sim --dry-run` (no GPU)."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIM = os.path.join(ROOT, "pincosimulator_amd", "twmc_sim")
INI = os.path.join(ROOT, "tests", "golden", "ini")


def dry(ini, *over):
    p = subprocess.run([SIM, "--dry-run", "-i", ini, *over], capture_output=True, text=True, timeout=60)
    return p.returncode, (json.loads(p.stdout) if p.returncode == 0 else p.stderr)


def test_host_library_and_driver_exist():
    assert os.path.exists(SIM)
    assert os.path.exists(os.path.join(ROOT, "pincosimulator_amd", "libpinco_host.so"))


def test_square_lattice_with_gaussian_F_noise():
    rc, d = dry(os.path.join(INI, "square4x4_noisyF.ini"))
    assert rc == 0
    assert d["solver"] == "TWMCBase" and (d["nx"], d["ny"]) == (4, 4) and d["dimension"] == 2
    assert d["F_noise"] == 2 and d["F0"] == [3.0, 0.0] and d["J"] == 0.45
    s = oracle.schedule(0.0, 20.0, 0.01, frames_freq=1.0)
    assert (d["frame_steps"], d["nFramesTot"], d["t_end"], d["n_dt"]) == (
        s["frame_steps"], s["n_frames_tot"], s["t_end"], s["n_dt"])


def test_lieb_with_site_files_and_E_noise():
    rc, d = dry(os.path.join(INI, "lieb24_open.ini"))
    assert rc == 0
    assert d["solver"] == "TWMCLieb" and d["cellSz"] == 3
    # a 24 x 1 lattice is stored as 1 x 24 (TWMCSystemData.cpp:25-37)
    assert (d["nx"], d["ny"]) == (1, 24) and d["dimension"] == 1
    assert d["E_noise"] == 2 and d["E0"] == [-1.0, 0.0] and d["F0"] == [1.5, 0.0]
    s = oracle.schedule(0.0, 700.0, 0.0005, n_frames=150)
    assert (d["frame_steps"], d["nFramesTot"], d["t_end"]) == (s["frame_steps"], s["n_frames_tot"], s["t_end"])


def test_matrix_forms_tables_and_overrides():
    rc, d = dry(os.path.join(INI, "chain16_tables.ini"), "t_start", "0", "t_end", "10", "timestep", "0.005",
                "frames_period", "0.5", "PBC", "true", "gamma", "1", "SEED", "77")
    assert rc == 0
    assert d["solver"] == "TWMCBase" and (d["nx"], d["ny"]) == (1, 16)
    assert d["omega0"] == [0.5, 0.0] and d["U0"] == [0.1, 0.0] and d["beta_sigma0"] == [0.5, 0.0]
    assert d["F0"] == [0.0, 0.0]  # first value of the column-major file
    assert d["frame_steps"] == 100 and d["seed"] == 77 and d["gamma"] == 1.0


def test_no_eligible_solver():
    # Manager = "TWMCThread" forces PBC (TWMCManager.cpp:41-44); with "TWMC"
    # the chain is open and no solver is eligible
    args = ("t_end", "1", "timestep", "0.01", "frames_freq", "1")
    rc, d = dry(os.path.join(INI, "chain16_tables.ini"), *args)
    assert rc == 0 and d["PBC"] is True
    rc, err = dry(os.path.join(INI, "chain16_tables.ini"), *args, "Manager", "'TWMC'")
    assert rc == 3 and "No solver available" in err
    rc, err = dry(os.path.join(INI, "chain16_timedep.ini"))
    assert rc == 3


def test_missing_matrix_is_an_error():
    rc, err = dry(os.path.join(INI, "square4x4_noisyF.ini"), "F_Noise_Val_0", "'missing.dat'")
    assert rc != 0


def test_bad_usage():
    p = subprocess.run([SIM], capture_output=True, text=True, timeout=60)
    assert p.returncode != 0 and "usage" in p.stderr
