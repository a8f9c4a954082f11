"""Shared test helpers: one parameter spec drives both the HIP engine and the oracle."""
import numpy as np

import oracle
from pincosimulator_amd import Engine, System


def c1_spec(**kw):
    """1D Bose-Hubbard chain, N=64, PBC (the 1D uniform configuration)."""
    s = dict(nx=1, ny=64, dt=0.005, gamma=1.0, J=1.0, U=0.1, omega=0.5, F=1.0,
             beta_init=0.0, beta_sigma=0.5, frame_steps=200, seed=1234, t_start=0.0, t_end=100.0)
    s.update(kw)
    return s


def c2_spec(**kw):
    """2D 64x64 PBC square lattice (the headline configuration)."""
    s = dict(nx=64, ny=64, dt=0.002, gamma=1.0, J=1.0, U=0.05, omega=1.0, F=2.0,
             beta_init=0.0, beta_sigma=0.5, frame_steps=500, seed=2024, t_start=0.0, t_end=100.0)
    s.update(kw)
    return s


def system(spec):
    keys = ("nx", "ny", "dt", "gamma", "J", "omega", "U", "F", "F_t", "beta_init", "beta_sigma", "frame_steps",
            "seed", "t_start", "t_end", "lattice", "coupling", "detuning", "onsite")
    return System(**{k: spec[k] for k in keys if k in spec})


def oracle_params(spec):
    p = dict(spec)
    if spec.get("lattice", "square") == "matrix":
        p["lattice"] = oracle.LAT_DENSE
        p["E"] = spec.get("omega", 0.0)
    else:
        p["lattice"] = oracle.LAT_FFT
    return p


def oracle_traj(spec, ids, **kw):
    p = oracle_params(spec)
    frames, finals, tfs = [], [], []
    for t in ids:
        fr, bf, tf = oracle.run_trajectory(p, int(t), **kw)
        frames.append(fr)
        finals.append(bf)
        tfs.append(tf)
    return np.array(frames), np.array(finals), tfs


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# Parity bars of the FP32 paths, each about ten times the largest error measured
# on MI355X for that path (FP32 rounding only: the FP32 and FP64 runs share the
# Philox stream), so that a systematic FP32 regression of ~1e-5 relative fails.
FP32_BAR = 2e-5       # FP32 kernels vs the FP64 oracle, per trajectory, relative to max |beta| (<= 1000 steps)
MFMA_BAR = 5e-5       # BF16 hi/lo x3 MFMA hopping (FP32 state) vs the FP64 oracle, per trajectory
OBS_SIGMA_BAR = 0.05  # FP32 vs FP64 ensemble observables, in standard errors of the FP64 ensemble
