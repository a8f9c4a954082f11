"""CPU oracle for the TWMC trajectory update -- TEST INFRASTRUCTURE ONLY.

ctypes bindings to ``oracle/libtwmc_oracle.so`` (built from
``oracle/twmc_oracle.cpp`` by ``oracle/Makefile``).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import
this package, and only as the checker / CPU timing baseline.  The product
(``pincosimulator_amd``) never imports it.

Functions restate (reference paths relative to the reference repo):
  * ``run_trajectory``  TWMCBaseSolver::Compute (src/TWMC/Solvers/TWMCBaseSolver.cpp:94-283)
                        and TWMCLiebSolver::Compute (src/TWMC/Solvers/TWMCLiebSolver.cpp:109-299)
  * ``schedule``        TWMCSystemData (src/TWMC/TWMCSystemData.cpp:67-159)
  * ``fft``             FFTW dft_1d/dft_2d as planned at TWMCBaseSolver.cpp:44-60
  * ``noisy_matrix``    NoisyMatrix::Generate (src/Base/NoisyMatrix.cpp:118-160)
  * ``lieb_coupling``   TWMCLiebSolver::Setup (src/TWMC/Solvers/TWMCLiebSolver.cpp:52-85)
  * ``philox``          Philox4x32-10 (Salmon et al., SC'11); no reference counterpart
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtwmc_oracle.so")
_lib = None

TAG_WIENER, TAG_INIT, TAG_U, TAG_OMEGA, TAG_F, TAG_BINIT, TAG_BSIGMA, TAG_BOND, TAG_GAUSS = range(9)
NOISE_NONE, NOISE_UNIFORM, NOISE_GAUSSIAN = 0, 1, 2
LAT_FFT, LAT_LIEB, LAT_DENSE = 0, 1, 2
INIT_SETTINGS, INIT_FIXEDPOINT, INIT_PREVIOUS = 0, 1, 2


class OrcMatrix(C.Structure):
    _fields_ = [
        ("base", C.POINTER(C.c_double)),
        ("value", C.c_double * 2),
        ("noise_type", C.c_int32),
        ("amp", C.POINTER(C.c_double)),
        ("amp_value", C.c_double * 2),
    ]


class OrcSystem(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("lattice", C.c_int32), ("pbc", C.c_int32),
        ("dt", C.c_double), ("gamma", C.c_double), ("J", C.c_double),
        ("J_AB", C.c_double), ("J_BC", C.c_double), ("detuning", C.c_double),
        ("omega", OrcMatrix), ("U", OrcMatrix), ("F", OrcMatrix),
        ("beta_init", OrcMatrix), ("beta_sigma", OrcMatrix),
        ("coupling", C.POINTER(C.c_double)),
        ("frame_steps", C.c_uint64),
        ("seed", C.c_uint64),
        ("F_n_times", C.c_uint64), ("F_times", C.POINTER(C.c_double)), ("F_table", C.POINTER(C.c_double)),
    ]


class OrcSchedule(C.Structure):
    _fields_ = [
        ("t_start", C.c_double), ("t_end", C.c_double), ("timestep", C.c_double),
        ("frames_freq", C.c_double), ("frames_period", C.c_double),
        ("n_frames_in", C.c_uint64),
        ("frame_steps", C.c_uint64), ("dt_obs", C.c_double), ("n_frames_tot", C.c_uint64),
        ("t_end_out", C.c_double), ("n_dt", C.c_uint64), ("error", C.c_int32),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        u32p = C.POINTER(C.c_uint32)
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_counter.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, u32p, u32p]
        L.orc_normal.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, dp]
        L.orc_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, dp]
        L.orc_box_muller.argtypes = [C.c_uint32, C.c_uint32, dp]
        L.orc_fft.argtypes = [C.c_int, C.c_int, dp, dp, C.c_int]
        L.orc_fft.restype = C.c_int
        L.orc_dft_direct.argtypes = [C.c_int, C.c_int, dp, dp, C.c_int]
        L.orc_schedule_compute.argtypes = [C.POINTER(OrcSchedule)]
        L.orc_compute_n_frames.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_compute_n_frames.restype = C.c_uint64
        L.orc_stored_times.argtypes = [C.POINTER(OrcSchedule), dp]
        L.orc_run_trajectory.argtypes = [C.POINTER(OrcSystem), C.c_uint64, C.c_int32, dp, C.c_int32,
                                         C.c_uint64, C.c_double, C.c_double, dp, C.c_uint64, dp, dp]
        L.orc_run_trajectory.restype = C.c_int64
        L.orc_noisy_matrix.argtypes = [C.POINTER(OrcMatrix), C.c_uint64, C.c_uint64, C.c_uint64,
                                       C.c_uint32, dp]
        L.orc_lieb_coupling.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_int32, dp]
        L.orc_time_batch.argtypes = [C.POINTER(OrcSystem), C.c_uint64, C.c_int32, C.c_uint64, C.c_int32]
        L.orc_time_batch.restype = C.c_double
        L.orc_gaussian_state.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_uint64, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    k = (C.c_uint32 * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def normal(seed, traj, tag, step, site):
    o = (C.c_double * 2)()
    lib().orc_normal(seed, traj, tag, step, site, o)
    return complex(o[0], o[1])


def uniform(seed, traj, tag, step, site):
    o = (C.c_double * 2)()
    lib().orc_uniform(seed, traj, tag, step, site, o)
    return (o[0], o[1])


def box_muller(a, b):
    o = (C.c_double * 2)()
    lib().orc_box_muller(a, b, o)
    return complex(o[0], o[1])


def fft(x, sign=-1):
    """Unnormalised DFT of a 2D complex array (rows x cols); 1 x N is 1D."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    nx, ny = x.shape
    out = np.empty_like(x)
    rc = lib().orc_fft(nx, ny, _dp(x.view(np.float64)), _dp(out.view(np.float64)), sign)
    assert rc == 0
    return out


def gaussian_state(nx, ny, seed, traj, mean, nk):
    """Warm-start Gaussian sample (nx, ny) complex for one trajectory id."""
    m = np.array([complex(mean).real, complex(mean).imag])
    nk = np.ascontiguousarray(np.asarray(nk, dtype=np.float64).reshape(-1))
    assert nk.size == nx * ny
    out = np.empty((nx, ny), dtype=np.complex128)
    lib().orc_gaussian_state(nx, ny, seed, traj, _dp(m), _dp(nk), _dp(out.view(np.float64)))
    return out


def dft_direct(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    nx, ny = x.shape
    out = np.empty_like(x)
    lib().orc_dft_direct(nx, ny, _dp(x.view(np.float64)), _dp(out.view(np.float64)), sign)
    return out


def schedule(t_start, t_end, timestep, frames_freq=0.0, frames_period=0.0, n_frames=0):
    s = OrcSchedule(t_start=t_start, t_end=t_end, timestep=timestep, frames_freq=frames_freq,
                    frames_period=frames_period, n_frames_in=n_frames)
    lib().orc_schedule_compute(C.byref(s))
    return {
        "frame_steps": int(s.frame_steps), "dt_obs": s.dt_obs, "n_frames_tot": int(s.n_frames_tot),
        "t_end": s.t_end_out, "n_dt": int(s.n_dt), "error": int(s.error), "_raw": s,
    }


def stored_times(sched):
    s = sched["_raw"]
    out = np.empty(int(s.n_frames_tot) + 1)
    lib().orc_stored_times(C.byref(s), _dp(out))
    return out


def compute_n_frames(dt_obs, t0, t1):
    return int(lib().orc_compute_n_frames(dt_obs, t0, t1))


class _Keep:
    """Holds numpy buffers alive while a ctypes struct points at them."""

    def __init__(self):
        self.bufs = []

    def arr(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128).ravel()
        self.bufs.append(a)
        return _dp(a.view(np.float64))


def _matrix(keep, spec):
    """spec: scalar | complex | ndarray | dict(base=..., noise=('gaussian'|'uniform', amp))."""
    m = OrcMatrix()
    if isinstance(spec, dict):
        base = spec.get("base", 0.0)
        noise = spec.get("noise")
    else:
        base, noise = spec, None
    if np.ndim(base) == 0:
        b = complex(base)
        m.base = C.POINTER(C.c_double)()
        m.value[0], m.value[1] = b.real, b.imag
    else:
        m.base = keep.arr(base)
    m.noise_type = NOISE_NONE
    m.amp = C.POINTER(C.c_double)()
    if noise is not None:
        kind, amp = noise
        m.noise_type = NOISE_GAUSSIAN if kind.lower().startswith("g") else NOISE_UNIFORM
        if np.ndim(amp) == 0:
            a = complex(amp)
            m.amp_value[0], m.amp_value[1] = a.real, a.imag
        else:
            m.amp = keep.arr(amp)
    return m


def make_system(p, keep):
    """p: dict with nx, ny, lattice, dt, gamma, J, omega, U, F, beta_init, beta_sigma,
    frame_steps, seed and optionally pbc, J_AB, J_BC, detuning, E, coupling,
    F_t = (times, table n_times x N) for a time-dependent drive."""
    s = OrcSystem()
    s.nx, s.ny = int(p["nx"]), int(p["ny"])
    s.lattice = int(p.get("lattice", LAT_FFT))
    s.pbc = int(p.get("pbc", 1))
    s.dt, s.gamma, s.J = float(p["dt"]), float(p["gamma"]), float(p.get("J", 0.0))
    s.J_AB, s.J_BC = float(p.get("J_AB", 0.0)), float(p.get("J_BC", 0.0))
    s.detuning = float(p.get("detuning", 0.0))
    onsite = p.get("E", 0.0) if s.lattice != LAT_FFT else p.get("omega", 0.0)
    s.omega = _matrix(keep, onsite)
    s.U = _matrix(keep, p.get("U", 0.0))
    s.F = _matrix(keep, p.get("F", 0.0))
    s.beta_init = _matrix(keep, p.get("beta_init", 0.0))
    s.beta_sigma = _matrix(keep, p.get("beta_sigma", 0.5))
    if s.lattice == LAT_DENSE:
        c = np.ascontiguousarray(p["coupling"], dtype=np.float64)
        keep.bufs.append(c)
        s.coupling = _dp(c)
    else:
        s.coupling = C.POINTER(C.c_double)()
    s.frame_steps = int(p["frame_steps"])
    s.seed = int(p["seed"])
    if p.get("F_t") is not None:
        times, table = p["F_t"]
        times = np.ascontiguousarray(times, dtype=np.float64)
        table = np.ascontiguousarray(np.asarray(table, dtype=np.complex128).reshape(times.size, -1))
        assert table.shape[1] == s.nx * s.ny * (3 if s.lattice == LAT_LIEB else 1)
        keep.bufs += [times, table]
        s.F_n_times, s.F_times, s.F_table = times.size, _dp(times), _dp(table.view(np.float64))
    else:
        s.F_n_times, s.F_times, s.F_table = 0, C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
    return s


def run_trajectory(p, traj, t_start=None, t_end=None, init_mode=INIT_SETTINGS, beta_prev=None,
                   store_initial=True, step0=0):
    """Returns (frames [n_frames, N] complex128, beta_final [N], t_final)."""
    keep = _Keep()
    s = make_system(p, keep)
    n = s.nx * s.ny * (3 if s.lattice == LAT_LIEB else 1)
    t_start = p.get("t_start", 0.0) if t_start is None else t_start
    t_end = p["t_end"] if t_end is None else t_end
    dt_obs = s.dt * s.frame_steps
    nf = compute_n_frames(dt_obs, t_start, t_end) + (1 if store_initial else 0)
    out = np.zeros((max(nf, 1), n), dtype=np.complex128)
    bf = np.zeros(n, dtype=np.complex128)
    tf = C.c_double(0.0)
    bp = keep.arr(beta_prev) if beta_prev is not None else C.POINTER(C.c_double)()
    got = lib().orc_run_trajectory(C.byref(s), traj, init_mode, bp, int(store_initial), step0,
                                   t_start, t_end, _dp(out.view(np.float64)), nf,
                                   _dp(bf.view(np.float64)), C.byref(tf))
    if got < 0:
        raise RuntimeError(f"orc_run_trajectory failed: {got}")
    return out[:got], bf, tf.value


def noisy_matrix(spec, n, seed, traj, tag):
    keep = _Keep()
    m = _matrix(keep, spec)
    out = np.zeros(n, dtype=np.complex128)
    lib().orc_noisy_matrix(C.byref(m), n, seed, traj, tag, _dp(out.view(np.float64)))
    return out


def lieb_coupling(ny, J_AB, J_BC, pbc):
    n = 3 * ny
    out = np.zeros((n, n))
    lib().orc_lieb_coupling(ny, J_AB, J_BC, int(pbc), _dp(out))
    return out


def time_batch(p, n_traj, n_threads, n_steps, faithful=True):
    keep = _Keep()
    s = make_system(p, keep)
    return float(lib().orc_time_batch(C.byref(s), n_traj, n_threads, n_steps, int(faithful)))
