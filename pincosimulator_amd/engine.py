"""Python front end of the HIP TWMC engine (libtwmc_hip.so).

``System`` mirrors the reference's TWMCSystemData (src/TWMC/TWMCSystemData.cpp:14-115):
lattice shape with the 1D normalisation (an N x 1 lattice becomes 1 x N,
TWMCSystemData.cpp:25-37), site-resolved parameters (NoisyMatrix), and the
frame schedule.  ``Engine`` is one solver context on one GPU: ``run`` plays the
role of Solver::Compute for a batch of trajectories (src/Base/Solver.hpp:23).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import TWMCError, check

_KERNELS = {"auto": _lib.KERNEL_AUTO, "generic": _lib.KERNEL_GENERIC, "fused64": _lib.KERNEL_FUSED64,
            "mfma": _lib.KERNEL_MFMA, "stencil64": _lib.KERNEL_STENCIL64, "band": _lib.KERNEL_BAND,
            "stack": _lib.KERNEL_STACK}
_INITS = {"settings": _lib.INIT_SETTINGS, "fixedpoint": _lib.INIT_FIXEDPOINT,
          "previous": _lib.INIT_PREVIOUS, "gaussian": _lib.INIT_GAUSSIAN}


def frame_schedule(t_start, t_end, timestep, frames_freq=0.0, frames_period=0.0, n_frames=0):
    """TWMCSystemData's frame schedule (src/TWMC/TWMCSystemData.cpp:67-115).

    Returns dict(frame_steps, dt_obs, n_frames_tot, t_end, n_dt)."""
    freq = frames_freq
    if frames_freq == 0 and frames_period == 0:
        if n_frames == 0:
            raise ValueError("Must have a non-zero, integer number of frames")
        freq = float(n_frames) / (t_end - t_start)
    elif frames_freq == 0 and frames_period != 0:
        freq = 1.0 / frames_period
    frame_steps = max(1, int(math.ceil(1.0 / (freq * timestep))))
    dt_obs = timestep * frame_steps
    n_frames_tot = int(math.ceil((t_end - t_start) / dt_obs))
    t_end = n_frames_tot * dt_obs
    n_dt, t = 0, t_start
    while t <= t_end:
        n_dt += 1
        t += timestep
    if dt_obs < timestep:
        dt_obs, frame_steps = timestep, 1
    return {"frame_steps": frame_steps, "dt_obs": dt_obs, "n_frames_tot": n_frames_tot,
            "t_end": t_end, "n_dt": n_dt}


def _as_matrix(spec, n, keep):
    """Converts a parameter spec into a twmc_matrix.

    spec: scalar | complex | array of N (or nx x ny) | dict(base=..., noise=(kind, amp))
    with kind 'gaussian' or 'uniform' (NoisyMatrix, src/Base/NoisyMatrix.cpp:64-85)."""
    m = _lib.Matrix()
    base, noise = (spec.get("base", 0.0), spec.get("noise")) if isinstance(spec, dict) else (spec, None)
    null = C.POINTER(C.c_double)()
    if np.ndim(base) == 0:
        b = complex(base)
        m.base = null
        m.value[0], m.value[1] = b.real, b.imag
    else:
        arr = np.ascontiguousarray(np.asarray(base, dtype=np.complex128).reshape(-1))
        if arr.size != n:
            raise ValueError(f"site-resolved parameter has {arr.size} entries, lattice has {n}")
        keep.append(arr)
        m.base = arr.ctypes.data_as(C.POINTER(C.c_double))
    m.noise_type = _lib.NOISE_NONE
    m.amp = null
    if noise is not None:
        kind, amp = noise
        k = kind.lower()
        if k in ("gaussian", "g"):
            m.noise_type = _lib.NOISE_GAUSSIAN
        elif k in ("uniform", "u"):
            m.noise_type = _lib.NOISE_UNIFORM
        else:
            raise ValueError(f"unsupported noise type {kind}")
        if np.ndim(amp) == 0:
            a = complex(amp)
            m.amp_value[0], m.amp_value[1] = a.real, a.imag
        else:
            arr = np.ascontiguousarray(np.asarray(amp, dtype=np.complex128).reshape(-1))
            if arr.size != n:
                raise ValueError(f"site-resolved noise amplitude has {arr.size} entries, lattice has {n}")
            keep.append(arr)
            m.amp = arr.ctypes.data_as(C.POINTER(C.c_double))
    return m


@dataclass
class System:
    """Parameters of one TWMC run (the reference's TWMCSystemData).

    lattice: 'square' (uniform hopping through the FFT, TWMCBaseSolver) or
    'matrix' (hopping h = -i beta.C with a given real coupling C, as
    TWMCLiebSolver does with its sparse Lieb coupling)."""

    nx: int
    ny: int
    dt: float
    gamma: float
    J: float = 0.0
    omega: object = 0.0
    U: object = 0.0
    F: object = 0.0
    beta_init: object = 0.0
    beta_sigma: object = 0.5
    frame_steps: int = 1
    seed: int = 0
    t_start: float = 0.0
    t_end: float = 0.0
    lattice: str = "square"
    coupling: np.ndarray | None = None
    detuning: float = 0.0
    onsite: str = "omega"     # 'omega' (Base) or 'detuning-E' (Lieb: omega holds E)
    F_t: tuple | None = None  # (times, table n_times x N): time-dependent drive, replaces F
    _swapped: bool = field(default=False, repr=False)

    def __post_init__(self):
        # a 1D N x 1 lattice becomes 1 x N (TWMCSystemData.cpp:25-37)
        if self.nx > 1 and self.ny == 1:
            self.nx, self.ny = 1, self.nx
            self._swapped = True

    @property
    def n_sites(self):
        return self.nx * self.ny

    @property
    def dt_obs(self):
        return self.dt * self.frame_steps


@dataclass
class Result:
    n_frames: int
    n_steps: int
    t_final: float
    truncated: bool
    kernel: str
    n_traj: int
    shape: tuple
    sum_n: np.ndarray
    sum_n2: np.ndarray
    sum_nk: np.ndarray
    sum_beta: np.ndarray
    traj_m2: np.ndarray
    traj_m4: np.ndarray
    traj: np.ndarray | None = None
    beta_final: np.ndarray | None = None


class Engine:
    """One solver context (a TWMC system on one GPU)."""

    def __init__(self, system: System, precision: int = 32, kernel: str = "auto", device: int = 0):
        L = _lib.load()
        self.system = system
        self.precision = precision
        keep = []
        p = _lib.Params()
        p.nx, p.ny = system.nx, system.ny
        n = system.n_sites
        p.hopping = _lib.HOP_FFT if system.lattice == "square" else _lib.HOP_MATRIX
        p.onsite = _lib.ONSITE_OMEGA if system.onsite == "omega" else _lib.ONSITE_DETUNING_MINUS_E
        p.precision = precision
        p.kernel = _KERNELS[kernel]
        p.dt, p.gamma, p.J, p.detuning = system.dt, system.gamma, system.J, system.detuning
        p.omega = _as_matrix(system.omega, n, keep)
        p.U = _as_matrix(system.U, n, keep)
        p.F = _as_matrix(system.F, n, keep)
        p.beta_init = _as_matrix(system.beta_init, n, keep)
        p.beta_sigma = _as_matrix(system.beta_sigma, n, keep)
        if p.hopping == _lib.HOP_MATRIX:
            c = np.ascontiguousarray(np.asarray(system.coupling, dtype=np.float64))
            if c.shape != (n, n):
                raise ValueError("coupling must be N x N")
            keep.append(c)
            p.coupling = c.ctypes.data_as(C.POINTER(C.c_double))
        else:
            p.coupling = C.POINTER(C.c_double)()
        p.seed = int(system.seed) & 0xFFFFFFFFFFFFFFFF
        p.frame_steps = int(system.frame_steps)
        if system.F_t is not None:
            times = np.ascontiguousarray(np.asarray(system.F_t[0], dtype=np.float64).reshape(-1))
            table = np.ascontiguousarray(np.asarray(system.F_t[1], dtype=np.complex128).reshape(times.size, n))
            keep += [times, table]
            p.F_n_times = times.size
            p.F_times = times.ctypes.data_as(C.POINTER(C.c_double))
            p.F_table = table.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))
        ctx = C.c_void_p()
        check(L.twmc_create(C.byref(p), device, C.byref(ctx)))
        self._ctx = ctx
        self._L = L

    def close(self):
        if getattr(self, "_ctx", None):
            self._L.twmc_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- staged interface (device-resident; used by the benchmark) -------------
    def begin(self, traj_ids, t_start=None, t_end=None, init="settings", beta_prev=None,
              store_initial=True, step0=0, traj=False, final=False, gauss_mean=None, gauss_nk=None):
        """init="gaussian" samples DFT(beta)_k = N gauss_mean [k=0] + sqrt(N (n_k + 1/2) / 2) xi
        with gauss_nk = n_k over the (nx, ny) k-grid in DFT order."""
        s = self.system
        ids = np.ascontiguousarray(np.asarray(traj_ids, dtype=np.uint64).reshape(-1))
        d = _lib.BatchDesc()
        self._ids = ids
        d.traj_ids = ids.ctypes.data_as(C.POINTER(C.c_uint64))
        d.n_traj = ids.size
        d.init_mode = _INITS[init]
        d.store_initial = int(bool(store_initial))
        if beta_prev is not None:
            bp = np.ascontiguousarray(np.asarray(beta_prev, dtype=np.complex128).reshape(ids.size, -1))
            if bp.shape[1] != s.n_sites:
                raise ValueError(f"beta_prev has {bp.shape[1]} sites per trajectory, lattice has {s.n_sites}")
            self._bp = bp
            d.beta_prev = bp.ctypes.data_as(C.POINTER(C.c_double))
        else:
            d.beta_prev = C.POINTER(C.c_double)()
        if init == "gaussian":
            self._gm = np.array([complex(gauss_mean).real, complex(gauss_mean).imag])
            self._gk = np.ascontiguousarray(np.asarray(gauss_nk, dtype=np.float64).reshape(-1))
            if self._gk.size != self.system.nx * self.system.ny:
                raise ValueError("gauss_nk needs one population per lattice mode")
            d.gauss_mean = self._gm.ctypes.data_as(C.POINTER(C.c_double))
            d.gauss_nk = self._gk.ctypes.data_as(C.POINTER(C.c_double))
        d.step0 = int(step0)
        d.t_start = s.t_start if t_start is None else t_start
        d.t_end = s.t_end if t_end is None else t_end
        d.outputs = _lib.OUT_OBS | (_lib.OUT_TRAJ if traj else 0) | (_lib.OUT_FINAL if final else 0)
        check(self._L.twmc_batch_begin(self._ctx, C.byref(d)), self._ctx)
        self._want_traj, self._want_final = traj, final
        return self.info()

    def info(self):
        i = _lib.BatchInfo()
        check(self._L.twmc_batch_info_get(self._ctx, C.byref(i)), self._ctx)
        return {"n_frames": i.n_frames, "n_steps": i.n_steps, "n_segments": i.n_segments,
                "segments_done": i.segments_done, "t_final": i.t_final,
                "kernel": _lib.KERNEL_NAMES[i.kernel], "truncated": bool(i.truncated)}

    def advance(self, n_segments=1):
        check(self._L.twmc_batch_advance(self._ctx, int(n_segments)), self._ctx)

    def sync(self):
        check(self._L.twmc_batch_sync(self._ctx), self._ctx)

    def set_timing(self, on=True):
        check(self._L.twmc_set_timing(self._ctx, int(on)), self._ctx)

    def stats(self):
        st = _lib.Stats()
        check(self._L.twmc_query(self._ctx, C.byref(st)), self._ctx)
        return {f: getattr(st, f) for f, _ in st._fields_}

    def reset_stats(self):
        check(self._L.twmc_reset_stats(self._ctx), self._ctx)

    def progress(self):
        return float(self._L.twmc_progress(self._ctx))

    def read(self):
        info = self.info()
        s = self.system
        n, nf, B = s.n_sites, int(info["n_frames"]), int(self._ids.size)
        arrs = {
            "sum_n": np.zeros((nf, n)), "sum_n2": np.zeros((nf, n)), "sum_nk": np.zeros((nf, n)),
            "sum_beta": np.zeros((nf, n), dtype=np.complex128), "traj_m2": np.zeros((B, nf)),
            "traj_m4": np.zeros((B, nf)),
        }
        o = _lib.Output()
        for k, a in arrs.items():
            setattr(o, k, a.ctypes.data_as(C.POINTER(C.c_double)))
        traj = final = None
        if self._want_traj:
            traj = np.zeros((B, nf, n), dtype=np.complex128)
            o.traj = traj.ctypes.data_as(C.POINTER(C.c_double))
        if self._want_final:
            final = np.zeros((B, n), dtype=np.complex128)
            o.beta_final = final.ctypes.data_as(C.POINTER(C.c_double))
        check(self._L.twmc_batch_read(self._ctx, C.byref(o)), self._ctx)
        return Result(n_frames=nf, n_steps=int(info["n_steps"]), t_final=info["t_final"],
                      truncated=info["truncated"], kernel=info["kernel"], n_traj=B,
                      shape=(s.nx, s.ny), traj=traj, beta_final=final, **arrs)

    # --- one-call interface (Solver::Compute) ------------------------------------
    def run(self, traj_ids, **kw):
        info = self.begin(traj_ids, **kw)
        self.advance(info["n_segments"])
        return self.read()


def noisy_matrix(spec, n, seed, traj, tag):
    """Host realisation of a NoisyMatrix with the kernels' Philox layout."""
    L = _lib.load()
    keep = []
    m = _as_matrix(spec, n, keep)
    out = np.zeros(n, dtype=np.complex128)
    check(L.twmc_noisy_matrix(C.byref(m), n, seed, traj, tag, out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


class Sweep:
    """Many uneven systems in one persistent launch per frame (libtwmc_hip.so
    twmc_sweep_*).  points: list of dicts with nx, ny, n_traj, traj_id0, J,
    gamma, omega, U, F, beta_init, beta_sigma (uniform, complex allowed)."""

    def __init__(self, points, dt, frame_steps, seed, device=0):
        L = _lib.load()
        arr = (_lib.SweepPoint * len(points))()
        for k, p in enumerate(points):
            q = arr[k]
            q.nx, q.ny = int(p["nx"]), int(p["ny"])
            q.n_traj, q.traj_id0 = int(p["n_traj"]), int(p.get("traj_id0", 0))
            q.J, q.gamma = float(p.get("J", 0.0)), float(p["gamma"])
            for name, default in (("omega", 0.0), ("U", 0.0), ("F", 0.0), ("beta_init", 0.0), ("beta_sigma", 0.5)):
                v = complex(p.get(name, default))
                getattr(q, name)[0], getattr(q, name)[1] = v.real, v.imag
        self.points = points
        self.n_traj = [int(p["n_traj"]) for p in points]
        ctx = C.c_void_p()
        check(L.twmc_sweep_create(arr, len(points), dt, frame_steps, seed, device, C.byref(ctx)), None, sweep=True)
        self._ctx, self._L = ctx, L
        self.dt, self.frame_steps = dt, frame_steps

    def close(self):
        if getattr(self, "_ctx", None):
            self._L.twmc_sweep_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def batching(self):
        tb = (C.c_int32 * len(self.points))()
        check(self._L.twmc_sweep_get_batching(self._ctx, tb), self._ctx, sweep=True)
        return list(tb)

    def set_batching(self, tb):
        arr = (C.c_int32 * len(self.points))(*[int(v) for v in tb])
        check(self._L.twmc_sweep_set_batching(self._ctx, arr), self._ctx, sweep=True)

    def autotune(self, steps=20):
        tb = (C.c_int32 * len(self.points))()
        check(self._L.twmc_sweep_autotune(self._ctx, steps, tb), self._ctx, sweep=True)
        return list(tb)

    def run(self, t_end, t_start=0.0, graph=True):
        """Returns dict(n_frames, kernel_ms, m2, m4, ksum): m2/m4 are per point
        arrays (n_traj, n_frames); ksum is (n_points, n_frames, 34) with the
        32 radial-bin sums of |DFT beta|^2 and the sum of beta (Re, Im)."""
        dt_obs = self.dt * self.frame_steps
        nf_max = int(round((t_end - t_start) / dt_obs)) + 2
        tot = sum(self.n_traj)
        m2 = np.zeros(tot * nf_max)
        m4 = np.zeros(tot * nf_max)
        ks = np.zeros(len(self.points) * nf_max * 34)
        nf = C.c_uint64(0)
        ms = C.c_double(0.0)
        dp = C.POINTER(C.c_double)
        check(self._L.twmc_sweep_run(self._ctx, t_start, t_end, int(bool(graph)), m2.ctypes.data_as(dp),
                                     m4.ctypes.data_as(dp), ks.ctypes.data_as(dp), C.byref(nf), C.byref(ms)),
              self._ctx, sweep=True)
        f = int(nf.value)
        m2, m4 = m2[:tot * f].reshape(tot, f), m4[:tot * f].reshape(tot, f)
        out2, out4, o = [], [], 0
        for n in self.n_traj:
            out2.append(m2[o:o + n])
            out4.append(m4[o:o + n])
            o += n
        ks = ks[:len(self.points) * f * 34].reshape(len(self.points), f, 34)
        return {"n_frames": f, "kernel_ms": ms.value, "m2": out2, "m4": out4, "ksum": ks}

    def mode_bins(self, point):
        p = self.points[point]
        nx, ny = int(p["nx"]), int(p["ny"])
        if nx > 1 and ny == 1:
            nx, ny = 1, nx
        out = np.zeros(nx * ny, dtype=np.uint8)
        check(self._L.twmc_sweep_mode_bins(self._ctx, point, out.ctypes.data_as(C.POINTER(C.c_uint8))),
              self._ctx, sweep=True)
        return out.reshape(nx, ny)

    def stats(self):
        st = _lib.Stats()
        check(self._L.twmc_sweep_query(self._ctx, C.byref(st)), self._ctx, sweep=True)
        return {f: getattr(st, f) for f, _ in st._fields_}
