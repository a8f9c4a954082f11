"""Pins the oracle's restatement of TWMCBaseSolver / TWMCLiebSolver / TWMCSystemData
against closed forms (CPU only).  None of these retypes the update formula:
each checks a mathematical consequence of it."""
import math

import numpy as np
import pytest

import oracle
from helpers import c1_spec
from pincosimulator_amd.engine import frame_schedule


def base(**kw):
    p = dict(nx=1, ny=16, lattice=oracle.LAT_FFT, dt=0.01, gamma=0.0, J=1.0, omega=0.5, U=0.0, F=0.0,
             beta_init=0.0, beta_sigma=0.0, frame_steps=10, seed=1, t_start=0.0, t_end=1.0)
    p.update(kw)
    return p


@pytest.mark.parametrize("nx,ny,kx,ky", [(1, 16, 0, 3), (1, 16, 0, 0), (8, 8, 2, 5), (12, 10, 7, 1)])
def test_plane_wave_is_eigenmode_with_reference_sign(nx, ny, kx, ky):
    """gamma = 0 (no noise), U = 0, F = 0: a plane wave e^{i k.r} is multiplied by
    exactly (1 + dt a_k) per Euler step with a_k = -i (omega + 2J ([nx!=1]cos kx + [ny!=1]cos ky))."""
    x, y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    pw = np.exp(2j * np.pi * (kx * x / nx + ky * y / ny)).reshape(-1)
    p = base(nx=nx, ny=ny, beta_init=pw, frame_steps=5, t_end=0.25, dt=0.01)
    frames, _, _ = oracle.run_trajectory(p, 0, init_mode=oracle.INIT_FIXEDPOINT)
    eps = 2 * p["J"] * ((nx != 1) * math.cos(2 * math.pi * kx / nx) + (ny != 1) * math.cos(2 * math.pi * ky / ny))
    g = 1 + p["dt"] * (-1j * (p["omega"] + eps))
    for f in range(frames.shape[0]):
        expect = g ** (5 * f) * pw
        assert np.max(np.abs(frames[f] - expect)) < 1e-12


def test_euler_fixed_point_is_exact():
    """gamma = 0, U = 0: beta* = F / (omega + eps_0) is a fixed point of the Euler map."""
    J, om, F = 1.0, 0.5, 1.7
    bstar = F / (om + 2 * J)  # 1D, k = 0
    p = base(J=J, omega=om, F=F, beta_init=bstar, t_end=2.0)
    frames, _, _ = oracle.run_trajectory(p, 0, init_mode=oracle.INIT_FIXEDPOINT)
    assert np.max(np.abs(frames - bstar)) < 1e-13


def test_single_site_nonlinear_recursion():
    """1x1 lattice, gamma = 0: the scalar recursion in long double (brute force)."""
    U, om, F, dt = 0.3, -0.4, 0.8, 0.01
    b0 = 0.3 + 0.1j
    p = base(nx=1, ny=1, U=U, omega=om, F=F, dt=dt, beta_init=b0, frame_steps=50, t_end=2.0)
    frames, _, _ = oracle.run_trajectory(p, 0, init_mode=oracle.INIT_FIXEDPOINT)
    br, bi = np.longdouble(b0.real), np.longdouble(b0.imag)
    out = [(br, bi)]
    for s in range(1, 201):
        n2 = br * br + bi * bi
        # a = dt [(-i om + i U (n2 - 1)) b + i F]
        w = np.longdouble(-om) + np.longdouble(U) * (n2 - 1)
        ar, ai = -w * bi, w * br + np.longdouble(F)
        br, bi = br + dt * ar, bi + dt * ai
        if s % 50 == 0:
            out.append((br, bi))
    ref = np.array([complex(float(a), float(b)) for a, b in out])
    assert np.max(np.abs(frames[:, 0] - ref)) < 1e-12


def test_ou_stationary_variance_per_mode():
    """U = 0, F = 0, vacuum start: the Euler OU process has, per k mode, the exact
    stationary variance E|b_k|^2 / N = (gamma dt / 2) / (1 - |1 + a_k dt|^2) with
    a_k = -gamma/2 - i (omega + eps_k).  Pins sqrt(gamma dt / 4) and the O(dt) bias."""
    nx, ny, dt, gam, om, J = 1, 8, 0.05, 1.0, 0.5, 1.0
    p = base(nx=nx, ny=ny, dt=dt, gamma=gam, omega=om, J=J, beta_sigma=0.5, frame_steps=20, t_end=40.0, seed=3)
    B = 600
    acc = np.zeros(ny)
    cnt = 0
    for t in range(B):
        frames, _, _ = oracle.run_trajectory(p, t)
        fk = np.fft.fft(frames[10:], axis=1)  # t >= 10: ~ 5 relaxation times
        acc += (np.abs(fk) ** 2).sum(axis=0) / ny
        cnt += fk.shape[0]
    meas = acc / cnt
    k = np.arange(ny)
    eps = 2 * J * np.cos(2 * np.pi * k / ny)
    g2 = (1 - gam * dt / 2) ** 2 + ((om + eps) * dt) ** 2
    pred = (gam * dt / 2) / (1 - g2)
    # frames within a trajectory are correlated; 8% covers > 4 sigma of the effective sample
    assert np.all(np.abs(meas / pred - 1) < 0.08), (meas, pred)
    # the half-variance error a wrong sqrt(gamma dt / 4) factor would make is far outside
    assert abs(np.mean(meas / pred) - 1) < 0.03


def test_mean_field_follows_linear_recursion():
    """U = 0: the ensemble mean obeys the noiseless Euler recursion of the k = 0 mode."""
    dt, gam, om, J, F = 0.02, 1.0, 1.0, 1.0, 2.0
    p = base(nx=4, ny=4, dt=dt, gamma=gam, omega=om, J=J, F=F, beta_sigma=0.5, frame_steps=25, t_end=3.0, seed=11)
    B = 400
    mean = 0
    for t in range(B):
        frames, _, _ = oracle.run_trajectory(p, t)
        mean = mean + frames.mean(axis=1)
    mean /= B
    a = -gam / 2 - 1j * (om + 4 * J)
    m, ref = 0j, [0j]
    for s in range(1, 151):
        m = m + dt * (a * m + 1j * F)
        if s % 25 == 0:
            ref.append(m)
    ref = np.array(ref)
    # per-frame SE of the site-averaged mean ~ sqrt(0.5 / (B N)) < 0.01
    assert np.max(np.abs(mean - ref)) < 0.05


@pytest.mark.parametrize("cfg,args,expect", [
    ("c1", (0.0, 100.0, 0.005, 1.0), (200, 100, 100.0)),
    ("c2", (0.0, 100.0, 0.002, 1.0), (500, 100, 100.0)),
    ("c3", (0.0, 100.0, 0.005, 1.0), (200, 100, 100.0)),
    ("c4", (0.0, 60.0, 0.002, 1.0), (500, 60, 60.0)),
    ("sample-sim.ini", (0.0, 20.0, 0.01, 1.0), (100, 20, 20.0)),
])
def test_schedule_configs(cfg, args, expect):
    t0, t1, dt, freq = args
    s = oracle.schedule(t0, t1, dt, frames_freq=freq)
    assert (s["frame_steps"], s["n_frames_tot"], s["t_end"]) == expect
    n_frames = oracle.compute_n_frames(s["dt_obs"], t0, s["t_end"]) + 1
    assert n_frames == expect[1] + 1
    assert frame_schedule(t0, t1, dt, frames_freq=freq) == {k: s[k] for k in
                                                           ("frame_steps", "dt_obs", "n_frames_tot", "t_end", "n_dt")}


def test_schedule_quirks():
    # t_end := nFramesTot * dt_obs, without t_start (TWMCSystemData.cpp:104-105)
    s = oracle.schedule(10.0, 30.0, 0.01, frames_freq=1.0)
    assert s["t_end"] == 20.0 and s["n_frames_tot"] == 20
    # n_frames given instead of a frequency; period form
    s = oracle.schedule(0.0, 700.0, 0.0005, n_frames=150)
    assert s["frame_steps"] == math.ceil(1 / (150 / 700 * 0.0005))
    s2 = oracle.schedule(0.0, 10.0, 0.01, frames_period=0.5)
    assert s2["frame_steps"] == 50
    times = oracle.stored_times(s2)
    assert len(times) == 21 and times[0] == 0.0 and abs(times[-1] - 10.0) < 1e-9
    # zero frames and no frequency is an error
    assert oracle.schedule(0.0, 1.0, 0.1)["error"] == 1


def test_schedule_python_mirror_random():
    rng = np.random.default_rng(0)
    for _ in range(200):
        t0 = float(rng.choice([0.0, 1.5, 3.0]))
        t1 = t0 + float(rng.uniform(0.5, 20))
        dt = float(rng.choice([0.001, 0.002, 0.005, 0.01, 0.0123]))
        freq = float(rng.uniform(0.05, 5.0))
        s = oracle.schedule(t0, t1, dt, frames_freq=freq)
        m = frame_schedule(t0, t1, dt, frames_freq=freq)
        assert m == {k: s[k] for k in ("frame_steps", "dt_obs", "n_frames_tot", "t_end", "n_dt")}


@pytest.mark.parametrize("ny,pbc", [(1, False), (1, True), (2, False), (2, True), (3, False), (3, True), (128, False)])
def test_lieb_coupling_enumeration(ny, pbc):
    C = oracle.lieb_coupling(ny, 1.0, 2.0, pbc)
    assert np.array_equal(C, C.T)
    nnz = np.count_nonzero(C)
    if ny == 1 and pbc:
        # the closing bond C(2)-B(1) coincides with the internal B-C bond
        assert nnz == 4 and C[1, 2] == 4.0
    else:
        assert nnz == 4 * ny + 2 * (ny - 1) + 2 * int(pbc)
    for i in range(ny):
        assert C[3 * i, 3 * i + 1] == 1.0
        assert C[3 * i + 1, 3 * i + 2] >= 2.0
        if i < ny - 1:
            assert C[3 * i + 2, 3 * (i + 1) + 1] == 2.0
        assert C[3 * i, 3 * i] == 0.0


def test_noisy_matrix_statistics():
    n = 20000
    g = oracle.noisy_matrix({"base": 1.0 + 2.0j, "noise": ("gaussian", 0.5 + 0.25j)}, n, 7, 0, oracle.TAG_U)
    assert abs(g.real.mean() - 1.0) < 5 * 0.5 / math.sqrt(n)
    assert abs(g.real.std() - 0.5) < 0.02 and abs(g.imag.std() - 0.25) < 0.01
    u = oracle.noisy_matrix({"base": 0.0, "noise": ("uniform", 2.0 + 0j)}, n, 7, 0, oracle.TAG_F)
    assert u.real.min() > 0 and u.real.max() < 2.0 and np.all(u.imag == 0)
    plain = oracle.noisy_matrix(3.0, 10, 7, 0, oracle.TAG_F)
    assert np.all(plain == 3.0)
    # independent per trajectory, reproducible per (seed, traj, tag)
    a = oracle.noisy_matrix({"base": 0.0, "noise": ("gaussian", 1.0)}, 64, 7, 1, oracle.TAG_U)
    b = oracle.noisy_matrix({"base": 0.0, "noise": ("gaussian", 1.0)}, 64, 7, 2, oracle.TAG_U)
    a2 = oracle.noisy_matrix({"base": 0.0, "noise": ("gaussian", 1.0)}, 64, 7, 1, oracle.TAG_U)
    assert np.array_equal(a, a2) and not np.array_equal(a, b)


def test_truncated_loop_reports_fewer_frames():
    """The time loop can end before the last frame (TWMCBaseSolver.cpp:268-274)."""
    p = base(dt=0.0123, frame_steps=37, t_end=1.2345, gamma=1.0, beta_sigma=0.5)
    frames, _, tf = oracle.run_trajectory(p, 0)
    assert frames.shape[0] == 3  # 4 expected, the loop stops after 101 steps
    assert tf > 1.2345


def _ft_spec(**kw):
    s = dict(nx=1, ny=1, lattice=oracle.LAT_FFT, dt=0.01, gamma=0.0, J=0.0, omega=0.7, U=0.0, F=0.0,
             beta_init=0.3 + 0.1j, beta_sigma=0.0, frame_steps=1, seed=5, t_start=0.0, t_end=3.0)
    s.update(kw)
    return s


def test_time_dependent_drive_single_site_recursion():
    """gamma = 0 (no noise), U = 0, one site: the Euler map with F(t) is
    b_{n+1} = b_n + dt (-i omega b_n + i F(t_n)), F piecewise linear with
    constant ends -- np.interp is an independent implementation of that rule."""
    times = np.array([0.5, 1.0, 2.2])
    vals = np.array([1.0 + 0.5j, -2.0 + 0.0j, 0.25 - 1.0j])
    spec = _ft_spec(F_t=(times, vals.reshape(3, 1)))
    fr, bf, tf = oracle.run_trajectory(spec, 0)
    b = complex(0.3, 0.1)
    t = 0.0
    ref = [b]
    while t <= 3.0 and len(ref) < fr.shape[0]:
        F = np.interp(t, times, vals.real) + 1j * np.interp(t, times, vals.imag)
        b = b + 0.01 * (-1j * 0.7 * b + 1j * F)
        t += 0.01
        ref.append(b)
    np.testing.assert_allclose(fr[:, 0], np.array(ref[:fr.shape[0]]), rtol=0, atol=1e-13)


def test_constant_drive_table_equals_static_drive():
    spec = dict(c1_spec(t_end=2.0, frame_steps=50), lattice=oracle.LAT_FFT)
    F = 1.0  # c1 drive
    a, _, _ = oracle.run_trajectory(spec, 3)
    b, _, _ = oracle.run_trajectory(dict(spec, F_t=(np.array([0.0]), np.full((1, 64), F, complex))), 3)
    np.testing.assert_array_equal(a, b)


def test_lieb_lattice_equals_dense_lieb_coupling():
    cells = 6
    base = dict(nx=1, dt=0.005, gamma=1.0, J=0.0, J_AB=1.5, J_BC=0.7, detuning=0.4, E=0.2, U=0.1, F=1.2,
                beta_init=0.0, beta_sigma=0.5, frame_steps=20, seed=9, t_start=0.0, t_end=0.4, pbc=1)
    a, _, _ = oracle.run_trajectory(dict(base, ny=cells, lattice=oracle.LAT_LIEB), 2)
    C = oracle.lieb_coupling(cells, 1.5, 0.7, True)
    b, _, _ = oracle.run_trajectory(dict(base, ny=3 * cells, lattice=oracle.LAT_DENSE, coupling=C), 2)
    np.testing.assert_array_equal(a, b)


# ---------------------------------------------------------------------------
# The matrix-hopping branch (TWMCLiebSolver.cpp:44-45 delta = -E + detuning,
# :259-263 a = dt[(rsl + iU(|b|^2-1)) b - i (b.C) + iF]) pinned against
# consequences of the mathematics, not against a restatement of the formula.

def _ring(n):
    """Adjacency of a periodic ring of n >= 3 sites."""
    A = np.zeros((n, n))
    for i in range(n):
        A[i, (i + 1) % n] = A[(i + 1) % n, i] = 1.0
    return A


def _torus(nx, ny):
    """Adjacency of an nx x ny periodic square lattice (row-major sites), nx, ny >= 3."""
    n = nx * ny
    A = np.zeros((n, n))
    for x in range(nx):
        for y in range(ny):
            s = x * ny + y
            for t in (((x + 1) % nx) * ny + y, x * ny + (y + 1) % ny):
                A[s, t] += 1.0
                A[t, s] += 1.0
    return A


@pytest.mark.parametrize("nx,ny", [(1, 12), (6, 5), (4, 8)])
@pytest.mark.parametrize("detuning,E", [(0.5, 0.0), (0.9, 0.4), (-0.3, -1.1)])
def test_dense_nearest_neighbour_coupling_equals_fft_hopping(nx, ny, detuning, E):
    """The FFT hopping IDFT(K . DFT b)/N with K = -2iJ([nx!=1]cos kx + [ny!=1]cos ky) is, term by term,
    -iJ times the sum over the periodic nearest neighbours.  So the matrix branch with C = J x the ring
    (torus) adjacency and on-site detuning - E must reproduce the FFT branch with omega = detuning - E, on
    the same Philox increments, over a few hundred noisy, nonlinear, driven steps.  The FFT branch's
    sign and normalisation are pinned by the plane-wave eigenvalue test above; a flipped delta sign
    (E - detuning), a flipped hopping sign (+i b.C) or a dropped term fails here."""
    J = 0.8
    A = _ring(ny) if nx == 1 else _torus(nx, ny)
    common = dict(nx=nx, ny=ny, dt=0.005, gamma=1.0, J=J, U=0.15, F=0.9 - 0.3j, beta_init=0.2 + 0.1j,
                  beta_sigma=0.5, frame_steps=60, seed=4242, t_start=0.0, t_end=1.5)
    fft_p = dict(common, lattice=oracle.LAT_FFT, omega=detuning - E)
    dense_p = dict(common, lattice=oracle.LAT_DENSE, coupling=J * A, detuning=detuning, E=E)
    for traj in (0, 5):
        a, af, _ = oracle.run_trajectory(fft_p, traj)
        b, bf, _ = oracle.run_trajectory(dense_p, traj)
        assert a.shape == b.shape and a.shape[0] == 6
        scale = np.max(np.abs(a))
        assert np.max(np.abs(a - b)) < 1e-13 * max(scale, 1.0) * 10, np.max(np.abs(a - b))
        assert np.max(np.abs(af - bf)) < 1e-12


def test_dense_site_resolved_detuning_equals_fft_site_resolved_omega():
    """Site-resolved E (a base array) enters as delta_i = detuning - E_i: the FFT branch with the
    site-resolved omega_i = detuning - E_i must agree."""
    nx, ny, J = 1, 10, 1.0
    rng = np.random.default_rng(3)
    E = rng.normal(size=nx * ny)
    common = dict(nx=nx, ny=ny, dt=0.004, gamma=1.0, J=J, U=0.1, F=1.0, beta_init=0.0, beta_sigma=0.5,
                  frame_steps=50, seed=77, t_start=0.0, t_end=1.0)
    a, _, _ = oracle.run_trajectory(dict(common, lattice=oracle.LAT_FFT, omega=0.6 - E), 1)
    b, _, _ = oracle.run_trajectory(dict(common, lattice=oracle.LAT_DENSE, coupling=J * _ring(ny),
                                         detuning=0.6, E=E), 1)
    assert np.max(np.abs(a - b)) < 1e-12


@pytest.mark.parametrize("cells,pbc", [(5, False), (6, True)])
def test_lieb_eigenvector_evolves_geometrically(cells, pbc):
    """U = gamma = F = 0 (so no noise): for an eigenvector v of the Lieb coupling, C v = lambda v, the
    Euler map is b -> b + dt(-i delta b - i b.C) = (1 - i dt (delta + lambda)) b, so
    b(n) = (1 - i dt (detuning - E + lambda))^n v exactly.  lambda and v come from numpy.linalg.eigh of
    the coupling the solver builds (itself pinned by the enumeration test)."""
    J_AB, J_BC, detuning, E, dt, fs = 1.3, 0.6, 0.45, 0.2, 0.01, 25
    C = oracle.lieb_coupling(cells, J_AB, J_BC, pbc)
    lam, V = np.linalg.eigh(C)
    for k in (0, len(lam) // 2, len(lam) - 1):
        v = V[:, k].astype(np.complex128) * (0.7 + 0.2j)
        p = dict(nx=1, ny=cells, lattice=oracle.LAT_LIEB, pbc=int(pbc), dt=dt, gamma=0.0, J_AB=J_AB, J_BC=J_BC,
                 detuning=detuning, E=E, U=0.0, F=0.0, beta_init=v, beta_sigma=0.0, frame_steps=fs, seed=1,
                 t_start=0.0, t_end=2.0)
        frames, _, _ = oracle.run_trajectory(p, 0, init_mode=oracle.INIT_FIXEDPOINT)
        g = 1 - 1j * dt * (detuning - E + lam[k])
        for f in range(frames.shape[0]):
            np.testing.assert_allclose(frames[f], g ** (fs * f) * v, rtol=0, atol=1e-12)


def test_dense_left_eigenvector_pins_row_vector_product():
    """The reference multiplies the row vector b by C from the right (b.C, TWMCLiebSolver.cpp:262).
    For an asymmetric banded C the left eigenvector w (w.C = mu w) evolves geometrically under
    U = gamma = F = 0; the right eigenvector (C u = mu u) does not, so C.b in place of b.C fails."""
    n, dt, fs, detuning, E = 9, 0.01, 20, 0.35, -0.15
    C = np.zeros((n, n))
    for i in range(n - 1):
        C[i, i + 1], C[i + 1, i] = 1.4, 0.5       # products > 0: real spectrum
    for i in range(n - 2):
        C[i, i + 2], C[i + 2, i] = 0.3, 0.12
    mu, W = np.linalg.eig(C.T)                   # C^T w = mu w  <=>  w.C = mu w
    order = np.argsort(mu.real)
    mu_r, U_r = np.linalg.eig(C)

    def run(v):
        p = dict(nx=1, ny=n, lattice=oracle.LAT_DENSE, coupling=C, dt=dt, gamma=0.0, J=0.0, detuning=detuning,
                 E=E, U=0.0, F=0.0, beta_init=v, beta_sigma=0.0, frame_steps=fs, seed=2, t_start=0.0, t_end=1.0)
        return oracle.run_trajectory(p, 0, init_mode=oracle.INIT_FIXEDPOINT)[0]

    for k in (order[0], order[-1]):
        w = W[:, k].astype(np.complex128)
        frames = run(w)
        g = 1 - 1j * dt * (detuning - E + mu[k])
        for f in range(frames.shape[0]):
            np.testing.assert_allclose(frames[f], g ** (fs * f) * w, rtol=0, atol=1e-12)
    # discrimination: a right eigenvector is not transported geometrically by b.C
    k = int(np.argmax(mu_r.real))
    u = U_r[:, k].astype(np.complex128)
    frames = run(u)
    g = 1 - 1j * dt * (detuning - E + mu_r[k])
    assert np.max(np.abs(frames[-1] - g ** (fs * (frames.shape[0] - 1)) * u)) > 1e-3
