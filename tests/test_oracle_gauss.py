"""The warm-start Gaussian sampler of the oracle, pinned by what the
mathematics fixes: an independent FFT (numpy) recovers the k-space sample, and
an ensemble of samples has the requested mean field and per-mode populations
(to 4 standard errors), no anomalous pairing, and the real-space variance
mean(n_k) + 1/2."""
import numpy as np

import oracle

SEED = 4242


def profile(nx, ny):
    kx = 2 * np.pi * np.fft.fftfreq(nx)[:, None]
    ky = 2 * np.pi * np.fft.fftfreq(ny)[None, :]
    nk = 0.4 * np.exp(-(kx ** 2 + ky ** 2)) + 0.02
    nk[0, 0] = 0.15
    return nk


def test_kspace_sample_recovered_by_independent_fft():
    nx, ny, mean = 6, 10, 0.8 - 0.35j
    nk = profile(nx, ny)
    b = oracle.gaussian_state(nx, ny, SEED, 77, mean, nk)
    k = np.fft.fft2(b)  # forward, unnormalised
    n = nx * ny
    amp = np.sqrt(n * (nk + 0.5) / 2)
    z = (k - n * mean * (np.arange(n).reshape(nx, ny) == 0)) / amp
    for idx in range(n):
        assert abs(z.reshape(-1)[idx] - oracle.normal(SEED, 77, oracle.TAG_GAUSS, 0, idx)) < 1e-12


def test_ensemble_moments():
    nx, ny, mean = 8, 8, 0.7 - 0.3j
    n = nx * ny
    nk = profile(nx, ny)
    S = 3000
    bs = np.array([oracle.gaussian_state(nx, ny, SEED, t, mean, nk) for t in range(S)])
    # mean field: sample mean of beta over sites and samples
    mf = bs.mean(axis=(1, 2))
    assert abs(mf.mean() - mean) < 4 * mf.std() / np.sqrt(S)
    k = np.fft.fft2(bs, axes=(1, 2))
    d = k.copy()
    d[:, 0, 0] -= n * mean
    pop = np.abs(d) ** 2 / n  # per sample: exponential with mean n_k + 1/2
    est = pop.mean(axis=0) - 0.5
    se = pop.std(axis=0, ddof=1) / np.sqrt(S)
    assert np.all(np.abs(est - nk) < 4.5 * se), np.max(np.abs(est - nk) / se)
    # the populations observable with the k = 0 coherent part removed
    n0 = (np.abs(k[:, 0, 0]) ** 2 / n).mean() - 0.5 - n * abs(mean) ** 2
    assert abs(n0 - nk[0, 0]) < 4 * np.sqrt(n) * abs(mean) * np.sqrt(2 * (nk[0, 0] + 0.5)) / np.sqrt(S) + 0.05
    # no anomalous pairing <d_k d_-k>
    minus = d[:, (-np.arange(nx)) % nx][:, :, (-np.arange(ny)) % ny]
    an = (d * minus).mean(axis=0) / n
    assert np.max(np.abs(an)) < 4.5 * np.max(nk + 0.5) * np.sqrt(2.0 / S)
    # real-space fluctuation variance per site
    var = (np.abs(bs - mean) ** 2).mean()
    assert abs(var - (nk.mean() + 0.5)) < 0.02 * (nk.mean() + 0.5)


def test_vacuum_profile_reduces_to_vacuum_start():
    """n_k = 0, mean 0 is the vacuum Wigner state: every quadrature has
    variance 1/4, as the reference's beta_init_sigma = 0.5 start."""
    nx, ny = 4, 16
    bs = np.array([oracle.gaussian_state(nx, ny, SEED, t, 0.0, np.zeros((nx, ny))) for t in range(2000)])
    v = bs.real.var()
    assert abs(v - 0.25) < 0.01 and abs(bs.imag.var() - 0.25) < 0.01
