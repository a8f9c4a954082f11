"""Host logic of the learned warm start (no GPU): the steady-state targets
recover a known Gaussian ensemble's mean field and radial populations, the
radial bin table has the properties its definition fixes, and the
time-to-steady-state rule finds the last exit from the 3 sigma band."""
import numpy as np

import oracle
from pincosimulator_amd import warmstart as ws


def test_radial_bins_properties():
    for nx, ny in ((16, 16), (64, 64), (12, 10)):
        b = ws.radial_bins(nx, ny)
        assert b[0, 0] == 0 and np.count_nonzero(b == 0) == 1 and b.max() <= 31
        # inversion symmetry k -> -k
        assert np.array_equal(b, b[(-np.arange(nx)) % nx][:, (-np.arange(ny)) % ny])
    b = ws.radial_bins(64, 64)
    # nearest neighbours of k = 0 have |k| = 2 pi / 64 < bin width: bin 1
    assert b[1, 0] == 1 and b[0, 1] == 1
    # the corner (pi, pi) has |k| = pi sqrt 2: last bin
    assert b[32, 32] == 31
    assert set(np.unique(b)) == set(range(32))  # 64 x 64 fills every bin


def test_steady_targets_recover_known_ensemble():
    nx = ny = 16
    n = nx * ny
    mean = 0.8 + 0.3j
    bins = ws.radial_bins(nx, ny)
    nb_true = 0.05 + 0.2 * np.exp(-np.arange(32) / 8.0)
    nk = nb_true[bins]
    S = 800
    frames = 3
    ks = np.zeros((frames, 34))
    for f in range(frames):
        bs = np.array([oracle.gaussian_state(nx, ny, 11, 1000 * f + t, mean, nk) for t in range(S)])
        k2 = np.abs(np.fft.fft2(bs, axes=(1, 2))) ** 2
        for b in range(32):
            ks[f, b] = k2[:, bins == b].sum()
        ks[f, 32], ks[f, 33] = bs.sum().real, bs.sum().imag
    t, w = ws.steady_targets(ks, S, nx, ny, np.arange(frames))
    assert abs(t[0] + 1j * t[1] - mean) < 0.01
    count = np.bincount(bins.reshape(-1), minlength=32)
    for b in range(32):
        if count[b] == 0:
            assert w[2 + b] == 0
            continue
        se = (nk[bins == b].mean() + 0.5) / np.sqrt(S * frames * count[b])
        assert abs(t[2 + b] - nb_true[b]) < 5 * se, b


def test_time_to_steady_rule():
    x = np.array([5.0, 3.0, 1.2, 1.01, 0.99, 1.3, 1.0, 1.0])
    se = np.full_like(x, 0.05)
    # inside the band: F F F T T F T T
    assert ws.time_to_steady(x, se, 1.0, 0.0, hold=2) == 3
    assert ws.time_to_steady(x, se, 1.0, 0.0, hold=3) == 6  # frame 5 breaks the hold from 3
    assert ws.time_to_steady(x, se, 1.0, 0.0, hold=10) == 6  # holds to the end of the run
    assert ws.time_to_steady(np.ones(4), se[:4], 1.0, 0.0) == 0
    assert ws.time_to_steady(np.array([1.0, 2.0]), se[:2], 1.0, 0.0) is None


def test_features_scale_the_training_ranges_to_order_one():
    F, D, L = ws.training_points(4096, 99)
    x = ws.features(F, D, 0.05, 1.0, L)
    assert x.shape == (4096, 5) and np.all(np.abs(x) <= 1.0 + 1e-6)
    assert set(L) == {16, 32, 48, 64}
