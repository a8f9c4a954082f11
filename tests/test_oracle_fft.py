"""Pins the oracle FFT against a long-double direct DFT and closed forms."""
import numpy as np
import pytest

import oracle

SIZES_1D = [1, 2, 3, 4, 5, 6, 7, 8, 11, 12, 16, 32, 48, 49, 64, 67, 96, 97, 127, 128, 256, 384, 512]
SIZES_2D = [(2, 2), (4, 4), (8, 8), (12, 10), (16, 15), (16, 16), (3, 5), (48, 48), (5, 7), (13, 11)]


def _rand(shape, seed):
    g = np.random.default_rng(seed)
    return g.standard_normal(shape) + 1j * g.standard_normal(shape)


@pytest.mark.parametrize("n", SIZES_1D)
def test_1d_vs_direct(n):
    x = _rand((1, n), n)
    for sign in (-1, +1):
        ref = oracle.dft_direct(x, sign)
        got = oracle.fft(x, sign)
        assert np.max(np.abs(got - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("shape", SIZES_2D)
def test_2d_vs_direct(shape):
    x = _rand(shape, shape[0] * 100 + shape[1])
    for sign in (-1, +1):
        ref = oracle.dft_direct(x, sign)
        got = oracle.fft(x, sign)
        assert np.max(np.abs(got - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))


def test_closed_forms():
    nx, ny = 16, 12
    # delta -> constant
    d = np.zeros((nx, ny), complex)
    d[0, 0] = 1
    assert np.allclose(oracle.fft(d), np.ones((nx, ny)), atol=1e-15)
    # plane wave e^{+i k.r} -> N delta at k (forward sign e^{-i k.r})
    kx, ky = 3, 5
    x, y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    pw = np.exp(2j * np.pi * (kx * x / nx + ky * y / ny))
    f = oracle.fft(pw, -1)
    expect = np.zeros((nx, ny), complex)
    expect[kx, ky] = nx * ny
    assert np.max(np.abs(f - expect)) < 1e-12
    # Parseval, linearity, backward(forward(x)) = N x
    a, b = _rand((nx, ny), 1), _rand((nx, ny), 2)
    fa, fb = oracle.fft(a), oracle.fft(b)
    assert abs(np.sum(np.abs(fa) ** 2) / (nx * ny) - np.sum(np.abs(a) ** 2)) < 1e-10
    assert np.max(np.abs(oracle.fft(2 * a + 3j * b) - (2 * fa + 3j * fb))) < 1e-12
    assert np.max(np.abs(oracle.fft(fa, +1) / (nx * ny) - a)) < 1e-13
