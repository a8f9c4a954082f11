"""Allocation-free stepping, twmc_advance_noalloc, called through the C ABI on
device buffers the test owns, against the allocating entry point of the same
context (a batch begun from the same state with init PREVIOUS and the same
step0, its final state and its one stored frame's ensemble sums).  Two calls
in a row on one context, the second with fewer trajectories and a later step0
(the array the library keeps between calls holds the first call's partial
sums then).  Bar: the state bit for bit (the same kernel on the same inputs),
the FP64 sums to 1e-12 relative, and the sums also against the host's own sum
over the returned state.  Also: a context with per-trajectory disorder is
rejected before any device work.  Device buffers come from the HIP runtime
the library itself uses (pincosimulator_amd.device), not from a second one."""
import numpy as np
import pytest

from pincosimulator_amd import Engine, System, _lib
from pincosimulator_amd._lib import check
from pincosimulator_amd.device import DeviceBuffer
from test_gpu_mfma import c3_spec

pytestmark = pytest.mark.gpu


def square(nx, ny, steps):
    return dict(nx=nx, ny=ny, dt=0.004, gamma=1.0, J=0.9, omega=0.6, U=0.07, F=1.4, beta_init=0.0,
                beta_sigma=0.5, frame_steps=steps, seed=9090, t_start=0.0, t_end=1.0)


CASES = [("stencil64", 32, square(64, 64, 30), 96, 40),
         ("generic", 64, square(12, 10, 30), 20, 7),
         ("generic", 64, square(64, 64, 30), 300, 40),
         ("generic", 32, square(17, 5, 30), 33, 9),
         ("band", 32, dict(c3_spec(cells=128, t_end=1.0, frame_steps=30)), 100, 33),
         ("mfma", 32, dict(c3_spec(cells=128, t_end=1.0, frame_steps=30)), 70, 33)]


def allocating(eng, ids, b0, step0, steps, dt):
    return eng.run(ids, init="previous", beta_prev=b0, store_initial=False, step0=step0,
                   t_start=step0 * dt, t_end=(step0 + steps) * dt, final=True)


@pytest.mark.parametrize("kernel,prec,spec,b1,b2", CASES,
                         ids=[f"{c[0]}-fp{c[1]}-{c[2]['nx']}x{c[2]['ny']}" for c in CASES])
def test_noalloc_equals_the_allocating_entry_point(kernel, prec, spec, b1, b2):
    steps, dt = 30, spec["dt"]
    n = spec["nx"] * spec["ny"]
    eng = Engine(System(**spec), precision=prec, kernel=kernel)
    L = _lib.load()
    rng = np.random.default_rng(b1 + b2)
    cdt = np.complex64 if prec == 32 else np.complex128
    for B, step0 in ((b1, 0), (b2, steps)):
        print(f"{kernel} FP{prec}: {B} trajectories from step {step0}", flush=True)
        ids = (np.arange(B, dtype=np.uint64) * 3 + 5).astype(np.uint64)
        b0 = (rng.normal(size=(B, n)) + 1j * rng.normal(size=(B, n))) * 0.7
        b0 = b0.astype(cdt).astype(np.complex128)  # the state both paths start from
        ref = allocating(eng, ids, b0, step0, steps, dt)
        assert ref.kernel == kernel and ref.n_frames == 1 and ref.n_steps == steps

        d_ids = DeviceBuffer.from_array(ids)
        d_beta = DeviceBuffer.from_array(b0.astype(cdt))
        d_sums = DeviceBuffer.from_array(np.full(5 * n, np.nan))
        check(L.twmc_advance_noalloc(eng._ctx, d_ids.ptr, B, step0, steps, d_beta.ptr, d_sums.ptr, None), eng._ctx)
        check(L.twmc_batch_sync(eng._ctx), eng._ctx)
        got = d_beta.read(cdt, (B, n)).astype(np.complex128)
        sums = d_sums.read(np.float64, (5, n))
        for b in (d_ids, d_beta, d_sums):
            b.free()
        assert np.array_equal(got, ref.beta_final), np.max(np.abs(got - ref.beta_final))
        for row, want in ((0, ref.sum_n[0]), (1, ref.sum_n2[0]), (3, ref.sum_beta[0].real), (4, ref.sum_beta[0].imag)):
            np.testing.assert_allclose(sums[row], want, rtol=1e-12, atol=1e-12 * np.max(np.abs(want)))
        assert np.all(sums[2] == 0.0)
        nn = np.abs(got) ** 2
        np.testing.assert_allclose(sums[0], nn.sum(axis=0), rtol=1e-12)
        np.testing.assert_allclose(sums[1], (nn ** 2).sum(axis=0), rtol=1e-12)
        np.testing.assert_allclose(sums[3] + 1j * sums[4], got.sum(axis=0), rtol=1e-12, atol=1e-12)
    eng.close()


def test_noalloc_rejects_per_trajectory_disorder():
    spec = dict(square(8, 8, 10), U={"base": 0.07, "noise": ("gaussian", 0.01)})
    eng = Engine(System(**spec), precision=32, kernel="generic")
    L = _lib.load()
    d = DeviceBuffer(64 * 8)
    rc = L.twmc_advance_noalloc(eng._ctx, d.ptr, 1, 0, 1, d.ptr, None, None)
    assert rc == -5 and b"site-shared" in L.twmc_last_error(eng._ctx)
    d.free()
    eng.close()
