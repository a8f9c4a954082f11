"""FP32 uniform FFT-hopping lattices on the stacked stencil (the sweep kernel
run as one point of the main API, trajectories stacked per work item, the
batch's own trajectory ids): per trajectory against the FP64 oracle for
chains and 2D shapes with scattered, unsorted ids, against the LDS kernel,
ensemble sums against the dump, and the AUTO selection.  FP32 bar: FP32_BAR of
max |beta| per trajectory."""
import numpy as np
import pytest

import oracle
from helpers import FP32_BAR, rel_err, system
from pincosimulator_amd import Engine
from pincosimulator_amd._lib import TWMCError
from pincosimulator_amd.observables import from_trajectories

pytestmark = pytest.mark.gpu

IDS = np.array([3, 1000, 17, 5, 123456789012, 6, 7, 8, 9, 99], dtype=np.uint64)


def spec(nx, ny, **kw):
    s = dict(nx=nx, ny=ny, dt=0.004, gamma=1.0, J=0.9, omega=0.6, U=0.07, F=1.4, beta_init=0.1, beta_sigma=0.5,
             frame_steps=50, seed=4711, t_start=0.0, t_end=0.8)
    s.update(kw)
    return s


@pytest.mark.parametrize("nx,ny", [(1, 64), (1, 500), (16, 16), (12, 10), (48, 48), (2, 8), (32, 2), (64, 32)])
def test_stack_vs_fp64_oracle(nx, ny):
    sp = spec(nx, ny)
    print(f"stack {nx}x{ny}", flush=True)
    eng = Engine(system(sp), precision=32, kernel="stack")
    res = eng.run(IDS, traj=True)
    eng.close()
    assert res.kernel == "stack"
    err = max(rel_err(res.traj[j], oracle.run_trajectory(dict(sp, lattice=oracle.LAT_FFT), int(t))[0])
              for j, t in enumerate(IDS))
    print(f"  max relative deviation {err:.2e}")
    assert err < FP32_BAR


@pytest.mark.parametrize("nx,ny,n_traj", [(24, 24, 300), (32, 48, 3000), (48, 48, 3000)])
def test_stack_vs_lds_kernel_and_sums(nx, ny, n_traj):
    """Frames of the stacked kernel come from the LDS kernel's capture launch:
    256 threads at 24x24, 1024 threads with one and two site pairs per thread
    at 32x48 and 48x48, where the batch gives each workgroup several
    trajectories to sum in registers."""
    sp = spec(nx, ny, t_end=0.4)
    ids = np.arange(n_traj, dtype=np.uint64) * 7 + 1
    print(f"stack sums {nx}x{ny}, {n_traj} trajectories", flush=True)
    a = Engine(system(sp), precision=32, kernel="stack").run(ids, traj=True)
    b = Engine(system(sp), precision=32, kernel="generic").run(ids, traj=True)
    assert rel_err(a.traj, b.traj) < 2e-5
    ref = from_trajectories(a.traj, (nx, ny))
    for k in ("sum_n", "sum_n2", "sum_beta", "traj_m2", "traj_m4"):
        assert rel_err(getattr(a, k), ref[k]) < 1e-9, k
    assert rel_err(a.sum_nk, ref["sum_nk"]) < 1e-5


def test_stack_resume_continues_the_stream():
    """Two segments with the second started from the first's final state
    (init previous, step0 carried) equal one unbroken run."""
    sp = spec(16, 16, t_end=0.4, frame_steps=25)
    ids = np.arange(40, dtype=np.uint64)
    full = Engine(system(sp), precision=32).run(ids, final=True)
    eng = Engine(system(dict(sp, t_end=0.2)), precision=32)
    half = eng.run(ids, final=True)
    eng.close()
    eng = Engine(system(dict(sp, t_start=0.2, t_end=0.4)), precision=32)
    rest = eng.run(ids, init="previous", beta_prev=half.beta_final, final=True, step0=half.n_steps)
    eng.close()
    assert rel_err(rest.beta_final, full.beta_final) < 1e-6


def test_auto_selection():
    # up to 64 sites the LDS kernel (packed trajectories) is faster than the stack
    for sp, prec, want in [(spec(32, 32), 32, "stack"), (spec(64, 64, t_end=0.2), 32, "stencil64"),
                           (spec(1, 64), 32, "generic"), (spec(8, 8), 32, "generic"), (spec(1, 66), 32, "stack"),
                           (spec(32, 32), 64, "generic"), (spec(5, 7), 32, "generic"),
                           (spec(32, 32, F_t=(np.array([0.0, 1.0]), np.ones((2, 1024), complex))), 32, "generic")]:
        eng = Engine(system(sp), precision=prec)
        assert eng.run(np.arange(4)).kernel == want, (sp["nx"], sp["ny"], prec, want)
        eng.close()
    with pytest.raises(TWMCError):
        Engine(system(spec(5, 7)), precision=32, kernel="stack")
