"""The archive twmc_sim writes, checked byte for byte against the reference's
write sequence (tests/pinco_bytes.py, assembled independently of the
project's own reader), and the worker pool on the GPU.  A new run folder's
first chunk is 0 and a later run's is the largest used id + 1
(PincoFormatDataStore.cpp:40-57; the reference's max over an empty set is
undefined, this writer starts at 0).  The payload is checked against the FP64
oracle to the FP64 bar (1e-9)."""
import os
import struct

import numpy as np
import pytest

import oracle
from helpers import rel_err
from pinco_bytes import ENTRY_BYTES, entry, header
from test_gpu_cli import CHAIN, run_sim

pytestmark = pytest.mark.gpu

N, DT, FS = 16, 0.005, 40


def spec(t_end):
    return dict(nx=1, ny=N, lattice=oracle.LAT_FFT, dt=DT, gamma=1.0, J=1.0, omega=0.5, U=0.1, F=1.0,
                beta_init=0.0, beta_sigma=0.5, frame_steps=FS, seed=42, t_start=0.0, t_end=t_end)


def check_chunk(folder, chunk, ids, frames_list):
    """index_<chunk>.bin byte for byte; variable0_<chunk>.cnk layout byte for byte, payload vs oracle."""
    idx = bytearray()
    cnk = open(os.path.join(folder, "variable0_%d.cnk" % chunk), "rb").read()
    off = 0
    for tid, frames in zip(ids, frames_list):
        nbytes = frames.size * 16
        idx += struct.pack("<3Q", tid, frames.shape[0], off)
        assert struct.unpack_from("<Q", cnk, off)[0] == nbytes
        got = np.frombuffer(cnk[off + 8:off + 8 + nbytes], dtype=np.complex128).reshape(frames.shape)
        assert rel_err(got, frames) < 1e-9
        off += 8 + nbytes
    assert len(cnk) == off
    assert open(os.path.join(folder, "index_%d.bin" % chunk), "rb").read() == bytes(idx)


def test_register_chunks_and_index_bytes_with_a_continuation(tmp_path):
    folder = str(tmp_path / "bytes")
    ini = CHAIN.replace("n_traj = 6", "n_traj = 2")
    run_sim(folder, ini.format(t_end=1.0))
    f1, tf1 = {}, {}
    for tid in (1, 2):
        fr, bf, tf = oracle.run_trajectory(spec(1.0), tid)
        f1[tid], tf1[tid] = (fr, bf), tf
    hdr = header(["traj"], [[1, N]], [22])
    begin = len(hdr)
    reg = open(os.path.join(folder, "_register.bin"), "rb").read()
    # store = traj writes the reference's dataset list only: "traj" (TWMCResults.cpp:26-38)
    assert reg == hdr + entry(1, 0, 0, 0, 0.0, tf1[1]) + entry(2, 0, 1, 0, 0.0, tf1[2])
    check_chunk(folder, 0, [1, 2], [f1[1][0], f1[2][0]])
    assert sorted(os.listdir(folder)) == sorted(["_register.bin", "variable0_0.cnk", "index_0.bin", "_t.dat",
                                                 "_observables.dat", "_n_k.dat", "run.ini"])

    # continue both trajectories to t = 2: new entries in chunk 1, the old ones point at them
    run_sim(folder, ini.format(t_end=2.0))
    f2, tf2 = {}, {}
    for tid in (1, 2):
        prev_final = f1[tid][0][-1]
        step0 = int(round(tf1[tid] / DT))
        fr, _, tf = oracle.run_trajectory(spec(2.0), tid, t_start=tf1[tid], t_end=2.0,
                                          init_mode=oracle.INIT_PREVIOUS, beta_prev=prev_final,
                                          store_initial=False, step0=step0)
        f2[tid], tf2[tid] = fr, tf
    esz = ENTRY_BYTES
    reg = open(os.path.join(folder, "_register.bin"), "rb").read()
    assert reg == (hdr + entry(1, 0, 0, begin + 2 * esz, 0.0, tf1[1]) + entry(2, 0, 1, begin + 3 * esz, 0.0, tf1[2])
                   + entry(1, 1, 0, 0, tf1[1], tf2[1]) + entry(2, 1, 1, 0, tf1[2], tf2[2]))
    check_chunk(folder, 0, [1, 2], [f1[1][0], f1[2][0]])
    check_chunk(folder, 1, [1, 2], [f2[1], f2[2]])
    # the unbroken oracle run equals the two segments
    full, _, _ = oracle.run_trajectory(spec(2.0), 1)
    assert rel_err(np.concatenate([f1[1][0], f2[1]]), full) < 1e-12


def test_store_both_adds_the_site_means(tmp_path):
    folder = str(tmp_path / "both")
    run_sim(folder, CHAIN.replace("n_traj = 6", "n_traj = 2").format(t_end=0.4) + 'store = "both"\n')
    reg = open(os.path.join(folder, "_register.bin"), "rb").read()
    assert reg.startswith(header(["traj", "mean_n2", "mean_n4"], [[1, N], [], []], [22, 11, 11]))


def test_two_workers_write_the_same_trajectories_as_one(tmp_path):
    """gpu_workers = 2 on GPU 0 (two contexts, two streams, batches of 2 from one
    queue) against one worker: the same trajectories per id, bit for bit, the
    same ensemble observables up to summation order, and the reporter's lines."""
    from pincosimulator_amd.pinco_format import Archive
    ini = CHAIN.format(t_end=1.0) + "gpu_batch = 2\nprogress_interval = 0.001\n"
    one = run_sim(str(tmp_path / "one"), ini)
    two = run_sim(str(tmp_path / "two"), ini, "gpu_workers", "2")
    assert one.ids() == two.ids() == [1, 2, 3, 4, 5, 6]
    for tid in one.ids():
        a, b = one.load(tid), two.load(tid)
        assert np.array_equal(a["traj"], b["traj"]) and a["t"] == b["t"], tid
    oa = np.loadtxt(str(tmp_path / "one" / "_observables.dat"), comments="%")
    ob = np.loadtxt(str(tmp_path / "two" / "_observables.dat"), comments="%")
    np.testing.assert_allclose(oa, ob, rtol=1e-12, atol=1e-14)
    assert isinstance(two, Archive)


def test_progress_lines_reach_all_trajectories(tmp_path):
    import subprocess
    from test_gpu_cli import SIM
    folder = str(tmp_path / "prog")
    os.makedirs(folder)
    ini = os.path.join(folder, "run.ini")
    with open(ini, "w") as f:
        f.write(CHAIN.format(t_end=2.0) + "gpu_batch = 3\nprogress_interval = 0.001\n")
    p = subprocess.run([SIM, "-i", ini, "-o", folder, "gpu_workers", "2"], capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 0, p.stderr
    lines = [ln for ln in p.stdout.splitlines() if ln.startswith("progress:")]
    assert lines and lines[-1].startswith("progress: 6/6 trajectories computed, 6 saved, 100.0%"), lines[-3:]
