"""PincoFormat bytes written by the project's writer (host/pinco_format.cpp,
run on the CPU by a small native program) against bytes assembled from the
reference's write sequence (tests/pinco_bytes.py): register header and
entries with continuation offsets, chunk files and chunk indices, for a first
run and a continuation run of the same folder."""
import os
import subprocess

import numpy as np

from pinco_bytes import ENTRY_BYTES, chunk_record, entry, header, index_record

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "pincosimulator_amd", "host")


def values(tid, count, scale):
    return scale * tid + 0.125 * np.arange(count, dtype=np.float64)


def test_register_chunks_and_index_bytes(tmp_path):
    exe = str(tmp_path / "pinco_format_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", HOST, os.path.join(ROOT, "tests", "native", "pinco_format_check.cpp"),
                    os.path.join(HOST, "pinco_format.cpp"), "-o", exe], check=True, capture_output=True, text=True)
    folder = str(tmp_path / "arc")
    r = subprocess.run([exe, folder], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    hdr = header(["traj", "mean_n2"], [[3, 2], []], [22, 11])
    begin = len(hdr)
    reg = open(os.path.join(folder, "_register.bin"), "rb").read()
    assert reg == (hdr + entry(1, 0, 0, begin + 2 * ENTRY_BYTES, 0.0, 1.5)
                   + entry(2, 0, 1, begin + 3 * ENTRY_BYTES, 0.0, 1.25)
                   + entry(1, 1, 0, 0, 1.5, 2.5) + entry(2, 1, 1, 0, 1.25, 2.5))
    for chunk, frames in ((0, 3), (1, 2)):
        recs = b""
        cnk = [b"", b""]
        for tid in (1, 2):
            traj = values(tid, frames * 12, 100.0).tobytes()
            m2 = values(tid, frames, -10.0).tobytes()
            recs += index_record(tid, [(frames, len(cnk[0])), (frames, len(cnk[1]))])
            cnk[0] += chunk_record(traj)
            cnk[1] += chunk_record(m2)
        assert open(os.path.join(folder, "index_%d.bin" % chunk), "rb").read() == recs
        for i in (0, 1):
            assert open(os.path.join(folder, "variable%d_%d.cnk" % (i, chunk)), "rb").read() == cnk[i], (chunk, i)
