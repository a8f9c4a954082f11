"""Reader for the PincoFormat archive written by twmc_sim (libpinco_host.so).

Layout (the reference's src/Base/FileFormats/PincoFormat, written by this
project's pincosimulator_amd/host/pinco_format.cpp):
  _register.bin  magic, version, dataset names / frame dims / formats, entry size,
                 then entries {id, chunk, index in chunk, continuation offset, t0, t1}
  variable<i>_<chunk>.cnk  size-prefixed raw data per entry
  index_<chunk>.bin        {id, (frames, offset) per dataset} per entry
"""
from __future__ import annotations

import os
import struct

import numpy as np

MAGIC = bytes([0x89, ord("P"), ord("N"), ord("C"), 0x0D, 0x0A, 0x1A, 0x0A])
_DTYPES = {22: np.complex128, 11: np.float64}


class Archive:
    def __init__(self, folder):
        self.folder = folder
        with open(os.path.join(folder, "_register.bin"), "rb") as f:
            raw = f.read()
        if raw[:8] != MAGIC:
            raise ValueError("not a PincoFormat register (bad magic)")
        self.version = raw[8]
        p = 9
        nd = raw[p]
        p += 1
        self.names = []
        for _ in range(nd):
            ln = raw[p]
            p += 1
            self.names.append(raw[p:p + ln].decode())
            p += ln
        self.dims = []
        for _ in range(nd):
            (D,) = struct.unpack_from("<Q", raw, p)
            p += 8
            self.dims.append(list(struct.unpack_from("<%dQ" % D, raw, p)))
            p += 8 * D
        self.formats = list(raw[p:p + nd])
        p += nd
        self.extra_size, self.entries_begin = struct.unpack_from("<Qq", raw, p)
        esz = 32 + self.extra_size
        self.entries = []
        q = self.entries_begin
        while q + esz <= len(raw):
            tid, chunk, index, cont = struct.unpack_from("<4Q", raw, q)
            t0, t1 = struct.unpack_from("<2d", raw, q + 32)
            self.entries.append(dict(id=tid, chunk=chunk, index=index, continuation=cont, t0=t0, t1=t1, pos=q))
            q += esz

    def ids(self):
        return sorted({e["id"] for e in self.entries})

    def load(self, tid, last_frame_only=False):
        """dict name -> array (frames, *dims) concatenated over continuations, plus 't'."""
        chain = [e for e in self.entries if e["id"] == tid]
        if not chain:
            raise KeyError(tid)
        if last_frame_only:
            chain = chain[-1:]
        out = {n: [] for n in self.names}
        nrec = 1 + 2 * len(self.names)
        for e in chain:
            with open(os.path.join(self.folder, "index_%d.bin" % e["chunk"]), "rb") as f:
                f.seek(e["index"] * nrec * 8)
                rec = struct.unpack("<%dQ" % nrec, f.read(nrec * 8))
            for i, name in enumerate(self.names):
                frames, off = rec[1 + 2 * i], rec[2 + 2 * i]
                with open(os.path.join(self.folder, "variable%d_%d.cnk" % (i, e["chunk"])), "rb") as f:
                    f.seek(off)
                    (nbytes,) = struct.unpack("<Q", f.read(8))
                    data = np.frombuffer(f.read(nbytes), dtype=_DTYPES[self.formats[i]])
                shape = (frames,) + tuple(self.dims[i]) if frames else (0,)
                arr = data.reshape(shape) if data.size else data
                if last_frame_only and frames > 1:
                    arr = arr[-1:]
                out[name].append(arr)
        res = {n: (np.concatenate(v) if len(v) > 1 else v[0]) for n, v in out.items()}
        res["t"] = (chain[0]["t0"], chain[-1]["t1"])
        return res
