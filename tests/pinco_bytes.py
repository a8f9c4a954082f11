"""Expected PincoFormat bytes assembled from the reference's write sequence
(independent of the project's writer and of pincosimulator_amd/pinco_format.py):
  * _register.bin header, ChunkRegister::InitializeRegisterHeader
    (src/Base/FileFormats/PincoFormat/ChunkRegister.cpp:152-233): 8 magic bytes,
    the version byte, fputc(dataset count), per dataset fputc(name length) +
    name, per dataset size_t D + D size_t dims, per dataset fputc(format), long
    extra-data size, long offset of the first entry (the position right after
    that field);
  * register entries, ChunkRegister::RegisterStoredData (:312-381): 4 size_t
    {id, chunk id, index in the chunk, continuation offset} + the extra data
    {t_start, t_final};
  * index_<chunk>.bin records, ChunkFileSet::WriteToChunk (:111-145, record size
    :93-94): size_t id + {elements, offset} per dataset;
  * variable<i>_<chunk>.cnk, ChunkFileSet::WriteDataToChunk (:147-160): size_t
    byte count + raw bytes per entry."""
import struct

MAGIC = bytes([0x89, 0x50, 0x4E, 0x43, 0x0D, 0x0A, 0x1A, 0x0A])
ENTRY_BYTES = 4 * 8 + 16


def header(names, dims, formats):
    b = bytearray(MAGIC) + bytes([1, len(names)])
    for n in names:
        b += bytes([len(n)]) + n.encode()
    for d in dims:
        b += struct.pack("<Q", len(d)) + struct.pack("<%dQ" % len(d), *d)
    b += bytes(formats)
    b += struct.pack("<q", 16)
    b += struct.pack("<q", len(b) + 8)
    return bytes(b)


def entry(tid, chunk, index, cont, t0, t1):
    return struct.pack("<4Q2d", tid, chunk, index, cont, t0, t1)


def index_record(tid, elements_offsets):
    b = struct.pack("<Q", tid)
    for el, off in elements_offsets:
        b += struct.pack("<2Q", el, off)
    return b


def chunk_record(payload):
    return struct.pack("<Q", len(payload)) + payload
