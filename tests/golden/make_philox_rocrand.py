"""Writes philox_rocrand.json: Philox4x32-10 outputs from rocRAND's header-only
engine (an independent implementation of Salmon et al.'s generator), compiled
for the host from /opt/rocm/include/rocrand/rocrand_philox4x32_10.h.

rocRAND maps seed -> key (lo, hi), subsequence -> counter words (z, w) and
offset/4 -> counter words (x, y), so seeding with (key, (c3 << 32) | c2,
4 * ((c1 << 32) | c0)) makes its first next4() return philox(ctr, key).
The offset is a 64-bit count of numbers, so c1 stays below 2^30 here."""
import json
import os
import subprocess
import tempfile

SRC = r'''
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>
int main() {
  const uint32_t cases[][6] = {
    {0u, 0u, 0u, 0u, 0u, 0u},
    {0xffffffffu, 0x3fffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
    {0x243f6a88u, 0x05a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
    {5u, 1000u, 42u, 0u, 1234u, 0u},
    {2047u, 49999u, 16383u, 0x01000000u, 0xdeadbeefu, 0x12345678u},
    {1u, 0u, 7u, 0x07000000u, 7u, 0u},
  };
  std::printf("[");
  for (int i = 0; i < 6; ++i) {
    const uint32_t* c = cases[i];
    unsigned long long seed = ((unsigned long long)c[5] << 32) | c[4];
    unsigned long long sub = ((unsigned long long)c[3] << 32) | c[2];
    unsigned long long off = 4ull * (((unsigned long long)c[1] << 32) | c[0]);
    rocrand_device::philox4x32_10_engine e(seed, sub, off);
    uint4 r = e.next4();
    std::printf("%s{\"ctr\":[%u,%u,%u,%u],\"key\":[%u,%u],\"out\":[%u,%u,%u,%u]}", i ? "," : "",
                c[0], c[1], c[2], c[3], c[4], c[5], r.x, r.y, r.z, r.w);
  }
  std::printf("]\n");
}
'''

def main():
    here = os.path.dirname(os.path.abspath(__file__))
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "p.cpp")
        exe = os.path.join(d, "p")
        open(src, "w").write(SRC)
        subprocess.check_call(["/opt/rocm/bin/hipcc", "-O1", "-x", "hip", "--offload-arch=gfx950", src, "-o", exe])
        out = subprocess.check_output([exe]).decode()
    data = json.loads(out)
    with open(os.path.join(here, "philox_rocrand.json"), "w") as f:
        json.dump({"source": "rocRAND philox4x32_10_engine (ROCm 7.2 headers), host build", "cases": data}, f, indent=1)


if __name__ == "__main__":
    main()
