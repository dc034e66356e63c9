#!/usr/bin/env python3
"""Write tests/golden/philox_vectors.json from rocRAND's Philox4x32-10 engine.

rocRAND's ``philox4x32_10_engine`` (``/opt/rocm/include/rocrand/
rocrand_philox4x32_10.h``, an open-source header) is host-callable.  Seeded
with ``key = (k0, k1)``, sub-sequence ``c2 + 2^32 c3`` and offset
``4 (c0 + 2^32 c1)``, its first ``next4()`` is the Philox block of counter
``(c0, c1, c2, c3)``.  This script compiles a small host program against that
header, evaluates it on the Random123 known-answer inputs and on counters of
the shapes the sweep engine uses, and stores the outputs as a fixture.  The
fixture is data: the tests compare the oracle's and the engine's Philox with it
and never need ROCm to do so.

Run from the repository root:  python tests/golden/make_philox_vectors.py
"""

import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = [
    # Random123 known-answer inputs (kat_vectors: philox4x32 10 rounds)
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000)),
    ((0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0)),
    # engine-shaped counters: (pair, sweep, lattice, tag) under assorted seeds
    ((0, 0, 0, 0), (0x9E3779B9, 0x7F4A7C15)),
    ((511, 21999, 1279, 0), (0x12345678, 0x9ABCDEF0)),
    ((524287, 4999, 191, 0), (0xDEADBEEF, 0x00C0FFEE)),
    ((3, 7, 11, 1), (0x0BADF00D, 0x0D15EA5E)),
    ((8191, 0, 4095, 1), (1, 0)),
    ((123456, 7654321, 42, 0), (0xFFFFFFFF, 0)),
]

SOURCE = r"""
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
int main() {
  unsigned int in[][6] = { %s };
  const int n = sizeof(in) / sizeof(in[0]);
  for (int i = 0; i < n; ++i) {
    unsigned long long seed = (unsigned long long)in[i][4] | ((unsigned long long)in[i][5] << 32);
    unsigned long long sub = (unsigned long long)in[i][2] | ((unsigned long long)in[i][3] << 32);
    unsigned long long blocks = (unsigned long long)in[i][0] | ((unsigned long long)in[i][1] << 32);
    rocrand_device::philox4x32_10_engine e(seed, sub, 0ull);
    while (blocks > 0) {
      unsigned long long step = blocks < (1ull << 61) ? blocks : (1ull << 61);
      e.discard(4ull * step);
      blocks -= step;
    }
    uint4 r = e.next4();
    std::printf("%%u %%u %%u %%u\n", r.x, r.y, r.z, r.w);
  }
  return 0;
}
"""


def main():
    rows = ",".join("{%s}" % ",".join(str(v) for v in (*c, *k)) for c, k in CASES)
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "philox_kat.hip")
        exe = os.path.join(tmp, "philox_kat")
        with open(src, "w") as fh:
            fh.write(SOURCE % rows)
        subprocess.run(["hipcc", "--offload-arch=gfx950", "-O1", src, "-o", exe],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    vectors = []
    for (ctr, key), line in zip(CASES, out):
        vectors.append({"counter": list(ctr), "key": list(key),
                        "output": [int(v) for v in line.split()]})
    with open(os.path.join(HERE, "philox_vectors.json"), "w") as fh:
        json.dump({"source": "rocRAND philox4x32_10_engine, ROCm 7.2 header",
                   "vectors": vectors}, fh, indent=1)


if __name__ == "__main__":
    main()
