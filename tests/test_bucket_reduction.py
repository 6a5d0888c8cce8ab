"""The probe's bucket index (hash mod num_buckets) is computed on the device
by Barrett reduction with a precomputed reciprocal (wk_common.hpp:bucket_of)
instead of a 64-bit division.  The host store places keys with the plain
`%` (store.cpp:bucketize, as KVStore::get_slot_local does with
`hash % num_buckets`, core/store/kvstore.hpp).  This test compiles the shared
header on the host and checks bucket_of against `%` over edge cases and
random 64-bit hashes for divisors across the whole range."""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "wukong-cube_amd", "csrc")

HARNESS = r"""
#include <cstdio>
#include <cstdlib>
#include "wk_common.hpp"
using namespace wk;
int main() {
  uint64_t h, d, bad = 0, n = 0;
  while (std::scanf("%lu %lu", &h, &d) == 2) {
    const uint64_t got = bucket_of(h, d, bucket_magic(d));
    if (got != h % d) {
      if (bad < 5) std::printf("mismatch h=%lu d=%lu got=%lu want=%lu\n", h, d, got, h % d);
      bad++;
    }
    n++;
  }
  std::printf("checked %lu bad %lu\n", n, bad);
  return bad != 0;
}
"""


def _cases():
    rng = np.random.default_rng(11)
    u64 = np.iinfo(np.uint64).max
    divs = [1, 2, 3, 5, 7, 8, 255, 256, 257, 1 << 20, (1 << 31) - 1, 1 << 31, (1 << 32) - 1, 1 << 32,
            (1 << 32) + 1, 460814056, 87475533, (1 << 63) - 1, 1 << 63, (1 << 63) + 1, u64 - 1, u64]
    divs += [int(x) for x in rng.integers(1, 1 << 34, size=200, dtype=np.uint64)]
    divs += [int(x) for x in rng.integers(1, u64, size=50, dtype=np.uint64)]
    hs_fixed = [0, 1, 2, u64, u64 - 1, 1 << 63, (1 << 63) - 1, (1 << 32) - 1, 1 << 32]
    out = []
    for d in divs:
        hs = hs_fixed + [int(x) for x in rng.integers(0, u64, size=300, dtype=np.uint64, endpoint=True)]
        # values next to multiples of d, where a quotient estimate off by one shows
        k = u64 // d
        for j in (k, k - 1, k // 2, 1):
            if j > 0:
                for e in (-1, 0, 1):
                    v = j * d + e
                    if 0 <= v <= u64:
                        hs.append(v)
        out += [(h, d) for h in hs]
    # small divisors exhaustively over a window of hashes
    for d in range(1, 65):
        for h in list(range(0, 300)) + list(range(u64 - 300, u64 + 1)):
            out.append((h, d))
    return out


def test_bucket_of_matches_modulo():
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "h.cpp")
        exe = os.path.join(tmp, "h")
        with open(src, "w") as f:
            f.write(HARNESS)
        subprocess.run(["g++", "-O2", "-std=c++17", "-I" + CSRC, src, "-o", exe], check=True)
        data = "\n".join("%d %d" % c for c in _cases())
        r = subprocess.run([exe], input=data, capture_output=True, text=True)
        assert r.returncode == 0, r.stdout
        assert "bad 0" in r.stdout
