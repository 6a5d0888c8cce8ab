"""wk_table_partition_noalloc through the C ABI: the split of caller-owned
device buffers with the histogram in the library's fixed device scratch
equals wk_table_partition's (and the oracle's) row for row, on two calls in a
row, the second smaller, so the second reads nothing the first left behind."""
import ctypes as C

import numpy as np
import pytest

import oracle
import wukong_gpu as wg

pytestmark = pytest.mark.gpu

HIP = C.CDLL("/opt/rocm/lib/libamdhip64.so")
H2D, D2H = 1, 2


def _alloc(nbytes):
    p = C.c_void_p()
    assert HIP.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 4))) == 0
    return p


def _split(dev, d_in, d_out, t, col, nparts):
    sizes = np.zeros(nparts, np.uint64)
    heads = np.zeros(nparts, np.uint64)
    rows, cols = t.shape
    if t.size:
        assert HIP.hipMemcpy(d_in, C.c_void_p(t.ctypes.data), C.c_size_t(t.nbytes), H2D) == 0
    rc = wg.lib().wk_table_partition_noalloc(dev._h, d_in, rows, cols, col, nparts, d_out,
                                             sizes.ctypes.data_as(wg.U64P), heads.ctypes.data_as(wg.U64P))
    got = np.zeros_like(t)
    if rc == 0 and t.size:
        assert HIP.hipMemcpy(C.c_void_p(got.ctypes.data), d_out, C.c_size_t(t.nbytes), D2H) == 0
    return rc, got, sizes, heads


def test_partition_noalloc_matches_partition(univ10_dev):
    rng = np.random.default_rng(77)
    cases = [(300_000, 3, 1, 7), (4_100, 2, 0, 64)]   # the second call is smaller
    cap = max(r * c for r, c, _, _ in cases) * 4
    d_in, d_out = _alloc(cap), _alloc(cap)
    try:
        for rows, cols, col, nparts in cases:
            t = rng.integers(0, 1 << 30, size=(rows, cols), dtype=np.uint32)
            rc, got, sizes, heads = _split(univ10_dev, d_in, d_out, t, col, nparts)
            assert rc == 0, wg.STATUS.get(rc, rc)
            want, wsizes = oracle.partition_rows(t, col, nparts)
            assert np.array_equal(sizes, wsizes)
            assert np.array_equal(heads, np.concatenate([[0], np.cumsum(wsizes)[:-1]]).astype(np.uint64))
            assert np.array_equal(got, want)
            tab = univ10_dev.upload(t)
            ref, rsizes, rheads = univ10_dev.partition(tab, col, nparts)
            assert np.array_equal(got, ref.numpy())
            assert np.array_equal(sizes, rsizes) and np.array_equal(heads, rheads)
            ref.free()
            tab.free()
        # an empty table, bad arguments and a histogram beyond the fixed scratch
        rc, _, sizes, heads = _split(univ10_dev, d_in, d_out, np.zeros((0, 2), np.uint32), 0, 5)
        assert rc == 0 and not sizes.any() and not heads.any()
        L = wg.lib()
        s = np.zeros(64, np.uint64)
        h = np.zeros(64, np.uint64)
        sp, hp = s.ctypes.data_as(wg.U64P), h.ctypes.data_as(wg.U64P)
        assert L.wk_table_partition_noalloc(univ10_dev._h, d_in, 10, 2, 2, 4, d_out, sp, hp) == wg.WK_E_INVALID_ARG
        assert L.wk_table_partition_noalloc(univ10_dev._h, d_in, 10, 2, 0, 65, d_out, sp, hp) == wg.WK_E_INVALID_ARG
        assert L.wk_table_partition_noalloc(univ10_dev._h, d_in, 1 << 31, 2, 0, 64, d_out, sp,
                                            hp) == wg.WK_E_SCRATCH_LIMIT
    finally:
        HIP.hipFree(d_in)
        HIP.hipFree(d_out)
