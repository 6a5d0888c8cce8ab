"""CPU oracle of the Wukong query path (ctypes over wk_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never as part of the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
_LIB = None
U32P = C.POINTER(C.c_uint32)
I32P = C.POINTER(C.c_int32)
U64P = C.POINTER(C.c_uint64)

DIR_CODES = {"<": 0, ">": 1, "<<": 2, ">>": 3}


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            from . import build as _b
            _b.build()
        L = C.CDLL(LIB_PATH)
        L.orc_hash_u64.restype = C.c_uint64
        L.orc_hash_u64.argtypes = [C.c_uint64]
        L.orc_inverse_hash_u64.restype = C.c_uint64
        L.orc_inverse_hash_u64.argtypes = [C.c_uint64]
        L.orc_store_build.restype = C.c_int
        L.orc_store_build.argtypes = [U32P, C.c_uint64, U32P, C.c_uint64, C.POINTER(C.c_void_p)]
        L.orc_store_free.argtypes = [C.c_void_p]
        L.orc_store_num_out.restype = C.c_uint64
        L.orc_store_num_out.argtypes = [C.c_void_p]
        L.orc_store_num_in.restype = C.c_uint64
        L.orc_store_num_in.argtypes = [C.c_void_p]
        L.orc_get_triples.restype = U32P
        L.orc_get_triples.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, U64P]
        L.orc_get_index.restype = U32P
        L.orc_get_index.argtypes = [C.c_void_p, C.c_uint32, C.c_int, U64P]
        L.orc_run_query.restype = C.c_int
        L.orc_run_query.argtypes = [C.c_void_p, I32P, C.c_uint32, C.c_int32, C.POINTER(U32P), U64P,
                                    C.POINTER(C.c_uint32), I32P]
        L.orc_run_query_from.restype = C.c_int
        L.orc_run_query_from.argtypes = [C.c_void_p, I32P, C.c_uint32, C.c_int32, C.c_uint32, U32P, C.c_uint64,
                                         C.c_uint32, I32P, C.POINTER(U32P), U64P, C.POINTER(C.c_uint32), I32P]
        L.orc_run_query_mt.restype = C.c_int
        L.orc_run_query_mt.argtypes = [C.c_void_p, I32P, C.c_uint32, C.c_int32, C.c_int, C.POINTER(U32P), U64P,
                                       C.POINTER(C.c_uint32), I32P]
        L.orc_run_stream.restype = C.c_int
        L.orc_run_stream.argtypes = [C.c_void_p, I32P, U64P, I32P, C.c_uint64, C.c_int, U64P, U64P, I32P]
        L.orc_table_checksum.restype = C.c_uint64
        L.orc_table_checksum.argtypes = [U32P, C.c_uint64, C.c_uint32]
        L.orc_set_direction.argtypes = [I32P, I32P, I32P, C.c_uint32, I32P]
        L.orc_final_process.restype = C.c_int
        L.orc_final_process.argtypes = [U32P, C.c_uint64, C.c_uint32, I32P, C.c_int32, I32P, C.c_uint32, C.c_int,
                                        I32P, I32P, C.c_uint32, C.c_int64, C.c_int64, C.POINTER(U32P), U64P]
        L.orc_free.argtypes = [C.c_void_p]
        _LIB = L
    return _LIB


def hash_u64(x):
    return int(lib().orc_hash_u64(int(x)))


def inverse_hash_u64(x):
    return int(lib().orc_inverse_hash_u64(int(x)))


def set_direction(raw_patterns, plan):
    """raw_patterns: [(s,p,o,dir)], plan: [(order, '<'|'>'|'<<'|'>>')] -> rewritten patterns."""
    raw = np.asarray(raw_patterns, dtype=np.int32).reshape(-1)
    orders = np.asarray([o for o, _ in plan], dtype=np.int32)
    dirs = np.asarray([DIR_CODES[d] for _, d in plan], dtype=np.int32)
    out = np.zeros(4 * len(plan), np.int32)
    lib().orc_set_direction(raw.ctypes.data_as(I32P), orders.ctypes.data_as(I32P), dirs.ctypes.data_as(I32P),
                            len(plan), out.ctypes.data_as(I32P))
    return [tuple(int(v) for v in out[4 * i:4 * i + 4]) for i in range(len(plan))]


class Store:
    def __init__(self, pso, pos=None):
        a = np.ascontiguousarray(np.asarray(pso).reshape(-1), dtype=np.uint32)
        b = a if pos is None else np.ascontiguousarray(np.asarray(pos).reshape(-1), dtype=np.uint32)
        self._keep = (a, b)
        self._h = C.c_void_p()
        rc = lib().orc_store_build(a.ctypes.data_as(U32P), a.size // 3, b.ctypes.data_as(U32P), b.size // 3,
                                   C.byref(self._h))
        if rc:
            raise ValueError("oracle store build failed with status %d" % rc)
        self._keep = None

    def __del__(self):
        try:
            if self._h:
                lib().orc_store_free(self._h)
        except Exception:
            pass

    def num_out(self):
        return int(lib().orc_store_num_out(self._h))

    def get_triples(self, vid, pid, d):
        n = C.c_uint64()
        p = lib().orc_get_triples(self._h, vid, pid, d, C.byref(n))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint32)

    def get_index(self, tpid, d):
        n = C.c_uint64()
        p = lib().orc_get_index(self._h, tpid, d, C.byref(n))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint32)

    def query(self, patterns, nvars):
        """-> (table (rows, cols) uint32, v2c list, status)."""
        pats = np.asarray(patterns, dtype=np.int32).reshape(-1)
        tp, rows, cols = U32P(), C.c_uint64(), C.c_uint32()
        v2c = np.zeros(max(1, nvars), np.int32)
        rc = lib().orc_run_query(self._h, pats.ctypes.data_as(I32P), len(patterns), nvars, C.byref(tp),
                                 C.byref(rows), C.byref(cols), v2c.ctypes.data_as(I32P))
        return self._take(rc, tp, rows, cols, v2c, nvars)

    def query_from(self, patterns, nvars, step, table, v2c_in):
        """Patterns [step, n) over a carried table with var->column map v2c_in."""
        pats = np.asarray(patterns, dtype=np.int32).reshape(-1)
        t = np.ascontiguousarray(table, dtype=np.uint32)
        vin = np.full(max(1, nvars), -1, np.int32)
        vin[:len(v2c_in)] = v2c_in
        tp, rows, cols = U32P(), C.c_uint64(), C.c_uint32()
        v2c = np.zeros(max(1, nvars), np.int32)
        rc = lib().orc_run_query_from(self._h, pats.ctypes.data_as(I32P), len(patterns), nvars, step,
                                      t.ctypes.data_as(U32P), t.shape[0], t.shape[1], vin.ctypes.data_as(I32P),
                                      C.byref(tp), C.byref(rows), C.byref(cols), v2c.ctypes.data_as(I32P))
        return self._take(rc, tp, rows, cols, v2c, nvars)

    def query_mt(self, patterns, nvars, nthreads):
        """A query from an index split over nthreads threads (mt_factor)."""
        pats = np.asarray(patterns, dtype=np.int32).reshape(-1)
        tp, rows, cols = U32P(), C.c_uint64(), C.c_uint32()
        v2c = np.zeros(max(1, nvars), np.int32)
        rc = lib().orc_run_query_mt(self._h, pats.ctypes.data_as(I32P), len(patterns), nvars, nthreads,
                                    C.byref(tp), C.byref(rows), C.byref(cols), v2c.ctypes.data_as(I32P))
        return self._take(rc, tp, rows, cols, v2c, nvars)

    def run_stream(self, queries, nthreads):
        """[(patterns, nvars)] on nthreads workers sharing one queue ->
        (rows, checksums, status) arrays."""
        nq = len(queries)
        off = np.zeros(nq + 1, np.uint64)
        off[1:] = np.cumsum([len(p) for p, _ in queries])
        pats = np.asarray([x for p, _ in queries for pat in p for x in pat] or [0], dtype=np.int32)
        nv = np.asarray([n for _, n in queries] or [0], dtype=np.int32)
        rows = np.zeros(max(1, nq), np.uint64)
        ck = np.zeros(max(1, nq), np.uint64)
        st = np.zeros(max(1, nq), np.int32)
        lib().orc_run_stream(self._h, pats.ctypes.data_as(I32P), off.ctypes.data_as(U64P), nv.ctypes.data_as(I32P),
                             nq, nthreads, rows.ctypes.data_as(U64P), ck.ctypes.data_as(U64P),
                             st.ctypes.data_as(I32P))
        return rows[:nq], ck[:nq], st[:nq]

    def _take(self, rc, tp, rows, cols, v2c, nvars):
        if rc:
            return np.zeros((0, 0), np.uint32), [], rc
        n = rows.value * cols.value
        t = np.ctypeslib.as_array(tp, shape=(n,)).copy().reshape(rows.value, cols.value) if n else \
            np.zeros((rows.value, cols.value), np.uint32)
        if tp:
            lib().orc_free(C.cast(tp, C.c_void_p))
        return t, [int(x) for x in v2c[:nvars]], 0


def table_checksum(table):
    """Order-independent fingerprint of a table (sum of per-row hashes)."""
    t = np.ascontiguousarray(table, dtype=np.uint32)
    if t.ndim == 1:
        t = t.reshape(-1, 1)
    return int(lib().orc_table_checksum(t.ctypes.data_as(U32P), t.shape[0], t.shape[1]))


def final_process(table, v2c, required, distinct=False, orders=(), offset=0, limit=-1):
    """orders: [(var, descending)] -> projected table."""
    t = np.ascontiguousarray(table, dtype=np.uint32)
    rows, cols = t.shape
    v = np.asarray(v2c, dtype=np.int32)
    req = np.asarray(required, dtype=np.int32)
    ov = np.asarray([o for o, _ in orders] or [0], dtype=np.int32)
    od = np.asarray([1 if d else 0 for _, d in orders] or [0], dtype=np.int32)
    op, n = U32P(), C.c_uint64()
    rc = lib().orc_final_process(t.ctypes.data_as(U32P), rows, cols, v.ctypes.data_as(I32P), len(v2c),
                                 req.ctypes.data_as(I32P), len(required), 1 if distinct else 0,
                                 ov.ctypes.data_as(I32P), od.ctypes.data_as(I32P), len(orders), offset, limit,
                                 C.byref(op), C.byref(n))
    if rc:
        raise ValueError("final_process status %d" % rc)
    if rows == 0:
        return t
    out = np.ctypeslib.as_array(op, shape=(n.value * len(required),)).copy().reshape(n.value, len(required)) \
        if n.value else np.zeros((0, len(required)), np.uint32)
    lib().orc_free(C.cast(op, C.c_void_p))
    return out


def partition_rows(table, col, nparts):
    """Split of a history table among nparts owners, as
    SPARQLEngine::generate_sub_query does it (core/engine/sparql.hpp:741-793):
    each row, in order, is appended to owner PARTITION(row[col]) =
    row[col] % nparts (core/store/vertex.hpp:59, utils/math.hpp:52-56).
    Returns the owners' tables concatenated in owner order, with the sizes."""
    t = np.ascontiguousarray(table, dtype=np.uint32)
    owner = t[:, col].astype(np.uint64) % nparts if t.shape[0] else np.zeros(0, np.uint64)
    order = np.argsort(owner, kind="stable")
    sizes = np.bincount(owner.astype(np.int64), minlength=nparts).astype(np.uint64)
    return t[order], sizes


def sorted_rows(t):
    """Canonical form for multiset comparison: rows sorted lexicographically."""
    t = np.asarray(t, dtype=np.uint32)
    if t.size == 0:
        return t.reshape(-1, t.shape[1] if t.ndim == 2 else 0)
    return t[np.lexsort(t.T[::-1])]
