"""Python binding of the MI355X Wukong GPU query path (ctypes over include/wk_gpu.h).

The host engine, store builder and kernels are native (libwukong_gpu.so);
this module only marshals arrays.  Loading fails loudly when the library has
not been built: there is no Python fallback.

Mirrors the reference's GPU engine flow: a `Store` is the host gstore
(SegmentRDFGraph), a `Device` owns the HBM gstore cache and history tables
(GPUMem + GPUCache + GPUEngine), `Device.query` runs a planned pattern list
as one query (GPUAgent::execute_sparql_query).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwukong_gpu.so")

IN, OUT = 0, 1
PREDICATE_ID, TYPE_ID = 0, 1
MAX_VARS = 64

STATUS = {
    0: "WK_OK", 1: "WK_E_UNKNOWN", 2: "WK_E_SYNTAX", 3: "WK_E_UNKNOWN_PATTERN",
    4: "WK_E_ATTR_DISABLE", 5: "WK_E_NO_REQUIRED_VAR", 6: "WK_E_UNSUPPORT_UNION",
    7: "WK_E_OBJ_ERROR", 8: "WK_E_VERTEX_INVALID", 9: "WK_E_UNKNOWN_SUB",
    10: "WK_E_SETTING", 11: "WK_E_FIRST_PATTERN", 12: "WK_E_UNKNOWN_FILTER",
    13: "WK_E_FILE_NOT_FOUND", 100: "WK_E_INVALID_ARG", 101: "WK_E_INVALID_TRIPLE",
    102: "WK_E_NOMEM_HOST", 103: "WK_E_NOMEM_BUDGET", 104: "WK_E_HIP",
    105: "WK_E_NO_DEVICE", 106: "WK_E_OVERFLOW", 107: "WK_E_SEG_TOO_LARGE",
    108: "WK_E_STATE", 109: "WK_E_TOO_MANY_VARS", 110: "WK_E_WIRE", 111: "WK_E_SCRATCH_LIMIT",
}
globals().update({v: k for k, v in STATUS.items()})


class WkError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = "%s (%d): %s" % (STATUS.get(code, "?"), code, _strerror(code))
        if code == 104:
            msg += " [hip error %d]" % lib().wk_last_hip_error()
        super().__init__(msg + (" in " + what if what else ""))


class Pattern(C.Structure):
    _fields_ = [("subject", C.c_int32), ("predicate", C.c_int32), ("object", C.c_int32),
                ("direction", C.c_int32)]


class Query(C.Structure):
    _fields_ = [("patterns", C.POINTER(Pattern)), ("npatterns", C.c_uint32), ("nvars", C.c_int32)]


class StoreOpts(C.Structure):
    _fields_ = [("pinned", C.c_int32), ("load_factor_pct", C.c_int32)]


class StoreInfo(C.Structure):
    _fields_ = [("num_triples_out", C.c_uint64), ("num_triples_in", C.c_uint64),
                ("num_keys", C.c_uint64), ("num_buckets", C.c_uint64), ("num_values", C.c_uint64),
                ("num_segments", C.c_uint32), ("pinned", C.c_int32)]


class SegInfo(C.Structure):
    _fields_ = [("index", C.c_uint32), ("pid", C.c_uint32), ("dir", C.c_uint32), ("pad", C.c_uint32),
                ("num_keys", C.c_uint64), ("num_buckets", C.c_uint64), ("num_ext_buckets", C.c_uint64),
                ("bucket_start", C.c_uint64), ("num_edges", C.c_uint64), ("edge_start", C.c_uint64)]


class DevOpts(C.Structure):
    _fields_ = [("device", C.c_int32), ("gstore_budget", C.c_uint64), ("table_budget", C.c_uint64),
                ("key_block_shift", C.c_uint32), ("value_block_shift", C.c_uint32),
                ("enable_pipeline", C.c_int32), ("probe_width", C.c_int32),
                ("rows_per_wave", C.c_int32), ("profile", C.c_int32), ("dense_index", C.c_int32),
                ("batch_lanes", C.c_int32)]


class DevStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "h2d_bytes", "segments_loaded", "segments_evicted", "key_blocks_total", "key_blocks_free",
        "value_blocks_total", "value_blocks_free", "table_capacity_elems", "queries_run", "dense_segments",
        "dense_bytes", "mapped_segments", "mapped_bytes")]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint64), ("total_ms", C.c_double),
                ("alg_bytes", C.c_double)]


class BatchResult(C.Structure):
    _fields_ = [("rows", C.POINTER(C.c_uint64)), ("cols", C.POINTER(C.c_uint32)),
                ("status", C.POINTER(C.c_int32)), ("offset", C.POINTER(C.c_uint64)),
                ("v2c", C.POINTER(C.c_int32))]


class ServeOpts(C.Structure):
    _fields_ = [("max_inflight", C.c_uint32), ("footprint_budget", C.c_uint64),
                ("per_query", C.c_int32), ("keep_results", C.c_int32), ("heavy_footprint", C.c_uint64)]


class ServeResult(C.Structure):
    _fields_ = [("latency_us", C.POINTER(C.c_double)), ("start_us", C.POINTER(C.c_double)),
                ("rows", C.POINTER(C.c_uint64)), ("status", C.POINTER(C.c_int32)),
                ("checksum", C.POINTER(C.c_uint64)), ("wall_s", C.c_double),
                ("batches", C.c_uint64)]


class WirePattern(C.Structure):
    _fields_ = [("subject", C.c_int32), ("predicate", C.c_int32), ("object", C.c_int32),
                ("direction", C.c_int32), ("pred_type", C.c_int32)]


# scalar members of wk_sparql_msg, in struct order
_MSG_HEAD = ("q_type", "qid", "pqid", "pg_type", "state", "dev_type", "job_type", "priority",
             "mt_factor", "mt_tid", "pattern_step", "local_var", "corun_enabled", "corun_step",
             "fetch_step", "union_done", "optional_step", "limit")
_MSG_RESULT = ("col_num", "row_num", "attr_col_num", "status_code", "blind", "nvars")


class SparqlMsg(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in _MSG_HEAD] + [("offset", C.c_uint32), ("distinct", C.c_int32)] +
                [("patterns", C.POINTER(WirePattern)), ("npatterns", C.c_uint32),
                 ("new_vars", C.POINTER(C.c_int32)), ("nnew_vars", C.c_uint32),
                 ("order_vars", C.POINTER(C.c_int32)), ("order_desc", C.POINTER(C.c_int32)),
                 ("norders", C.c_uint32)] +
                [(n, C.c_int32) for n in _MSG_RESULT] +
                [("required_vars", C.POINTER(C.c_int32)), ("nrequired", C.c_uint32),
                 ("v2c_map", C.POINTER(C.c_int32)), ("nv2c", C.c_uint32),
                 ("optional_matched", C.POINTER(C.c_uint8)), ("noptional_matched", C.c_uint64),
                 ("table", C.POINTER(C.c_uint32)), ("ntable", C.c_uint64),
                 ("gpu_nelems", C.c_uint64), ("gpu_col_num", C.c_int32)])


class WireOpts(C.Structure):
    _fields_ = [("library_version", C.c_uint32), ("use_gpu", C.c_int32)]


_LIB = None
P = C.c_void_p
U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)
I32P = C.POINTER(C.c_int32)

# name -> (restype, argtypes); every symbol declared in include/wk_gpu.h
SIGNATURES = {
    "wk_store_build": (C.c_int, [U32P, C.c_uint64, U32P, C.c_uint64, C.POINTER(StoreOpts), C.POINTER(P)]),
    "wk_store_load_dir": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.POINTER(StoreOpts), C.POINTER(P), U64P]),
    "wk_store_destroy": (None, [P]),
    "wk_store_info_get": (C.c_int, [P, C.POINTER(StoreInfo)]),
    "wk_store_segment": (C.c_int, [P, C.c_uint32, C.POINTER(SegInfo)]),
    "wk_store_raw": (C.c_int, [P, C.POINTER(P), U64P, C.POINTER(U32P), U64P]),
    "wk_store_get_triples": (C.c_int, [P, C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(U32P), U64P]),
    "wk_store_get_index": (C.c_int, [P, C.c_uint32, C.c_int32, C.POINTER(U32P), U64P]),
    "wk_dev_create": (C.c_int, [P, C.POINTER(DevOpts), C.POINTER(P)]),
    "wk_dev_destroy": (None, [P]),
    "wk_dev_stats_get": (C.c_int, [P, C.POINTER(DevStats)]),
    "wk_dev_kernel_stats": (C.c_int, [P, C.POINTER(KernelStat), C.c_uint32, U32P]),
    "wk_dev_reset_kernel_stats": (C.c_int, [P]),
    "wk_dev_set_tuning": (C.c_int, [P, C.c_int32, C.c_int32]),
    "wk_dev_set_profile": (C.c_int, [P, C.c_int32]),
    "wk_table_partition": (C.c_int, [P, P, C.c_uint32, C.c_uint32, C.POINTER(P), U64P, U64P]),
    "wk_table_partition_noalloc": (C.c_int, [P, P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, P, U64P,
                                             U64P]),
    "wk_table_final_process": (C.c_int, [P, P, I32P, C.c_int32, I32P, C.c_uint32, C.c_int32, I32P, I32P,
                                         C.c_uint32, C.c_int64, C.c_int64, C.POINTER(P)]),
    "wk_dev_sync": (C.c_int, [P]),
    "wk_dev_load_all": (C.c_int, [P]),
    "wk_seg_ensure": (C.c_int, [P, C.c_uint32, C.c_uint32, C.c_int32]),
    "wk_seg_resident": (C.c_int, [P, C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(C.c_int32)]),
    "wk_table_upload": (C.c_int, [P, U32P, C.c_uint64, C.c_uint32, C.POINTER(P)]),
    "wk_table_shape": (C.c_int, [P, U64P, U32P]),
    "wk_table_download": (C.c_int, [P, P, U32P, C.c_uint64]),
    "wk_table_free": (C.c_int, [P, P]),
    "wk_start_i2u": (C.c_int, [P, C.c_uint32, C.c_int32, C.POINTER(P)]),
    "wk_start_c2u": (C.c_int, [P, C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(P)]),
    "wk_expand_k2u": (C.c_int, [P, P, C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(P)]),
    "wk_expand_k2k": (C.c_int, [P, P, C.c_uint32, C.c_uint32, C.c_int32, C.c_uint32, C.POINTER(P)]),
    "wk_expand_k2c": (C.c_int, [P, P, C.c_uint32, C.c_uint32, C.c_int32, C.c_uint32, C.POINTER(P)]),
    "wk_expand_c2k": (C.c_int, [P, P, C.c_uint32, C.c_uint32, C.c_int32, C.c_uint32, C.POINTER(P)]),
    "wk_query_run": (C.c_int, [P, C.POINTER(Query), U64P, U32P, C.POINTER(C.c_int32)]),
    "wk_result_download": (C.c_int, [P, U32P, C.c_uint64]),
    "wk_batch_run": (C.c_int, [P, C.POINTER(Query), C.c_uint32, C.c_int32, C.POINTER(BatchResult)]),
    "wk_dev_autotune": (C.c_int, [P, C.POINTER(Query), C.c_uint32, C.c_int32, C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "wk_batch_download": (C.c_int, [P, U32P, C.c_uint64, U64P]),
    "wk_batch_submit": (C.c_int, [P, C.POINTER(Query), C.c_uint32, C.c_int32, C.POINTER(BatchResult),
                                  C.POINTER(C.c_uint32)]),
    "wk_batch_wait": (C.c_int, [P, C.c_uint32]),
    "wk_batch_fetch": (C.c_int, [P, C.c_uint32, U32P, C.c_uint64, U64P]),
    "wk_batch_traffic": (C.c_int, [P, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double), U64P]),
    "wk_serve_replay": (C.c_int, [P, C.POINTER(Query), C.POINTER(C.c_double), C.c_uint32,
                                  C.POINTER(ServeOpts), C.POINTER(ServeResult)]),
    "wk_serve_result_rows": (C.c_int, [P, C.c_uint32, U32P, C.c_uint64, U64P, U32P]),
    "wk_table_checksum": (C.c_uint64, [U32P, C.c_uint64, C.c_uint32]),
    "wk_wire_encode": (C.c_int, [C.POINTER(SparqlMsg), C.POINTER(WireOpts), C.POINTER(C.c_uint8), C.c_uint64,
                                 U64P]),
    "wk_wire_decode": (C.c_int, [C.POINTER(C.c_uint8), C.c_uint64, C.POINTER(WireOpts),
                                 C.POINTER(C.POINTER(SparqlMsg))]),
    "wk_wire_free": (None, [C.POINTER(SparqlMsg)]),
    "wk_agent_execute": (C.c_int, [P, C.POINTER(C.c_uint8), C.c_uint64, C.POINTER(WireOpts), U64P]),
    "wk_agent_reply": (C.c_int, [P, C.POINTER(C.c_uint8), C.c_uint64]),
    "wk_gen_university": (C.c_int, [C.c_uint32, C.c_uint32, C.POINTER(U32P), U64P]),
    "wk_univ_id": (C.c_uint32, [C.c_char_p]),
    "wk_gen_powerlaw": (C.c_int, [C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double,
                                  C.c_double, C.c_uint64, C.POINTER(U32P), U64P]),
    "wk_free": (None, [P]),
    "wk_strerror": (C.c_char_p, [C.c_int]),
    "wk_last_hip_error": (C.c_int, []),
    "wk_device_count": (C.c_int, [C.POINTER(C.c_int)]),
}


def lib():
    """Load libwukong_gpu.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libwukong_gpu.so is not built (run __graft_entry__.build()): " + LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _strerror(code):
    return lib().wk_strerror(code).decode()


def _check(rc, what=""):
    if rc != 0:
        raise WkError(rc, what)


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(U32P)


def device_count():
    n = C.c_int(0)
    lib().wk_device_count(C.byref(n))
    return n.value


# ------------------------------------------------------------------ generators

def univ_id(name):
    return int(lib().wk_univ_id(name.encode()))


def _take_triples(ptr, n):
    if n == 0:
        lib().wk_free(C.cast(ptr, P))
        return np.zeros((0, 3), np.uint32)
    arr = np.ctypeslib.as_array(ptr, shape=(n * 3,)).copy().reshape(n, 3)
    lib().wk_free(C.cast(ptr, P))
    return arr


def gen_university(seed, n_univ):
    """(n,3) uint32 triples of the seeded university-schema generator."""
    ptr, n = U32P(), C.c_uint64(0)
    _check(lib().wk_gen_university(seed, n_univ, C.byref(ptr), C.byref(n)), "wk_gen_university")
    return _take_triples(ptr, n.value)


def gen_powerlaw(seed, n_vertices, n_triples, n_preds=256, n_types=64, zipf_s=1.1, alpha=2.0,
                 max_degree=10**6):
    ptr, n = U32P(), C.c_uint64(0)
    _check(lib().wk_gen_powerlaw(seed, n_vertices, n_triples, n_preds, n_types, zipf_s, alpha, max_degree,
                                 C.byref(ptr), C.byref(n)), "wk_gen_powerlaw")
    return _take_triples(ptr, n.value)


def checksum(table):
    t = np.ascontiguousarray(table, dtype=np.uint32)
    if t.ndim == 1:
        t = t.reshape(-1, 1)
    rows, cols = t.shape
    return int(lib().wk_table_checksum(t.ctypes.data_as(U32P), rows, cols))


# ----------------------------------------------------------------------- store

class Store:
    """Host gstore (SegmentRDFGraph::init_gstore over pso/pos triples)."""

    def __init__(self, pso, pos=None, pinned=False, load_factor_pct=0):
        pso_a, pso_p = _u32(np.asarray(pso).reshape(-1))
        if pos is None:
            pos_a, pos_p = pso_a, pso_p
        else:
            pos_a, pos_p = _u32(np.asarray(pos).reshape(-1))
        self._h = P()
        opts = StoreOpts(1 if pinned else 0, load_factor_pct)
        _check(lib().wk_store_build(pso_p, pso_a.size // 3, pos_p, pos_a.size // 3, C.byref(opts),
                                    C.byref(self._h)), "wk_store_build")

    @classmethod
    def from_dir(cls, path, server=0, nservers=1, pinned=False, load_factor_pct=0):
        """Store of server `server` of `nservers` from a directory of Wukong
        ID-triple files (id_*).  Returns (store, triples read from the files)."""
        st = cls.__new__(cls)
        st._h = P()
        n = C.c_uint64()
        opts = StoreOpts(1 if pinned else 0, load_factor_pct)
        _check(lib().wk_store_load_dir(os.fsencode(path), server, nservers, C.byref(opts), C.byref(st._h),
                                       C.byref(n)), "wk_store_load_dir")
        return st, n.value

    def close(self):
        if self._h:
            lib().wk_store_destroy(self._h)
            self._h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self):
        i = StoreInfo()
        _check(lib().wk_store_info_get(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in StoreInfo._fields_}

    def segments(self):
        out = []
        for k in range(self.info()["num_segments"]):
            s = SegInfo()
            _check(lib().wk_store_segment(self._h, k, C.byref(s)))
            out.append({f: getattr(s, f) for f, _ in SegInfo._fields_ if f != "pad"})
        return out

    def raw(self):
        """(slots as (buckets*8, 2) uint64 view, values uint32 view) of the host arrays."""
        sp, nb, vp, nv = P(), C.c_uint64(), U32P(), C.c_uint64()
        _check(lib().wk_store_raw(self._h, C.byref(sp), C.byref(nb), C.byref(vp), C.byref(nv)))
        slots = np.ctypeslib.as_array(C.cast(sp, U64P), shape=(nb.value * 8 * 2,)).reshape(-1, 2)
        vals = np.ctypeslib.as_array(vp, shape=(nv.value,)) if nv.value else np.zeros(0, np.uint32)
        return slots, vals

    def get_triples(self, vid, pid, d):
        p, n = U32P(), C.c_uint64()
        _check(lib().wk_store_get_triples(self._h, vid, pid, d, C.byref(p), C.byref(n)))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint32)

    def get_index(self, tpid, d):
        p, n = U32P(), C.c_uint64()
        _check(lib().wk_store_get_index(self._h, tpid, d, C.byref(p), C.byref(n)))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint32)


def write_id_triples(path, triples, nfiles=1, prefix="id_"):
    """Write (s, p, o) triples as Wukong ID-triple text files
    (datagen/README.md): one "s p o" line each, split over nfiles files
    named <prefix>0.nt, <prefix>1.nt, ..."""
    t = np.ascontiguousarray(triples, dtype=np.uint32).reshape(-1, 3)
    os.makedirs(path, exist_ok=True)
    for k, part in enumerate(np.array_split(t, max(1, nfiles))):
        with open(os.path.join(path, "%s%d.nt" % (prefix, k)), "w") as f:
            if len(part):
                f.write("\n".join("%d\t%d\t%d" % tuple(r) for r in part.tolist()))
                f.write("\n")


def load_str_index(path):
    """Name -> id of a dataset's str_index and str_normal files, read as
    StringServer::load_normal_or_index_file does (whitespace-separated
    "string id" pairs up to the first that fails; stringserver/string_server.hpp:214-225)."""
    names = {}
    for fname in ("str_index", "str_normal"):
        fp = os.path.join(path, fname)
        if not os.path.exists(fp):
            continue
        with open(fp) as f:
            tok = f.read().split()
        for i in range(0, len(tok) - 1, 2):
            if not tok[i + 1].isdigit() or int(tok[i + 1]) >= 1 << 32:
                break
            names[tok[i]] = int(tok[i + 1])
    return names


# ----------------------------------------------------------------- wire form

def query_msg(patterns, nvars, required, dev_type=1, **kw):
    """A SPARQLQuery request as the proxy sends it, as a dict: the members'
    default values (core/sparql/query.hpp:710-750, 370-397), BASIC group,
    `patterns` as (s, p, o, d) in plan order, `required` variable ids.
    Keyword arguments override members (limit, offset, distinct, orders,
    blind, ...)."""
    m = {n: 0 for n in _MSG_HEAD + _MSG_RESULT}
    m.update(qid=-1, pqid=-1, mt_factor=1, limit=-1, offset=0, distinct=0, dev_type=dev_type,
             nvars=nvars, patterns=[tuple(p) + (0,) * (5 - len(p)) for p in patterns], new_vars=[], orders=[],
             required_vars=list(required), v2c_map=[], optional_matched=[],
             table=np.zeros(0, np.uint32), gpu_nelems=0, gpu_col_num=0)
    m.update(kw)
    return m


def _wire_opts(use_gpu, library_version):
    return WireOpts(library_version, 1 if use_gpu else 0)


def wire_encode(m, use_gpu=False, library_version=0):
    """Bundle string (type word + boost binary archive) of a message dict."""
    pats = (WirePattern * max(1, len(m["patterns"])))(*[WirePattern(*p) for p in m["patterns"]])
    arr = lambda xs: (C.c_int32 * max(1, len(xs)))(*xs)
    nv, ov = arr(m["new_vars"]), arr([o for o, _ in m["orders"]])
    od = arr([1 if d else 0 for _, d in m["orders"]])
    req, v2c = arr(m["required_vars"]), arr(m["v2c_map"])
    om = (C.c_uint8 * max(1, len(m["optional_matched"])))(*[1 if b else 0 for b in m["optional_matched"]])
    tab = np.ascontiguousarray(m["table"], dtype=np.uint32).reshape(-1)
    msg = SparqlMsg(*[int(m[n]) for n in _MSG_HEAD], int(m["offset"]), int(m["distinct"]),
                    pats, len(m["patterns"]), nv, len(m["new_vars"]), ov, od, len(m["orders"]),
                    *[int(m[n]) for n in _MSG_RESULT], req, len(m["required_vars"]), v2c,
                    len(m["v2c_map"]), om, len(m["optional_matched"]), tab.ctypes.data_as(U32P),
                    tab.size, int(m["gpu_nelems"]), int(m["gpu_col_num"]))
    o = _wire_opts(use_gpu, library_version)
    n = C.c_uint64()
    lib().wk_wire_encode(C.byref(msg), C.byref(o), None, 0, C.byref(n))
    buf = (C.c_uint8 * max(1, n.value))()
    _check(lib().wk_wire_encode(C.byref(msg), C.byref(o), buf, n.value, C.byref(n)), "wk_wire_encode")
    return bytes(buf[:n.value])


def wire_decode(data, use_gpu=False, library_version=0):
    """Message dict of a Bundle string; raises WkError (WK_E_WIRE, ...)."""
    src = (C.c_uint8 * max(1, len(data))).from_buffer_copy(data if data else b"\0")
    o = _wire_opts(use_gpu, library_version)
    h = C.POINTER(SparqlMsg)()
    _check(lib().wk_wire_decode(src, len(data), C.byref(o), C.byref(h)), "wk_wire_decode")
    try:
        s = h.contents
        m = {n: getattr(s, n) for n in _MSG_HEAD + _MSG_RESULT + ("offset", "distinct")}
        m["patterns"] = [(p.subject, p.predicate, p.object, p.direction, p.pred_type)
                         for p in s.patterns[:s.npatterns]]
        m["new_vars"] = list(s.new_vars[:s.nnew_vars])
        m["orders"] = list(zip(s.order_vars[:s.norders], [bool(d) for d in s.order_desc[:s.norders]]))
        m["required_vars"] = list(s.required_vars[:s.nrequired])
        m["v2c_map"] = list(s.v2c_map[:s.nv2c])
        m["optional_matched"] = [bool(b) for b in s.optional_matched[:s.noptional_matched]]
        m["table"] = (np.ctypeslib.as_array(s.table, (s.ntable,)).copy() if s.ntable
                      else np.zeros(0, np.uint32))
        m["gpu_nelems"], m["gpu_col_num"] = s.gpu_nelems, s.gpu_col_num
        return m
    finally:
        lib().wk_wire_free(h)


def make_query(patterns, nvars):
    """Pack [(s, p, o, dir), ...] into a wk_query (keeps the array alive on the object)."""
    arr = (Pattern * max(1, len(patterns)))()
    for i, (s, p, o, d) in enumerate(patterns):
        arr[i] = Pattern(s, p, o, d)
    q = Query(C.cast(arr, C.POINTER(Pattern)), len(patterns), nvars)
    q._keep = arr
    return q


# ---------------------------------------------------------------------- device

class Table:
    def __init__(self, dev, h):
        self.dev, self._h = dev, h

    @property
    def shape(self):
        r, c = C.c_uint64(), C.c_uint32()
        _check(lib().wk_table_shape(self._h, C.byref(r), C.byref(c)))
        return r.value, c.value

    def numpy(self):
        r, c = self.shape
        out = np.zeros((r, c), np.uint32)
        _check(lib().wk_table_download(self.dev._h, self._h, out.ctypes.data_as(U32P), out.size),
               "wk_table_download")
        return out

    def free(self):
        if self._h:
            lib().wk_table_free(self.dev._h, self._h)
            self._h = P()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PreparedBatch:
    """A batch of queries packed once into wk_query structs, with its result
    arrays: repeated Device.batch calls then cost only the C call."""

    def __init__(self, queries):
        self.n = len(queries)
        self.qs = (Query * max(1, self.n))()
        self._keep = []
        self.nvars = []
        for i, (pats, nv) in enumerate(queries):
            q = make_query(pats, nv)
            self._keep.append(q)
            self.qs[i] = q
            self.nvars.append(nv)
        self.rows = np.zeros(self.n, np.uint64)
        self.cols = np.zeros(self.n, np.uint32)
        self.status = np.zeros(self.n, np.int32)
        self.offs = np.zeros(self.n, np.uint64)
        self.v2c = np.zeros(self.n * MAX_VARS, np.int32)
        self.res = BatchResult(self.rows.ctypes.data_as(U64P), self.cols.ctypes.data_as(U32P),
                               self.status.ctypes.data_as(C.POINTER(C.c_int32)), self.offs.ctypes.data_as(U64P),
                               self.v2c.ctypes.data_as(C.POINTER(C.c_int32)))
        # the same arrays without the var->column maps (not filled by the library)
        self.res_rows_only = BatchResult(self.rows.ctypes.data_as(U64P), self.cols.ctypes.data_as(U32P),
                                         self.status.ctypes.data_as(C.POINTER(C.c_int32)),
                                         self.offs.ctypes.data_as(U64P), None)


class Device:
    """HBM gstore cache + history tables + expansion kernels on one GPU."""

    def __init__(self, store, device=0, gstore_budget=0, table_budget=0, key_block_shift=0,
                 value_block_shift=0, enable_pipeline=1, probe_width=0, rows_per_wave=0, profile=0,
                 dense_index=0, batch_lanes=0):
        self.store = store
        self._h = P()
        o = DevOpts(device, gstore_budget, table_budget, key_block_shift, value_block_shift,
                    enable_pipeline, probe_width, rows_per_wave, profile, dense_index, batch_lanes)
        _check(lib().wk_dev_create(store.handle, C.byref(o), C.byref(self._h)), "wk_dev_create")

    def close(self):
        if self._h:
            lib().wk_dev_destroy(self._h)
            self._h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(lib().wk_dev_sync(self._h))

    def load_all(self):
        _check(lib().wk_dev_load_all(self._h), "wk_dev_load_all")

    def set_tuning(self, probe_width, rows_per_wave):
        _check(lib().wk_dev_set_tuning(self._h, probe_width, rows_per_wave))

    def partition(self, table, col, nparts):
        """Rows of a device Table grouped by owner row[col] % nparts: (Table, sizes, heads)."""
        h = P()
        sizes = np.zeros(nparts, np.uint64)
        heads = np.zeros(nparts, np.uint64)
        _check(lib().wk_table_partition(self._h, table._h, col, nparts, C.byref(h),
                                        sizes.ctypes.data_as(U64P), heads.ctypes.data_as(U64P)),
               "wk_table_partition")
        return Table(self, h), sizes, heads

    def final_process(self, table, v2c, required, distinct=False, orders=(), offset=0, limit=-1):
        """SPARQLEngine::final_process of a device Table; orders: [(var, descending)].
        Returns a new device Table of the required variables' columns."""
        v = (C.c_int32 * max(1, len(v2c)))(*v2c)
        req = (C.c_int32 * max(1, len(required)))(*required)
        ov = (C.c_int32 * max(1, len(orders)))(*[o for o, _ in orders])
        od = (C.c_int32 * max(1, len(orders)))(*[1 if d else 0 for _, d in orders])
        h = P()
        _check(lib().wk_table_final_process(self._h, table._h, v, len(v2c), req, len(required),
                                            1 if distinct else 0, ov, od, len(orders), offset, limit,
                                            C.byref(h)), "wk_table_final_process")
        return Table(self, h)

    def set_profile(self, on):
        _check(lib().wk_dev_set_profile(self._h, 1 if on else 0))

    def stats(self):
        s = DevStats()
        _check(lib().wk_dev_stats_get(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in DevStats._fields_}

    def kernel_stats(self):
        n = C.c_uint32()
        _check(lib().wk_dev_kernel_stats(self._h, None, 0, C.byref(n)))
        arr = (KernelStat * max(1, n.value))()
        _check(lib().wk_dev_kernel_stats(self._h, arr, n.value, C.byref(n)))
        return {arr[i].name.decode(): {"launches": arr[i].launches, "total_ms": arr[i].total_ms,
                                       "alg_bytes": arr[i].alg_bytes} for i in range(n.value)}

    def reset_kernel_stats(self):
        _check(lib().wk_dev_reset_kernel_stats(self._h))

    def seg_ensure(self, index, pid, d):
        _check(lib().wk_seg_ensure(self._h, index, pid, d), "wk_seg_ensure")

    def seg_resident(self, index, pid, d):
        r = C.c_int32()
        _check(lib().wk_seg_resident(self._h, index, pid, d, C.byref(r)))
        return bool(r.value)

    # per-pattern API
    def upload(self, table):
        t = np.ascontiguousarray(table, dtype=np.uint32)
        if t.ndim == 1:
            t = t.reshape(-1, 1)
        h = P()
        _check(lib().wk_table_upload(self._h, t.ctypes.data_as(U32P), t.shape[0], t.shape[1], C.byref(h)))
        return Table(self, h)

    def i2u(self, tpid, d):
        h = P()
        _check(lib().wk_start_i2u(self._h, tpid, d, C.byref(h)), "wk_start_i2u")
        return Table(self, h)

    def c2u(self, vid, pid, d):
        h = P()
        _check(lib().wk_start_c2u(self._h, vid, pid, d, C.byref(h)), "wk_start_c2u")
        return Table(self, h)

    def k2u(self, t, col, pid, d):
        h = P()
        _check(lib().wk_expand_k2u(self._h, t._h, col, pid, d, C.byref(h)), "wk_expand_k2u")
        return Table(self, h)

    def k2k(self, t, col, pid, d, end_col):
        h = P()
        _check(lib().wk_expand_k2k(self._h, t._h, col, pid, d, end_col, C.byref(h)), "wk_expand_k2k")
        return Table(self, h)

    def k2c(self, t, col, pid, d, end):
        h = P()
        _check(lib().wk_expand_k2c(self._h, t._h, col, pid, d, end, C.byref(h)), "wk_expand_k2c")
        return Table(self, h)

    def c2k(self, t, vid, pid, d, end_col):
        h = P()
        _check(lib().wk_expand_c2k(self._h, t._h, vid, pid, d, end_col, C.byref(h)), "wk_expand_c2k")
        return Table(self, h)

    # whole queries
    def query(self, patterns, nvars, fetch=True):
        """Per-query execution; returns (table (rows, cols) uint32 or None, v2c list, status)."""
        q = make_query(patterns, nvars)
        rows, cols = C.c_uint64(), C.c_uint32()
        v2c = (C.c_int32 * max(1, nvars))()
        rc = lib().wk_query_run(self._h, C.byref(q), C.byref(rows), C.byref(cols), v2c)
        if rc not in (0,) and rc < 100:
            return np.zeros((0, 0), np.uint32), [], rc
        _check(rc, "wk_query_run")
        out = None
        if fetch:
            out = np.zeros((rows.value, cols.value), np.uint32)
            _check(lib().wk_result_download(self._h, out.ctypes.data_as(U32P), out.size))
        return out, list(v2c[:nvars]), 0

    def batch(self, queries, use_graph=True, fetch=True):
        """queries: list of (patterns, nvars) or a PreparedBatch.  Returns list of
        (table, v2c, status) or, without fetch, (rows, cols, status) arrays."""
        pb = queries if isinstance(queries, PreparedBatch) else PreparedBatch(queries)
        rows, cols, status, offs, v2c = pb.rows, pb.cols, pb.status, pb.offs, pb.v2c
        res = pb.res if fetch else pb.res_rows_only
        _check(lib().wk_batch_run(self._h, pb.qs, pb.n, 1 if use_graph else 0, C.byref(res)), "wk_batch_run")
        if not fetch:
            return rows.copy(), cols.copy(), status.copy()
        used = C.c_uint64()
        _check(lib().wk_batch_download(self._h, None, 0, C.byref(used)))
        arena = np.zeros(max(1, used.value), np.uint32)
        _check(lib().wk_batch_download(self._h, arena.ctypes.data_as(U32P), arena.size, C.byref(used)))
        return self._tables(pb, arena)

    def batch_submit(self, pb, use_graph=True, v2c=False):
        """Pipelined batch: launches the PreparedBatch `pb` and returns its
        ticket without waiting for the GPU (two may be in flight).  pb's result
        arrays are filled by batch_wait (or by any later call that completes
        it); pb must not be submitted again before then."""
        t = C.c_uint32()
        res = pb.res if v2c else pb.res_rows_only
        _check(lib().wk_batch_submit(self._h, pb.qs, pb.n, 1 if use_graph else 0, C.byref(res), C.byref(t)),
               "wk_batch_submit")
        return t.value

    def batch_wait(self, ticket):
        _check(lib().wk_batch_wait(self._h, ticket), "wk_batch_wait")

    def batch_fetch(self, ticket, pb):
        """The tables of a waited-for ticket (submitted with v2c=True), as
        Device.batch returns them."""
        used = C.c_uint64()
        _check(lib().wk_batch_fetch(self._h, ticket, None, 0, C.byref(used)), "wk_batch_fetch")
        arena = np.zeros(max(1, used.value), np.uint32)
        _check(lib().wk_batch_fetch(self._h, ticket, arena.ctypes.data_as(U32P), arena.size, C.byref(used)),
               "wk_batch_fetch")
        return self._tables(pb, arena)

    def _tables(self, pb, arena):
        out = []
        for i in range(pb.n):
            r, c, o = int(pb.rows[i]), int(pb.cols[i]), int(pb.offs[i])
            t = arena[o:o + r * c].reshape(r, c) if c else np.zeros((0, 0), np.uint32)
            out.append((t.copy(), list(pb.v2c[i * MAX_VARS:i * MAX_VARS + pb.nvars[i]]), int(pb.status[i])))
        return out

    def batch_traffic(self, enable):
        """(algorithmic bytes, edges, batches) counted since the last call;
        `enable` turns the counting of later batches on or off."""
        b, e, n = C.c_double(), C.c_double(), C.c_uint64()
        _check(lib().wk_batch_traffic(self._h, 1 if enable else 0, C.byref(b), C.byref(e), C.byref(n)))
        return b.value, e.value, n.value

    def autotune(self, queries, use_graph=True):
        """Native sweep of probe width x gather mapping on a batch; keeps the fastest.
        Returns (probe_width, rows_per_wave, batch_ms)."""
        pb = queries if isinstance(queries, PreparedBatch) else PreparedBatch(queries)
        pw, rpw, ms = C.c_int32(), C.c_int32(), C.c_double()
        _check(lib().wk_dev_autotune(self._h, pb.qs, pb.n, 1 if use_graph else 0, C.byref(pw), C.byref(rpw),
                                     C.byref(ms)), "wk_dev_autotune")
        return pw.value, rpw.value, ms.value

    def serve(self, queries, arrivals, max_inflight=4096, footprint_budget=0, per_query=False,
              keep_results=False, heavy_footprint=0):
        n = len(queries)
        qs = (Query * max(1, n))()
        keep = []
        for i, (pats, nv) in enumerate(queries):
            q = make_query(pats, nv)
            keep.append(q)
            qs[i] = q
        arr = np.ascontiguousarray(arrivals, dtype=np.float64)
        lat = np.zeros(n, np.float64)
        start = np.zeros(n, np.float64)
        rows = np.zeros(n, np.uint64)
        status = np.zeros(n, np.int32)
        ck = np.zeros(n, np.uint64)
        res = ServeResult(lat.ctypes.data_as(C.POINTER(C.c_double)), start.ctypes.data_as(C.POINTER(C.c_double)),
                          rows.ctypes.data_as(U64P), status.ctypes.data_as(C.POINTER(C.c_int32)),
                          ck.ctypes.data_as(U64P), 0.0, 0)
        o = ServeOpts(max_inflight, footprint_budget, 1 if per_query else 0, 1 if keep_results else 0,
                      heavy_footprint)
        _check(lib().wk_serve_replay(self._h, qs, arr.ctypes.data_as(C.POINTER(C.c_double)), n, C.byref(o),
                                     C.byref(res)), "wk_serve_replay")
        return {"latency_us": lat, "start_us": start, "rows": rows, "status": status, "checksum": ck,
                "wall_s": res.wall_s, "batches": res.batches}

    def agent_execute(self, bundle, use_gpu=False, library_version=0):
        """GPUAgent::execute_sparql_query: request Bundle bytes in, reply Bundle bytes out."""
        src = (C.c_uint8 * max(1, len(bundle))).from_buffer_copy(bundle if bundle else b"\0")
        o = _wire_opts(use_gpu, library_version)
        n = C.c_uint64()
        _check(lib().wk_agent_execute(self._h, src, len(bundle), C.byref(o), C.byref(n)), "wk_agent_execute")
        buf = (C.c_uint8 * max(1, n.value))()
        _check(lib().wk_agent_reply(self._h, buf, n.value), "wk_agent_reply")
        return bytes(buf[:n.value])

    def served_rows(self, i):
        r, c = C.c_uint64(), C.c_uint32()
        _check(lib().wk_serve_result_rows(self._h, i, None, 0, C.byref(r), C.byref(c)))
        out = np.zeros((r.value, c.value), np.uint32)
        _check(lib().wk_serve_result_rows(self._h, i, out.ctypes.data_as(U32P), out.size, C.byref(r),
                                          C.byref(c)))
        return out
