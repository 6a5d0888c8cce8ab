"""The SPARQLQuery wire form (Bundle + boost binary archive) of the C ABI
against the oracle's member-by-member restatement of query.hpp's save()
functions, on the CPU (no device call)."""
import struct

import numpy as np
import pytest

import wukong_gpu as wg
from oracle import wire as W


def random_msg(rng, with_table=None):
    nv = int(rng.integers(1, 9))
    np_ = int(rng.integers(0, 7))
    rows = int(rng.integers(0, 6)) if with_table is None else with_table
    cols = int(rng.integers(1, 5))
    word = lambda: int(rng.integers(-2**31, 2**31))
    m = wg.query_msg(
        [(word(), word(), word(), int(rng.integers(0, 3)), int(rng.integers(-128, 128))) for _ in range(np_)],
        nv, [int(x) for x in rng.integers(-nv, 0, size=int(rng.integers(0, 4)))])
    for k in W.HEAD + ("corun_step", "fetch_step", "optional_step", "limit", "col_num",
                       "attr_col_num", "status_code"):
        m[k] = word()
    for k in ("corun_enabled", "union_done", "distinct", "blind"):
        m[k] = int(rng.integers(0, 2))
    m["offset"] = int(rng.integers(0, 2**32))
    m["row_num"] = rows
    m["new_vars"] = [int(x) for x in rng.integers(-5, 0, size=int(rng.integers(0, 5)))]
    m["orders"] = [(int(rng.integers(-nv, 0)), bool(rng.integers(0, 2))) for _ in range(int(rng.integers(0, 3)))]
    m["v2c_map"] = [int(x) for x in rng.integers(0, 0x10000, size=nv)]
    m["optional_matched"] = [bool(x) for x in rng.integers(0, 2, size=int(rng.integers(0, 6)))]
    m["table"] = rng.integers(0, 2**32, size=rows * cols, dtype=np.uint64).astype(np.uint32)
    m["gpu_nelems"] = int(rng.integers(0, 2**63))
    m["gpu_col_num"] = word()
    return m


def same_msg(a, b, use_gpu=False):
    for k in a:
        if k in ("gpu_nelems", "gpu_col_num") and not use_gpu:
            continue  # Result::gpu is on the wire in USE_GPU builds only
        x, y = a[k], b[k]
        if k == "patterns":
            x = [tuple(p) + (0,) * (5 - len(p)) for p in x]
        if k == "new_vars":
            x = sorted(set(x))  # a std::set on the wire
        if k in ("corun_enabled", "union_done", "distinct", "blind"):
            x, y = bool(x), bool(y)
        if k == "table" and a["row_num"] <= 0:
            continue  # not written when there are no rows
        if isinstance(x, np.ndarray) or isinstance(y, np.ndarray):
            if not np.array_equal(np.asarray(x), np.asarray(y)):
                return k
        elif x != y:
            return k
    return None


@pytest.mark.parametrize("use_gpu", [False, True])
def test_encode_matches_oracle(use_gpu):
    rng = np.random.default_rng(11 + use_gpu)
    for i in range(200):
        m = random_msg(rng)
        lv = int(rng.choice([0, 7, 16, 17]))
        got = wg.wire_encode(m, use_gpu=use_gpu, library_version=lv)
        want = W.encode(m, use_gpu=use_gpu, library_version=lv or 16)
        assert got == want, i
        assert len(got) == W.encoded_length(m, use_gpu), i


@pytest.mark.parametrize("use_gpu", [False, True])
def test_decode_inverts_encode(use_gpu):
    rng = np.random.default_rng(21 + use_gpu)
    for i in range(200):
        m = random_msg(rng)
        d = wg.wire_decode(W.encode(m, use_gpu=use_gpu), use_gpu=use_gpu)
        assert same_msg(m, d, use_gpu) is None, (i, same_msg(m, d, use_gpu))
        assert wg.wire_encode(d, use_gpu=use_gpu) == W.encode(m, use_gpu=use_gpu), i


def test_header_layout():
    m = wg.query_msg([(1 << 17, 2, -1, 1)], 1, [-1], dev_type=1)
    b = wg.wire_encode(m)
    assert b[:4] == b"\0\0\0\0"                            # req_type SPARQL_QUERY
    assert struct.unpack("<Q", b[4:12])[0] == 22
    assert b[12:34] == b"serialization::archive"
    assert struct.unpack("<H", b[34:36])[0] == 16
    assert b[36:40] == bytes([4, 8, 4, 8])
    assert struct.unpack("<i", b[40:44])[0] == 1
    # SPARQLQuery's first members: q_type SELECT, qid -1, pqid -1, pg_type
    # BASIC, state SQ_PATTERN, dev_type GPU, job_type FULL_JOB, priority 0,
    # mt_factor 1 (query.hpp:710-728)
    assert struct.unpack("<9i", b[44:80]) == (0, -1, -1, 0, 0, 1, 0, 0, 1)


def test_truncated_and_foreign_bytes_are_refused():
    rng = np.random.default_rng(3)
    m = random_msg(rng, with_table=3)
    for use_gpu in (False, True):
        b = W.encode(m, use_gpu=use_gpu)
        for cut in range(len(b)):
            with pytest.raises(wg.WkError) as e:
                wg.wire_decode(b[:cut], use_gpu=use_gpu)
            assert e.value.code == wg.WK_E_WIRE, cut
    b = W.encode(m)
    for bad in (b"\1" + b[1:],                                  # req_type DYNAMIC_LOAD
                b[:12] + b"S" + b[13:],                         # signature
                b[:34] + struct.pack("<H", 6) + b[36:],         # library version 6
                b[:37] + b"\4" + b[38:],                        # sizeof(long) 4
                b[:40] + struct.pack("<i", 1 << 24) + b[44:]):  # byte order
        with pytest.raises(wg.WkError) as e:
            wg.wire_decode(bad)
        assert e.value.code == wg.WK_E_WIRE


def _flag_offset(m):
    """Offset of the filters flag of the top pattern group."""
    return len(W.encode(dict(m, orders=[], row_num=0))) - len(
        W.save_result(dict(m, row_num=0), False)) - 1 - 3


def test_filters_unions_optional_and_attributes_are_refused():
    m = wg.query_msg([(1 << 17, 2, -1, 1)], 1, [-1])
    b = W.encode(m)
    off = _flag_offset(m)
    assert b[off:off + 3] == bytes([1, 1, 1])
    for i, code in enumerate((wg.WK_E_UNKNOWN_FILTER, wg.WK_E_UNSUPPORT_UNION, wg.WK_E_UNSUPPORT_UNION)):
        bad = bytearray(b)
        bad[off + i] = 0
        with pytest.raises(wg.WkError) as e:
            wg.wire_decode(bytes(bad))
        assert e.value.code == code
    # a non-empty attr_res_table
    m2 = dict(m, row_num=1, col_num=1, table=np.array([7], np.uint32))
    b2 = W.encode(m2)
    at = b2.rfind(struct.pack("<QI", 0, 0))
    assert at == len(b2) - 12
    bad = b2[:at] + struct.pack("<QI", 1, 0) + struct.pack("<ii", 0, 5)
    with pytest.raises(wg.WkError) as e:
        wg.wire_decode(bad)
    assert e.value.code == wg.WK_E_ATTR_DISABLE


def test_any_nonzero_flag_reads_as_absent():
    """load() compares the flag with `occupied` only (query.hpp:1133-1153)."""
    m = wg.query_msg([(1 << 17, 2, -1, 1)], 1, [-1])
    b = bytearray(W.encode(m))
    off = _flag_offset(m)
    b[off:off + 3] = bytes([7, 255, 2])
    assert same_msg(m, wg.wire_decode(bytes(b))) is None


def test_encode_sizes_and_errors():
    m = wg.query_msg([(1 << 17, 2, -1, 1)], 1, [-1])
    L = wg.lib()
    n = wg.C.c_uint64()
    msg = wg.SparqlMsg()
    rc = L.wk_wire_encode(wg.C.byref(msg), None, None, 0, wg.C.byref(n))
    assert rc == wg.WK_E_INVALID_ARG and n.value == W.encoded_length(wg.query_msg([], 0, []))
    msg.npatterns = 3  # a count without its array
    assert L.wk_wire_encode(wg.C.byref(msg), None, None, 0, wg.C.byref(n)) == wg.WK_E_INVALID_ARG
    assert L.wk_wire_encode(None, None, None, 0, wg.C.byref(n)) == wg.WK_E_INVALID_ARG
    assert len(wg.wire_encode(m)) == W.encoded_length(m)


def test_mutated_archives_decode_or_fail_cleanly():
    """Random byte changes, insertions and cuts of valid archives either
    decode to a message whose own encoding decodes again, or are refused
    with a status code."""
    rng = np.random.default_rng(5)
    codes = {wg.WK_E_WIRE, wg.WK_E_UNKNOWN_FILTER, wg.WK_E_UNSUPPORT_UNION, wg.WK_E_ATTR_DISABLE}
    for i in range(3000):
        use_gpu = bool(i & 1)
        b = bytearray(W.encode(random_msg(rng), use_gpu=use_gpu))
        kind = i % 3
        at = int(rng.integers(0, len(b)))
        if kind == 0:
            b[at] = int(rng.integers(0, 256))
        elif kind == 1:
            b[at:at] = bytes(rng.integers(0, 256, size=int(rng.integers(1, 9)), dtype=np.uint8))
        else:
            b = b[:at] + b[at + int(rng.integers(1, 9)):]
        try:
            d = wg.wire_decode(bytes(b), use_gpu=use_gpu)
        except wg.WkError as e:
            assert e.code in codes, (i, e.code)
            continue
        again = wg.wire_encode(d, use_gpu=use_gpu)
        assert wg.wire_decode(again, use_gpu=use_gpu) is not None
