"""Oracle of the SPARQLQuery wire form: a Bundle string holding a boost
binary archive.

TEST INFRASTRUCTURE: imported only by tests/ as the checker of the library's
wk_wire_encode / wk_wire_decode / wk_agent_execute, never by the product.

A plain restatement, member by member, of the save() functions of
core/sparql/query.hpp:1089-1243 (Pattern, PatternGroup, Result, SPARQLQuery)
and of Bundle::to_str (core/common/bundle.hpp:155-164: the 4-byte req_type,
then the archive), as boost 1.67's binary_oarchive (deps/deps.sh:5) writes
them on x86-64 (little-endian, int 4 bytes, long 8):

- header: std::string signature "serialization::archive" (size_t length,
  then the bytes), library version (uint16), sizeof int/long/float/double as
  one byte each, then int 1;
- int, enum, unsigned: 4 bytes; bool, char: 1 byte;
- vector of an arithmetic type: size_t count, then the elements as an array
  (no item version);
- vector<bool>: size_t count, then one bool per element;
- vector of a class, std::set: size_t count, uint32 item version, elements;
- classes declared object_serializable (query.hpp:1288-1296): no preamble;
- Result::GPUResult (query.hpp:311-362, default class level): its first
  instance is preceded by tracking (bool) and class version (uint32).

Messages are dicts with the keys of wukong_gpu.query_msg.

Parity unpinned: boost is not in this image and the reference holds no
archive bytes, so the byte layout rests on this restatement of boost 1.67's
binary archive rules (the library version word defaults to 16) and on the
member order of query.hpp; the tests pin encode/decode symmetry, the closed
form of the length and the library against this independent restatement.
"""
import struct

import numpy as np

SIGNATURE = b"serialization::archive"
OCCUPIED, EMPTY = 0, 1  # query.hpp:1086-1087
HEAD = ("q_type", "qid", "pqid", "pg_type", "state", "dev_type", "job_type", "priority",
        "mt_factor", "mt_tid", "pattern_step", "local_var")


def _i(v):
    return struct.pack("<i", int(v))


def _b(v):
    return struct.pack("<B", 1 if v else 0)


def _size(n):
    return struct.pack("<Q", int(n))


def header(library_version=16):
    return (_size(len(SIGNATURE)) + SIGNATURE + struct.pack("<H", library_version) +
            bytes([4, 8, 4, 8]) + _i(1))


def save_pattern(p):  # query.hpp:1089-1100
    s, pr, o, d = p[:4]
    pt = p[4] if len(p) > 4 else 0
    return _i(s) + _i(pr) + _i(o) + _i(d) + struct.pack("<b", pt)


def save_pattern_group(m):  # query.hpp:1113-1132
    out = _size(len(m["patterns"])) + struct.pack("<I", 0)
    for p in m["patterns"]:
        out += save_pattern(p)
    nv = sorted(set(m["new_vars"]))  # std::set<ssid_t>
    out += _size(len(nv)) + struct.pack("<I", 0) + b"".join(_i(v) for v in nv)
    return out + bytes([EMPTY, EMPTY, EMPTY])  # filters, optional, unions


def save_result(m, use_gpu):  # query.hpp:1155-1181
    out = _i(m["col_num"]) + _i(m["row_num"]) + _i(m["attr_col_num"]) + _i(m["status_code"])
    out += _b(m["blind"]) + _i(m["nvars"])
    out += _size(len(m["required_vars"])) + b"".join(_i(v) for v in m["required_vars"])
    out += _size(len(m["v2c_map"])) + b"".join(_i(v) for v in m["v2c_map"])
    out += _size(len(m["optional_matched"])) + b"".join(_b(v) for v in m["optional_matched"])
    if m["row_num"] > 0:
        t = np.asarray(m["table"], dtype="<u4").reshape(-1)
        out += bytes([OCCUPIED]) + _size(t.size) + t.tobytes()
        out += _size(0) + struct.pack("<I", 0)  # attr_res_table, vector<variant>
    else:
        out += bytes([EMPTY])
    if use_gpu:
        out += _b(False) + struct.pack("<I", 0)  # GPUResult class info
        out += struct.pack("<Q", int(m["gpu_nelems"])) + _i(m["gpu_col_num"])
    return out


def encode(m, use_gpu=False, library_version=16):
    """Bundle string of a SPARQL_QUERY: req_type word, then the archive
    (save(SPARQLQuery), query.hpp:1208-1243)."""
    out = struct.pack("<I", 0) + header(library_version)
    out += b"".join(_i(m[k]) for k in HEAD)
    out += _b(m["corun_enabled"]) + _i(m["corun_step"]) + _i(m["fetch_step"])
    out += _b(m["union_done"]) + _i(m["optional_step"]) + _i(m["limit"])
    out += struct.pack("<I", int(m["offset"])) + _b(m["distinct"])
    out += save_pattern_group(m)
    if m["orders"]:
        out += bytes([OCCUPIED]) + _size(len(m["orders"])) + struct.pack("<I", 0)
        out += b"".join(_i(v) + _b(d) for v, d in m["orders"])
    else:
        out += bytes([EMPTY])
    return out + save_result(m, use_gpu)


def encoded_length(m, use_gpu=False):
    """Closed form of len(encode(m))."""
    n = 4 + 8 + len(SIGNATURE) + 2 + 4 + 4          # type word, header
    n += 12 * 4 + 1 + 4 + 4 + 1 + 4 + 4 + 4 + 1      # SPARQLQuery scalars
    n += 12 + 17 * len(m["patterns"]) + 12 + 4 * len(set(m["new_vars"])) + 3
    n += 1 + (12 + 5 * len(m["orders"]) if m["orders"] else 0)
    n += 4 * 4 + 1 + 4 + 8 + 4 * len(m["required_vars"]) + 8 + 4 * len(m["v2c_map"])
    n += 8 + len(m["optional_matched"]) + 1
    if m["row_num"] > 0:
        n += 8 + 4 * np.asarray(m["table"]).size + 12
    if use_gpu:
        n += 1 + 4 + 8 + 4
    return n
