"""The GPU agent behind the wire form: request Bundle in, reply Bundle out
(wk_agent_execute), checked against the oracle's query and final_process."""
import numpy as np
import pytest

import oracle
import wukong_gpu as wg
from oracle import wire as W
from wukong_gpu import queries as Q

from _graphs import same_rows

pytestmark = pytest.mark.gpu

NO_RESULT = 0xFFFF


def run(dev, m, use_gpu=False):
    reply = dev.agent_execute(W.encode(m, use_gpu=use_gpu), use_gpu=use_gpu)
    r = wg.wire_decode(reply, use_gpu=use_gpu)
    assert r["state"] == 5 and r["job_type"] == 0  # SQ_REPLY, FULL_JOB
    return r


def bound_vars(v2c):
    return [-(i + 1) for i, c in enumerate(v2c) if c >= 0]


@pytest.mark.parametrize("use_gpu", [False, True])
def test_agent_replies_match_oracle(univ10_dev, univ10_oracle, use_gpu):
    cases = [
        ("chain3", {}),
        ("lubm_q1", dict(distinct=1, orders=[(-2, True), (-1, False)], limit=50, offset=3)),
        ("type_start2", dict(orders=[(-1, False)], limit=1000)),
        ("lubm_q7", dict(distinct=1)),
    ]
    for name, kw in cases:
        pats, nv = getattr(Q, name)()
        want, wv2c, wrc = univ10_oracle.query(pats, nv)
        hist, v2c, rc = univ10_dev.query(pats, nv)
        assert rc == 0 and wrc == 0 and v2c == wv2c and same_rows(hist, want), name
        req = bound_vars(v2c)[::-1]
        r = run(univ10_dev, wg.query_msg(pats, nv, req, **kw), use_gpu)
        assert r["status_code"] == 0, name
        expect = oracle.final_process(hist, v2c, req, distinct=bool(kw.get("distinct")),
                                      orders=kw.get("orders", ()), offset=kw.get("offset", 0),
                                      limit=kw.get("limit", -1))
        got = r["table"].reshape(-1, len(req)) if r["row_num"] else np.zeros((0, len(req)), np.uint32)
        assert r["row_num"] == expect.shape[0] and r["col_num"] == len(req), name
        assert np.array_equal(got, expect), name
        assert r["pattern_step"] == len(pats) and r["required_vars"] == req
        assert r["v2c_map"] == [c if c >= 0 else NO_RESULT for c in v2c]
        assert r["qid"] > 0
        if use_gpu:
            assert r["gpu_nelems"] == hist.size and r["gpu_col_num"] == len(req)


def test_agent_blind_and_empty(univ10_dev, univ10_oracle):
    pats, nv = Q.type_start2()
    want, v2c, _ = univ10_oracle.query(pats, nv)
    r = run(univ10_dev, wg.query_msg(pats, nv, [-1], blind=1))
    # shrink(): rows and columns kept, table, patterns and required vars dropped
    assert r["status_code"] == 0 and r["row_num"] == want.shape[0] and r["col_num"] == want.shape[1]
    assert r["table"].size == 0 and r["patterns"] == [] and r["required_vars"] == []
    # a constant start with no edges: an empty result keeps its history columns
    vid = int(want[:, 0].max()) + 1_000_000
    r = run(univ10_dev, wg.query_msg([(vid, int(pats[1][1]), -1, 1)], 1, [-1]))
    assert r["status_code"] == 0 and r["row_num"] == 0 and r["col_num"] == 1 and r["table"].size == 0


def test_agent_continues_a_carried_table(univ10_dev, univ10_oracle):
    """A request at pattern_step k carrying its history table (the CPU
    engine's partial result) continues on the device."""
    for name in ("chain3", "lubm_q7", "lubm_q1"):
        pats, nv = getattr(Q, name)()
        full, v2c, _ = univ10_oracle.query(pats, nv)
        for k in range(1, len(pats)):
            part, pv2c, prc = univ10_oracle.query(pats[:k], nv)
            assert prc == 0
            req = bound_vars(v2c)
            m = wg.query_msg(pats, nv, req, pattern_step=k, col_num=part.shape[1], row_num=part.shape[0],
                             table=part.reshape(-1), v2c_map=[c if c >= 0 else NO_RESULT for c in pv2c])
            r = run(univ10_dev, m)
            assert r["status_code"] == 0, (name, k)
            expect = oracle.final_process(full, v2c, req)
            got = r["table"].reshape(-1, len(req)) if r["row_num"] else np.zeros((0, len(req)), np.uint32)
            assert r["row_num"] == expect.shape[0], (name, k)
            assert same_rows(got, expect), (name, k)


def test_agent_status_codes(univ10_dev):
    pats, nv = Q.chain3()
    for kw, code in [(dict(dev_type=0), wg.WK_E_SETTING),
                     (dict(state=5), wg.WK_E_SETTING),
                     (dict(required=[]), wg.WK_E_NO_REQUIRED_VAR),
                     (dict(patterns=[(p[0], -7, p[2], p[3]) if i == 1 else p for i, p in enumerate(pats)]),
                      wg.WK_E_UNKNOWN_PATTERN),
                     (dict(patterns=[(p[0], p[1], p[2], 2) if i == 2 else p for i, p in enumerate(pats)]),
                      wg.WK_E_UNKNOWN_PATTERN),
                     (dict(nvars=1), wg.WK_E_VERTEX_INVALID)]:
        req = kw.pop("required", [-1])
        p = kw.pop("patterns", pats)
        m = wg.query_msg(p, kw.pop("nvars", nv), req, **kw)
        r = run(univ10_dev, m)
        assert r["status_code"] == code, (kw, r["status_code"])
        assert r["row_num"] == 0 and r["table"].size == 0
    with pytest.raises(wg.WkError) as e:
        univ10_dev.agent_execute(b"\0" * 40)
    assert e.value.code == wg.WK_E_WIRE
    # the failed call leaves no reply behind
    assert wg.lib().wk_agent_reply(univ10_dev._h, None, 0) == wg.WK_E_STATE


def test_agent_index_to_known_on_a_carried_table(univ10_dev, univ10_oracle):
    """A request at pattern_step 0 that carries a history table, whose first
    pattern is an index pattern with its object bound (index_to_known,
    sparql.hpp:83-137): the device keeps the rows whose value is in the index
    list, each once, then goes on; replies equal the oracle's continuation of
    the same table.  A constant start with a bound object after the table is
    const_to_known (sparql.hpp:139-188)."""
    GS, US, MEM, SUB = (Q.U("ub:GraduateStudent"), Q.U("ub:UndergraduateStudent"), Q.U("ub:memberOf"),
                        Q.U("ub:subOrganizationOf"))
    g, _, _ = univ10_oracle.query([(GS, wg.TYPE_ID, -1, wg.IN), (-1, MEM, -2, wg.OUT)], 2)
    u, _, _ = univ10_oracle.query([(US, wg.TYPE_ID, -1, wg.IN), (-1, MEM, -2, wg.OUT)], 2)
    table = np.concatenate([u[:3000], g[:2000], u[:500]])  # duplicate rows stay duplicates
    dept = int(g[0, 1])
    cases = [
        [(GS, wg.TYPE_ID, -1, wg.IN), (-2, SUB, -3, wg.OUT)],      # type index, then k2u
        [(MEM, wg.PREDICATE_ID, -1, wg.IN)],                         # predicate index [0|p|IN]
        [(US, wg.TYPE_ID, -1, wg.IN), (dept, MEM, -1, wg.IN)],       # then const_to_known
    ]
    for pats in cases:
        want, wv, wrc = univ10_oracle.query_from(pats, 3, 0, table, [0, 1])
        assert wrc == 0 and want.shape[0] > 0, pats
        req = bound_vars(wv)
        m = wg.query_msg(pats, 3, req, pattern_step=0, col_num=2, row_num=table.shape[0],
                         table=table.reshape(-1), v2c_map=[0, 1, NO_RESULT])
        r = run(univ10_dev, m)
        assert r["status_code"] == 0, pats
        expect = oracle.final_process(want, wv, req)
        got = r["table"].reshape(-1, len(req))
        assert r["row_num"] == expect.shape[0] and same_rows(got, expect), pats
    # an unbound object after a table, and a constant object, are refused
    for pats, code in [([(GS, wg.TYPE_ID, -3, wg.IN)], wg.WK_E_FIRST_PATTERN),
                       ([(GS, wg.TYPE_ID, dept, wg.IN)], wg.WK_E_VERTEX_INVALID),
                       ([(GS, MEM, -1, wg.IN)], wg.WK_E_OBJ_ERROR)]:
        m = wg.query_msg(pats, 3, [-1], pattern_step=0, col_num=2, row_num=table.shape[0],
                         table=table.reshape(-1), v2c_map=[0, 1, NO_RESULT])
        assert run(univ10_dev, m)["status_code"] == code, pats
        assert univ10_oracle.query_from(pats, 3, 0, table, [0, 1])[2] == code, pats


def test_c2k_entry_point(univ10_dev, univ10_oracle):
    """wk_expand_c2k on an uploaded table: index lists (vid 0) and a vertex's
    edges, both directions."""
    GS, MEM = Q.U("ub:GraduateStudent"), Q.U("ub:memberOf")
    people = np.sort(np.concatenate([univ10_oracle.get_index(GS, wg.IN),
                                     univ10_oracle.get_index(Q.U("ub:UndergraduateStudent"), wg.IN)]))
    t = univ10_dev.upload(people.reshape(-1, 1))
    dept = int(univ10_oracle.get_triples(int(people[0]), MEM, wg.OUT)[0])
    for vid, pid, d in [(0, GS, wg.IN), (0, MEM, wg.IN), (dept, MEM, wg.IN)]:
        out = univ10_dev.c2k(t, vid, pid, d, 0)
        members = set((univ10_oracle.get_index(pid, d) if vid == 0 else
                       univ10_oracle.get_triples(vid, pid, d)).tolist())
        want = people[np.array([int(v) in members for v in people], bool)].reshape(-1, 1)
        assert want.shape[0] > 0 and same_rows(out.numpy(), want), (vid, pid, d)
        out.free()
    t.free()
