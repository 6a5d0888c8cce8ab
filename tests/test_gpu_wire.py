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
