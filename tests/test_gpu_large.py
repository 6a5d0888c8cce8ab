"""Tables of 10^7 to 10^9 rows on the device, past 2^31 elements.

A real query whose tables reach 5.3M rows (the students who share a course
with an undergraduate, and their departments) is compared row for row with
the oracle, alone, in chunks under a small table budget, and inside a batch.

The known_to_unknown entry point then runs over an input table that repeats
the 74,644 undergraduates 4,900 times (366M rows), so that its output of
1.08G rows passes 2^31 elements (8.6 GB).  The output of a repeated input is
the base output repeated, since each input row's output rows are contiguous
and in input order; the base output equals the oracle's.

Every device call goes through the C ABI; the oracle is only the checker."""
import numpy as np
import pytest

import wukong_gpu as wg
from wukong_gpu import queries as Q

from _graphs import same_rows

pytestmark = pytest.mark.gpu


def shared_course_query(dept=None):
    """?s type UndergraduateStudent . ?s takesCourse ?c . ?t takesCourse ?c .
    ?t memberOf ?d  (or ?t memberOf D, a known_to_const filter on the 5.3M
    rows of the step before)"""
    US, TC, MEM = Q.U("ub:UndergraduateStudent"), Q.U("ub:takesCourse"), Q.U("ub:memberOf")
    head = [(US, wg.TYPE_ID, -1, wg.IN), (-1, TC, -2, wg.OUT), (-2, TC, -3, wg.IN)]
    if dept is None:
        return head + [(-3, MEM, -4, wg.OUT)], 4
    return head + [(-3, MEM, dept, wg.OUT)], 3


@pytest.fixture(scope="module")
def shared_course(univ10_oracle):
    pats, nv = shared_course_query()
    want, wv, wrc = univ10_oracle.query(pats, nv)
    assert wrc == 0 and want.shape == (5_298_855, 4)
    return pats, nv, want, wv


@pytest.mark.parametrize("variant,budget", [("all", 2 << 30), ("dept", 2 << 30), ("dept", 16 << 20)])
def test_query_with_ten_million_row_tables(univ10_store, univ10_oracle, shared_course, variant, budget):
    """With a 2 GiB table budget the tables fit whole; with 16 MiB (599,186
    elements per table) the 16M-element step runs in output chunks."""
    if variant == "all":
        pats, nv, want, wv = shared_course
    else:
        dept = int(np.sort(univ10_oracle.get_index(Q.U("ub:Department"), wg.IN))[5])
        pats, nv = shared_course_query(dept)
        want, wv, wrc = univ10_oracle.query(pats, nv)
        assert wrc == 0 and 0 < want.shape[0] < 599_186 // 3
    dev = wg.Device(univ10_store, table_budget=budget)
    got, gv, grc = dev.query(pats, nv)
    assert grc == 0 and gv == wv
    assert same_rows(got, want)
    dev.close()


def test_batch_with_ten_million_row_tables(univ10_store, univ10_oracle, shared_course):
    """The query inside a batch of template queries: its step tables hold 21M
    elements beside the others'."""
    pats, nv, want, _ = shared_course
    stream = Q.template_stream(univ10_store, Q.univ_templates(), 60, seed=23)
    qs = [(p, n) for p, n, _ in stream]
    qs.insert(30, (pats, nv))
    dev = wg.Device(univ10_store, table_budget=2 << 30)
    res = dev.batch(qs, use_graph=False)
    for (p, n), (t, v2c, rc) in zip(qs, res):
        assert rc == 0, (p, rc)
        w = want if p is pats else univ10_oracle.query(p, n)[0]
        assert same_rows(t, w), p
    dev.close()


def test_k2u_output_past_two_to_the_31_elements(univ10_store, univ10_oracle):
    US, TC = Q.U("ub:UndergraduateStudent"), Q.U("ub:takesCourse")
    inst = np.sort(univ10_oracle.get_index(US, wg.IN)).astype(np.uint32)
    dev = wg.Device(univ10_store, table_budget=1 << 30)
    base_t = dev.upload(inst.reshape(-1, 1))
    base = dev.k2u(base_t, 0, TC, wg.OUT).numpy()
    want = np.concatenate([np.stack([np.full(e.size, v, np.uint32), e], axis=1)
                           for v in inst for e in [univ10_oracle.get_triples(int(v), TC, wg.OUT)]])
    assert same_rows(base, want)
    B = 4900
    assert B * base.size > 2**31
    big_t = dev.upload(np.tile(inst, B).reshape(-1, 1))
    base_t.free()
    out_t = dev.k2u(big_t, 0, TC, wg.OUT)
    big_t.free()
    assert out_t.shape == (B * base.shape[0], 2)
    out = out_t.numpy()
    out_t.free()
    reps = out.reshape(B, base.shape[0], 2)
    step = 256
    for b0 in range(0, B, step):
        assert (reps[b0:b0 + step] == base[None]).all(), b0
    dev.close()
