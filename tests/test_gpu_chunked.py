"""Steps whose tables do not fit the device's table budget, on the device.

With a 1 MiB table budget the tables hold 65,536 elements each.  The type
index start of UndergraduateStudent alone yields 74,644 rows at 10
universities, and the steps after it grow by 3x, so every step of
`chunk_query` runs in output chunks, nested three deep, including one input
row (the virtual start row) whose outputs span two chunks.  The results must
still equal the oracle's, through the per-query entry point, inside a batch
and while serving.  Also here: result rows placed past the output-tile map of
a batch step, and a final table too large for the arena ahead of small ones.

Every device call goes through the C ABI; the oracle is only the checker."""
import numpy as np
import pytest

import oracle
import wukong_gpu as wg
from wukong_gpu import queries as Q

from _graphs import same_rows

pytestmark = pytest.mark.gpu

SMALL = 1 << 20  # table budget: 65,536 elements per table (the minimum)


def chunk_query(orc):
    """?s type UndergraduateStudent . ?s takesCourse ?c . ?p teacherOf ?c .
    ?p worksFor D, for one department D."""
    US, TC, TO, WF = (Q.U("ub:UndergraduateStudent"), Q.U("ub:takesCourse"), Q.U("ub:teacherOf"),
                      Q.U("ub:worksFor"))
    dept = int(np.sort(orc.get_index(Q.U("ub:Department"), wg.IN))[3])
    return [(US, wg.TYPE_ID, -1, wg.IN), (-1, TC, -2, wg.OUT), (-2, TO, -3, wg.IN), (-3, WF, dept, wg.OUT)], 3


@pytest.fixture(scope="module")
def small_dev(univ10_store):
    dev = wg.Device(univ10_store, table_budget=SMALL, profile=1)
    assert dev.stats()["table_capacity_elems"] == 65536
    yield dev
    dev.close()


def test_chunked_query_matches_oracle(small_dev, univ10_oracle):
    pats, nv = chunk_query(univ10_oracle)
    want, wv, wrc = univ10_oracle.query(pats, nv)
    assert wrc == 0 and 0 < want.size <= 65536
    small_dev.reset_kernel_stats()
    got, gv, grc = small_dev.query(pats, nv)
    assert grc == 0 and gv == wv
    assert same_rows(got, want)
    assert small_dev.kernel_stats()["query:chunked_steps"]["launches"] >= 3


def test_chunked_query_result_larger_than_the_arena(small_dev, univ10_oracle):
    """Chunking bounds intermediate tables only: a result larger than the
    result arena is reported (WK_E_NOMEM_BUDGET), never truncated; the same
    steps under a selective last filter come back exact."""
    US, TC = Q.U("ub:UndergraduateStudent"), Q.U("ub:takesCourse")
    pats = [(US, wg.TYPE_ID, -1, wg.IN), (-1, TC, -2, wg.OUT), (-2, wg.TYPE_ID, Q.U("ub:Course"), wg.OUT)]
    want, _, _ = univ10_oracle.query(pats, 2)
    assert want.size > 65536
    with pytest.raises(wg.WkError) as e:
        small_dev.query(pats, 2)
    assert e.value.code == wg.WK_E_NOMEM_BUDGET
    courses = np.sort(univ10_oracle.get_index(Q.U("ub:Course"), wg.IN))
    c = int(courses[len(courses) // 2])
    pats2 = [(US, wg.TYPE_ID, -1, wg.IN), (-1, TC, -2, wg.OUT), (-1, TC, c, wg.OUT)]
    want2, _, _ = univ10_oracle.query(pats2, 2)
    assert want2.shape[0] > 0
    got2, _, rc2 = small_dev.query(pats2, 2)
    assert rc2 == 0 and same_rows(got2, want2)


def test_chunked_query_in_a_batch(small_dev, univ10_oracle, univ10_store):
    """Inside a batch the query overflows the step tables; it is split out and
    run alone, in chunks, with its result placed in the batch arena beside
    the others'."""
    pats, nv = chunk_query(univ10_oracle)
    stream = Q.template_stream(univ10_store, Q.univ_templates(), 40, seed=13)
    qs = [(pats, nv)] + [(p, n) for p, n, _ in stream]
    res = small_dev.batch(qs, use_graph=False)
    assert res[0][2] == 0
    ok = 0
    for (p, n), (t, v2c, rc) in zip(qs, res):
        if rc == wg.WK_E_NOMEM_BUDGET:
            continue  # its result did not fit what the arena had left
        want, wv, _ = univ10_oracle.query(p, n)
        assert rc == 0 and v2c == wv, p
        assert same_rows(t, want), p
        ok += 1
    assert ok >= 20


def test_chunked_query_while_serving(small_dev, univ10_oracle, univ10_store):
    pats, nv = chunk_query(univ10_oracle)
    stream = Q.template_stream(univ10_store, Q.univ_templates(), 60, seed=17)
    qs = [(p, n) for p, n, _ in stream]
    qs.insert(25, (pats, nv))
    arr = np.cumsum(np.random.default_rng(3).exponential(1 / 20000.0, size=len(qs)))
    # heavy = 1: every query on the second lane, where the chunked query's
    # tables overflow and it runs alone, in chunks, on that lane; heavy =
    # 4096: queries also move there after a step, with the table it made
    for heavy in (0, 1, 4096):
        r = small_dev.serve(qs, arr, max_inflight=32, keep_results=True, heavy_footprint=heavy)
        for i, (p, n) in enumerate(qs):
            want, _, _ = univ10_oracle.query(p, n)
            if want.size > 65536:
                assert r["status"][i] == wg.WK_E_NOMEM_BUDGET, i
                continue
            assert r["status"][i] == 0, (heavy, i, r["status"][i])
            assert same_rows(small_dev.served_rows(i), want), (heavy, i)
            assert int(r["checksum"][i]) == (wg.checksum(want) if want.size else 0), (heavy, i)


def test_pattern_entry_points_on_tables_above_the_row_scratch(small_dev, univ10_oracle):
    """known_to_unknown over an uploaded table of 150,000 rows (more than the
    65,536 rows of per-row scratch): probed in slices into one exact table."""
    US, TC = Q.U("ub:UndergraduateStudent"), Q.U("ub:takesCourse")
    inst = np.sort(univ10_oracle.get_index(US, wg.IN))
    rows = np.concatenate([inst, inst[::-1]])[:150_000].reshape(-1, 1).astype(np.uint32)
    t = small_dev.upload(rows)
    out = small_dev.k2u(t, 0, TC, wg.OUT)
    want = []
    for v in rows[:, 0]:
        e = univ10_oracle.get_triples(int(v), TC, wg.OUT)
        want.append(np.stack([np.full(e.size, v, np.uint32), e], axis=1))
    want = np.concatenate(want)
    assert out.shape == want.shape
    assert same_rows(out.numpy(), want)
    # a known_to_const filter over the same rows keeps each matching row once
    c = int(want[len(want) // 3, 1])
    f = small_dev.k2c(t, 0, TC, wg.OUT, c)
    fw = rows[np.isin(rows[:, 0], want[want[:, 1] == c, 0])]
    assert same_rows(f.numpy(), fw)
    out.free()
    f.free()
    t.free()


@pytest.mark.parametrize("extra", [3, 300])
def test_batch_rows_past_the_output_tile_map(univ10_store, univ10_oracle, extra):
    """Final tables dropped for lack of arena room still occupy output rows of
    the step; a table after them, past the step's output-tile map, is still
    written: the 2-pattern query at the end gets its exact rows."""
    dev = wg.Device(univ10_store, table_budget=4 << 20)  # 149,796 elements per table
    tcap = dev.stats()["table_capacity_elems"]
    GS, MEM = Q.U("ub:GraduateStudent"), Q.U("ub:memberOf")
    q = ([(GS, wg.TYPE_ID, -1, wg.IN)], 1)   # one pattern: 23,151 rows straight into the arena
    want1, _, _ = univ10_oracle.query(*q)
    n_fit = tcap // want1.size
    dept = int(np.sort(univ10_oracle.get_index(Q.U("ub:Department"), wg.IN))[0])
    tail = ([(GS, wg.TYPE_ID, -1, wg.IN), (-1, MEM, dept, wg.OUT)], 1)
    want_tail, _, _ = univ10_oracle.query(*tail)
    assert 0 < want_tail.size <= tcap - n_fit * want1.size
    qs = [q] * (n_fit + extra) + [tail]
    res = dev.batch(qs, use_graph=False)
    for i, (t, v2c, rc) in enumerate(res[:-1]):
        if i < n_fit:
            assert rc == 0 and same_rows(t, want1), i
        else:
            assert rc == wg.WK_E_NOMEM_BUDGET, (i, rc)
    t, _, rc = res[-1]
    assert rc == 0 and same_rows(t, want_tail)
    dev.close()


def test_oversized_final_table_does_not_crowd_out_later_ones(univ10_store, univ10_oracle):
    """A final table larger than the whole arena takes no arena room: the
    small final tables after it in the same step still fit."""
    dev = wg.Device(univ10_store, table_budget=SMALL)       # arena of 65,536 elements
    US, GS = Q.U("ub:UndergraduateStudent"), Q.U("ub:GraduateStudent")
    big = ([(US, wg.TYPE_ID, -1, wg.IN)], 1)                 # 74,644 elements
    small = ([(GS, wg.TYPE_ID, -1, wg.IN)], 1)               # 23,151 elements: two fit
    res = dev.batch([big, small, small], use_graph=False)
    assert res[0][2] == wg.WK_E_NOMEM_BUDGET
    want, _, _ = univ10_oracle.query(*small)
    for t, v2c, rc in res[1:]:
        assert rc == 0 and same_rows(t, want)
    dev.close()
