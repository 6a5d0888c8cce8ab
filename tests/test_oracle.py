"""Pinning the CPU oracle (CPU only).

* hash: the forward mix composed with the reference's own inverse, restated
  step for step from utils/math.hpp:70-104, must be the identity (this pins
  the forward mix against the reference's inverse, not against a second
  derivation);
* store: the reference's own store fixture (tests/test_graph.cc:142-226),
  restated with its asserted sizes and values;
* engine: brute-force BGP evaluation (all variable assignments over the
  triple set) and SQLite self-joins on tiny random graphs;
* planner: the .fmt rewrite rules (optimizer/planner.hpp:667-699);
* final_process: projection / DISTINCT / ORDER BY / OFFSET / LIMIT against
  direct Python restatements of the reference's steps.
"""
import itertools
import sqlite3

import numpy as np
import pytest

import oracle
from _graphs import VID0, random_graph, same_rows

IN, OUT, TYPE_ID = 0, 1, 1


# ------------------------------------------------------------------ hash

def test_hash_inverse_roundtrip():
    rng = np.random.default_rng(0)
    xs = [0, 1, 2, 2**63, 2**64 - 1, (1 << 18) | 3] + [int(v) for v in rng.integers(0, 2**63, 2000, dtype=np.uint64)]
    for x in xs:
        assert oracle.inverse_hash_u64(oracle.hash_u64(x)) == x
        assert oracle.hash_u64(oracle.inverse_hash_u64(x)) == x


def test_hash_is_a_bijection_on_a_sample():
    rng = np.random.default_rng(1)
    xs = set(int(v) for v in rng.integers(0, 2**63, 5000, dtype=np.uint64))
    assert len({oracle.hash_u64(x) for x in xs}) == len(xs)
    # keys differing in one bit land far apart (avalanche)
    flips = [bin(oracle.hash_u64(x) ^ oracle.hash_u64(x ^ 1)).count("1") for x in list(xs)[:500]]
    assert 20 < np.mean(flips) < 44


# --------------------------------------------------- reference store fixture

def mock_triples():
    """tests/test_graph.cc:142-151: 20 subjects, 3 OUT + 3 IN edges and 1 type."""
    VID_MIN = VID0
    pso, pos = [], []
    for s in range(VID_MIN, VID_MIN + 20):
        for i in range(1, 4):
            pso.append((s, (s % 10) + 11, (s % VID_MIN) * 2 + VID_MIN + i))
            pos.append(((s % VID_MIN) * 3 + VID_MIN + i + 1, ((s + 1) % 10) + 11, s))
        pso.append((s, TYPE_ID, (s % 9) + 2))
    return np.array(pso, np.uint32), np.array(pos, np.uint32)


def check_mock_store(get_triples):
    VID_MIN = VID0
    for s in range(VID_MIN, VID_MIN + 20):
        e = get_triples(s, (s % 10) + 11, OUT)            # test_graph.cc:164-173
        assert len(e) == 3
        assert list(e) == [(s % VID_MIN) * 2 + VID_MIN + i + 1 for i in range(3)]
        e = get_triples(s, ((s + 1) % 10) + 11, IN)       # :176-185
        assert len(e) == 3
        assert list(e) == [(s % VID_MIN) * 3 + VID_MIN + i + 2 for i in range(3)]
        assert len(get_triples(s, TYPE_ID, OUT)) == 1     # :190-195
    for p in range(11, 21):                                # :199-217
        assert len(get_triples(0, p, IN)) == 2
        assert len(get_triples(0, p, OUT)) == 2
    for t in range(2, 11):                                 # :221-226
        assert len(get_triples(0, t, IN)) >= 2


def test_reference_store_fixture_oracle():
    pso, pos = mock_triples()
    st = oracle.Store(pso, pos)
    check_mock_store(st.get_triples)


def test_reference_store_fixture_product_host_store():
    import wukong_gpu as wg
    pso, pos = mock_triples()
    st = wg.Store(pso, pos)
    check_mock_store(st.get_triples)


# ------------------------------------------------------- brute-force BGP

def brute_force(triples, pats, nvars):
    """All assignments of the query's variables over the triple set that
    satisfy every pattern (direction only selects which end is the start).
    Index starts (s tpid, p in {0,1}) bind the variable to the index list."""
    T = set(map(tuple, triples.tolist()))
    verts = sorted({t[0] for t in T} | {t[2] for t in T})
    sols = [dict()]
    for k, (s, p, o, d) in enumerate(pats):
        if k == 0 and 1 < s < (1 << 17) and p in (0, TYPE_ID):
            if p == TYPE_ID:   # [0|t|IN]: instances of type s
                vals = sorted({a for (a, b, c) in T if b == TYPE_ID and c == s})
            elif d == IN:      # [0|p|IN]: subjects of predicate s
                vals = sorted({a for (a, b, c) in T if b == s})
            else:              # [0|p|OUT]: objects of predicate s
                vals = sorted({c for (a, b, c) in T if b == s})
            sols = [{o: v} for v in vals]
            continue
        subj, obj = (s, o) if d == OUT else (o, s)  # back to triple orientation
        new = []
        for m in sols:
            a = m.get(subj, subj if subj >= 0 else None)
            c = m.get(obj, obj if obj >= 0 else None)
            cands_a = [a] if a is not None else verts
            cands_c = [c] if c is not None else verts
            for x in cands_a:
                for y in cands_c:
                    if (x, p, y) in T:
                        mm = dict(m)
                        if subj < 0:
                            mm[subj] = x
                        if obj < 0:
                            mm[obj] = y
                        new.append(mm)
        sols = new
    return sols


def tiny_query_set(info):
    t0, t1 = int(info["types"][0]), int(info["types"][1])
    p0, p1, p2 = (int(x) for x in info["preds"][:3])
    v0 = int(info["verts"][0])
    return [
        ([(t0, TYPE_ID, -1, IN), (-1, p0, -2, OUT)], 2),
        ([(t0, TYPE_ID, -1, IN), (-1, p0, -2, OUT), (-2, TYPE_ID, t1, OUT)], 2),
        ([(t1, TYPE_ID, -1, IN), (-1, p1, -2, IN), (-2, p2, -3, OUT), (-1, p0, -3, OUT)], 3),
        ([(v0, p0, -1, OUT), (-1, p1, -2, OUT), (-2, p0, -3, IN)], 3),
        ([(p0, 0, -1, IN), (-1, p0, -2, OUT), (-2, p0, -1, OUT)], 2),
        ([(p1, 0, -1, OUT), (-1, p1, -2, IN), (v0, p2, -2, OUT)], 2),
        # known_to_unknown with rdf:type IN: instances of ?t (get_index(cur, IN), sparql.hpp:342-343)
        ([(v0, TYPE_ID, -1, OUT), (-1, TYPE_ID, -2, IN)], 2),
        ([(t0, TYPE_ID, -1, IN), (-1, TYPE_ID, -2, OUT), (-2, TYPE_ID, -3, IN)], 3),
        # const_to_known, both directions (sparql.hpp:139-188)
        ([(t0, TYPE_ID, -1, IN), (v0, p0, -1, OUT)], 1),
        ([(p0, 0, -1, IN), (v0, p1, -1, IN)], 1),
        ([(t1, TYPE_ID, -1, IN), (-1, p0, -2, OUT), (v0, p0, -2, IN)], 2),
    ]


@pytest.mark.parametrize("seed", range(4))
def test_oracle_matches_brute_force(seed):
    tri, info = random_graph(seed, n_v=40, n_edges=160, tpid_objects=False)
    st = oracle.Store(tri)
    for pats, nv in tiny_query_set(info):
        t, v2c, rc = st.query(pats, nv)
        assert rc == 0
        sols = brute_force(tri, pats, nv)
        cols = sorted(range(nv), key=lambda i: v2c[i])
        cols = [c for c in cols if v2c[c] >= 0]
        want = np.array([[m[-(c + 1)] for c in cols] for m in sols], np.uint32).reshape(len(sols), len(cols))
        assert same_rows(t, want), pats


def _triple_set(tri):
    return set(map(tuple, tri.tolist()))


def _index_members(T, s, p, d):
    """The index list [0|s|d] from the triple set itself."""
    if p == TYPE_ID:
        return {a for (a, b, c) in T if b == TYPE_ID and c == s}
    if d == IN:
        return {a for (a, b, c) in T if b == s}
    return {c for (a, b, c) in T if b == s and c >= VID0}


@pytest.mark.parametrize("seed", range(3))
def test_index_to_known_on_a_carried_table(seed):
    """A request arriving with a history table whose first pattern is an index
    pattern with a bound object (index_to_known, sparql.hpp:83-137): each row
    is kept once when its value is in the index list, computed here from the
    triple set; then the query goes on from the filtered table."""
    tri, info = random_graph(20 + seed, n_v=50, n_edges=300)
    st = oracle.Store(tri)
    T = _triple_set(tri)
    rng = np.random.default_rng(seed)
    verts = info["verts"]
    table = np.stack([rng.choice(verts, 80), rng.choice(verts, 80)], axis=1).astype(np.uint32)
    table = np.concatenate([table, table[:5]])          # duplicate rows stay duplicates
    t0, p0 = int(info["types"][0]), int(info["preds"][0])
    for pat, col in [((t0, TYPE_ID, -1, IN), 0), ((p0, 0, -2, IN), 1), ((p0, 0, -1, OUT), 0)]:
        got, v2c, rc = st.query_from([pat], 2, 0, table, [0, 1])
        assert rc == 0 and v2c == [0, 1]
        members = _index_members(T, pat[0], pat[1], pat[3])
        want = table[np.array([int(v) in members for v in table[:, col]], bool)]
        assert np.array_equal(got, want), pat
    # then on to a known_to_unknown step
    got, v2c, rc = st.query_from([(t0, TYPE_ID, -1, IN), (-1, p0, -3, OUT)], 3, 0, table, [0, 1])
    kept = table[np.array([int(v) in _index_members(T, t0, TYPE_ID, IN) for v in table[:, 0]], bool)]
    want = [(a, b, c) for a, b in kept.tolist() for (x, y, c) in sorted(T) if x == a and y == p0]
    assert rc == 0 and v2c == [0, 1, 2]
    assert same_rows(got, np.array(want, np.uint32).reshape(-1, 3))
    # status codes: unbound object after a table, constant object, bad predicate
    assert st.query_from([(t0, TYPE_ID, -3, IN)], 3, 0, table, [0, 1])[2] == 11
    assert st.query_from([(t0, TYPE_ID, VID0, IN)], 2, 0, table, [0, 1])[2] == 8
    assert st.query_from([(t0, p0, -1, IN)], 2, 0, table, [0, 1])[2] == 7


def test_query_from_continues_a_query():
    """Running patterns [0, k) then [k, n) from the intermediate table equals
    running all n."""
    tri, info = random_graph(31, n_v=60, n_edges=500)
    st = oracle.Store(tri)
    for pats, nv in tiny_query_set(info):
        full, v2c, rc = st.query(pats, nv)
        for k in range(1, len(pats)):
            part, pv, prc = st.query(pats[:k], nv)
            assert prc == 0
            got, gv, grc = st.query_from(pats, nv, k, part if part.size else part.reshape(0, max(1, len(
                [x for x in pv if x >= 0]))), pv)
            assert grc == rc and gv == v2c and np.array_equal(got, full), (pats, k)


@pytest.mark.parametrize("nthreads", [1, 2, 3, 7, 64])
def test_mt_split_equals_one_thread(nthreads):
    """mt_factor split of index starts (sparql.hpp:214-224) over up to more
    threads than index entries: the parts' tables in thread order are the
    one-thread table row for row (each part keeps the list's order)."""
    tri, info = random_graph(41, n_v=120, n_edges=900)
    st = oracle.Store(tri)
    for pats, nv in tiny_query_set(info):
        want, wv, wrc = st.query(pats, nv)
        got, gv, grc = st.query_mt(pats, nv, nthreads)
        assert grc == wrc and gv == wv
        assert np.array_equal(got, want), pats


def test_run_stream_matches_queries():
    """The shared-queue stream runner gives each query's row count, status and
    fingerprint as the query run alone does, on 1 and 8 workers."""
    import wukong_gpu as wg
    from _graphs import random_query
    tri, info = random_graph(51, n_v=200, n_edges=2000)
    st = oracle.Store(tri)
    rng = np.random.default_rng(2)
    qs = [random_query(rng, info) for _ in range(300)]
    for nt in (1, 8):
        rows, ck, status = st.run_stream(qs, nt)
        for i, (pats, nv) in enumerate(qs):
            t, _, rc = st.query(pats, nv)
            assert status[i] == rc
            if rc == 0:
                assert rows[i] == t.shape[0]
                assert int(ck[i]) == oracle.table_checksum(t) == (wg.checksum(t) if t.size else 0)


def test_oracle_matches_sqlite():
    tri, info = random_graph(9, n_v=60, n_edges=400)
    db = sqlite3.connect(":memory:")
    db.execute("create table t(s integer, p integer, o integer)")
    db.executemany("insert into t values (?,?,?)", set(map(tuple, tri.tolist())))
    st = oracle.Store(tri)
    t0 = int(info["types"][0])
    p0, p1 = int(info["preds"][0]), int(info["preds"][1])
    # ?x type t0 . ?x p0 ?y . ?y p1 ?z
    got, v2c, rc = st.query([(t0, TYPE_ID, -1, IN), (-1, p0, -2, OUT), (-2, p1, -3, OUT)], 3)
    want = db.execute("select a.s, b.o, c.o from t a join t b on b.s=a.s join t c on c.s=b.o "
                      "where a.p=1 and a.o=? and b.p=? and c.p=?", (t0, p0, p1)).fetchall()
    assert rc == 0 and same_rows(got, np.array(want, np.uint32).reshape(-1, 3))
    # ?x p0 ?y . ?y p0 ?x  (cycle: known_to_known)
    got, v2c, rc = st.query([(p0, 0, -1, IN), (-1, p0, -2, OUT), (-2, p0, -1, OUT)], 2)
    want = db.execute("select a.s, a.o from t a join t b on b.s=a.o and b.o=a.s where a.p=? and b.p=?",
                      (p0, p0)).fetchall()
    assert rc == 0 and same_rows(got, np.array(want, np.uint32).reshape(-1, 2))


def test_in_direction_skips_index_objects():
    """IN keys are not built for objects that are type/predicate ids
    (segment_rdf_dgraph.hpp:581-584), so an IN lookup from such an id is empty."""
    tri = np.array([(VID0 + 1, 5, 3), (VID0 + 1, TYPE_ID, 3), (VID0 + 2, 5, VID0 + 1)], np.uint32)
    st = oracle.Store(tri)
    assert len(st.get_triples(3, 5, IN)) == 0
    assert list(st.get_triples(VID0 + 1, 5, OUT)) == [3]
    assert list(st.get_index(3, IN)) == [VID0 + 1]            # type index
    assert list(st.get_index(5, OUT)) == [VID0 + 1]           # objects of p=5 that are vertices


def test_invalid_triples_rejected():
    for bad in ([(VID0, 0, VID0)], [(VID0, TYPE_ID, VID0 + 5)], [(1, 5, VID0)], [(VID0, 5, 0)]):
        with pytest.raises(ValueError):
            oracle.Store(np.array(bad, np.uint32))


@pytest.mark.parametrize("pats,nv,code", [
    ([(-1, 5, -2, OUT)], 2, 9),                                   # unknown start: UNKNOWN_SUB
    ([(VID0, 5, VID0 + 1, OUT)], 0, 3),                           # const-const: UNKNOWN_PATTERN
    ([(3, 5, -1, IN)], 1, 7),                                     # index start, bad predicate: OBJ_ERROR
    ([(VID0, 5, -1, OUT), (VID0, 5, -2, OUT)], 2, 11),            # second const start: FIRST_PATTERN
    ([(VID0, 5, -7, OUT)], 1, 8),                                 # variable beyond nvars
    ([(VID0, -3, -1, OUT)], 3, 3),                                # variable predicate
])
def test_oracle_status_codes(pats, nv, code):
    tri = np.array([(VID0, 5, VID0 + 1), (VID0 + 1, TYPE_ID, 3)], np.uint32)
    st = oracle.Store(tri)
    assert st.query(pats, nv)[2] == code


# ---------------------------------------------------------------- planner

def test_set_direction_rules():
    raw = [(-1, 7, -2, OUT), (-2, TYPE_ID, 4, OUT)]
    assert oracle.set_direction(raw, [(1, "<")]) == [(-2, 7, -1, IN)]
    assert oracle.set_direction(raw, [(1, ">")]) == [(-1, 7, -2, OUT)]
    assert oracle.set_direction(raw, [(1, "<<")]) == [(7, 0, -1, IN)]
    assert oracle.set_direction(raw, [(1, ">>")]) == [(7, 0, -2, OUT)]
    assert oracle.set_direction(raw, [(2, "<"), (1, ">")]) == [(4, TYPE_ID, -2, IN), (-1, 7, -2, OUT)]


def test_product_planner_matches_oracle():
    import wukong_gpu  # noqa: F401
    from wukong_gpu import queries as Q
    raw = [(-2, TYPE_ID, 21, OUT), (-1, 4, -2, OUT), (-1, TYPE_ID, 25, OUT), (-1, 5, -3, OUT),
           (-3, TYPE_ID, 28, OUT), (-2, 6, -3, OUT)]
    for plan in ([(1, "<"), (2, "<"), (3, ">"), (4, ">"), (5, ">"), (6, "<")],
                 [(6, "<<"), (6, ">"), (2, "<"), (1, ">"), (4, ">>"), (3, ">")]):
        assert Q.set_plan(raw, plan) == oracle.set_direction(raw, plan)
    assert Q.parse_fmt("# c\n1 <\n2 >\n\n3 <<\n}\n4 >") == [(1, "<"), (2, ">"), (3, "<<")]


# ----------------------------------------------------------- final_process

def test_final_process_projection_distinct_order_limit():
    rng = np.random.default_rng(3)
    t = rng.integers(1, 5, size=(60, 3)).astype(np.uint32)
    v2c = [2, 0, 1]  # -1 -> col 2, -2 -> col 0, -3 -> col 1
    # projection only
    out = oracle.final_process(t, v2c, [-1, -2])
    assert np.array_equal(out, t[:, [2, 0]])
    # DISTINCT: sort rows by all columns (signed compare), drop neighbours equal on required cols
    out = oracle.final_process(t, v2c, [-2, -3], distinct=True)
    rows = sorted(map(tuple, t.astype(np.int32).tolist()))
    keep = [rows[0]]
    for r in rows[1:]:
        if (r[0], r[1]) != (keep[-1][0], keep[-1][1]):
            keep.append(r)
    assert out.tolist() == [[r[0], r[1]] for r in keep]
    # ORDER BY ?-1 DESC, OFFSET 5 LIMIT 10 (ties broken by full row)
    out = oracle.final_process(t, v2c, [-1, -3], orders=[(-1, True)], offset=5, limit=10)
    srt = sorted(t.tolist(), key=lambda r: (-r[2], r))
    assert out.tolist() == [[r[2], r[1]] for r in srt[5:15]]
    # empty table stays untouched
    e = np.zeros((0, 3), np.uint32)
    assert oracle.final_process(e, v2c, [-1]).shape == (0, 3)


@pytest.mark.parametrize("seed", range(4))
def test_distinct_properties(seed):
    """DISTINCT by properties.  The reference sorts the rows on all columns
    and drops a row when it equals the last kept row on the required columns
    (sparql.hpp:1441-1465), so: every projected tuple survives at least once,
    no two consecutive kept rows are equal on the required columns, and when
    the required variables cover every column the result is exactly the set
    of rows.  (With fewer required columns equal tuples need not be adjacent
    after the full-row sort, and the reference keeps each run of them.)"""
    rng = np.random.default_rng(seed)
    t = rng.integers(0, 4, size=(200, 4)).astype(np.uint32)
    t[:20, 1] |= np.uint32(1 << 31)  # words above 2^31 (signed order in the sort)
    v2c = [3, 1, 0, 2]
    for req in ([-1], [-2, -4], [-4, -1, -3], [-1, -2, -3, -4], [-3, -4, -2, -1]):
        out = oracle.final_process(t, v2c, req, distinct=True)
        proj = {tuple(r) for r in t[:, [v2c[-(v + 1)] for v in req]].tolist()}
        got = [tuple(r) for r in out.tolist()]
        assert set(got) == proj and len(got) <= len(t)
        assert all(a != b for a, b in zip(got, got[1:]))
        if len(req) == t.shape[1]:
            assert len(got) == len(proj)
    # the same rows in the order of the required columns: a true set
    out = oracle.final_process(t, [0, 1, 2, 3], [-1, -2, -3, -4], distinct=True)
    assert len(out) == len({tuple(r) for r in t.tolist()})


def test_partition_rows_matches_the_reference_loop():
    """oracle.partition_rows against a literal restatement of
    generate_sub_query's loop (append row i to owner row[col] % n, in order),
    including empty tables, one owner and owners with no rows."""
    rng = np.random.default_rng(5)
    for rows, cols, n in [(0, 2, 3), (1, 1, 1), (7, 3, 4), (1000, 2, 7), (3000, 4, 64)]:
        t = rng.integers(0, 1 << 20, size=(rows, cols), dtype=np.uint32)
        col = int(rng.integers(0, cols))
        got, sizes = oracle.partition_rows(t, col, n)
        subs = [[] for _ in range(n)]
        for r in range(rows):
            subs[int(t[r, col]) % n].append(t[r])
        want = np.array([row for s in subs for row in s], dtype=np.uint32).reshape(rows, cols)
        assert np.array_equal(got, want)
        assert [len(s) for s in subs] == sizes.tolist()


def test_partition_rows_invariants():
    """What any split of rows among owners must satisfy, on rows tagged with
    their input index: a permutation of the input, owner blocks in owner
    order with the owner's sizes, every row in the block of row[col] % n, and
    input order kept inside each block."""
    rng = np.random.default_rng(9)
    for rows, n in [(1, 1), (50, 3), (4096, 64), (10000, 5)]:
        t = np.stack([np.arange(rows, dtype=np.uint32),
                      rng.integers(0, 1 << 31, size=rows, dtype=np.uint32)], axis=1)
        got, sizes = oracle.partition_rows(t, 1, n)
        assert sorted(got[:, 0].tolist()) == list(range(rows))
        assert int(sizes.sum()) == rows
        start = 0
        for p in range(n):
            blk = got[start:start + int(sizes[p])]
            assert np.all(blk[:, 1] % n == p)
            assert np.all(np.diff(blk[:, 0].astype(np.int64)) > 0)
            assert np.array_equal(blk, t[blk[:, 0]])
            start += int(sizes[p])
