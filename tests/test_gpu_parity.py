"""Parity of the HIP query path with the CPU oracle (bit-exact multisets of rows).

Every call goes through the C ABI (libwukong_gpu.so); the oracle is only the
checker."""
import numpy as np
import pytest

import oracle
import wukong_gpu as wg
from wukong_gpu import queries as Q

from _graphs import random_graph, random_query, same_rows

pytestmark = pytest.mark.gpu


def check_query(dev, orc, pats, nvars):
    want, wv2c, wrc = orc.query(pats, nvars)
    got, gv2c, grc = dev.query(pats, nvars)
    assert grc == wrc, (pats, grc, wrc)
    if wrc == 0:
        assert gv2c == wv2c, (pats, gv2c, wv2c)
        assert got.shape == want.shape, (pats, got.shape, want.shape)
        assert same_rows(got, want), pats
    return want


def test_config1_chain(univ10_dev, univ10_oracle):
    pats, nv = Q.chain3()
    t = check_query(univ10_dev, univ10_oracle, pats, nv)
    assert t.shape[0] > 20000 and t.shape[1] == 3


@pytest.mark.parametrize("name", ["lubm_q7", "type_start2", "lubm_q1"])
def test_lubm_shapes(univ10_dev, univ10_oracle, name):
    pats, nv = getattr(Q, name)()
    t = check_query(univ10_dev, univ10_oracle, pats, nv)
    assert t.shape[0] > 0


def test_config1_fixtures(univ10_dev):
    """Device results of the committed config-1 fixtures: row count, var->col
    map and order-independent checksum."""
    import json
    import os
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "config1.json")))
    for name, want in golden["queries"].items():
        got, v2c, rc = univ10_dev.query([tuple(p) for p in want["patterns"]], want["nvars"])
        assert rc == want["status"], name
        assert got.shape[0] == want["rows"] and list(v2c) == want["v2c"], name
        assert str(wg.checksum(got) if got.size else 0) == want["checksum"], name


def test_per_pattern_api(univ10_dev, univ10_oracle):
    """i2u -> k2u -> k2c -> k2k through the per-pattern entry points."""
    dev, orc = univ10_dev, univ10_oracle
    GS, MEM, SUB, ADV, WF = (Q.U("ub:GraduateStudent"), Q.U("ub:memberOf"), Q.U("ub:subOrganizationOf"),
                             Q.U("ub:advisor"), Q.U("ub:worksFor"))
    t0 = dev.i2u(GS, wg.IN)
    a = t0.numpy()
    assert same_rows(a, orc.get_index(GS, wg.IN).reshape(-1, 1))
    t1 = dev.k2u(t0, 0, ADV, wg.OUT)          # ?x advisor ?p
    t2 = dev.k2u(t1, 1, WF, wg.OUT)           # ?p worksFor ?d
    t3 = dev.k2k(t2, 0, MEM, wg.OUT, 2)       # ?x memberOf ?d (known_to_known)
    dept = int(t3.numpy()[0, 2])
    t4 = dev.k2c(t3, 2, SUB, wg.OUT, int(orc.get_triples(dept, SUB, wg.OUT)[0]))
    raw = [(GS, wg.TYPE_ID, -1, wg.IN), (-1, ADV, -2, wg.OUT), (-2, WF, -3, wg.OUT), (-1, MEM, -3, wg.OUT),
           (-3, SUB, int(orc.get_triples(dept, SUB, wg.OUT)[0]), wg.OUT)]
    for k, t in enumerate([t1, t2, t3, t4]):
        want, _, rc = orc.query(raw[:k + 2], 3)
        assert rc == 0
        assert same_rows(t.numpy(), want), k
    # a table uploaded from the host expands the same way
    up = dev.upload(a)
    assert same_rows(dev.k2u(up, 0, ADV, wg.OUT).numpy(), t1.numpy())


def test_type_index_expansion(univ10_dev, univ10_oracle):
    """k2u with rdf:type IN reads the type index (sparql.hpp:342-343)."""
    v = int(univ10_oracle.get_index(Q.U("ub:FullProfessor"), wg.IN)[0])
    pats = [(v, wg.TYPE_ID, -1, wg.OUT), (-1, wg.TYPE_ID, -2, wg.IN)]
    t = check_query(univ10_dev, univ10_oracle, pats, 2)
    assert t.shape[0] > 100


@pytest.mark.parametrize("dense", [0, -1])
@pytest.mark.parametrize("seed", range(6))
def test_random_graphs(seed, dense):
    """Random graphs and queries; probes through the direct key indexes
    (dense = 0) and through the hash buckets only (dense = -1)."""
    tri, info = random_graph(seed, tpid_objects=(seed % 2 == 0))
    orc = oracle.Store(tri)
    st = wg.Store(tri, load_factor_pct=95 if seed % 3 == 0 else 0)  # 95%: long bucket chains
    dev = wg.Device(st, table_budget=64 << 20, dense_index=dense)
    assert (dev.stats()["dense_segments"] > 0) == (dense == 0)
    rng = np.random.default_rng(100 + seed)
    for _ in range(60):
        pats, nv = random_query(rng, info)
        check_query(dev, orc, pats, nv)
    dev.close()


@pytest.mark.parametrize("pw", [1, 2, 4, 8])
@pytest.mark.parametrize("rpw", [0, 1, 4, 16, 64])
def test_tuning_knobs(univ10_dev, univ10_oracle, pw, rpw):
    """Every probe width (lanes per key) and gather mapping (rows per
    wavefront) gives the oracle's rows, filters included."""
    univ10_dev.set_tuning(pw, rpw)
    try:
        for f in (Q.chain3, Q.type_start2, Q.lubm_q1, Q.lubm_q7):
            pats, nv = f()
            check_query(univ10_dev, univ10_oracle, pats, nv)
    finally:
        univ10_dev.set_tuning(2, 0)
    with pytest.raises(wg.WkError):
        univ10_dev.set_tuning(3, 0)


def test_autotune_picks_a_valid_setting(univ10_dev, univ10_oracle, univ10_store):
    """The native autotuner sweeps every probe width x gather mapping on a
    batch, keeps the fastest, and the batch stays exact under it."""
    stream = Q.template_stream(univ10_store, Q.univ_templates(), 120, seed=21)
    qs = [(p, nv) for p, nv, _ in stream]
    try:
        pw, rpw, ms = univ10_dev.autotune(qs, use_graph=True)
        assert pw in (1, 2, 4, 8) and rpw in (0, 1, 4, 16, 64) and ms > 0
        for (pats, nv), (t, v2c, rc) in zip(qs, univ10_dev.batch(qs, use_graph=True)):
            want, _, wrc = univ10_oracle.query(pats, nv)
            assert rc == wrc == 0 and same_rows(t, want), pats
    finally:
        univ10_dev.set_tuning(2, 0)


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_oracle(univ10_dev, univ10_oracle, univ10_store, use_graph):
    stream = Q.template_stream(univ10_store, Q.univ_templates(), 300, seed=5)
    qs = [(p, nv) for p, nv, _ in stream]
    for rep in range(3):  # graph replays with eager work in between
        res = univ10_dev.batch(qs, use_graph=use_graph)
        for (pats, nv), (t, v2c, rc) in zip(qs, res):
            want, wv2c, wrc = univ10_oracle.query(pats, nv)
            assert rc == wrc
            assert v2c == wv2c
            assert same_rows(t, want), pats
        pats, nv = Q.chain3()
        check_query(univ10_dev, univ10_oracle, pats, nv)


def test_batch_mixed_random(univ10_oracle):
    tri, info = random_graph(11)
    orc = oracle.Store(tri)
    dev = wg.Device(wg.Store(tri), table_budget=64 << 20)
    rng = np.random.default_rng(7)
    qs = [random_query(rng, info) for _ in range(200)]
    res = dev.batch(qs, use_graph=True)
    for (pats, nv), (t, v2c, rc) in zip(qs, res):
        want, wv2c, wrc = orc.query(pats, nv)
        assert rc == wrc, pats
        if rc == 0:
            assert v2c == wv2c
            assert same_rows(t, want), pats
    dev.close()


def test_batch_tiles_spanning_more_queries_than_a_block():
    """3,000 queries, many of them empty or invalid, so that a 1,024-row
    probe tile spans more than 1,024 queries (row starts searched in global
    memory) as well as fewer (query descriptors and segments staged in LDS)."""
    tri, info = random_graph(23)
    orc = oracle.Store(tri)
    dev = wg.Device(wg.Store(tri), table_budget=256 << 20)
    rng = np.random.default_rng(29)
    qs = [random_query(rng, info) for _ in range(3000)]
    ok = 0
    for (pats, nv), (t, v2c, rc) in zip(qs, dev.batch(qs, use_graph=True)):
        want, wv2c, wrc = orc.query(pats, nv)
        assert rc == wrc, pats
        if rc == 0:
            assert v2c == wv2c
            assert same_rows(t, want), pats
            ok += 1
    assert 0 < ok < len(qs)  # both valid and invalid queries took part
    dev.close()


@pytest.mark.parametrize("n,frac", [(2000, 0.3), (6000, 0.05)])
def test_batch_type_filter_steps(univ10_dev, univ10_oracle, univ10_store, n, frac):
    """Batched type-filter steps (the type bit of each row).  Queries ?X
    takesCourse C . ?X type GraduateStudent, with C a graduate course for a
    fraction of them and a department (no rows) for the rest: at 5% a
    1,024-row probe tile of the filter step spans more than 1,024 queries (row
    starts searched in memory), at 30% fewer (staged in LDS).  Then the same
    queries with every tenth one's second pattern an expansion (advisor), so
    that the step mixes filters and expansions."""
    U = Q.U
    courses = np.sort(univ10_store.get_index(U("ub:GraduateCourse"), Q.IN))
    depts = np.sort(univ10_store.get_index(U("ub:Department"), Q.IN))
    filt = [(-1, U("ub:takesCourse"), None, Q.OUT), (-1, Q.TYPE_ID, U("ub:GraduateStudent"), Q.OUT)]
    expand = [(-1, U("ub:takesCourse"), None, Q.OUT), (-1, U("ub:advisor"), -2, Q.OUT)]
    rng = np.random.default_rng(43)
    consts = [int(courses[rng.integers(len(courses))]) if rng.random() < frac else int(depts[rng.integers(len(depts))])
              for _ in range(n)]
    plan = [(1, "<"), (2, ">")]
    only = [(Q.set_plan(Q.instantiate(filt, c), plan), 1) for c in consts]
    mixed = [(Q.set_plan(Q.instantiate(expand if i % 10 == 0 else filt, c), plan), 2 if i % 10 == 0 else 1)
             for i, c in enumerate(consts)]
    for qs in (only, mixed):
        nonempty = 0
        for use_graph in (True, False):
            for (pats, nv), (t, v2c, rc) in zip(qs, univ10_dev.batch(qs, use_graph=use_graph)):
                want, wv2c, wrc = univ10_oracle.query(pats, nv)
                assert rc == wrc == 0, pats
                assert v2c == wv2c
                assert same_rows(t, want), pats
                nonempty += len(want) > 0
        assert nonempty > 0


def test_batch_overflow_splits(univ10_store, univ10_oracle):
    """A batch whose step table exceeds the table capacity is split and still
    exact; a query that alone exceeds it reports WK_E_NOMEM_BUDGET."""
    dev = wg.Device(univ10_store, table_budget=4 << 20)   # 149,796 elements per table
    tcap = dev.stats()["table_capacity_elems"]
    GS, MEM = Q.U("ub:GraduateStudent"), Q.U("ub:memberOf")
    depts = np.sort(univ10_store.get_index(Q.U("ub:Department"), wg.IN))[:8]
    # index start (23,151 rows each) then a known_to_const filter on one department
    qs = [([(GS, wg.TYPE_ID, -1, wg.IN), (-1, MEM, int(d), wg.OUT)], 1) for d in depts]
    assert 23151 * len(qs) > tcap  # step 0 of the whole batch cannot fit: the batch must split
    qs.append(Q.type_start2())     # 220,425 x 2 elements alone: cannot fit
    res = dev.batch(qs, use_graph=False)
    for (pats, nv), (t, v2c, rc) in zip(qs[:-1], res[:-1]):
        want, _, wrc = univ10_oracle.query(pats, nv)
        assert wrc == 0 and rc == 0 and want.shape[0] > 0
        assert same_rows(t, want), pats
    assert res[-1][2] == wg.WK_E_NOMEM_BUDGET
    dev.close()


def test_batch_arena_exhaustion(univ10_store, univ10_oracle):
    """Results kept in the batch arena beyond its capacity: every query is
    either exact or reports WK_E_NOMEM_BUDGET, never wrong rows."""
    dev = wg.Device(univ10_store, table_budget=4 << 20)
    stream = Q.template_stream(univ10_store, Q.univ_templates(), 200, seed=9)
    qs = [(p, nv) for p, nv, _ in stream]
    res = dev.batch(qs, use_graph=False)
    ok = 0
    for (pats, nv), (t, v2c, rc) in zip(qs, res):
        assert rc in (0, wg.WK_E_NOMEM_BUDGET), rc
        if rc == 0:
            want, _, _ = univ10_oracle.query(pats, nv)
            assert same_rows(t, want), pats
            ok += 1
    assert ok >= 50
    dev.close()


@pytest.mark.parametrize("pinned", [True, False])
def test_gstore_swapping(univ10, univ10_store, univ10_oracle, pinned):
    """A gstore budget below the store size evicts and reloads segments, from
    a pinned store (direct DMA) and from a pageable one (through the pinned
    staging buffers)."""
    if not pinned:
        univ10_store = wg.Store(univ10, pinned=False)
        assert univ10_store.info()["pinned"] == 0
    info = univ10_store.info()
    total = info["num_buckets"] * 128 + info["num_values"] * 4
    # the largest single segment (type OUT keys, index IN values) still fits;
    # no direct indexes or type bitmaps, so that the type filters read the
    # type segment as well
    dev = wg.Device(univ10_store, gstore_budget=total * 2 // 5, key_block_shift=10, value_block_shift=14,
                    table_budget=256 << 20, dense_index=-1)
    for f in (Q.chain3, Q.type_start2, Q.lubm_q1, Q.chain3):
        pats, nv = f()
        check_query(dev, univ10_oracle, pats, nv)
    s = dev.stats()
    assert s["segments_evicted"] > 0 and s["h2d_bytes"] > total // 3
    dev.close()


@pytest.mark.parametrize("pinned,dense", [(True, 0), (True, -1), (False, -1)])
def test_segments_larger_than_the_gstore_budget(univ10, univ10_store, univ10_oracle, pinned, dense):
    """A 4 MiB gstore budget is below the two largest key sets (6.7 MB of
    buckets each) and the index segment's values (4.2 MB): those segments are
    read in place from pinned host memory (the store's own, or a pinned copy
    of a pageable store's), the others swap through the budget, and every
    result is exact per query, batched and served."""
    store = univ10_store if pinned else wg.Store(univ10, pinned=False)
    dev = wg.Device(store, gstore_budget=4 << 20, key_block_shift=10, value_block_shift=14,
                    table_budget=256 << 20, dense_index=dense)
    for f in (Q.chain3, Q.type_start2, Q.lubm_q1, Q.lubm_q7):
        pats, nv = f()
        check_query(dev, univ10_oracle, pats, nv)
    stream = Q.template_stream(store, Q.univ_templates(), 200, seed=11)
    qs = [(p, nv) for p, nv, _ in stream]
    wants = [univ10_oracle.query(p, nv)[0] for p, nv in qs]
    res = dev.batch(qs, use_graph=False)
    for (pats, nv), (t, v2c, rc), want in zip(qs, res, wants):
        assert rc == 0, (pats, rc)
        assert same_rows(t, want), pats
    arr = np.cumsum(np.random.default_rng(2).exponential(1 / 20000.0, size=len(qs)))
    r = dev.serve(qs, arr, max_inflight=64)
    assert (r["status"] == 0).all()
    for i, want in enumerate(wants):
        assert int(r["checksum"][i]) == (wg.checksum(want) if want.size else 0), i
    s = dev.stats()
    # at least the type index segment, which every index start reads
    assert s["mapped_segments"] >= 1 and s["segments_evicted"] > 0
    dev.close()


@pytest.mark.parametrize("heavy", [0, 48 << 10])
def test_serve_results_alone_and_in_mix(univ10_dev, univ10_oracle, univ10_store, heavy):
    """Continuous batching: each query's rows equal the oracle's (alone) and
    its checksum equals the per-query path's; with heavy = 48 KiB the
    queries projected above it run as batches on the second lane while the
    others' iterations go on."""
    stream = Q.template_stream(univ10_store, Q.univ_templates(), 400, seed=3)
    qs = [(p, nv) for p, nv, _ in stream] + [Q.lubm_q1(), Q.chain3(), Q.lubm_q7()]
    arr = np.cumsum(np.random.default_rng(1).exponential(1 / 20000.0, size=len(qs)))
    r = univ10_dev.serve(qs, arr, max_inflight=64, keep_results=True, heavy_footprint=heavy)
    assert (r["status"] == 0).all()
    for i, (pats, nv) in enumerate(qs):
        want, _, _ = univ10_oracle.query(pats, nv)
        got = univ10_dev.served_rows(i)
        assert same_rows(got, want), i
        assert int(r["checksum"][i]) == (wg.checksum(want) if want.size else 0)
    pq = univ10_dev.serve(qs, arr, per_query=True)
    assert np.array_equal(pq["checksum"], r["checksum"])
    assert np.array_equal(pq["rows"], r["rows"])
    assert r["batches"] < len(qs)


def test_serve_restarts_queries_that_do_not_fit_beside_others(univ10_store, univ10_oracle):
    """Served queries whose step tables do not fit beside the other in-flight
    queries' tables are restarted alone and come out exact; only a query that
    cannot fit by itself reports WK_E_NOMEM_BUDGET."""
    dev = wg.Device(univ10_store, table_budget=4 << 20)   # 149,796 elements per table
    tcap = dev.stats()["table_capacity_elems"]
    GS, MEM = Q.U("ub:GraduateStudent"), Q.U("ub:memberOf")
    depts = np.sort(univ10_store.get_index(Q.U("ub:Department"), wg.IN))[:8]
    qs = [([(GS, wg.TYPE_ID, -1, wg.IN), (-1, MEM, int(d), wg.OUT)], 1) for d in depts]
    assert 23151 * len(qs) > tcap and 23151 * 2 < tcap
    qs.insert(3, Q.type_start2())  # 220,425 x 2 elements alone: cannot fit
    arr = np.zeros(len(qs))
    r = dev.serve(qs, arr, max_inflight=64, footprint_budget=1 << 40, keep_results=True)
    for i, (pats, nv) in enumerate(qs):
        if i == 3:
            assert r["status"][i] == wg.WK_E_NOMEM_BUDGET
            continue
        assert r["status"][i] == 0, i
        want, _, _ = univ10_oracle.query(pats, nv)
        assert want.shape[0] > 0
        assert same_rows(dev.served_rows(i), want), i
        assert int(r["checksum"][i]) == wg.checksum(want)
    dev.close()


def test_known_to_known_on_long_edge_lists():
    """known_to_known membership tests over edge lists of 2,000-9,000 entries
    (11-13 levels of the binary search before the 4-entry tail), with probed
    ids equal to the list's first and last entries, between entries, and
    outside it."""
    rng = np.random.default_rng(21)
    V0 = 1 << 17
    B, D = 10, 11            # predicate ids
    hubs = V0 + np.arange(4, dtype=np.uint32)
    verts = V0 + 100 + np.arange(40000, dtype=np.uint32)
    rows = []
    lists = {}
    for j, h in enumerate(hubs):
        lst = np.sort(rng.choice(verts, size=2000 + 2300 * j, replace=False))
        lists[int(h)] = lst
        rows += [(h, B, v) for v in lst]
        probe_ids = np.concatenate([rng.choice(verts, size=1500), rng.choice(lst, size=1500),
                                    lst[:3], lst[-3:], lst[:3] - 1, lst[-3:] + 1, [verts[0], verts[-1]]])
        rows += [(h, D, int(v)) for v in np.unique(probe_ids)]
    tri = np.array(rows, dtype=np.uint32)
    tri = tri[rng.permutation(len(tri))]
    orc = oracle.Store(tri)
    dev = wg.Device(wg.Store(tri), table_budget=64 << 20)
    X, Y = -1, -2
    pats = [(D, wg.PREDICATE_ID, X, wg.IN), (X, D, Y, wg.OUT), (X, B, Y, wg.OUT)]
    want = check_query(dev, orc, pats, 2)
    # the expected rows from the lists themselves
    exp = 0
    for h in hubs:
        ds = tri[(tri[:, 0] == h) & (tri[:, 1] == D), 2]
        exp += int(np.isin(ds, lists[int(h)]).sum())
    assert want.shape[0] == exp and exp > 4000
    dev.close()


@pytest.mark.parametrize("extra", [3, 300])  # 300: the multi-block query scan
def test_batch_drops_only_the_final_tables_that_do_not_fit(univ10_store, univ10_oracle, extra):
    """Final tables past the result arena's capacity are dropped one query at a
    time (WK_E_NOMEM_BUDGET) while the batch runs once: the queries before
    them keep exact results and nothing is rerun."""
    dev = wg.Device(univ10_store, table_budget=4 << 20)  # 149,796 elements per table
    tcap = dev.stats()["table_capacity_elems"]
    GS = Q.U("ub:GraduateStudent")
    q = ([(GS, wg.TYPE_ID, -1, wg.IN)], 1)   # one pattern: 23,151 rows straight into the arena
    want, _, _ = univ10_oracle.query(*q)
    n_fit = tcap // want.size
    qs = [q] * (n_fit + extra)
    before = dev.stats()["queries_run"]
    res = dev.batch(qs, use_graph=False)
    assert dev.stats()["queries_run"] - before == len(qs)   # one pass, no halving
    for i, (t, v2c, rc) in enumerate(res):
        if i < n_fit:
            assert rc == 0 and same_rows(t, want), i
        else:
            assert rc == wg.WK_E_NOMEM_BUDGET, (i, rc)
    dev.close()


def test_profile_switch(univ10_store, univ10_oracle):
    """Per-kernel timing follows wk_dev_set_profile; results do not."""
    dev = wg.Device(univ10_store, table_budget=64 << 20, profile=1)
    pats, nv = Q.chain3()
    check_query(dev, univ10_oracle, pats, nv)
    assert any(k.startswith("k_probe") for k in dev.kernel_stats())
    dev.set_profile(False)
    dev.reset_kernel_stats()
    check_query(dev, univ10_oracle, pats, nv)
    assert dev.kernel_stats() == {}
    dev.set_profile(True)
    check_query(dev, univ10_oracle, pats, nv)
    assert dev.kernel_stats()["k_probe"]["launches"] == 3
    dev.close()


@pytest.mark.parametrize("rows,cols,nparts", [(0, 2, 3), (1, 1, 1), (1000, 3, 2), (1025, 2, 8), (5000, 4, 64),
                                              (70000, 3, 5), (3000, 7, 16), (2100, 11, 3), (40000, 13, 7),
                                              (200000, 2, 64)])
def test_partition_matches_oracle(univ10_dev, rows, cols, nparts):
    """Fork-join dispatch on the device equals the CPU engine's split row for
    row (same owners, same order within each owner), on ragged sizes around
    the 1,024-row tile, tables narrow enough to sort in LDS and wider ones
    (11 and 13 columns), and a histogram of several scan chunks (64 owners x 196
    tiles)."""
    rng = np.random.default_rng(rows + 31 * nparts)
    t = rng.integers(0, 1 << 24, size=(rows, cols), dtype=np.uint32)
    col = cols - 1
    want, wsizes = oracle.partition_rows(t, col, nparts)
    dt = univ10_dev.upload(t)
    out, sizes, heads = univ10_dev.partition(dt, col, nparts)
    assert out.shape == (rows, cols)
    assert np.array_equal(sizes, wsizes)
    assert np.array_equal(heads, np.concatenate([[0], np.cumsum(wsizes)[:-1]]).astype(np.uint64))
    if rows:
        assert np.array_equal(out.numpy(), want)
    out.free()
    dt.free()


def test_partition_rejects_bad_arguments(univ10_dev):
    dt = univ10_dev.upload(np.arange(12, dtype=np.uint32).reshape(4, 3))
    for col, n in [(3, 2), (0, 0), (0, 65)]:
        with pytest.raises(wg.WkError):
            univ10_dev.partition(dt, col, n)
    dt.free()


# ------------------------------------------------------------ final_process

def _fp_case(dev, t, v2c, required, **kw):
    want = oracle.final_process(t, v2c, required, **kw)
    dt = dev.upload(t)
    got = dev.final_process(dt, v2c, required, **kw)
    assert got.shape == want.shape, (got.shape, want.shape, kw)
    if want.size:
        assert np.array_equal(got.numpy(), want), kw
    got.free()
    dt.free()


@pytest.mark.parametrize("rows,cols,hi", [(1, 1, 4), (60, 3, 5), (1024, 2, 40), (1025, 4, 3), (5000, 3, 1 << 32),
                                          (70000, 5, 50)])
def test_final_process_matches_oracle(univ10_dev, rows, cols, hi):
    """Device final_process equals the oracle's row for row: projection,
    DISTINCT (signed full-row sort, words above 2^31 included), ORDER BY
    ASC/DESC on one and two keys with full-row tie order, OFFSET/LIMIT
    clamping, on small value ranges so that duplicates and ties abound."""
    rng = np.random.default_rng(rows + cols)
    t = rng.integers(0, hi, size=(rows, cols), dtype=np.uint64).astype(np.uint32)
    v2c = list(rng.permutation(cols)) + [-1]  # the last variable has no column
    v2c = [int(x) for x in v2c]
    allv = [-(i + 1) for i in range(cols)]
    last = -cols
    cases = [
        dict(required=allv),
        dict(required=[allv[0]], distinct=True),
        dict(required=allv[::-1], distinct=True),
        dict(required=[last, allv[0]], orders=[(last, True)]),
        dict(required=allv, orders=[(allv[0], False), (last, True)], offset=3, limit=7),
        dict(required=[allv[0]], distinct=True, orders=[(allv[0], True)], offset=1),
        dict(required=allv, distinct=True, orders=[(last, False)], limit=0),
        dict(required=allv, offset=rows + 5),
        dict(required=allv, distinct=True, offset=2, limit=1 << 40),
    ]
    for c in cases:
        req = c.pop("required")
        _fp_case(univ10_dev, t, v2c, req, **c)


def test_final_process_on_a_query_result(univ10_dev, univ10_oracle):
    """DISTINCT + ORDER BY + LIMIT over a real lubm_q1-shaped result."""
    pats, nv = Q.lubm_q1()
    t, v2c, rc = univ10_dev.query(pats, nv)
    assert rc == 0 and t.shape[0] > 0
    vars_ = [-(i + 1) for i in range(nv) if v2c[i] >= 0]
    _fp_case(univ10_dev, t, v2c, vars_[:2], distinct=True, orders=[(vars_[1], True), (vars_[0], False)], limit=100)
    _fp_case(univ10_dev, t, v2c, vars_, orders=[(vars_[-1], False)])


def test_final_process_empty_and_errors(univ10_dev):
    e = univ10_dev.upload(np.zeros((0, 3), np.uint32))
    out = univ10_dev.final_process(e, [0, 1, 2], [-1])
    assert out.shape == (0, 3)  # an empty result is left as it is
    out.free()
    e.free()
    dt = univ10_dev.upload(np.arange(12, dtype=np.uint32).reshape(4, 3))
    for req, orders, code in [([], (), wg.WK_E_NO_REQUIRED_VAR), ([-4], (), wg.WK_E_VERTEX_INVALID),
                              ([1], (), wg.WK_E_VERTEX_INVALID), ([-1], [(-4, False)], wg.WK_E_VERTEX_INVALID),
                              ([-1] * 65, (), wg.WK_E_TOO_MANY_VARS)]:
        with pytest.raises(wg.WkError) as ei:
            univ10_dev.final_process(dt, [0, 1, 2, -1], req, orders=orders)
        assert ei.value.code == code
    dt.free()


def test_direct_key_index_and_buckets_agree(univ10_store, univ10_oracle):
    """The same queries, per query and batched, with the direct key indexes
    of the dense segments (students' memberOf/takesCourse, rdf:type OUT, ...)
    and with every probe through the hash buckets: both equal the oracle."""
    stream = Q.template_stream(univ10_store, Q.univ_templates(), 300, seed=41)
    qs = [(p, nv) for p, nv, _ in stream] + [f() for f in (Q.chain3, Q.type_start2, Q.lubm_q1, Q.lubm_q7)]
    for dense in (0, -1):
        dev = wg.Device(univ10_store, table_budget=256 << 20, dense_index=dense)
        s = dev.stats()
        assert (s["dense_segments"] >= 4) if dense == 0 else (s["dense_segments"] == 0)
        for pats, nv in qs[-4:]:
            check_query(dev, univ10_oracle, pats, nv)
        for (pats, nv), (t, v2c, rc) in zip(qs, dev.batch(qs, use_graph=True)):
            want, wv, wrc = univ10_oracle.query(pats, nv)
            assert rc == wrc and v2c == wv and same_rows(t, want), (dense, pats)
        dev.close()


def _selective_filters(store):
    """Index starts followed by known_to_const filters that keep about 1/2,
    1/200 and 1/5,000 of their rows, and a known_to_known join: the gather's
    output tiles then span 2 to thousands of input rows each."""
    T = wg.TYPE_ID
    ug, gs = Q.U("ub:UndergraduateStudent"), Q.U("ub:GraduateStudent")
    depts = np.sort(store.get_index(Q.U("ub:Department"), wg.IN))
    courses = np.sort(store.get_index(Q.U("ub:Course"), wg.IN))
    univs = np.sort(store.get_index(Q.U("ub:University"), wg.IN))
    qs = []
    for d in depts[:3]:
        qs.append(([(ug, T, -1, wg.IN), (-1, Q.U("ub:memberOf"), int(d), wg.OUT)], 1))
    for c in courses[:3]:
        qs.append(([(ug, T, -1, wg.IN), (-1, Q.U("ub:takesCourse"), int(c), wg.OUT)], 1))
    for u in univs[:2]:
        qs.append(([(gs, T, -1, wg.IN), (-1, Q.U("ub:undergraduateDegreeFrom"), int(u), wg.OUT)], 1))
    # advisors who are full professors: a k2c by type bitmap
    qs.append(([(gs, T, -1, wg.IN), (-1, Q.U("ub:advisor"), -2, wg.OUT),
                (-2, T, Q.U("ub:FullProfessor"), wg.OUT)], 2))
    # courses taught by the student's advisor: a selective known_to_known join
    qs.append(([(gs, T, -1, wg.IN), (-1, Q.U("ub:advisor"), -2, wg.OUT), (-1, Q.U("ub:takesCourse"), -3, wg.OUT),
                (-2, Q.U("ub:teacherOf"), -3, wg.OUT)], 3))
    return qs


def test_selective_filters_alone_and_batched(univ10_dev, univ10_oracle, univ10_store):
    """Filters of every selectivity, alone and in a batch mixed with the
    template stream, row for row against the oracle."""
    qs = _selective_filters(univ10_store)
    wants = [univ10_oracle.query(p, nv) for p, nv in qs]
    kept = [w[0].shape[0] for w in wants]
    assert min(kept) > 0 and max(kept) > 1_000
    for (pats, nv), (want, wv, wrc) in zip(qs, wants):
        got, gv, grc = univ10_dev.query(pats, nv)
        assert grc == wrc == 0 and gv == wv
        assert same_rows(got, want), pats
    stream = Q.template_stream(univ10_store, Q.univ_templates(), 200, seed=41)
    mix = [(p, nv) for p, nv, _ in stream]
    mix = mix[:100] + qs + mix[100:] + qs[::-1]
    for rep in range(2):
        for (pats, nv), (t, v2c, rc) in zip(mix, univ10_dev.batch(mix, use_graph=bool(rep))):
            want, wv, wrc = univ10_oracle.query(pats, nv)
            assert rc == wrc == 0 and v2c == wv
            assert same_rows(t, want), pats
