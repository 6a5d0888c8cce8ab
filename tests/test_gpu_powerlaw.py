"""Power-law graphs on the device (the shape of the 1.2B / 2.5B-triple
configurations at a size the oracle covers): the generator with seed 7,
Zipf(1.1) predicates, Pareto(2) out-degree and 64 types; star, 5-hop path
and snowflake templates with constant starts drawn in proportion to degree.
Rows are compared with the oracle row for row (as sorted multisets), per
query, batched, batched under a gstore budget that swaps segments, and served
as a 95% light / 5% heavy mix under that budget.

Every device call goes through the C ABI; the oracle is only the checker."""
import numpy as np
import pytest

import oracle
import wukong_gpu as wg
from wukong_gpu import queries as Q

from _graphs import same_rows

pytestmark = pytest.mark.gpu

N_VERTICES, N_TRIPLES = 1_000_000, 12_000_000


@pytest.fixture(scope="module")
def pl():
    tri = wg.gen_powerlaw(7, N_VERTICES, N_TRIPLES)
    store = wg.Store(tri, pinned=True)
    return tri, oracle.Store(tri), store


@pytest.fixture(scope="module")
def pl_stream(pl):
    tri, orc, _ = pl
    stream = Q.powerlaw_stream(tri, Q.powerlaw_templates(), 600, seed=107)
    qs = [(p, nv) for p, nv, _ in stream]
    wants = [orc.query(p, nv) for p, nv in qs]
    return qs, [k for _, _, k in stream], wants


def _budget(store, qs):
    """A gstore budget (in 128-KiB key and 64-KiB value blocks) below the
    segments the queries read: half of them, but room for the largest twice.
    Returns (budget, bytes of the segments read)."""
    segs = {(s["index"], s["pid"], s["dir"]): s for s in store.segments()}
    kb, vb = 1 << 10, 1 << 14  # buckets / edges per block
    blocks = lambda s: (-(-(s["num_buckets"] + s["num_ext_buckets"]) // kb) * kb * 128 +
                        -(-s["num_edges"] // vb) * vb * 4)
    used = {(0, p, d) for pats, _ in qs for (s, p, o, d) in pats} & set(segs)
    sizes = [blocks(segs[u]) for u in used]
    budget = max(sum(sizes) // 2, 2 * max(sizes))
    assert budget < sum(sizes)
    return budget, sum(sizes)


def test_powerlaw_stream_shape(pl_stream):
    """The stream mixes all three templates, empty and heavy results."""
    qs, kinds, wants = pl_stream
    rows = np.array([w[0].shape[0] for w in wants])
    assert all(w[2] == 0 for w in wants)
    assert set(kinds) == {0, 1, 2}
    assert (rows == 0).any() and (rows <= 100).sum() > 400 and rows.max() >= 100_000


def test_powerlaw_per_query(pl, pl_stream):
    _, _, store = pl
    qs, _, wants = pl_stream
    dev = wg.Device(store, table_budget=256 << 20)
    dev.load_all()
    for (pats, nv), (want, wv, wrc) in zip(qs[:150], wants[:150]):
        got, gv, grc = dev.query(pats, nv)
        assert grc == wrc == 0 and gv == wv
        assert same_rows(got, want), pats
    dev.close()


@pytest.mark.parametrize("pw,dense", [(1, 0), (2, 0), (8, 0), (2, -1)])
def test_powerlaw_batch_matches_oracle(pl, pl_stream, pw, dense):
    """The 600-query stream in one batch (one set of launches per pattern
    step over a concatenated table), replayed from a hipGraph three times,
    with and without direct key indexes."""
    _, _, store = pl
    qs, _, wants = pl_stream
    dev = wg.Device(store, table_budget=512 << 20, probe_width=pw, dense_index=dense)
    dev.load_all()
    for rep in range(3):
        res = dev.batch(qs, use_graph=True)
        for (pats, nv), (t, v2c, rc), (want, wv, _) in zip(qs, res, wants):
            assert rc == 0, (pats, rc)
            assert v2c == wv
            assert same_rows(t, want), pats
    dev.close()


def test_powerlaw_batch_under_swapping_budget(pl, pl_stream):
    """A gstore budget of half the segments the stream reads: the batch's
    segments cannot all be resident, so it runs in parts that each fit,
    evicting and reloading segments, and every result stays exact."""
    _, _, store = pl
    qs, _, wants = pl_stream
    budget, total = _budget(store, qs)
    dev = wg.Device(store, gstore_budget=budget, key_block_shift=10, value_block_shift=14,
                    table_budget=512 << 20)
    res = dev.batch(qs, use_graph=False)
    for (pats, nv), (t, v2c, rc), (want, wv, _) in zip(qs, res, wants):
        assert rc == 0, (pats, rc)
        assert same_rows(t, want), pats
    res = dev.batch(qs[::-1], use_graph=True)
    for (pats, nv), (t, v2c, rc), (want, wv, _) in zip(qs[::-1], res, wants[::-1]):
        assert rc == 0 and same_rows(t, want), pats
    s = dev.stats()
    assert s["segments_evicted"] > 0 and s["h2d_bytes"] > total
    dev.close()


@pytest.mark.parametrize("heavy_fp", [0, 1 << 20, 64 << 10])
def test_powerlaw_serve_light_heavy_mix_swapping(pl, pl_stream, heavy_fp):
    """Serving a 95% light (<= 100 rows) / 5% heavy (the largest results) mix
    at a high offered rate under a swapping gstore budget: each served result
    equals the oracle's rows, and its checksum equals the per-query path's
    (the same query alone).  With the lower heavy thresholds, queries whose
    tables grow past the threshold continue on the second lane from the
    table made on the main lane."""
    _, _, store = pl
    qs, _, wants = pl_stream
    rows = np.array([w[0].shape[0] for w in wants])
    light = [i for i in np.argsort(rows, kind="stable") if rows[i] <= 100][:380]
    heavy = [int(i) for i in np.argsort(rows)[::-1][:20]]
    assert rows[heavy].min() >= 1000 and len(light) == 380
    rng = np.random.default_rng(5)
    order = rng.permutation(np.array(light + heavy))
    mix = [qs[i] for i in order]
    mix_want = [wants[i] for i in order]
    budget, total = _budget(store, mix)
    dev = wg.Device(store, gstore_budget=budget, key_block_shift=10, value_block_shift=14,
                    table_budget=512 << 20)
    arr = np.cumsum(rng.exponential(1 / 50000.0, size=len(mix)))
    dev.set_profile(True)
    r = dev.serve(mix, arr, max_inflight=256, keep_results=True, heavy_footprint=heavy_fp)
    ks = dev.kernel_stats()
    dev.set_profile(False)
    assert (r["status"] == 0).all(), np.unique(r["status"])
    if heavy_fp:
        assert ks["serve:heavy_continued"]["launches"] > 0
    for i, (want, _, _) in enumerate(mix_want):
        assert int(r["rows"][i]) == want.shape[0], i
        assert same_rows(dev.served_rows(i), want), i
        assert int(r["checksum"][i]) == (wg.checksum(want) if want.size else 0), i
    pq = dev.serve(mix, arr, per_query=True)
    assert np.array_equal(pq["checksum"], r["checksum"])
    assert np.array_equal(pq["rows"], r["rows"])
    assert r["batches"] < len(mix)
    assert dev.stats()["segments_evicted"] > 0
    dev.close()
