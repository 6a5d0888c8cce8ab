"""Host gstore of the product path vs the oracle, and the store layout
invariants the device probe relies on (CPU only)."""
import re
import os

import numpy as np
import pytest

import oracle
import wukong_gpu as wg
from _graphs import VID0, random_graph

IN, OUT, TYPE_ID = 0, 1, 1
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def slot_fields(slots):
    key, ptr = slots[:, 0], slots[:, 1]
    return key, ptr & ((1 << 28) - 1), (ptr >> 28) & ((1 << 34) - 1)


def check_store(st, orc, triples):
    """Every key's edge list equals the oracle's; every (s,p)/(o,p) of the
    data is a key; index lists equal as sets; layout invariants hold."""
    slots, vals = st.raw()
    segs = st.segments()
    seen = 0
    for s in segs:
        b0, nb, ne = s["bucket_start"], s["num_buckets"], s["num_ext_buckets"]
        sl = slots[b0 * 8:(b0 + nb + ne) * 8]
        key, size, off = slot_fields(sl)
        data = np.ones(len(sl), bool)
        data[7::8] = False
        used = data & (key != 0)
        assert used.sum() == s["num_keys"]
        assert (off[used] >= s["edge_start"]).all()
        assert (off[used] + size[used] <= s["edge_start"] + s["num_edges"]).all()
        assert size[used].sum() == s["num_edges"]
        # every key lives in its segment and is reachable from its hash bucket
        kk = key[used]
        assert (((kk >> 18) == 0) == bool(s["index"])).all()
        if not s["index"]:
            assert (((kk >> 1) & ((1 << 17) - 1)) == s["pid"]).all()
        assert ((kk & 1) == s["dir"]).all()
        links = key[7::8]
        lk = links[links != 0] >> 18
        assert ((lk >= nb) & (lk < nb + ne)).all()
        for k in kk[:: max(1, len(kk) // 300)]:
            vid, pid, d = int(k) >> 18, (int(k) >> 1) & ((1 << 17) - 1), int(k) & 1
            a = st.get_triples(vid, pid, d)
            b = orc.get_triples(vid, pid, d)
            assert np.array_equal(np.sort(a), np.sort(b)), (vid, pid, d)
            assert (np.diff(a.astype(np.int64)) > 0).all()  # sorted, unique
        seen += len(kk)
    assert seen == st.info()["num_keys"]
    # data-driven: each (s,p) has an OUT key, each (o,p) with a vertex object an IN key
    t = np.unique(triples, axis=0)
    for s_, p_, o_ in t[:: max(1, len(t) // 400)]:
        assert int(o_) in set(st.get_triples(int(s_), int(p_), OUT).tolist())
        if o_ >= VID0:
            assert int(s_) in set(st.get_triples(int(o_), int(p_), IN).tolist())
        else:
            assert len(st.get_triples(int(o_), int(p_), IN)) == 0
    # absent keys
    assert len(st.get_triples(VID0 + 10**8, int(t[0, 1]), OUT)) == 0
    assert len(st.get_triples(int(t[0, 0]), 60000, OUT)) == 0


def test_store_config1(univ10, univ10_oracle):
    st = wg.Store(univ10)
    check_store(st, univ10_oracle, univ10)
    info = st.info()
    assert info["num_triples_out"] == univ10_oracle.num_out()
    types = [wg.univ_id(n) for n in ("ub:GraduateStudent", "ub:Course", "ub:University", "ub:Department")]
    preds = [wg.univ_id(n) for n in ("ub:memberOf", "ub:takesCourse", "ub:advisor")]
    for t in types:
        assert np.array_equal(np.sort(st.get_index(t, IN)), np.sort(univ10_oracle.get_index(t, IN)))
    for p in preds:
        for d in (IN, OUT):
            assert np.array_equal(np.sort(st.get_index(p, d)), np.sort(univ10_oracle.get_index(p, d)))


@pytest.mark.parametrize("seed,lf", [(0, 0), (1, 95), (2, 99), (3, 30)])
def test_store_random(seed, lf):
    tri, info = random_graph(seed, tpid_objects=True)
    st = wg.Store(tri, load_factor_pct=lf)
    check_store(st, oracle.Store(tri), tri)
    if lf >= 95:  # high load factor: clusters spill into chained buckets
        assert sum(s["num_ext_buckets"] for s in st.segments()) > 0


def test_store_build_is_deterministic():
    tri, _ = random_graph(5)
    a, b = wg.Store(tri), wg.Store(tri[::-1].copy())
    sa, va = a.raw()
    sb, vb = b.raw()
    assert np.array_equal(sa, sb) and np.array_equal(va, vb)


def test_store_rejects_invalid_triples():
    for bad in ([(VID0, 0, VID0)], [(VID0, TYPE_ID, VID0 + 5)], [(1, 5, VID0)], [(VID0, 1 << 17, VID0)]):
        with pytest.raises(wg.WkError) as e:
            wg.Store(np.array(bad, np.uint32))
        assert e.value.code == wg.WK_E_INVALID_TRIPLE


def test_generator_deterministic_and_sized():
    a = wg.gen_university(1, 2)
    b = wg.gen_university(1, 2)
    c = wg.gen_university(2, 2)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert 200_000 < len(a) < 330_000            # ~133K triples per university
    p = set(np.unique(a[:, 1]).tolist())
    assert len(p) == 18                          # 18 predicates incl. rdf:type
    assert len(np.unique(a[a[:, 1] == TYPE_ID][:, 2])) == 15
    assert a[:, [0, 2]].min() >= 2 and (a[:, 1] < (1 << 17)).all()


def test_powerlaw_generator():
    t = wg.gen_powerlaw(7, 20000, 200000, n_preds=16, n_types=8)
    assert abs(len(t) - 200000) < 20000
    deg = np.bincount(t[t[:, 1] != TYPE_ID][:, 0] - VID0)
    assert deg.max() > 20 * max(1.0, deg.mean())  # heavy tail
    pc = np.bincount(t[t[:, 1] != TYPE_ID][:, 1])
    assert pc[2] > pc[2 + 8]                      # Zipf head
    assert np.array_equal(t, wg.gen_powerlaw(7, 20000, 200000, n_preds=16, n_types=8))


def test_checksum_is_order_independent():
    rng = np.random.default_rng(0)
    t = rng.integers(0, 2**32, size=(500, 4), dtype=np.uint64).astype(np.uint32)
    assert wg.checksum(t) == wg.checksum(t[rng.permutation(500)])
    u = t.copy()
    u[3, 2] ^= 1
    assert wg.checksum(u) != wg.checksum(t)
    assert wg.checksum(t[:, [1, 0, 2, 3]]) != wg.checksum(t)


def test_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(REPO, "include", "wk_gpu.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|uint32_t|uint64_t|const char\*)\s+(wk_\w+)\(", hdr, re.M))
    assert len(declared) >= 40
    lib = wg.lib()
    for name in declared:
        assert hasattr(lib, name), name
    assert set(wg.SIGNATURES) == declared


def test_abi_errors_without_gpu():
    if wg.device_count() > 0:
        pytest.skip("a GPU is present")
    st = wg.Store(np.array([(VID0, 5, VID0 + 1)], np.uint32))
    with pytest.raises(wg.WkError) as e:
        wg.Device(st)
    assert e.value.code == wg.WK_E_NO_DEVICE
    assert wg.lib().wk_strerror(wg.WK_E_NOMEM_BUDGET) == b"device budget exceeded"
