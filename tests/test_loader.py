"""Wukong ID-triple directories (loader/base_loader.hpp:536-576,
loader/posix_loader.hpp:49-69) through wk_store_load_dir, on the CPU.

A directory's store must equal the store wk_store_build makes from the
triples that `while (file >> s >> p >> o)` reads.  Here that reading is
restated with a regular expression: one unsigned number per read, after
optional whitespace, and reading stops at the first read that fails."""
import os
import re

import numpy as np
import pytest

import wukong_gpu as wg
from _graphs import random_graph

NUM = re.compile(r"\s*\+?(\d+)")


def istream_triples(text):
    vals, pos = [], 0
    while True:
        m = NUM.match(text, pos)
        if not m or int(m.group(1)) >= 1 << 32:
            break
        vals.append(int(m.group(1)))
        pos = m.end()
    vals = vals[:len(vals) // 3 * 3]
    return np.array(vals, dtype=np.uint32).reshape(-1, 3)


def same_store(a, b):
    ia, ib = a.info(), b.info()
    assert ia == ib
    sa, va = a.raw()
    sb, vb = b.raw()
    assert np.array_equal(sa, sb) and np.array_equal(va, vb)
    assert a.segments() == b.segments()


def test_directory_equals_built_store(tmp_path):
    tri, _ = random_graph(3, n_v=300, n_edges=4000)
    d = str(tmp_path)
    wg.write_id_triples(d, tri, nfiles=3)
    # entries the loader must ignore: no id_ prefix, hidden, attribute file
    (tmp_path / "str_index").write_text("<p0> 2\n<p1> 3\n")
    (tmp_path / ".id_hidden").write_text("1 2 3\n")
    (tmp_path / "attr_0.nt").write_text("5 6 1 7\n")
    (tmp_path / "notes_id_x").write_text("1 2 3\n")
    st, n = wg.Store.from_dir(d)
    assert n == len(tri)
    same_store(st, wg.Store(tri))
    assert wg.load_str_index(d) == {"<p0>": 2, "<p1>": 3}


def test_reading_stops_like_operator_extraction(tmp_path):
    # vertex ids are >= 2^17 (NBITS_IDX); predicates 2 and 3
    texts = {
        "id_a.nt": "200010 2 200011\t\n 200012  3   200013\r\n200014 2 200015 200016 2\n",  # incomplete triple
        "id_b.nt": "200020 2 200021 200022 2 x 200023 2 200024\n",                 # stops at x
        "id_c.nt": "200030 2 200031 200032 3 4294967296 200033 2 200034\n",        # 2^32 fails the read
        "id_d.nt": "200040 2 200041+200042 3 200043 200044 2 200045abc 200046 2 200047\n",  # +n is read
        "id_e.nt": "",
    }
    want = []
    for name, t in sorted(texts.items()):
        (tmp_path / name).write_text(t)
        want.append(istream_triples(t))
    want = np.concatenate(want)
    assert want.tolist() == [[200010, 2, 200011], [200012, 3, 200013], [200014, 2, 200015],
                             [200020, 2, 200021], [200030, 2, 200031], [200040, 2, 200041],
                             [200042, 3, 200043], [200044, 2, 200045]]
    st, n = wg.Store.from_dir(str(tmp_path))
    assert n == len(want)
    same_store(st, wg.Store(want))


def test_server_shares(tmp_path):
    tri, _ = random_graph(4, n_v=200, n_edges=3000)
    wg.write_id_triples(str(tmp_path), tri, nfiles=2)
    for nsrv in (2, 3):
        for sid in range(nsrv):
            st, n = wg.Store.from_dir(str(tmp_path), server=sid, nservers=nsrv)
            assert n == len(tri)
            pso = tri[tri[:, 0] % nsrv == sid]
            pos = tri[tri[:, 2] % nsrv == sid]
            same_store(st, wg.Store(pso, pos))


def test_errors(tmp_path):
    with pytest.raises(wg.WkError) as e:
        wg.Store.from_dir(str(tmp_path / "missing"))
    assert e.value.code == wg.WK_E_FILE_NOT_FOUND
    for sid, nsrv in [(1, 1), (0, 0)]:
        with pytest.raises(wg.WkError) as e:
            wg.Store.from_dir(str(tmp_path), server=sid, nservers=nsrv)
        assert e.value.code == wg.WK_E_INVALID_ARG
    st, n = wg.Store.from_dir(str(tmp_path))  # no id_ files: an empty store
    assert n == 0 and st.info()["num_triples_out"] == 0
