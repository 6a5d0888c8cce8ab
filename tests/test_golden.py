"""The committed config-1 fixtures (tests/golden/config1.json, written by
tests/golden/make_golden.py) against the generator, the host store and the
oracle of this tree: any drift in one of them shows here."""
import json
import os

import pytest

import oracle
import wukong_gpu as wg

from golden.make_golden import SHAPES, triples_digest

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "config1.json")))


def test_generator_and_store_shape(univ10):
    g = GOLDEN["graph"]
    assert len(univ10) == g["triples"]
    assert triples_digest(univ10) == g["sha256_sorted_spo"]
    info = wg.Store(univ10).info()
    for k, v in GOLDEN["store"].items():
        assert info[k] == v, k


@pytest.mark.parametrize("name", SHAPES)
def test_oracle_matches_fixture(univ10_oracle, name):
    want = GOLDEN["queries"][name]
    rows, v2c, rc = univ10_oracle.query([tuple(p) for p in want["patterns"]], want["nvars"])
    assert rc == want["status"]
    assert rows.shape[0] == want["rows"] and list(v2c) == want["v2c"]
    assert str(wg.checksum(rows) if rows.size else 0) == want["checksum"]
