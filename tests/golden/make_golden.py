"""Write tests/golden/config1.json: fixtures of the 10-university graph.

For the seeded university graph (seed 1, 10 universities) this records the
generator's output digest, the host store's shape, and for each query shape
(plan-ordered patterns) the oracle's row count, var->column map and the
order-independent row checksum (sum of per-row hashes, the same function
the serving path computes on the device).  The oracle is the CPU restatement
of the reference engine (core/engine/sparql.hpp:197-552).

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [REPO, os.path.join(REPO, "wukong-cube_amd")]

import numpy as np  # noqa: E402

import oracle  # noqa: E402
import wukong_gpu as wg  # noqa: E402
from wukong_gpu import queries as Q  # noqa: E402

SHAPES = ("chain3", "type_start2", "lubm_q1", "lubm_q7")
OUT = os.path.join(HERE, "config1.json")


def triples_digest(tri):
    """sha256 of the (s, p, o) rows sorted lexicographically (order-free)."""
    t = np.ascontiguousarray(tri, dtype=np.uint32)
    order = np.lexsort((t[:, 2], t[:, 1], t[:, 0]))
    return hashlib.sha256(np.ascontiguousarray(t[order]).tobytes()).hexdigest()


def compute():
    tri = wg.gen_university(1, 10)
    st = wg.Store(tri)
    info = st.info()
    orc = oracle.Store(tri)
    out = {
        "graph": {"generator": "university", "seed": 1, "universities": 10, "triples": int(len(tri)),
                  "sha256_sorted_spo": triples_digest(tri)},
        "store": {k: int(info[k]) for k in ("num_triples_out", "num_triples_in", "num_keys", "num_values",
                                            "num_segments")},
        "queries": {},
    }
    for name in SHAPES:
        pats, nv = getattr(Q, name)()
        rows, v2c, rc = orc.query(pats, nv)
        out["queries"][name] = {"patterns": [list(map(int, p)) for p in pats], "nvars": nv, "status": int(rc),
                                "rows": int(rows.shape[0]), "cols": int(rows.shape[1]) if rows.ndim == 2 else 0,
                                "v2c": [int(x) for x in v2c],
                                "checksum": str(wg.checksum(rows) if rows.size else 0)}
    return out


if __name__ == "__main__":
    data = compute()
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote", OUT)
