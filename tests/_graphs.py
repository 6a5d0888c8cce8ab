"""Seeded random graphs and random planned queries for parity tests."""
import numpy as np

VID0 = 1 << 17
TYPE_ID = 1
IN, OUT = 0, 1


def random_graph(seed, n_v=400, n_types=4, n_preds=6, n_edges=3000, tpid_objects=False, skew=True):
    """(triples (n,3) uint32, info dict).  Types are ids 2..; predicates after."""
    rng = np.random.default_rng(seed)
    types = np.arange(2, 2 + n_types, dtype=np.uint32)
    preds = np.arange(2 + n_types, 2 + n_types + n_preds, dtype=np.uint32)
    verts = np.arange(VID0, VID0 + n_v, dtype=np.uint32)
    rows = []
    for v in verts:
        for t in rng.choice(types, size=int(rng.integers(1, 3)), replace=False):
            rows.append((v, TYPE_ID, t))
    if skew:
        w = 1.0 / np.arange(1, n_v + 1) ** 0.9
        w /= w.sum()
        s = rng.choice(verts, size=n_edges, p=w)
        o = rng.choice(verts, size=n_edges, p=w[::-1])
    else:
        s = rng.choice(verts, size=n_edges)
        o = rng.choice(verts, size=n_edges)
    p = rng.choice(preds, size=n_edges)
    for a, b, c in zip(s, p, o):
        rows.append((a, b, c))
    # duplicates exercise the loader's de-duplication
    for i in rng.integers(0, len(rows), size=len(rows) // 20):
        rows.append(rows[int(i)])
    if tpid_objects:
        for _ in range(50):
            rows.append((int(rng.choice(verts)), int(preds[0]), int(rng.choice(types))))
    t = np.array(rows, dtype=np.uint32)
    t = t[rng.permutation(len(t))]
    return t, {"types": types, "preds": preds, "verts": verts}


def random_query(rng, info, max_steps=5, allow_invalid=True):
    """A planned pattern list (s,p,o,dir) and nvars."""
    types, preds, verts = info["types"], info["preds"], info["verts"]
    nvars = 0
    known = []
    pats = []

    def newvar():
        nonlocal nvars
        nvars += 1
        return -nvars

    kind = rng.integers(4)
    if kind == 0:  # type index start
        x = newvar()
        pats.append((int(rng.choice(types)), TYPE_ID, x, IN))
        known.append(x)
    elif kind == 1:  # predicate index start
        x = newvar()
        pats.append((int(rng.choice(preds)), 0, x, int(rng.integers(2))))
        known.append(x)
    elif kind == 2:  # const start
        x = newvar()
        pats.append((int(rng.choice(verts)), int(rng.choice(preds)), x, int(rng.integers(2))))
        known.append(x)
    else:  # types of a vertex, then instances of those types (TYPE_ID IN expansion)
        t, x = newvar(), newvar()
        pats.append((int(rng.choice(verts)), TYPE_ID, t, OUT))
        pats.append((t, TYPE_ID, x, IN))
        known += [t, x]
    steps = int(rng.integers(1, max_steps))
    for _ in range(steps):
        op = rng.integers(10)
        v = int(rng.choice(known))
        if op < 5:  # k2u
            y = newvar()
            pats.append((v, int(rng.choice(preds)), y, int(rng.integers(2))))
            known.append(y)
        elif op < 7:  # k2c
            if rng.integers(2):
                pats.append((v, TYPE_ID, int(rng.choice(types)), OUT))
            else:
                pats.append((v, int(rng.choice(preds)), int(rng.choice(verts)), int(rng.integers(2))))
        elif op < 9:  # k2k
            w = int(rng.choice(known))
            pats.append((v, int(rng.choice(preds)), w, int(rng.integers(2))))
        else:  # const_to_known
            w = int(rng.choice(known))
            pats.append((int(rng.choice(verts)), int(rng.choice(preds)), w, int(rng.integers(2))))
    if allow_invalid and rng.integers(12) == 0:
        bad = rng.integers(5)
        if bad == 0:
            pats.append((-(nvars + 1), int(rng.choice(preds)), -(nvars + 2), OUT))  # unknown start
            nvars += 2
        elif bad == 1:
            pats.append((int(rng.choice(verts)), int(rng.choice(preds)), int(rng.choice(verts)), OUT))
        elif bad == 2:
            pats.append((int(rng.choice(verts)), int(rng.choice(preds)), -(nvars + 1), OUT))  # 2nd const start
            nvars += 1
        elif bad == 3:
            pats.append((known[0], int(rng.choice(preds)), -(nvars + 5), OUT))  # var beyond nvars
        else:
            pats.insert(0, (int(rng.choice(types)), int(rng.choice(preds)), -(nvars + 1), IN))  # index start, bad p
            nvars += 1
    return pats, nvars


def same_rows(a, b):
    """Multiset equality of uint32 rows."""
    a = np.asarray(a, dtype=np.uint32)
    b = np.asarray(b, dtype=np.uint32)
    if a.shape != b.shape:
        return False
    if a.size == 0:
        return True
    ka = a[np.lexsort(a.T[::-1])]
    kb = b[np.lexsort(b.T[::-1])]
    return np.array_equal(ka, kb)
