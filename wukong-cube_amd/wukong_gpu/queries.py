"""Query shapes, plans and seeded query streams for the benchmark configurations.

Patterns are (subject, predicate, object, direction) with variables -1, -2, ...
Raw patterns are written as in the SPARQL text (direction OUT) and rewritten
by a plan with the semantics of Planner::set_direction
(optimizer/planner.hpp:667-699): "<" swaps subject/object and reads the IN
edges, ">" reads OUT, "<<"/">>" turn the pattern into a predicate-index start.
"""
import numpy as np

from . import IN, OUT, PREDICATE_ID, TYPE_ID, univ_id


def set_plan(raw, plan):
    """Rewrite raw patterns by a .fmt plan [(1-based order, dir), ...]."""
    out = []
    for order, d in plan:
        s, p, o, _ = raw[order - 1]
        if d == "<":
            out.append((o, p, s, IN))
        elif d == ">":
            out.append((s, p, o, OUT))
        elif d == "<<":
            out.append((p, PREDICATE_ID, s, IN))
        elif d == ">>":
            out.append((p, PREDICATE_ID, o, OUT))
        else:
            raise ValueError("bad plan direction %r" % d)
    return out


def parse_fmt(text):
    """Pattern order/direction lines of a .fmt plan file (comments skipped)."""
    plan = []
    for line in text.splitlines():
        line = line.strip()
        if not line or line.startswith("#") or line == "{":
            continue
        if line == "}":
            break
        parts = line.split()
        plan.append((int(parts[0]), parts[1] if len(parts) > 1 else ">"))
    return plan


def U(name):
    v = univ_id(name)
    if not v:
        raise KeyError(name)
    return v


# ----------------------------------------------------------- university graph

def chain3():
    """?X type GraduateStudent . ?X memberOf ?D . ?D subOrganizationOf ?U; plan 1< 2> 3>."""
    raw = [(-1, TYPE_ID, U("ub:GraduateStudent"), OUT), (-1, U("ub:memberOf"), -2, OUT),
           (-2, U("ub:subOrganizationOf"), -3, OUT)]
    return set_plan(raw, [(1, "<"), (2, ">"), (3, ">")]), 3


def lubm_q7():
    """6-pattern cyclic query, scripts/sparql_query/lubm/basic/lubm_q7 with the
    OSDI16 plan (osdi16_plan/lubm_q7.fmt: 1< 2< 3> 4> 5> 6<); X=-1 Y=-2 Z=-3."""
    raw = [(-2, TYPE_ID, U("ub:FullProfessor"), OUT), (-1, U("ub:advisor"), -2, OUT),
           (-1, TYPE_ID, U("ub:UndergraduateStudent"), OUT), (-1, U("ub:takesCourse"), -3, OUT),
           (-3, TYPE_ID, U("ub:Course"), OUT), (-2, U("ub:teacherOf"), -3, OUT)]
    return set_plan(raw, [(1, "<"), (2, "<"), (3, ">"), (4, ">"), (5, ">"), (6, "<")]), 3


def type_start2():
    """?X type UndergraduateStudent . ?X takesCourse ?C; plan 1< 2>."""
    raw = [(-1, TYPE_ID, U("ub:UndergraduateStudent"), OUT), (-1, U("ub:takesCourse"), -2, OUT)]
    return set_plan(raw, [(1, "<"), (2, ">")]), 2


def lubm_q1():
    """scripts/sparql_query/lubm/basic/lubm_q1 with the OSDI16 plan (3< 4> 2> 1> 5< 6>)."""
    X, Y, Z = -1, -2, -3
    raw = [(Y, TYPE_ID, U("ub:University"), OUT), (X, U("ub:undergraduateDegreeFrom"), Y, OUT),
           (X, TYPE_ID, U("ub:GraduateStudent"), OUT), (X, U("ub:memberOf"), Z, OUT),
           (Z, U("ub:subOrganizationOf"), Y, OUT), (Z, TYPE_ID, U("ub:Department"), OUT)]
    return set_plan(raw, [(3, "<"), (4, ">"), (2, ">"), (1, ">"), (5, "<"), (6, ">")]), 3


# Emulator-style templates with one constant start (%Type placeholder filled
# from the type index, as ConsoleProxy::fill_template does,
# src/client/console_proxy.hpp:50-114).  Each: (placeholder type, raw patterns
# with the constant as None, plan, nvars).
def univ_templates():
    T = TYPE_ID
    return [
        # q1: ?X takesCourse %GraduateCourse . ?X type GraduateStudent
        ("ub:GraduateCourse", [(-1, U("ub:takesCourse"), None, OUT), (-1, T, U("ub:GraduateStudent"), OUT)],
         [(1, "<"), (2, ">")], 1),
        # q4: ?X worksFor %Department . ?X type FullProfessor . name/email/telephone
        ("ub:Department", [(-1, U("ub:worksFor"), None, OUT), (-1, T, U("ub:FullProfessor"), OUT),
                           (-1, U("ub:name"), -2, OUT), (-1, U("ub:emailAddress"), -3, OUT),
                           (-1, U("ub:telephone"), -4, OUT)],
         [(1, "<"), (2, ">"), (3, ">"), (4, ">"), (5, ">")], 4),
        # q6: ?X subOrganizationOf %University . ?X type Department . ?Y worksFor ?X . ?Y type FullProfessor
        ("ub:University", [(-1, U("ub:subOrganizationOf"), None, OUT), (-1, T, U("ub:Department"), OUT),
                           (-2, U("ub:worksFor"), -1, OUT), (-2, T, U("ub:FullProfessor"), OUT)],
         [(1, "<"), (2, ">"), (3, "<"), (4, ">")], 2),
        # ?X memberOf %Department . ?X type UndergraduateStudent . ?X takesCourse ?C
        ("ub:Department", [(-1, U("ub:memberOf"), None, OUT), (-1, T, U("ub:UndergraduateStudent"), OUT),
                           (-1, U("ub:takesCourse"), -2, OUT)],
         [(1, "<"), (2, ">"), (3, ">")], 2),
        # ?X undergraduateDegreeFrom %University . ?X type GraduateStudent
        ("ub:University", [(-1, U("ub:undergraduateDegreeFrom"), None, OUT),
                           (-1, T, U("ub:GraduateStudent"), OUT)],
         [(1, "<"), (2, ">")], 1),
    ]


def instantiate(raw, const):
    return [(s if s is not None else const, p, o if o is not None else const, d) for s, p, o, d in raw]


def template_stream(store, templates, n, seed):
    """n queries: template uniform, constant uniform over the placeholder type's
    instances (type index [0|t|IN]).  Returns [(patterns, nvars, template_id)]."""
    rng = np.random.default_rng(seed)
    cands = [np.sort(store.get_index(U(t), IN)) for t, _, _, _ in templates]
    out = []
    for _ in range(n):
        k = int(rng.integers(len(templates)))
        t, raw, plan, nv = templates[k]
        c = int(cands[k][rng.integers(len(cands[k]))])
        out.append((set_plan(instantiate(raw, c), plan), nv, k))
    return out


# ------------------------------------------------------------ power-law graph

def powerlaw_templates(n_preds=256, pred_base=2):
    """star (4 patterns on one subject), 5-hop path, snowflake (8 patterns with
    two known_to_known joins); predicates are the most frequent (Zipf head).
    The snowflake's joins check two expanded edges again from their far end
    (IN direction), so they hold for every row while still probing the IN
    segments and testing membership on the device."""
    p = [pred_base + i for i in range(n_preds)]
    star = ([(-1, p[0], None, OUT), (-1, p[1], -2, OUT), (-1, p[2], -3, OUT), (-1, p[3], -4, OUT)],
            [(1, "<"), (2, ">"), (3, ">"), (4, ">")], 4)
    path = ([(None, p[0], -1, OUT), (-1, p[1], -2, OUT), (-2, p[0], -3, OUT), (-3, p[2], -4, OUT),
             (-4, p[1], -5, OUT)], [(1, ">"), (2, ">"), (3, ">"), (4, ">"), (5, ">")], 5)
    snow = ([(None, p[0], -1, OUT), (-1, p[1], -2, OUT), (-1, p[2], -3, OUT), (-1, p[1], -2, OUT),
             (-1, p[2], -3, OUT), (-2, p[0], -4, OUT), (-3, p[3], -5, OUT), (-1, p[4], -6, OUT)],
            [(1, ">"), (2, ">"), (3, ">"), (4, "<"), (5, "<"), (6, ">"), (7, ">"), (8, ">")], 6)
    return [star, path, snow]


def powerlaw_stream(triples, templates, n, seed):
    """Constant starts sampled in proportion to degree: a uniform pick over the
    triples of the first pattern's predicate, taking the endpoint that the
    constant stands for (the subject of (C, p, ?x), the object of (?x, p, C))."""
    rng = np.random.default_rng(seed)
    cand = []
    for raw, _, _ in templates:
        s0, p0, o0, _ = raw[0]
        col = 0 if s0 is None else 2
        cand.append(triples[triples[:, 1] == p0][:, col])
    out = []
    for _ in range(n):
        k = int(rng.integers(len(templates)))
        raw, plan, nv = templates[k]
        c = int(cand[k][rng.integers(len(cand[k]))])
        out.append((set_plan(instantiate(raw, c), plan), nv, k))
    return out
