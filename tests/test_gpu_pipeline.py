"""Pipelined batches (wk_batch_submit / wk_batch_wait / wk_batch_fetch): two
batches in flight on the two lanes, each exact against the oracle and equal
to the synchronous wk_batch_run; completion by another call; the fallbacks
(step overflow, segments that do not fit beside the other lane's) inside a
pipelined batch."""
import numpy as np
import pytest

import wukong_gpu as wg
from wukong_gpu import queries as Q

from _graphs import same_rows

pytestmark = pytest.mark.gpu


def _slices(store, n, size, seed):
    stream = Q.template_stream(store, Q.univ_templates(), n * size, seed=seed)
    qs = [(p, nv) for p, nv, _ in stream]
    return [qs[k * size:(k + 1) * size] for k in range(n)]


def _check(orc, qs, tables, allow_nomem=False):
    ok = 0
    for (pats, nv), (t, v2c, rc) in zip(qs, tables):
        want, wv2c, wrc = orc.query(pats, nv)
        assert wrc == 0 and v2c == wv2c, pats
        if allow_nomem and rc == wg.WK_E_NOMEM_BUDGET:
            continue  # its table did not fit the rest of the arena
        assert rc == 0, (pats, rc)
        assert same_rows(t, want), pats
        ok += 1
    assert ok > 0


@pytest.mark.parametrize("use_graph", [False, True])
def test_two_batches_in_flight(univ10_dev, univ10_oracle, univ10_store, use_graph):
    dev = univ10_dev
    sl = _slices(univ10_store, 5, 160, seed=41)
    pbs = [wg.PreparedBatch(s) for s in sl]
    t0 = dev.batch_submit(pbs[0], use_graph, v2c=True)
    t1 = dev.batch_submit(pbs[1], use_graph, v2c=True)
    with pytest.raises(wg.WkError) as e:   # both lanes busy
        dev.batch_submit(pbs[2], use_graph, v2c=True)
    assert e.value.code == wg.WK_E_STATE
    with pytest.raises(wg.WkError):        # not waited for yet
        dev.batch_fetch(t0, pbs[0])
    got = {}
    dev.batch_wait(t0)
    got[0] = dev.batch_fetch(t0, pbs[0])
    t2 = dev.batch_submit(pbs[2], use_graph, v2c=True)
    dev.batch_wait(t1)
    got[1] = dev.batch_fetch(t1, pbs[1])
    t3 = dev.batch_submit(pbs[3], use_graph, v2c=True)
    dev.batch_wait(t2)
    got[2] = dev.batch_fetch(t2, pbs[2])
    t4 = dev.batch_submit(pbs[4], use_graph, v2c=True)
    dev.batch_wait(t3)
    dev.batch_wait(t3)                      # a second wait is a no-op
    got[3] = dev.batch_fetch(t3, pbs[3])
    dev.batch_wait(t4)
    got[4] = dev.batch_fetch(t4, pbs[4])
    with pytest.raises(wg.WkError):        # lane 0 has run t2 and t4 since
        dev.batch_fetch(t0, pbs[0])
    for k in range(5):
        _check(univ10_oracle, sl[k], got[k])
        sync = dev.batch(sl[k], use_graph=use_graph)
        for (a, av, ar), (b, bv, br) in zip(got[k], sync):
            assert ar == br and av == bv and np.array_equal(a, b)


def test_batch_completed_by_another_call(univ10_dev, univ10_oracle, univ10_store):
    """A call that uses a lane first completes the batches in flight: their
    result arrays are filled and their wait returns at once; only the arena
    of the lane it reuses is lost."""
    dev = univ10_dev
    sl = _slices(univ10_store, 2, 100, seed=43)
    pbs = [wg.PreparedBatch(s) for s in sl]
    for pb in pbs:
        pb.status[:] = -7
    ta = dev.batch_submit(pbs[0], v2c=True)
    tb = dev.batch_submit(pbs[1], v2c=True)
    pats, nv = Q.chain3()
    got, _, rc = dev.query(pats, nv)
    want, _, _ = univ10_oracle.query(pats, nv)
    assert rc == 0 and same_rows(got, want)
    for pb in pbs:
        assert (pb.status == 0).all()
    dev.batch_wait(ta)
    dev.batch_wait(tb)
    for (pats, nv), r in zip(sl[0] + sl[1], np.concatenate([pbs[0].rows, pbs[1].rows])):
        assert int(r) == univ10_oracle.query(pats, nv)[0].shape[0], pats
    # the query reused the first lane's arena: ta's tables are not fetched;
    # tb's lane was not reused
    with pytest.raises(wg.WkError) as e:
        dev.batch_fetch(ta, pbs[0])
    assert e.value.code == wg.WK_E_STATE
    _check(univ10_oracle, sl[1], dev.batch_fetch(tb, pbs[1]))


def test_pipelined_overflow_and_budget_fallbacks(univ10_store, univ10_oracle):
    """Within pipelined batches: a step that overflows the table capacity
    splits at the wait, a query that alone exceeds it reports
    WK_E_NOMEM_BUDGET, and under a gstore budget smaller than the store
    segments swap between the two lanes' batches."""
    dev = wg.Device(univ10_store, table_budget=4 << 20)
    GS, MEM = Q.U("ub:GraduateStudent"), Q.U("ub:memberOf")
    depts = np.sort(univ10_store.get_index(Q.U("ub:Department"), wg.IN))[:8]
    qs = [([(GS, wg.TYPE_ID, -1, wg.IN), (-1, MEM, int(d), wg.OUT)], 1) for d in depts]
    qs.append(Q.type_start2())
    small = _slices(univ10_store, 1, 50, seed=47)[0]
    pa, pb = wg.PreparedBatch(qs), wg.PreparedBatch(small)
    ta = dev.batch_submit(pa, use_graph=False, v2c=True)
    tb = dev.batch_submit(pb, use_graph=True, v2c=True)
    dev.batch_wait(ta)
    dev.batch_wait(tb)
    ra = dev.batch_fetch(ta, pa)
    _check(univ10_oracle, qs[:-1], ra[:-1])
    assert ra[-1][2] == wg.WK_E_NOMEM_BUDGET
    _check(univ10_oracle, small, dev.batch_fetch(tb, pb), allow_nomem=True)
    dev.close()

    info = univ10_store.info()
    total = info["num_buckets"] * 128 + info["num_values"] * 4
    dev = wg.Device(univ10_store, gstore_budget=total * 2 // 5, key_block_shift=10, value_block_shift=14,
                    table_budget=256 << 20, dense_index=-1)
    sl = _slices(univ10_store, 4, 120, seed=53)
    sl[1] = [Q.type_start2(), Q.lubm_q1()] + sl[1]
    sl[2] = [Q.chain3()] + sl[2]
    pbs = [wg.PreparedBatch(s) for s in sl]
    tickets = [dev.batch_submit(pbs[0], v2c=True), dev.batch_submit(pbs[1], v2c=True)]
    for k in range(4):
        dev.batch_wait(tickets[k])
        _check(univ10_oracle, sl[k], dev.batch_fetch(tickets[k], pbs[k]))
        if k + 2 < 4:
            tickets.append(dev.batch_submit(pbs[k + 2], v2c=True))
    assert dev.stats()["segments_evicted"] > 0
    dev.close()


def test_tickets_completed_by_submits_and_runs_under_swapping(univ10_store, univ10_oracle):
    """Under a gstore budget smaller than the store a submit first completes
    the other lane's batch, and a synchronous batch between submit and wait
    completes the rest and reuses the first lane: every ticket's wait still
    returns its batch's status, with its result arrays exact, and a lane
    whose finished batch was not waited for takes no new submit."""
    info = univ10_store.info()
    total = info["num_buckets"] * 128 + info["num_values"] * 4
    dev = wg.Device(univ10_store, gstore_budget=total * 2 // 5, key_block_shift=10, value_block_shift=14,
                    table_budget=256 << 20, dense_index=-1)
    sl = _slices(univ10_store, 4, 100, seed=61)
    pbs = [wg.PreparedBatch(s) for s in sl]
    for pb in pbs:
        pb.status[:] = -7
    ta = dev.batch_submit(pbs[0])
    tb = dev.batch_submit(pbs[1], v2c=True)  # completes ta first
    assert (pbs[0].status == 0).all()
    with pytest.raises(wg.WkError) as e:    # ta's lane is idle but ta not waited for
        dev.batch_submit(pbs[2])
    assert e.value.code == wg.WK_E_STATE
    dev.batch_wait(ta)
    tc = dev.batch_submit(pbs[2])           # ta's lane again; completes tb
    sync = dev.batch(sl[3])                 # completes tc, then runs on tc's lane
    dev.batch_wait(tb)
    dev.batch_wait(tc)
    for k in range(3):
        assert (pbs[k].status == 0).all()
        for (pats, nv), r in zip(sl[k], pbs[k].rows):
            assert int(r) == univ10_oracle.query(pats, nv)[0].shape[0], pats
    _check(univ10_oracle, sl[3], sync)
    with pytest.raises(wg.WkError) as e:    # tc's arena was reused by the synchronous batch
        dev.batch_fetch(tc, pbs[2])
    assert e.value.code == wg.WK_E_STATE
    _check(univ10_oracle, sl[1], dev.batch_fetch(tb, pbs[1]))  # tb's lane was not reused
    td = dev.batch_submit(pbs[0], v2c=True)
    dev.batch_wait(td)
    _check(univ10_oracle, sl[0], dev.batch_fetch(td, pbs[0]))
    dev.close()


def test_three_lanes(univ10_store, univ10_oracle):
    """batch_lanes = 3: three batches in flight, a fourth submit refused."""
    dev = wg.Device(univ10_store, table_budget=256 << 20, batch_lanes=3)
    sl = _slices(univ10_store, 6, 80, seed=59)
    pbs = [wg.PreparedBatch(s) for s in sl]
    ts = [dev.batch_submit(pbs[k], v2c=True) for k in range(3)]
    with pytest.raises(wg.WkError) as e:
        dev.batch_submit(pbs[3], v2c=True)
    assert e.value.code == wg.WK_E_STATE
    for k in range(6):
        dev.batch_wait(ts[k])
        _check(univ10_oracle, sl[k], dev.batch_fetch(ts[k], pbs[k]))
        if k + 3 < 6:
            ts.append(dev.batch_submit(pbs[k + 3], v2c=True))
    dev.close()
    with pytest.raises(wg.WkError):
        wg.Device(univ10_store, batch_lanes=5)
