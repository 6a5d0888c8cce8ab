"""Benchmark of the MI355X Wukong GPU query path.

Default workload: the university-schema graph at 2,560 universities (seed 1,
~339M triples, gstore resident in HBM):
  * single-query kernels: Q1 = lubm_q7 (6-pattern cyclic query with its
    OSDI16 plan and a known_to_known closing edge), Q2 = the 2-pattern
    type-start query, and lubm_q1.  Per-kernel HIP-event timing gives the
    roofline of the dominant kernel;
  * one "step" = one batched pass over a distinct slice of 4,096 selective
    queries of a stream from 5 templates (constants drawn from the type
    index), so no timed step replays an earlier batch.  Each pattern step of
    all the slice's queries is one set of launches over a concatenated table.
    The native autotuner first picks the probe width and the gather mapping.
    The steps' algorithmic bytes, counted from the device's own row counts in
    a separate pass, give the batch's roofline;
  * the same stream one query at a time (wk_query_run) for comparison;
  * open-loop Poisson serving replay, continuous batching vs per-query, p50/p99;
  * the C oracle on the host cores as the CPU baseline: one worker per thread
    pulling queries from a shared queue over the timed slices, and the single
    queries on one thread and split over all threads (mt_factor).

`--workload powerlaw`: the power-law graph (seed 7, 100M vertices, 1.2B
triples, 256 Zipf(1.1) predicates, Pareto(2) out-degree, 64 types) with the
star / 5-hop path / snowflake stream, starts drawn in proportion to degree.
`--gstore-budget-gb` below the store size makes segments swap from host
memory.  The serving replay runs at the offered rates given by `--rates`.

Every timed result is checked against the oracle: bit-exact row multisets for
the single queries and for the first 512 batched queries, row counts and
order-independent row fingerprints for every batched and served query the
oracle's time budget covers.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "wukong-cube_amd"))
sys.path.insert(0, REPO)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` as measured by the
    rocprofv3 FETCH_SIZE/WRITE_SIZE passes summarised under profiles/ (a
    `traffic: kernel=... workload=... bytes_per_launch=...` line), or None
    when no measurement of this kernel and workload is there."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(REPO, "profiles", "*hbm_traffic*.md"))):
        for line in open(path):
            m = re.match(r"traffic: kernel=(\S+) workload=(\S+) bytes_per_launch=(\d+)", line.strip())
            if m and m.group(1) == kernel and m.group(2) == workload:
                return int(m.group(3))
    return None


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=("university", "powerlaw"), default="university")
    ap.add_argument("--univ", type=int, default=2560)
    ap.add_argument("--pl-vertices", type=int, default=100_000_000)
    ap.add_argument("--pl-triples", type=int, default=1_200_000_000)
    ap.add_argument("--ids-dir", type=str, default="",
                    help="build the host store from a directory of Wukong ID-triple files (id_*); it must "
                         "hold the graph of --univ/--seed (e.g. written by wukong_gpu.write_id_triples), "
                         "since the oracle checks against the generated triples")
    ap.add_argument("--nq", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=None, help="default 1 (university), 7 (power law)")
    ap.add_argument("--gstore-budget-gb", type=float, default=0, help="0: the whole store")
    ap.add_argument("--table-budget-gb", type=float, default=48)
    ap.add_argument("--rates", type=str, default="", help="serving offered rates, q/s, comma separated")
    ap.add_argument("--mix", type=str, default="",
                    help="serving stream as LIGHT/HEAVY percent, e.g. 95/5: light queries have at most "
                         "--light-rows result rows, heavy ones at least --heavy-rows (classified by the batched "
                         "run's row counts, drawn from the first --mix-slices timed slices); default: the stream "
                         "as generated")
    ap.add_argument("--light-rows", type=int, default=100)
    ap.add_argument("--heavy-rows", type=int, default=1_000_000)
    ap.add_argument("--mix-slices", type=int, default=4)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--heavy-footprint", type=int, default=0,
                    help="serving: projected table bytes from which a query runs on the second lane "
                         "(0 = library default, -1 = no second lane)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the native autotuner (probe width x gather mapping) before the batch")
    ap.add_argument("--probe-width", type=int, default=0,
                    help="lanes per key in the probe (1, 2, 4, 8; 0 = library default)")
    ap.add_argument("--rows-per-wave", type=int, default=0, help="gather mapping: 0 = output-balanced")
    ap.add_argument("--dense-index", type=int, default=0,
                    help="direct key indexes of dense segments: 0 = build (default), -1 = probe the buckets only")
    ap.add_argument("--no-graph", action="store_true", help="launch batched steps directly (no hipGraph)")
    ap.add_argument("--batch-lanes", type=int, default=0,
                    help="lanes the pipelined batches rotate over, 1..4 (0 = library default, 2)")
    ap.add_argument("--single-only", action="store_true",
                    help="only the single-query kernels and their roofline (the launches a kernel trace "
                         "of this command should cover)")
    ap.add_argument("--batch-only", action="store_true", help="only the batched stream (for kernel traces)")
    args = ap.parse_args()
    if args.seed is None:
        args.seed = 1 if args.workload == "university" else 7
    return args


def make_workload(args, wg, Q):
    """Triples, single queries, a stream-building function and a description."""
    t0 = time.time()
    if args.workload == "university":
        tri = wg.gen_university(args.seed, args.univ)
        singles = {name: getattr(Q, name)() for name in ("lubm_q7", "type_start2", "lubm_q1")}
        desc = ("university graph, %d universities (seed %d, %d triples): %d-query template stream batched "
                "per step, a distinct slice of the stream each step; single queries Q1 lubm_q7 (6-pattern "
                "cyclic), Q2 type-start, and lubm_q1" % (args.univ, args.seed, len(tri), args.nq))
        cfg = {"universities": args.univ}

        def stream(store, n):
            return Q.template_stream(store, Q.univ_templates(), n, seed=args.seed + 100)
    else:
        tri = wg.gen_powerlaw(args.seed, args.pl_vertices, args.pl_triples)
        singles = {}
        desc = ("power-law graph (seed %d, %d vertices, %d triples, 256 Zipf(1.1) predicates, Pareto(2) "
                "out-degree): %d-query star/path/snowflake stream, degree-proportional starts" % (
                    args.seed, args.pl_vertices, len(tri), args.nq))
        cfg = {"vertices": args.pl_vertices}

        def stream(store, n):
            return Q.powerlaw_stream(tri, Q.powerlaw_templates(), n, seed=args.seed + 100)
    log("generated %d triples in %.1fs" % (len(tri), time.time() - t0))
    cfg.update({"triples": int(len(tri)), "queries_per_step": args.nq})
    return tri, singles, stream, desc, cfg


def run_single(dev, singles, args):
    """Each single query `steps` times with profiling off (latency), then
    `steps` times with every kernel timed (per-kernel times and bytes)."""
    single = {}
    for name, (pats, nv) in singles.items():
        dev.set_profile(False)
        for _ in range(args.warmup):
            dev.query(pats, nv, fetch=False)
        dev.sync()
        lat = []
        for _ in range(args.steps):
            t = time.perf_counter()
            _, _, rc = dev.query(pats, nv, fetch=False)
            lat.append((time.perf_counter() - t) * 1e6)
            assert rc == 0
        dev.set_profile(True)
        dev.reset_kernel_stats()
        for _ in range(args.steps):
            dev.query(pats, nv, fetch=False)
        ks = dev.kernel_stats()
        table, v2c, rc = dev.query(pats, nv, fetch=True)
        single[name] = {"patterns": len(pats), "rows": int(table.shape[0]), "cols": int(table.shape[1]),
                        "latency_us_mean": float(np.mean(lat)), "latency_us_p50": float(np.median(lat)),
                        "kernels": ks, "_table": table, "_pats": pats, "_nv": nv, "_v2c": v2c}
        log("%s: %d rows, %.0f us/query" % (name, table.shape[0], np.mean(lat)))
    return single


def run_partition(dev, single, args, owners=(8, 64)):
    """Fork-join split of the largest single-query result among 8 and 64
    owners (row[0] % n, input order kept within each owner).  Minimal bytes
    are one read and one write of the table; the timed call includes the
    output allocation and the owners' heads coming back to the host."""
    if not single:
        return None
    name = max(single, key=lambda k: single[k]["rows"])
    host = single[name]["_table"]
    tbl = dev.upload(host)
    res = {"table": name, "rows": int(host.shape[0]), "cols": int(host.shape[1]), "alg_bytes": 2.0 * host.size * 4,
           "owners": {}}
    for n in owners:
        for _ in range(args.warmup):
            out, _, _ = dev.partition(tbl, 0, n)
            out.free()
        dev.sync()
        lat = []
        for _ in range(args.steps):
            t = time.perf_counter()
            out, sizes, heads = dev.partition(tbl, 0, n)
            lat.append((time.perf_counter() - t) * 1e3)
            out.free()
        assert int(sizes.sum()) == host.shape[0]
        ms = float(np.median(lat))
        gbps = res["alg_bytes"] / (ms * 1e-3) / 1e9
        res["owners"][str(n)] = {"ms_p50": ms, "GBps": gbps, "max_owner_rows": int(sizes.max())}
        log("split of %s (%d x %d) among %d owners: %.3f ms, %.0f GB/s" % (name, host.shape[0], host.shape[1], n,
                                                                          ms, gbps))
    tbl.free()
    return res


def run_final_process(dev, single, args, cpu_rows=2_000_000):
    """final_process on the device over the largest single-query result:
    SELECT DISTINCT of all its variables ORDER BY the last one DESC LIMIT
    100, timed per call (sort, dedup, projection, output allocation).  The
    oracle runs the same on the first cpu_rows rows, one thread."""
    if not single:
        return None
    import oracle
    name = max(single, key=lambda k: single[k]["rows"])
    r = single[name]
    host, v2c, nv = r["_table"], r["_v2c"], r["_nv"]
    req = [-(i + 1) for i in range(nv) if v2c[i] >= 0]
    kw = dict(distinct=True, orders=[(req[-1], True)], limit=100)
    tbl = dev.upload(host)
    for _ in range(args.warmup):
        dev.final_process(tbl, v2c, req, **kw).free()
    dev.sync()
    lat = []
    for _ in range(args.steps):
        t = time.perf_counter()
        out = dev.final_process(tbl, v2c, req, **kw)
        lat.append((time.perf_counter() - t) * 1e3)
        out.free()
    tbl.free()
    ms = float(np.median(lat))
    res = {"table": name, "rows": int(host.shape[0]), "cols": int(host.shape[1]), "query": "DISTINCT, ORDER BY "
           "last variable DESC, LIMIT 100", "ms_p50": ms, "rows_per_s": host.shape[0] / (ms * 1e-3)}
    if not args.no_cpu:
        sample = host[:cpu_rows]
        t = time.perf_counter()
        oracle.final_process(sample, v2c, req, **kw)
        cs = time.perf_counter() - t
        res["cpu_oracle"] = {"rows": int(sample.shape[0]), "s": cs, "rows_per_s": sample.shape[0] / cs,
                             "threads": 1}
    log("final_process of %s (%d rows): %.2f ms on the device%s" % (
        name, host.shape[0], ms, "; oracle %.0f rows/s" % res["cpu_oracle"]["rows_per_s"] if "cpu_oracle" in res
        else ""))
    return res


def roofline_of(single, args):
    """The dominant kernel over the single queries: algorithmic bytes per
    launch over its average launch time."""
    agg = {}
    for r in single.values():
        for k, v in r["kernels"].items():
            a = agg.setdefault(k, {"launches": 0, "total_ms": 0.0, "alg_bytes": 0.0})
            for f in a:
                a[f] += v[f]
    if not agg:
        return None
    dom = max(agg, key=lambda k: agg[k]["total_ms"])
    d = agg[dom]
    avg_ms = d["total_ms"] / d["launches"]
    bytes_per_launch = d["alg_bytes"] / d["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom, "univ%d-single" % args.univ),
            "alg_bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_ms, "launches": d["launches"]}


def cpu_threads():
    """Host threads for the CPU baseline: this process's CPU share."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(env))) if env and env.isdigit() else n


def cpu_baseline_and_gate(args, wg, dev, tri, single, slices, tq, rows_t, st_t, serve):
    """The oracle on the host cores of the same box, as the CPU baseline, and
    the parity gate of everything timed above:
      * single queries: exact row multisets; CPU latency on one thread and
        with the index start split over all threads (mt_factor);
      * the batched stream: slice after timed slice, the oracle runs the
        slice on one worker per thread pulling from a shared queue (timed),
        the device runs it again and its result arena comes back; every
        query's row count and row fingerprint must match, and the first 512
        queries' rows must be identical as multisets;
      * served results: fingerprints equal the oracle's.
    The oracle's time budget bounds the sample (--cpu-seconds)."""
    import oracle
    T = cpu_threads()
    gate = {}
    t0 = time.time()
    orc = oracle.Store(tri)
    log("oracle store built in %.1fs; %d host threads" % (time.time() - t0, T))
    for name, r in single.items():
        t = time.perf_counter()
        want, _, rc = orc.query(r["_pats"], r["_nv"])
        r["cpu_latency_us"] = (time.perf_counter() - t) * 1e6
        t = time.perf_counter()
        want_mt, _, rc_mt = orc.query_mt(r["_pats"], r["_nv"], T)
        r["cpu_latency_us_%dthreads" % T] = (time.perf_counter() - t) * 1e6
        g = oracle.sorted_rows(r["_table"])
        w = oracle.sorted_rows(want)
        gate[name] = bool(rc == 0 and rc_mt == 0 and g.shape == w.shape and np.array_equal(g, w)
                          and np.array_equal(oracle.sorted_rows(want_mt), w))
    # the stream, slice by slice
    ora_s = 0.0
    n = 0
    counts_ok = cks_ok = True
    checked_rows = 0
    ck_of = {}
    k = 0
    while k * args.nq < len(tq) and ora_s < args.cpu_seconds:
        sl = tq[k * args.nq:(k + 1) * args.nq]
        t = time.perf_counter()
        o_rows, o_ck, o_st = orc.run_stream(sl, T)
        ora_s += time.perf_counter() - t
        res = dev.batch(slices[args.warmup + k], use_graph=True, fetch=True)
        for i, (tab, _, rc) in enumerate(res):
            q = k * args.nq + i
            ck_of[q] = (int(o_ck[i]), int(o_rows[i]))
            if o_st[i] != 0:
                counts_ok = False
                continue
            if rc != 0:
                continue  # over the table budget in this round: its rows come in a later round
            if int(rows_t[q]) != o_rows[i] or tab.shape[0] != o_rows[i]:
                counts_ok = False
            if (wg.checksum(tab) if tab.size else 0) != int(o_ck[i]):
                cks_ok = False
            if q < 512:
                want, _, _ = orc.query(*sl[i])
                if not np.array_equal(oracle.sorted_rows(tab), oracle.sorted_rows(want)):
                    cks_ok = False
                checked_rows += 1
        n += len(sl)
        k += 1
    served = 0
    for s in serve.values():
        for i, q in enumerate(s["_idx"]):
            if s["_st"][i] == 0 and int(q) in ck_of:
                served += 1
                if (int(s["_ck"][i]), int(s["_rows"][i])) != ck_of[int(q)]:
                    cks_ok = False
    gate["stream_rows"] = bool(counts_ok)
    gate["stream_fingerprints"] = bool(cks_ok)
    # one thread over the first slice, for scale
    m1 = min(len(tq), 2048)
    t = time.perf_counter()
    orc.run_stream(tq[:m1], 1)
    one = m1 / (time.perf_counter() - t)
    cpu = {"value": n / ora_s, "unit": "queries/s", "cores": T, "kind": "port",
           "sample": "the first %d queries of the timed stream (%d slices), oracle engine with %d workers "
                     "pulling from a shared queue (%.1fs); one worker: %.0f q/s" % (n, k, T, ora_s, one),
           "single_thread_qps": one}
    log("cpu oracle: %.0f q/s on %d threads (%.0f on one) over %d queries; %d rows-exact, %d served checked; "
        "gate %s" % (n / ora_s, T, one, n, checked_rows, served, gate))
    return cpu, gate


def main():
    args = parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import wukong_gpu as wg
    from wukong_gpu import queries as Q
    if not os.path.exists(wg.LIB_PATH):
        raise SystemExit("libwukong_gpu.so is not built: run python __graft_entry__.py")

    tri, singles, make_stream, desc, cfg = make_workload(args, wg, Q)
    t0 = time.time()
    if args.ids_dir:
        store, nread = wg.Store.from_dir(args.ids_dir, pinned=True)
        cfg["ids_dir"] = {"triples_read": nread, "load_s": time.time() - t0}
        log("host gstore loaded from %s in %.1fs (%d triples read)" % (args.ids_dir, time.time() - t0, nread))
    else:
        store = wg.Store(tri, pinned=True)
    info = store.info()
    log("host gstore built in %.1fs: %s" % (time.time() - t0, info))
    dev = wg.Device(store, gstore_budget=int(args.gstore_budget_gb * (1 << 30)),
                    table_budget=int(args.table_budget_gb * (1 << 30)), profile=1,
                    probe_width=args.probe_width, rows_per_wave=args.rows_per_wave, dense_index=args.dense_index,
                    batch_lanes=args.batch_lanes)
    t0 = time.time()
    dev.load_all()
    load_s = time.time() - t0
    st = dev.stats()
    # swaps come by DMA from a pinned store, else through the pinned staging ring
    gstore = {"host_bytes": info["num_buckets"] * 128 + info["num_values"] * 4, "pinned": int(info["pinned"]),
              "staging_ring": int(not info["pinned"]), "load_all_s": load_s,
              "load_all_h2d_bytes": st["h2d_bytes"], "load_all_h2d_gbps": st["h2d_bytes"] / load_s / 1e9,
              "mapped_segments": st["mapped_segments"]}
    log("gstore loaded in %.1fs (%.2f GB H2D, %.1f GB/s, %s; %d key / %d value blocks free; direct key indexes "
        "for %d segments, %.2f GB; %d segments read in place from host memory)" % (
            load_s, st["h2d_bytes"] / 1e9, gstore["load_all_h2d_gbps"],
            "pinned store" if info["pinned"] else "pageable store through the staging ring",
            st["key_blocks_free"], st["value_blocks_free"], st["dense_segments"], st["dense_bytes"] / 1e9,
            st["mapped_segments"]))

    out = {"metric": "queries/sec + p50/p99 latency (batched vs per-query GPU); edges/s; HBM GB/s"}

    # ---------------------------------------------------------------- single
    single = {} if args.batch_only else run_single(dev, singles, args)
    roofline = roofline_of(single, args)
    partition = run_partition(dev, single, args)
    final = run_final_process(dev, single, args)
    edges = sum(r["rows"] for r in single.values())
    if args.single_only:
        for r in single.values():
            for k in ("_table", "_pats", "_nv", "_v2c"):
                r.pop(k, None)
        print(json.dumps({"metric": "single-query kernels", "single_query": single, "roofline": roofline,
                          "partition": partition, "final_process": final,
                          "config": cfg}))
        dev.close()
        return

    # ---------------------------------------------------------------- batch
    # one distinct slice of nq queries per warmup and timed step, so no timed
    # step replays a batch run before it
    nslices = args.warmup + args.steps
    t0 = time.time()
    stream = make_stream(store, nslices * args.nq)
    qs = [(p, nv) for p, nv, _ in stream]
    slices = [wg.PreparedBatch(qs[k * args.nq:(k + 1) * args.nq]) for k in range(nslices)]
    log("stream of %d queries in %d slices packed in %.1fs" % (len(qs), nslices, time.time() - t0))
    if not args.no_tune:
        pw, rpw, tms = dev.autotune(slices[0], use_graph=not args.no_graph)
        log("autotune: probe_width=%d rows_per_wave=%d (%.3f ms per batch in the sweep)" % (pw, rpw, tms))
        out["tuning"] = {"probe_width": pw, "rows_per_wave": rpw, "sweep_ms_per_batch": tms}
    rows_all = np.zeros(len(qs), np.uint64)
    st_all = np.zeros(len(qs), np.int32)
    rounds_all = []

    def run_slice(k):
        """The whole job of slice k, one call at a time: all its queries, then
        those whose results did not fit the rest of the result arena, in
        further rounds that reuse the arena.  Results stay on the device, as
        on the per-query path."""
        r, _, st = dev.batch(slices[k], use_graph=not args.no_graph, fetch=False)
        return more_rounds(k, r, st)

    def more_rounds(k, r, st):
        """Queries of slice k whose results did not fit the rest of the
        result arena, in further synchronous rounds until a round completes
        none (their results alone exceed it)."""
        base = k * args.nq
        idx = np.arange(args.nq)
        nr = 1
        while True:
            fail = idx[st[idx] == wg.WK_E_NOMEM_BUDGET]
            if len(fail) == 0 or len(fail) == len(idx):
                break
            sub = wg.PreparedBatch([qs[base + i] for i in fail])
            r2, _, s2 = dev.batch(sub, use_graph=not args.no_graph, fetch=False)
            r[fail], st[fail] = r2, s2
            idx = fail
            nr += 1
        rows_all[base:base + args.nq] = r
        st_all[base:base + args.nq] = st
        return nr

    inflight = args.batch_lanes or 2  # slices in flight: one per lane

    def run_pipelined(ks):
        """Slices ks with one in flight per lane: slice k + 1 is compiled,
        uploaded and launched on the next lane while slice k runs on the GPU.
        Returns the rounds each slice took."""
        pend, nr = [], []

        def complete(k, tk):
            dev.batch_wait(tk)
            pb = slices[k]
            nr.append(more_rounds(k, pb.rows.copy(), pb.status.copy()))

        for k in ks:
            pend.append((k, dev.batch_submit(slices[k], use_graph=not args.no_graph)))
            if len(pend) == inflight:
                complete(*pend.pop(0))
        for kt in pend:
            complete(*kt)
        return nr

    for k in range(args.warmup):
        run_slice(k)
    dev.sync()
    # one slice after the other, each waited for before the next is prepared
    dev.reset_kernel_stats()
    t = time.perf_counter()
    for k in range(args.warmup, nslices):
        rounds_all.append(run_slice(k))
    dev.sync()
    sync_s = (time.perf_counter() - t) / args.steps
    # host phases of wk_batch_run (the device's profile mode times them)
    batch_host = {k[5:]: v["total_ms"] / args.steps for k, v in dev.kernel_stats().items()
                  if k.startswith("host:")}
    out["batch_host_ms"] = batch_host
    if batch_host:
        log("batch host phases (ms per batch): " + ", ".join("%s %.3f" % kv for kv in sorted(batch_host.items())))
    log("batched, one slice at a time: %.3f ms per %d-query slice -> %.0f q/s" % (
        sync_s * 1e3, args.nq, args.nq / sync_s))
    # the headline: the same slices pipelined, two in flight (value)
    rows_all[:] = 0
    st_all[:] = -1
    run_pipelined(range(args.warmup))
    dev.sync()
    t = time.perf_counter()
    rounds_all = run_pipelined(range(args.warmup, nslices))
    dev.sync()
    batch_s = (time.perf_counter() - t) / args.steps
    out["batch_one_at_a_time"] = {"ms_per_step": sync_s * 1e3, "queries_per_s": args.nq / sync_s,
                                  "note": "wk_batch_run per slice: host preparation and GPU run in sequence"}
    timed = slice(args.warmup * args.nq, nslices * args.nq)
    out["batch_rounds_max"] = int(max(rounds_all))
    ok_b = st_all[timed] == 0
    qps_batched = args.nq / batch_s
    log("batched, pipelined (%d slices in flight): %.3f ms per %d-query slice (up to %d rounds) -> %.0f q/s "
        "(%d ok, %d over the table budget, %d result rows over %d slices)" % (
            inflight, batch_s * 1e3, args.nq, max(rounds_all), qps_batched, ok_b.sum(),
            (st_all[timed] == wg.WK_E_NOMEM_BUDGET).sum(), int(rows_all[timed][ok_b].sum()), args.steps))
    assert ((st_all == 0) | (st_all == wg.WK_E_NOMEM_BUDGET)).all(), np.unique(st_all)

    # algorithmic bytes of the timed slices' pattern steps, from the device's
    # own row counts, counted in a pass of its own (the counting reads the
    # step descriptors back after each batch)
    dev.set_profile(False)
    dev.batch_traffic(True)
    for k in range(args.warmup, nslices):
        run_slice(k)
    alg_bytes, alg_edges, nb = dev.batch_traffic(False)
    dev.set_profile(True)
    per_step = alg_bytes / args.steps
    achieved = per_step / batch_s / 1e9
    out["roofline_batch"] = {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
        "alg_bytes_per_step": per_step, "edges_per_step": alg_edges / args.steps, "ms_per_step": batch_s * 1e3,
        "edges_per_s": alg_edges / args.steps / batch_s,
        "note": "whole batched step (all launches and host work) against the per-pattern byte formulas; "
                "a batch of selective queries is bound by launch and dependent-load latency, not HBM bandwidth"}
    log("batch roofline: %.1f MB of algorithmic bytes per step -> %.0f GB/s = %.2f%% of peak" % (
        per_step / 1e6, achieved, 100 * achieved / HBM_PEAK_GBS))

    # the same slices with every round's result arena copied to pinned host
    # memory after the round, as a caller that wants the rows on the host
    # does: the PCIe-inclusive rate, reported beside the HBM-resident one
    L = wg.lib()
    hip = wg.C.CDLL("/opt/rocm/lib/libamdhip64.so")  # hipHostMalloc / hipHostFree for the pinned buffer
    host_buf = [np.zeros(1, np.uint32)]
    host_ptr = [None]

    def to_host(fetch):
        """Copies a result arena to the pinned buffer through fetch(dst, cap,
        used) (wk_batch_fetch of a ticket, or wk_batch_download); returns the
        bytes copied."""
        used = wg.C.c_uint64()
        wg._check(fetch(None, 0, wg.C.byref(used)))
        if host_buf[0].size < used.value:
            if host_ptr[0]:
                hip.hipHostFree(host_ptr[0])
            n = int(used.value) * 2
            ptr = wg.C.c_void_p()
            assert hip.hipHostMalloc(wg.C.byref(ptr), wg.C.c_size_t(n * 4), 0) == 0
            host_ptr[0] = ptr
            host_buf[0] = np.ctypeslib.as_array(wg.C.cast(ptr, wg.U32P), (n,))
        wg._check(fetch(host_buf[0].ctypes.data_as(wg.U32P), host_buf[0].size, wg.C.byref(used)))
        return used.value * 4

    def run_fetch_pipelined(ks):
        """run_pipelined with each slice's arena copied to the host after its
        wait (while the next slice runs).  The queries whose tables did not
        fit a slice's arena run in further synchronous rounds, each arena
        copied after its round, once every pipelined slice is fetched: a
        synchronous round reuses the first lane's arena, which may hold a
        pipelined slice not yet fetched."""
        pend, nbytes, st_of = [], [0], {}

        def complete(k, tk):
            dev.batch_wait(tk)
            nbytes[0] += to_host(lambda d, c, u: L.wk_batch_fetch(dev._h, tk, d, c, u))
            st_of[k] = slices[k].status.copy()

        for k in ks:
            pend.append((k, dev.batch_submit(slices[k], use_graph=not args.no_graph)))
            if len(pend) == inflight:
                complete(*pend.pop(0))
        for kt in pend:
            complete(*kt)
        for k in ks:
            base = k * args.nq
            st = st_of[k]
            idx = np.arange(args.nq)
            while True:
                fail = idx[st[idx] == wg.WK_E_NOMEM_BUDGET]
                if len(fail) == 0 or len(fail) == len(idx):
                    break
                _, _, s2 = dev.batch(wg.PreparedBatch([qs[base + i] for i in fail]),
                                     use_graph=not args.no_graph, fetch=False)
                nbytes[0] += to_host(lambda d, c, u: L.wk_batch_download(dev._h, d, c, u))
                st[fail] = s2
                idx = fail
        return nbytes[0]

    run_fetch_pipelined(range(args.warmup))
    dev.sync()
    t = time.perf_counter()
    fetched = run_fetch_pipelined(range(args.warmup, nslices))
    dev.sync()
    fetch_s = (time.perf_counter() - t) / args.steps
    out["pcie_inclusive"] = {"queries_per_s": args.nq / fetch_s, "ms_per_step": fetch_s * 1e3,
                             "result_bytes_per_step": fetched / args.steps,
                             "note": "pipelined as the headline, each slice's result arena copied to pinned "
                                     "host memory after its wait, while the next slice runs"}
    log("batched with results copied to the host: %.3f ms per slice (%.1f MB) -> %.0f q/s"
        % (fetch_s * 1e3, fetched / args.steps / 1e6, args.nq / fetch_s))
    host_buf[0] = None
    hip.hipHostFree(host_ptr[0])
    if args.batch_only:
        print(json.dumps({"metric": "batched queries/s", "value": qps_batched, "ms_per_step": batch_s * 1e3,
                          "tuning": out.get("tuning"), "roofline_batch": out["roofline_batch"], "config": cfg}))
        dev.close()
        return
    # the timed slices' queries, in stream order, with their batched results
    tq = qs[args.warmup * args.nq:]
    rows_t, st_t = rows_all[timed], st_all[timed]

    # per-query execution of the same stream (bounded sample), untimed per kernel
    dev.set_profile(False)
    npq = min(args.nq, 1024)
    t = time.perf_counter()
    for pats, nv in tq[:npq]:
        dev.query(pats, nv, fetch=False)
    pq_s = (time.perf_counter() - t) / npq
    qps_perq = 1.0 / pq_s
    log("per-query: %.1f us/query -> %.0f q/s" % (pq_s * 1e6, qps_perq))

    # ---------------------------------------------------------------- serve
    serve = {}
    mix_info = None
    if not args.no_serve and args.mix:
        # the served stream in the given light/heavy proportion, from the
        # first timed slices' queries as the batched run classified them
        pl, ph = (float(x) for x in args.mix.split("/"))
        span = min(len(tq), args.mix_slices * args.nq)
        ok_t = st_t[:span] == 0
        light = np.flatnonzero(ok_t & (rows_t[:span] <= args.light_rows))
        heavy = np.flatnonzero(ok_t & (rows_t[:span] >= args.heavy_rows))
        if len(light) == 0 or len(heavy) == 0:
            raise SystemExit("--mix: %d light and %d heavy queries in the first %d timed queries" % (
                len(light), len(heavy), span))
        mix_info = {"light_pct": pl, "heavy_pct": ph, "light_rows_max": args.light_rows,
                    "heavy_rows_min": args.heavy_rows, "light_candidates": int(len(light)),
                    "heavy_candidates": int(len(heavy)), "drawn_from_first_queries": int(span)}
        log("serving mix %s: %d light (<= %d rows) and %d heavy (>= %d rows) candidates in the first %d queries" % (
            args.mix, len(light), args.light_rows, len(heavy), args.heavy_rows, span))
    if not args.no_serve:
        rng = np.random.default_rng(args.seed + 7)
        if args.rates:
            rates = [float(r) for r in args.rates.split(",")]
            plan = [("per_query", min(rates[0], 0.7 * qps_perq))] + [("batched@%g" % r, r) for r in rates]
        else:
            plan = [("per_query", 0.7 * qps_perq), ("batched", 10 * 0.7 * qps_perq)]
        for mode, rate in plan:
            if mix_info:
                n = int(rate * 2.0) + 256
                mrng = np.random.default_rng(args.seed + 11)
                is_heavy = mrng.random(n) < ph / (pl + ph)
                idx = np.where(is_heavy, heavy[mrng.integers(len(heavy), size=n)],
                               light[mrng.integers(len(light), size=n)])
            else:
                n = min(len(tq), int(rate * 2.0) + 256)
                idx = np.arange(n)
            sq = [tq[int(i)] for i in idx]
            arr = np.cumsum(rng.exponential(1.0 / rate, size=n))
            dev.reset_kernel_stats()
            # only the continuous-batching loop keeps its (host-side) iteration timer
            dev.set_profile(mode != "per_query")
            r = dev.serve(sq, arr, max_inflight=4096, per_query=(mode == "per_query"),
                          heavy_footprint=(2**64 - 1) if args.heavy_footprint < 0 else args.heavy_footprint)
            phases = {k[6:]: {"count": v["launches"], "ms": round(v["total_ms"], 3)}
                      for k, v in dev.kernel_stats().items() if k.startswith("serve:")}
            lat = r["latency_us"]
            ok = r["status"] == 0
            serve[mode] = {"offered_qps": rate, "queries": n, "achieved_qps": n / r["wall_s"],
                           "p50_us": float(np.percentile(lat[ok], 50)) if ok.any() else None,
                           "p99_us": float(np.percentile(lat[ok], 99)) if ok.any() else None,
                           "ok": int(ok.sum()), "iterations": int(r["batches"]), "phases": phases,
                           "_ck": r["checksum"], "_st": r["status"], "_rows": r["rows"], "_idx": idx}
            if mix_info:
                hv = rows_t[idx] >= args.heavy_rows
                serve[mode]["heavy_share"] = float(np.mean(hv))
                for cls, m in (("light", ok & ~hv), ("heavy", ok & hv)):
                    if m.any():
                        serve[mode][cls] = {"p50_us": float(np.percentile(lat[m], 50)),
                                            "p99_us": float(np.percentile(lat[m], 99)), "ok": int(m.sum())}
                log("  light p50/p99 %s / %s us, heavy p50/p99 %s / %s us" % (
                    *(round(serve[mode].get("light", {}).get(k, -1)) for k in ("p50_us", "p99_us")),
                    *(round(serve[mode].get("heavy", {}).get(k, -1)) for k in ("p50_us", "p99_us"))))
            log("serve %s @ %.0f q/s offered: %.0f q/s, p50 %s us, p99 %s us, %d ok; %s" % (
                mode, rate, n / r["wall_s"], serve[mode]["p50_us"], serve[mode]["p99_us"], ok.sum(), phases))

    # the per-query sample through the wire boundary (after the serving legs,
    # whose timings it must not disturb): request bundles (boost
    # binary archives, as the proxy sends them) in, reply bundles with the
    # projected rows out, one query at a time
    nag = min(npq, 256)
    bundles = [wg.wire_encode(wg.query_msg(p, nv, [-(i + 1) for i in range(nv)])) for p, nv in tq[:nag]]
    t = time.perf_counter()
    replies = [dev.agent_execute(b) for b in bundles]
    ag_s = (time.perf_counter() - t) / nag
    ag_rows = 0
    for i in range(nag):
        m = wg.wire_decode(replies[i])
        replies[i] = None
        want = int(rows_t[i]) if st_t[i] == 0 else None
        assert m["status_code"] == 0 and (want is None or m["row_num"] == want), (i, m["status_code"])
        ag_rows += m["row_num"]
    del replies
    out["agent"] = {"queries": nag, "queries_per_s": 1.0 / ag_s, "us_per_query": ag_s * 1e6,
                    "reply_rows": ag_rows, "request_bytes_mean": float(np.mean([len(b) for b in bundles])),
                    "note": "wk_agent_execute: decode, device run, final_process, download, encode"}
    log("agent (bundle in, bundle out): %.1f us/query -> %.0f q/s, %d reply rows"
        % (ag_s * 1e6, 1.0 / ag_s, ag_rows))

    # ------------------------------------------------------ CPU baseline + gate
    cpu = None
    gate = {}
    if not args.no_cpu:
        cpu, gate = cpu_baseline_and_gate(args, wg, dev, tri, single, slices, tq, rows_t, st_t, serve)
    for r in single.values():
        for k in ("_table", "_pats", "_nv", "_v2c"):
            r.pop(k, None)
    for r in serve.values():
        for k in ("_ck", "_st", "_rows", "_idx"):
            r.pop(k, None)

    out.update({
        "value": qps_batched, "unit": "queries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": batch_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic (seeded %s generator)" % args.workload,
        "config": dict({"workload": desc}, **cfg),
        "batch_ok": int(ok_b.sum()), "batch_result_rows": int(rows_t[ok_b].sum()),
        "per_query_qps": qps_perq, "batched_vs_per_query": qps_batched / qps_perq,
        "single_query": single,
        "edges_per_s": (edges / sum(r["latency_us_mean"] * 1e-6 for r in single.values())) if single else None,
        "serve": serve, "serve_mix": mix_info, "partition": partition, "final_process": final, "roofline": roofline,
        "cpu_baseline": cpu, "parity_gate": gate,
        "gstore": dict(gstore, h2d_bytes=dev.stats()["h2d_bytes"], segments_evicted=dev.stats()["segments_evicted"]),
    })
    dev.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
