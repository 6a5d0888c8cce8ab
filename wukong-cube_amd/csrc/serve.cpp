// Serving: open-loop replay of an arrival trace with continuous batching.
//
// The model is the reference's emulator and engine loop (run_query_emu,
// src/client/console_proxy.hpp:287-439; Engine::run, core/engine/engine.hpp:132-233)
// moved onto the device: every scheduler iteration advances each in-flight
// query by one pattern step in one set of launches; queries that arrived are
// admitted whenever an in-flight slot and projected table footprint are free,
// and finished queries leave at the end of the iteration that completed them.
#include <algorithm>
#include <chrono>
#include <thread>

#include "device.hpp"

namespace wk {
namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kHeavyBatch = 8;  // heavy queries launched together on the second lane

struct QState {
  uint32_t k = 0;      // next step
  uint64_t rows = 0;   // rows of the current table
  uint64_t off = 0;    // element offset of the current table in the input buffer
  double admit = 0;
  uint32_t retries = 0;  // restarts after its step did not fit beside others
};

// Projected history-table bytes of a query from step k on, given the rows of
// its table before that step: the exact cardinality of a constant start from
// the host store, then the average degree of each expanded segment.  At
// admission (k = 0, one virtual start row) this underestimates queries that
// walk through high-degree vertices; the measured table after each step
// corrects it.
uint64_t project_footprint(const Store& st, const Plan& p, size_t k0 = 0, uint64_t rows0 = 1) {
  double rows = double(rows0), best = 0;
  for (size_t k = k0; k < p.steps.size(); k++) {
    const StepDesc& sd = p.steps[k];
    const double in = rows;
    if (sd.op == OP_EXPAND) {
      if (sd.key_mode == KEY_CONST) {
        uint64_t sz = 0;
        st.get_triples(sd.kstart, sd.pid, int(sd.dir), sz);
        rows = double(sz);
      } else if (sd.seg >= 0) {
        const SegMeta& m = st.segs[size_t(sd.seg)];
        rows = in * (m.num_keys ? double(m.num_edges) / double(m.num_keys) : 0.0);
      } else {
        rows = 0;
      }
    }
    const double bytes = (in * sd.in_cols + rows * sd.out_cols) * 4.0 + in * 16.0;
    best = std::max(best, bytes);
  }
  return uint64_t(best) + 64;
}

struct Server {
  wk_dev* dev;
  const std::vector<Plan>& plans;
  std::vector<QState>& qs;

  Server(wk_dev* d, const std::vector<Plan>& p, std::vector<QState>& s) : dev(d), plans(p), qs(s) {}

  // One step for queries F[a, b): outputs from tout_base / arena_base.  On
  // overflow the range is halved; a single query that cannot fit fails.
  // `res` receives the step descriptor of each query after the step.
  int step_range(const std::vector<uint32_t>& F, size_t a, size_t b, int cur, uint64_t tout_base,
                 uint64_t arena_base, uint64_t* used_t, uint64_t* used_a, std::vector<StepDesc>& res,
                 std::vector<int>& failed, bool checksum_now) {
    const uint32_t n = uint32_t(b - a);
    *used_t = tout_base;
    *used_a = arena_base;
    if (n == 0) return WK_OK;
    // two descriptor sets of at least 2 queries hold the staging block below
    int rc = reserve_steps(dev, 1, std::max<uint32_t>(n, 2));
    if (rc) return rc;
    // one staging block each way: [descriptors n][row starts n+1][state][sums n]
    // in the step-descriptor buffers (sized for two steps of nq_cap queries),
    // so an iteration is one upload and one download
    const size_t off_rs = size_t(n) * sizeof(StepDesc);
    const size_t off_st = (off_rs + size_t(n + 1) * sizeof(uint64_t) + 15) & ~size_t(15);
    const size_t off_ck = off_st + sizeof(StepState);
    const size_t end_ck = off_ck + size_t(n) * sizeof(uint64_t);
    uint8_t* const hb = reinterpret_cast<uint8_t*>(dev->h_desc);
    uint8_t* const db = reinterpret_cast<uint8_t*>(dev->d_desc);
    StepDesc* const hdesc = dev->h_desc;
    uint64_t* const hrs = reinterpret_cast<uint64_t*>(hb + off_rs);
    StepState* const hst = reinterpret_cast<StepState*>(hb + off_st);
    const uint64_t* const hsum = reinterpret_cast<const uint64_t*>(hb + off_ck);
    std::vector<int> segs;
    // segments already listed, one byte each (a linear search of `segs` per
    // query grew with in-flight queries times segments)
    std::vector<uint8_t> listed(dev->seg.size(), 0);
    uint64_t R = 0;
    uint32_t nfinal = 0;
    for (uint32_t i = 0; i < n; i++) {
      const uint32_t q = F[a + i];
      const Plan& p = plans[q];
      StepDesc d = p.steps[qs[q].k];
      d.final_step = (qs[q].k + 1 == p.steps.size()) ? 1 : 0;
      d.in_rows = qs[q].k == 0 ? 1 : qs[q].rows;
      d.in_off = qs[q].k == 0 ? 0 : qs[q].off;
      hdesc[i] = d;
      nfinal += d.final_step;
      hrs[i] = R;
      R += d.in_rows;
      if (d.seg >= 0 && !listed[size_t(d.seg)]) {
        listed[size_t(d.seg)] = 1;
        segs.push_back(d.seg);
      }
    }
    hrs[n] = R;
    if (R > dev->rcap) {
      if (n == 1) {
        failed.push_back(int(a));
        return WK_OK;
      }
      goto split;
    }
    {
      *hst = StepState{};
      hst->tout_base = tout_base;
      hst->arena_base = arena_base;
      rc = ensure_segments(dev, segs);
      if (rc == WK_E_NOMEM_BUDGET || rc == WK_E_SEG_TOO_LARGE) {
        // the step's segments do not fit the gstore budget together
        if (n == 1) {
          failed.push_back(int(a));
          return WK_OK;
        }
        goto split;
      }
      if (rc) return rc;
      const uint32_t* tin = dev->tbuf[cur];
      uint32_t* tout = dev->tbuf[cur ^ 1];
      if (hip_failed(hipMemcpyAsync(db, hb, off_ck, hipMemcpyHostToDevice, dev->stream))) {
        unpin_segments(dev, segs);
        return WK_E_HIP;
      }
      StepArgs args = make_args(dev, 0, n, tin, tout, dev->tcap, false);
      args.row_start = reinterpret_cast<uint64_t*>(db + off_rs);
      args.st = reinterpret_cast<StepState*>(db + off_st);
      // the gather sums the row hashes of each final table as it writes it
      // (the query scan zeroes the sums first)
      const bool ck = checksum_now && nfinal > 0;
      if (ck) args.ck = reinterpret_cast<unsigned long long*>(db + off_ck);
      launch_probe(args, dev->opts.probe_width, R, dev->stream);
      launch_expand_tail(args, dev->opts.rows_per_wave, dev->stream);
      if (hip_failed(hipMemcpyAsync(hb, db, ck ? end_ck : off_ck, hipMemcpyDeviceToHost, dev->stream)) ||
          hip_failed(hipStreamSynchronize(dev->stream))) {
        unpin_segments(dev, segs);
        return WK_E_HIP;
      }
      unpin_segments(dev, segs);
      if (hst->status == 0) {
        for (uint32_t i = 0; i < n; i++) res[a + i] = hdesc[i];
        *used_t = tout_base + hst->out_elems;
        uint64_t ua = arena_base;
        for (uint32_t i = 0; i < n; i++) {
          if (!hdesc[i].final_step) continue;
          if (hdesc[i].out_off == kNoSpace) {
            failed.push_back(int(a + i));  // its final table did not fit the arena
            continue;
          }
          ua = std::max<uint64_t>(ua, hdesc[i].out_off + hdesc[i].out_rows * hdesc[i].out_cols);
        }
        *used_a = ua;
        if (ck)
          for (uint32_t i = 0; i < n; i++)
            if (hdesc[i].final_step) res[a + i].in_off = hsum[i];  // carry the checksum
        return WK_OK;
      }
      if (n == 1) {
        failed.push_back(int(a));
        return WK_OK;
      }
    }
  split : {
    if (dev->opts.profile) dev->kstats["serve:splits"].launches++;
    const size_t m = a + n / 2;
    uint64_t t1 = 0, a1 = 0;
    rc = step_range(F, a, m, cur, tout_base, arena_base, &t1, &a1, res, failed, checksum_now);
    if (rc) return rc;
    return step_range(F, m, b, cur, t1, a1, used_t, used_a, res, failed, checksum_now);
  }
  }
};

// Heavy queries on the second lane: a batch of them runs there, all its
// pattern steps launched at once on that lane's stream, while the light
// queries' iterations go on on the main lane; the scheduler polls the batch's
// completion event.  A query enters the lane at its first step (projected
// heavy at admission) or, once a main-lane step has made a table whose
// projection is heavy, at its next step, continuing from a copy of that
// table.  Each query's final table and its row-hash sum (summed by the
// gather as it writes the rows) stay on the lane until finished.
struct HeavyLane {
  // A query of a batch: its first step on the lane and, when k0 > 0, the rows
  // and element offset of its table in the main lane's current input buffer.
  struct Entry {
    uint32_t q, k0;
    uint64_t rows, off;
  };

  wk_dev* dev;
  const std::vector<Plan>& plans;
  bool busy = false;
  std::vector<uint32_t> q;   // the batch's queries
  std::vector<int> segs;     // its segments, pinned while it runs
  uint32_t S = 0;
  hipEvent_t done = nullptr;
  hipEvent_t seeded = nullptr;

  HeavyLane(wk_dev* d, const std::vector<Plan>& p) : dev(d), plans(p) {}
  ~HeavyLane() {
    if (done) (void)hipEventDestroy(done);
    if (seeded) (void)hipEventDestroy(seeded);
  }

  // The second lane's stream, tables and scratch, made once per device.
  int ready() {
    if (!done) WK_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    if (!seeded) WK_HIP(hipEventCreateWithFlags(&seeded, hipEventDisableTiming));
    return ensure_alt_lane(dev);
  }

  // Elements of the lane's first input table (the continued tables).
  uint64_t seed_cap() const { return dev->alt[0].tcap; }

  // Launch the batch `b` on the second lane; `main_tin` is the main lane's
  // current input buffer, which holds the continued tables.  Returns
  // WK_E_NOMEM_BUDGET when its segments do not fit the gstore beside the
  // pinned ones (nothing is launched then).
  int launch(const std::vector<Entry>& b, const uint32_t* main_tin) {
    hipStream_t main = dev->stream;
    swap_lanes(dev);
    const int rc = launch_swapped(b, main_tin, main);
    swap_lanes(dev);
    return rc;
  }

  int launch_swapped(const std::vector<Entry>& b, const uint32_t* main_tin, hipStream_t main) {
    const uint32_t n = uint32_t(b.size());
    S = 0;
    for (const Entry& e : b) S = std::max<uint32_t>(S, uint32_t(plans[e.q].steps.size()) - e.k0);
    int rc = reserve_steps(dev, S, n);
    if (rc) return rc;
    const size_t qc = dev->nq_cap;
    std::vector<uint8_t> listed(dev->seg.size(), 0);
    std::vector<uint64_t> seed_off(n, 0);
    segs.clear();
    StepDesc skip{};
    skip.op = OP_SKIP;
    skip.seg = -1;
    uint64_t R = 0, seed = 0;
    for (uint32_t i = 0; i < n; i++) {
      const Entry& e = b[i];
      const Plan& p = plans[e.q];
      const uint32_t ns = uint32_t(p.steps.size()) - e.k0;
      if (e.k0) {
        const uint64_t elems = e.rows * p.steps[e.k0].in_cols;
        if (seed + elems > dev->tcap) return WK_E_INVALID_ARG;  // the caller bounds the seeds
        seed_off[i] = seed;
        seed += elems;
      }
      dev->h_rs[i] = R;
      R += e.k0 ? e.rows : 1;
      for (uint32_t k = 0; k < S; k++) {
        StepDesc* d = dev->h_desc + k * qc + i;
        if (k < ns) {
          *d = p.steps[e.k0 + k];
          d->final_step = (k + 1 == ns) ? 1 : 0;
          d->in_rows = k ? 0 : e.k0 ? e.rows : 1;
          d->in_off = k ? 0 : seed_off[i];
          d->out_rows = d->out_off = 0;
          if (d->seg >= 0 && !listed[size_t(d->seg)]) {
            listed[size_t(d->seg)] = 1;
            segs.push_back(d->seg);
          }
        } else {
          *d = skip;
        }
      }
    }
    dev->h_rs[n] = R;
    if (R > dev->rcap) return WK_E_INVALID_ARG;
    rc = ensure_segments(dev, segs);
    if (rc) return rc;
    for (uint32_t k = 0; k <= S; k++) dev->h_state[k] = StepState{};
    bool bad = hip_failed(hipMemcpyAsync(dev->d_desc, dev->h_desc, ((S - 1) * qc + n) * sizeof(StepDesc),
                                         hipMemcpyHostToDevice, dev->stream)) ||
               hip_failed(hipMemcpyAsync(dev->d_rs, dev->h_rs, (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice,
                                         dev->stream)) ||
               hip_failed(hipMemcpyAsync(dev->d_state, dev->h_state, (S + 1) * sizeof(StepState),
                                         hipMemcpyHostToDevice, dev->stream));
    if (seed && !bad) {
      // the continued tables into the lane's first input buffer; the main
      // lane's next work waits for the copies, as its iteration after next
      // overwrites that buffer
      for (uint32_t i = 0; i < n && !bad; i++)
        if (b[i].k0)
          bad = hip_failed(hipMemcpyAsync(dev->tbuf[1] + seed_off[i], main_tin + b[i].off,
                                          b[i].rows * plans[b[i].q].steps[b[i].k0].in_cols * sizeof(uint32_t),
                                          hipMemcpyDeviceToDevice, dev->stream));
      bad = bad || hip_failed(hipEventRecord(seeded, dev->stream)) || hip_failed(hipStreamWaitEvent(main, seeded, 0));
    }
    if (bad) {
      unpin_segments(dev, segs);
      return WK_E_HIP;
    }
    for (uint32_t k = 0; k < S; k++) {
      StepArgs a = make_args(dev, k, n, dev->tbuf[(k + 1) & 1], dev->tbuf[k & 1], dev->tcap, k + 1 < S);
      a.qres = dev->d_qres;
      a.ck = reinterpret_cast<unsigned long long*>(dev->d_ck);
      launch_probe(a, dev->opts.probe_width, 0, dev->stream);
      launch_expand_tail(a, dev->opts.rows_per_wave, dev->stream);
    }
    if (hip_failed(hipMemcpyAsync(dev->h_qres, dev->d_qres, size_t(n) * 2 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  dev->stream)) ||
        hip_failed(hipMemcpyAsync(dev->h_state, dev->d_state, S * sizeof(StepState), hipMemcpyDeviceToHost,
                                  dev->stream)) ||
        hip_failed(hipMemcpyAsync(dev->h_ck, dev->d_ck, size_t(n) * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  dev->stream)) ||
        hip_failed(hipEventRecord(done, dev->stream))) {
      unpin_segments(dev, segs);
      return WK_E_HIP;
    }
    q.clear();
    for (const Entry& e : b) q.push_back(e.q);
    busy = true;
    return WK_OK;
  }

  bool finished() const { return busy && hipEventQuery(done) == hipSuccess; }
};

}  // namespace
}  // namespace wk

using namespace wk;

extern "C" int wk_serve_replay(wk_dev* dev, const wk_query* queries, const double* arrival_s, uint32_t nq,
                               const wk_serve_opts* opts, wk_serve_result* out) {
  if (!dev || (!queries && nq) || (!arrival_s && nq) || !opts || !out) return WK_E_INVALID_ARG;
  if (int drc = drain_batches(dev, 2)) return drc;  // the main lane and the heavy-query lane
  for (uint32_t i = 1; i < nq; i++)
    if (arrival_s[i] < arrival_s[i - 1]) return WK_E_INVALID_ARG;
  std::vector<Plan> plans(nq);
  for (uint32_t i = 0; i < nq; i++) compile_query(*dev->st, queries[i], plans[i], dev->type_slot.data());
  std::vector<QState> qs(nq);
  std::vector<uint64_t> fp(nq);
  for (uint32_t i = 0; i < nq; i++) fp[i] = plans[i].status ? 0 : project_footprint(*dev->st, plans[i]);
  // projected bytes per row of a table entering step k (k >= 1, where the
  // projection is linear in the rows: only a first step starts from a
  // constant), for the check after each main-lane step
  constexpr uint64_t kGainRows = uint64_t(1) << 20;
  std::vector<std::vector<float>> gain(nq);
  for (uint32_t i = 0; i < nq; i++) {
    if (plans[i].status) continue;
    gain[i].assign(plans[i].steps.size(), 0.f);
    for (size_t k = 1; k < plans[i].steps.size(); k++)
      gain[i][k] = float(double(project_footprint(*dev->st, plans[i], k, kGainRows)) / double(kGainRows));
  }
  const uint32_t max_inflight = opts->per_query ? 1 : std::max<uint32_t>(1, opts->max_inflight);
  // queries projected above heavy_footprint run on the second lane
  const uint64_t heavy_fp = opts->heavy_footprint ? opts->heavy_footprint : (uint64_t(1) << 20);
  const bool use_heavy = !opts->per_query && heavy_fp != ~uint64_t(0);
  const uint64_t budget = opts->footprint_budget ? opts->footprint_budget : dev->tcap * 4;
  if (opts->keep_results) {
    dev->served_rows.assign(nq, {});
    dev->served_cols.assign(nq, 0);
  } else {
    dev->served_rows.clear();
    dev->served_cols.clear();
  }
  Server srv(dev, plans, qs);
  const auto t0 = Clock::now();
  auto now_s = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };
  auto finish = [&](uint32_t q, double t, uint64_t rows, int status, uint64_t ck) {
    out->latency_us[q] = (t - arrival_s[q]) * 1e6;
    out->start_us[q] = (qs[q].admit - arrival_s[q]) * 1e6;
    out->rows[q] = rows;
    out->status[q] = status;
    out->checksum[q] = ck;
  };
  auto keep = [&](uint32_t q, const uint32_t* dsrc, uint64_t rows, uint32_t cols) -> int {
    if (!opts->keep_results) return WK_OK;
    dev->served_rows[q].resize(rows * cols);
    dev->served_cols[q] = cols;
    if (rows * cols)
      WK_HIP(hipMemcpy(dev->served_rows[q].data(), dsrc, rows * cols * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return WK_OK;
  };
  uint32_t next = 0, ndone = 0;
  uint64_t batches = 0;
  int cur = 0;
  uint64_t fp_used = 0;
  std::vector<uint32_t> inflight;
  // queries whose step did not fit beside the other in-flight queries'
  // steps (tables or segments): rerun from their first pattern, alone, on
  // the per-query path, which runs steps too large for the tables in chunks
  std::vector<uint32_t> retry;
  std::vector<StepDesc> res;
  std::vector<int> failed;
  CkDesc* d_ck = nullptr;
  uint64_t* d_sum = nullptr;
  HeavyLane heavy(dev, plans);
  std::vector<uint32_t> heavy_queue;
  if (use_heavy) {
    int hr = heavy.ready();
    if (hr) return hr;
  }
  // one query by itself (its lane has nothing else in flight), its result
  // into the (then unused) result arena
  auto run_alone = [&](uint32_t q) -> int {
    uint64_t rows = 0;
    const int st = run_one(dev, plans[q], &rows);
    if (st == WK_E_HIP) return st;
    uint64_t ck = 0;
    if (st == WK_OK && rows) {
      CkDesc c{0, rows, 0, plans[q].cols, 0};
      WK_HIP(hipMemcpyAsync(d_ck, &c, sizeof(c), hipMemcpyHostToDevice, dev->stream));
      launch_fill_u64(d_sum, 1, 0, dev->stream);
      launch_checksum(dev->arena, d_ck, 1, rows, d_sum, dev->stream);
      WK_HIP(hipMemcpyAsync(&ck, d_sum, sizeof(uint64_t), hipMemcpyDeviceToHost, dev->stream));
      WK_HIP(hipStreamSynchronize(dev->stream));
      const int kr = keep(q, dev->arena, rows, plans[q].cols);
      if (kr) return kr;
    } else if (opts->keep_results) {
      dev->served_cols[q] = plans[q].cols;
    }
    finish(q, now_s(), st == WK_OK ? rows : 0, st, ck);
    ndone++;
    return WK_OK;
  };
  WK_HIP(hipMalloc(&d_ck, sizeof(CkDesc)));
  WK_HIP(hipMalloc(&d_sum, sizeof(uint64_t)));
  // a completed heavy batch: each query's rows and checksum, or, if its
  // tables overflowed the lane's or its final table did not fit the lane's
  // arena, the query alone on the chunking path (on that lane)
  auto finish_heavy = [&]() -> int {
    swap_lanes(dev);
    unpin_segments(dev, heavy.segs);
    heavy.busy = false;
    bool overflow = false;
    for (uint32_t k = 0; k < heavy.S; k++) overflow |= dev->h_state[k].status != 0;
    const double t = now_s();
    int r = WK_OK;
    std::vector<uint32_t> alone;
    for (uint32_t i = 0; i < heavy.q.size() && r == WK_OK; i++) {
      const uint32_t q = heavy.q[i];
      const uint64_t off = dev->h_qres[2 * size_t(i) + 1];
      if (overflow || off == kNoSpace) {
        alone.push_back(q);
        continue;
      }
      const uint64_t rows = dev->h_qres[2 * size_t(i)];
      r = keep(q, dev->arena + off, rows, plans[q].cols);
      finish(q, t, rows, WK_OK, rows ? dev->h_ck[i] : 0);
      ndone++;
    }
    for (uint32_t q : alone)
      if (r == WK_OK) r = run_alone(q);
    swap_lanes(dev);
    return r;
  };
  // Launch a heavy batch: the continued queries `b`, then the oldest waiting
  // heavy queries up to kHeavyBatch.  Queries whose segments do not fit the
  // gstore beside the pinned ones go back: continued ones to the main lane
  // (their tables are still there), waiting ones to run alone later.
  auto launch_heavy = [&](std::vector<HeavyLane::Entry> b) -> int {
    const size_t nc = b.size();
    const size_t nw = std::min(heavy_queue.size(), kHeavyBatch - std::min(kHeavyBatch, nc));
    for (size_t i = 0; i < nw; i++) b.push_back({heavy_queue[i], 0, 0, 0});
    const int lr = heavy.launch(b, dev->tbuf[cur]);
    if (lr && lr != WK_E_NOMEM_BUDGET && lr != WK_E_SEG_TOO_LARGE) return lr;
    heavy_queue.erase(heavy_queue.begin(), heavy_queue.begin() + nw);
    if (lr) {
      for (size_t i = 0; i < nc; i++) {
        inflight.push_back(b[i].q);
        fp_used += fp[b[i].q];
      }
      for (size_t i = nc; i < b.size(); i++) retry.push_back(b[i].q);
    } else if (dev->opts.profile) {
      dev->kstats["serve:heavy_batches"].launches++;
      dev->kstats["serve:heavy_queries"].launches += b.size();
      dev->kstats["serve:heavy_continued"].launches += nc;
    }
    return WK_OK;
  };
  int rc = WK_OK;
  while (ndone < nq && rc == WK_OK) {
    if (heavy.finished()) {
      rc = finish_heavy();
      if (rc) break;
    }
    double now = now_s();
    if (!retry.empty() && inflight.empty()) {
      const uint32_t q = retry.front();
      retry.erase(retry.begin());
      batches++;
      rc = run_alone(q);
      continue;
    }
    while (retry.empty() && next < nq && arrival_s[next] <= now && inflight.size() < max_inflight) {
      const uint32_t q = next;
      if (plans[q].status || plans[q].steps.empty()) {
        qs[q].admit = now;
        finish(q, now, 0, plans[q].status, 0);
        if (opts->keep_results) dev->served_cols[q] = plans[q].cols;
        ndone++;
        next++;
        continue;
      }
      if (use_heavy && fp[q] >= heavy_fp) {
        qs[q] = QState{};
        qs[q].admit = now;
        heavy_queue.push_back(q);
        next++;
        continue;
      }
      if (!inflight.empty() && fp_used + fp[q] > budget) break;
      qs[q] = QState{};
      qs[q].admit = now;
      fp_used += fp[q];
      inflight.push_back(q);
      next++;
    }
    if (!heavy.busy && !heavy_queue.empty()) {
      rc = launch_heavy({});
      if (rc) break;
    }
    if (inflight.empty()) {
      double wait = next < nq ? arrival_s[next] - now_s() : 1.0;
      if (heavy.busy) wait = std::min(wait, 50e-6);  // poll the heavy lane
      if (wait > 200e-6) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 100e-6));
      else if (heavy.busy && wait > 0) std::this_thread::yield();
      continue;
    }
    batches++;
    if (opts->per_query) {
      // the per-query baseline: one query at a time, FIFO
      const uint32_t q = inflight[0];
      inflight.clear();
      fp_used -= fp[q];
      rc = run_alone(q);
      continue;
    }
    res.assign(inflight.size(), StepDesc{});
    failed.clear();
    uint64_t used_t = 0, used_a = 0;
    const double t_it = host_now();
    rc = srv.step_range(inflight, 0, inflight.size(), cur, 0, 0, &used_t, &used_a, res, failed, true);
    if (rc) break;
    host_phase(dev, "serve:iteration", t_it);
    const double t = now_s();
    std::vector<uint32_t> still;
    for (size_t i = 0; i < inflight.size(); i++) {
      const uint32_t q = inflight[i];
      if (std::find(failed.begin(), failed.end(), int(i)) != failed.end()) {
        // rerun from its first pattern, alone, once the others have drained
        fp_used -= fp[q];
        qs[q].retries++;
        retry.push_back(q);
        continue;
      }
      const StepDesc& d = res[i];
      if (d.final_step || d.out_rows == 0) {
        // an empty table stays empty through the remaining patterns
        uint64_t ck = d.final_step ? d.in_off : 0;
        if (d.final_step) rc = keep(q, dev->arena + d.out_off, d.out_rows, d.out_cols);
        else if (opts->keep_results) dev->served_cols[q] = plans[q].cols;
        finish(q, t, d.final_step ? d.out_rows : 0, WK_OK, ck);
        fp_used -= fp[q];
        ndone++;
      } else {
        qs[q].rows = d.out_rows;
        qs[q].off = d.out_off;
        qs[q].k++;
        still.push_back(q);
      }
    }
    inflight.swap(still);
    cur ^= 1;
    if (use_heavy && !heavy.busy && !inflight.empty()) {
      // queries whose measured table projects past heavy_footprint continue
      // on the second lane from that table (now in tbuf[cur])
      std::vector<HeavyLane::Entry> b;
      uint64_t seed = 0;
      still.clear();
      for (uint32_t q : inflight) {
        const uint32_t k = qs[q].k;
        const uint64_t elems = qs[q].rows * plans[q].steps[k].in_cols;
        if (b.size() < kHeavyBatch && double(qs[q].rows) * gain[q][k] >= double(heavy_fp) &&
            seed + elems + kHeavyBatch <= heavy.seed_cap()) {  // + the waiting queries' start rows
          b.push_back({q, k, qs[q].rows, qs[q].off});
          seed += elems;
          fp_used -= fp[q];
        } else {
          still.push_back(q);
        }
      }
      if (!b.empty()) {
        inflight.swap(still);
        rc = launch_heavy(std::move(b));
      }
    }
  }
  if (heavy.busy) {
    // stopped by an error: let the heavy batch finish before its segments
    // are released
    (void)hipEventSynchronize(heavy.done);
    unpin_segments(dev, heavy.segs);
  }
  if (d_ck) (void)hipFree(d_ck);
  if (d_sum) (void)hipFree(d_sum);
  out->wall_s = now_s();
  out->batches = batches;
  return rc;
}

extern "C" int wk_serve_result_rows(wk_dev* dev, uint32_t i, uint32_t* dst, uint64_t cap_elems, uint64_t* rows,
                                    uint32_t* cols) {
  if (!dev || i >= dev->served_cols.size()) return WK_E_STATE;
  const uint32_t c = dev->served_cols[i];
  const uint64_t n = dev->served_rows[i].size();
  if (rows) *rows = c ? n / c : 0;
  if (cols) *cols = c;
  if (dst) {
    if (cap_elems < n) return WK_E_INVALID_ARG;
    std::copy(dev->served_rows[i].begin(), dev->served_rows[i].end(), dst);
  }
  return WK_OK;
}
