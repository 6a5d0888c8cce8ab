// Batched execution: all queries of a batch advance one pattern step per set
// of launches over one concatenated history table with per-query offsets.
// Sizes stay on the device between steps, so the whole multi-step sequence
// is launched without host round trips and can be replayed from a hipGraph.
#include <algorithm>

#include "device.hpp"

namespace wk {

namespace {

// Pattern steps per captured graph.  With six kernels per step, a single
// graph of a 5-step batch (30 kernels) showed a gap of up to 90 us before its
// 16th kernel in every batch, and graphs of two steps were used.  With five
// kernels per step, one graph of the 5-step batch (25 kernels) is 1-4%
// faster than three graphs of at most two steps (0.280-0.284 against
// 0.284-0.287 ms per pipelined 4,096-query batch; 0.473 against 0.483-0.494
// ms one at a time), and no slower on the 8-step power-law stream.
constexpr uint32_t kStepsPerGraph = 5;

void launch_batch_steps(wk_dev* dev, uint32_t k0, uint32_t k1, uint32_t S, uint32_t nq, bool timed) {
  for (uint32_t k = k0; k < k1; k++) {
    const uint32_t* tin = dev->tbuf[(k + 1) & 1];
    uint32_t* tout = dev->tbuf[k & 1];
    StepArgs a = make_args(dev, k, nq, tin, tout, dev->tcap, k + 1 < S);
    a.qres = dev->d_qres;
    hipEvent_t ev = nullptr;
    int w = dev->opts.probe_width;
    if (nq == 1) {
      // a key read from a direct index is one 8-B word, so lanes sharing a
      // key would only repeat it: one lane per key (as the per-query path)
      const StepDesc& d = dev->h_desc[size_t(k) * dev->nq_cap];
      if (d.seg >= 0 && dev->h_segs[d.seg].dense && d.key_mode != KEY_TYPE_INDEX) w = 1;
    }
    if (timed) timed_begin(dev, &ev);
    launch_probe(a, w, 0, dev->stream, k == 0);  // step 0 reads one row per query
    if (timed) timed_end(dev, ev, "batch_probe", 0);
    if (timed) timed_begin(dev, &ev);
    launch_expand_tail(a, dev->opts.rows_per_wave, dev->stream);
    if (timed) timed_end(dev, ev, "batch_scan_gather", 0);
  }
}

// Algorithmic bytes of the S pattern steps of the batch just run, from the
// row counts its scans left in the step descriptors.
int count_traffic(wk_dev* dev, uint32_t S, uint32_t nq) {
  const size_t qc = dev->nq_cap;
  if (hip_failed(hipMemcpy(dev->h_desc, dev->d_desc, ((S - 1) * qc + nq) * sizeof(StepDesc), hipMemcpyDeviceToHost)))
    return WK_E_HIP;
  double bytes = 0, edges = 0;
  for (uint32_t k = 0; k < S; k++)
    for (uint32_t i = 0; i < nq; i++) {
      const StepDesc& d = dev->h_desc[k * qc + i];
      if (d.op == OP_SKIP || d.in_rows == 0) continue;
      const double R = double(d.in_rows), C = double(d.in_cols), E = double(d.out_rows), Co = double(d.out_cols);
      if (d.op == OP_EXPAND && d.key_mode == KEY_CONST) {
        bytes += 16 + 8 * E;
        edges += E;
      } else if (d.op == OP_EXPAND) {
        bytes += 4 * R * C + 16 * R + 4 * E + 4 * E * Co;
        edges += E;
      } else {
        bytes += 4 * R * C + 16 * R + 4 * R + 4 * E * Co;
        edges += R;
      }
    }
  dev->traffic_bytes += bytes;
  dev->traffic_edges += edges;
  dev->traffic_batches++;
  return WK_OK;
}

}  // namespace

int split_batch(wk_dev* dev, const std::vector<Plan>& plans, uint32_t q0, uint32_t q1, uint64_t arena_base,
                bool use_graph, uint64_t* rows, uint64_t* offsets, int32_t* status, uint64_t* arena_end) {
  if (dev->opts.profile) dev->kstats["batch:splits"].launches++;
  const uint32_t mid = q0 + (q1 - q0) / 2;
  uint64_t end1 = arena_base;
  int rc = run_batch(dev, plans, q0, mid, arena_base, use_graph, rows, offsets, status, &end1);
  if (rc) return rc;
  return run_batch(dev, plans, mid, q1, end1, use_graph, rows, offsets, status, arena_end);
}

int run_alone(wk_dev* dev, const Plan& plan, uint64_t arena_base, uint64_t* rows, uint64_t* offset, int32_t* status,
              uint64_t* arena_end) {
  *arena_end = arena_base;
  *rows = 0;
  *offset = 0;
  if (plan.status || plan.steps.empty()) {
    *status = plan.status;
    return WK_OK;
  }
  uint64_t r = 0;
  const int rc = run_query_to(dev, plan, dev->arena + arena_base, dev->arena_cap - std::min(arena_base, dev->arena_cap),
                              &r);
  if (rc == WK_E_HIP || rc == WK_E_NO_DEVICE) return rc;
  *status = rc;
  if (rc == WK_OK) {
    *rows = r;
    *offset = arena_base;
    *arena_end = arena_base + r * plan.cols;
  }
  dev->stats.queries_run++;
  return WK_OK;
}

int begin_batch(wk_dev* dev, const std::vector<Plan>& plans, uint32_t q0, uint32_t q1, uint64_t arena_base,
                bool use_graph, uint64_t* rows, uint64_t* offsets, int32_t* status, BatchLaunch& bl) {
  const uint32_t nq = q1 - q0;
  double t = host_now();
  bl.running = false;
  bl.S = 0;
  bl.arena_base = bl.arena_end = arena_base;
  bl.segs.clear();
  uint32_t S = 0;
  for (uint32_t q = q0; q < q1; q++) {
    rows[q] = 0;
    offsets[q] = 0;
    status[q] = plans[q].status;
    if (plans[q].status == WK_OK) S = std::max<uint32_t>(S, uint32_t(plans[q].steps.size()));
  }
  if (nq == 0 || S == 0) return WK_OK;
  int rc = reserve_steps(dev, S, nq);
  if (rc) return rc;
  // descriptors [step][query]: the query's own steps, then skips.  The same
  // pass marks the segments the batch reads (one byte each, with atomic
  // reads and writes), which are then made resident before anything is
  // launched.
  std::vector<uint8_t> seen(dev->seg.size(), 0);
  uint8_t* const mark = seen.data();
  const size_t qc = dev->nq_cap;
  StepDesc skip{};
  skip.op = OP_SKIP;
  skip.seg = -1;
#pragma omp parallel for schedule(static) num_threads(batch_threads()) if (nq >= 1024)
  for (uint32_t i = 0; i < nq; i++) {
    const Plan& p = plans[q0 + i];
    const uint32_t ns = p.status == WK_OK ? uint32_t(p.steps.size()) : 0;
    for (uint32_t k = 0; k < S; k++) {
      StepDesc* d = dev->h_desc + k * qc + i;
      if (k < ns) {
        const int sg = p.steps[k].seg;
        if (sg >= 0) {
          // read first: once a mark is set the line stays shared, unwritten
          uint8_t m;
#pragma omp atomic read
          m = mark[sg];
          if (!m) {
#pragma omp atomic write
            mark[sg] = 1;
          }
        }
        *d = p.steps[k];
        d->final_step = (k + 1 == ns) ? 1 : 0;
        d->in_rows = k == 0 ? 1 : 0;
        d->in_off = 0;
        d->out_rows = d->out_off = 0;
      } else {
        *d = skip;
      }
    }
  }
  // step 0: one input row (the constant start) per query that has steps
  uint64_t nvalid = 0;
  for (uint32_t i = 0; i < nq; i++) {
    dev->h_rs[i] = nvalid;
    const Plan& p = plans[q0 + i];
    nvalid += (p.status == WK_OK && !p.steps.empty()) ? 1 : 0;
  }
  dev->h_rs[nq] = nvalid;
  t = host_phase(dev, "host:batch_fill", t);
  std::vector<int>& segs = bl.segs;
  for (size_t i = 0; i < seen.size(); i++)
    if (seen[i]) segs.push_back(int(i));
  rc = ensure_segments(dev, segs);
  if (rc == WK_E_NOMEM_BUDGET || rc == WK_E_SEG_TOO_LARGE) {
    // the segments of all the batch's steps do not fit the gstore budget
    // together: run the halves one after the other, as for a table overflow
    segs.clear();
    if (nq > 1) return split_batch(dev, plans, q0, q1, arena_base, use_graph, rows, offsets, status, &bl.arena_end);
    return run_alone(dev, plans[q0], arena_base, &rows[q0], &offsets[q0], &status[q0], &bl.arena_end);
  }
  if (rc) {
    segs.clear();
    return rc;
  }
  t = host_phase(dev, "host:batch_segs", t);
  for (uint32_t k = 0; k <= S; k++) dev->h_state[k] = StepState{};
  dev->h_state[0].arena_base = arena_base;
  auto fail = [&](int code) {
    unpin_segments(dev, segs);
    segs.clear();
    return code;
  };
  // steps 0..S-1 are read; the last step writes no next-step descriptors
  if (hip_failed(hipMemcpyAsync(dev->d_desc, dev->h_desc, ((S - 1) * qc + nq) * sizeof(StepDesc), hipMemcpyHostToDevice,
                     dev->stream)) ||
      hip_failed(hipMemcpyAsync(dev->d_rs, dev->h_rs, (nq + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, dev->stream)) ||
      hip_failed(hipMemcpyAsync(dev->d_state, dev->h_state, (S + 1) * sizeof(StepState), hipMemcpyHostToDevice,
                     dev->stream)))
    return fail(WK_E_HIP);

  if (use_graph) {
    for (uint32_t k0 = 0; k0 < S; k0 += kStepsPerGraph) {
      const uint32_t k1 = std::min(S, k0 + kStepsPerGraph);
      auto key = std::make_pair((k0 << 16) | S, nq);
      auto it = dev->graphs.find(key);
      if (it == dev->graphs.end()) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        if (hip_failed(hipStreamBeginCapture(dev->stream, hipStreamCaptureModeRelaxed))) return fail(WK_E_HIP);
        launch_batch_steps(dev, k0, k1, S, nq, false);
        if (hip_failed(hipStreamEndCapture(dev->stream, &g))) return fail(WK_E_HIP);
        if (hip_failed(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0))) {
          (void)hipGraphDestroy(g);
          return fail(WK_E_HIP);
        }
        (void)hipGraphDestroy(g);
        it = dev->graphs.emplace(key, ge).first;
      }
      if (hip_failed(hipGraphLaunch(it->second, dev->stream))) return fail(WK_E_HIP);
    }
  } else {
    launch_batch_steps(dev, 0, S, S, nq, dev->opts.profile != 0);
  }
  // the per-query results and step states come back behind the steps
  if (hip_failed(hipMemcpyAsync(dev->h_qres, dev->d_qres, size_t(nq) * 2 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                dev->stream)) ||
      hip_failed(hipMemcpyAsync(dev->h_state, dev->d_state, S * sizeof(StepState), hipMemcpyDeviceToHost, dev->stream)))
    return fail(WK_E_HIP);
  host_phase(dev, "host:batch_upload_launch", t);
  bl.S = S;
  bl.running = true;
  return WK_OK;
}

int finish_batch(wk_dev* dev, const std::vector<Plan>& plans, uint32_t q0, uint32_t q1, bool use_graph,
                 uint64_t* rows, uint64_t* offsets, int32_t* status, BatchLaunch& bl) {
  if (!bl.running) return WK_OK;
  bl.running = false;
  const uint32_t nq = q1 - q0, S = bl.S;
  const uint64_t arena_base = bl.arena_base;
  double t = host_now();
  const bool ok = !hip_failed(hipStreamSynchronize(dev->stream));
  unpin_segments(dev, bl.segs);
  bl.segs.clear();
  if (!ok) return WK_E_HIP;
  host_phase(dev, "host:batch_wait", t);
  if (dev->traffic_on) {
    int trc = count_traffic(dev, S, nq);
    if (trc) return trc;
  }
  bool overflow = false;
  for (uint32_t k = 0; k < S; k++) overflow |= dev->h_state[k].status != 0;
  if (overflow) {
    // an intermediate table overflowed: split the batch; halves reuse the
    // tables and append to the arena
    if (nq == 1) return run_alone(dev, plans[q0], arena_base, &rows[q0], &offsets[q0], &status[q0], &bl.arena_end);
    return split_batch(dev, plans, q0, q1, arena_base, use_graph, rows, offsets, status, &bl.arena_end);
  }
  uint64_t end = arena_base;
  for (uint32_t i = 0; i < nq; i++) {
    const uint32_t q = q0 + i;
    const Plan& p = plans[q];
    if (p.status != WK_OK || p.steps.empty()) continue;
    if (dev->h_qres[2 * size_t(i) + 1] == kNoSpace) {
      // its final table did not fit the rest of the arena
      status[q] = WK_E_NOMEM_BUDGET;
      continue;
    }
    rows[q] = dev->h_qres[2 * size_t(i)];
    offsets[q] = dev->h_qres[2 * size_t(i) + 1];
    end = std::max<uint64_t>(end, offsets[q] + rows[q] * p.cols);
  }
  bl.arena_end = end;
  dev->stats.queries_run += nq;
  return WK_OK;
}

int run_batch(wk_dev* dev, const std::vector<Plan>& plans, uint32_t q0, uint32_t q1, uint64_t arena_base,
              bool use_graph, uint64_t* rows, uint64_t* offsets, int32_t* status, uint64_t* arena_end) {
  BatchLaunch bl;
  int rc = begin_batch(dev, plans, q0, q1, arena_base, use_graph, rows, offsets, status, bl);
  if (!rc) rc = finish_batch(dev, plans, q0, q1, use_graph, rows, offsets, status, bl);
  *arena_end = bl.arena_end;
  return rc;
}

int run_one(wk_dev* dev, const Plan& plan, uint64_t* rows) {
  *rows = 0;
  if (dev->opts.profile) return run_query_to(dev, plan, dev->arena, dev->arena_cap, rows);
  if (plan.status || plan.steps.empty()) return plan.status;
  dev->one_plan.resize(1);
  dev->one_plan[0] = plan;
  uint64_t offset = 0, end = 0;
  int32_t status = WK_OK;
  // direct launches: for one query they measured faster than a graph replay
  const int rc = run_batch(dev, dev->one_plan, 0, 1, 0, false, rows, &offset, &status, &end);
  if (rc) return rc;
  return status;  // a table that fits starts at offset 0: the query's arena base
}

}  // namespace wk
