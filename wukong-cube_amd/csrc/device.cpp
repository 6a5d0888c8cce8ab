// Host orchestration of the device query path: gstore block cache (GPUCache),
// table/scratch arena (GPUMem), query compilation (the dispatch of
// GPUEngine::execute_one_pattern) and per-query / batched execution.
#include "device.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace wk {

thread_local int g_last_hip_error = 0;

// ---------------------------------------------------------------- compile

namespace {
enum VStat { KNOWN_VAR = 0, UNKNOWN_VAR = 1, CONST_VAR = 2 };  // core/sparql/query.hpp:56
}

int compile_query(const Store& st, const wk_query& q, Plan& plan) {
  plan.steps.clear();
  plan.cols = 0;
  plan.status = WK_OK;
  if (q.nvars < 0 || q.nvars > WK_MAX_VARS) return plan.status = WK_E_TOO_MANY_VARS;
  if (q.npatterns && !q.patterns) return plan.status = WK_E_INVALID_ARG;
  plan.v2c.assign(size_t(q.nvars), -1);
  uint32_t cols = 0;
  int err = WK_OK;
  auto stat = [&](int32_t x) -> int {
    if (x >= 0) return CONST_VAR;
    const int64_t idx = -(int64_t(x) + 1);
    if (idx >= q.nvars) {
      err = WK_E_VERTEX_INVALID;
      return UNKNOWN_VAR;
    }
    return plan.v2c[size_t(idx)] < 0 ? UNKNOWN_VAR : KNOWN_VAR;
  };
  auto col_of = [&](int32_t x) { return uint32_t(plan.v2c[size_t(-(int64_t(x) + 1))]); };
  auto bind = [&](int32_t x, uint32_t c) { plan.v2c[size_t(-(int64_t(x) + 1))] = int32_t(c); };
  auto seg_of_key = [&](uint64_t vid, uint64_t pid, uint64_t dir) {
    const SegId s = segid_of_key(make_key(vid, pid, dir));
    return st.seg_index(s.index, s.pid, s.dir);
  };
  for (uint32_t k = 0; k < q.npatterns; k++) {
    const wk_pattern& p = q.patterns[k];
    StepDesc sd{};
    sd.in_cols = cols;
    if (p.direction != DIR_IN && p.direction != DIR_OUT) return plan.status = WK_E_UNKNOWN_PATTERN;
    sd.dir = uint32_t(p.direction);
    // variable ids outside [-nvars, -1] (Result::var2col, query.hpp:425-440)
    stat(p.subject);
    stat(p.object);
    if (err) return plan.status = err;
    if (k == 0 && is_tpid(p.subject)) {
      // index_to_unknown (start_from_index, core/sparql/query.hpp:816-838)
      if (p.predicate != int32_t(PREDICATE_ID) && p.predicate != int32_t(TYPE_ID))
        return plan.status = WK_E_OBJ_ERROR;
      if (p.object >= 0) return plan.status = WK_E_VERTEX_INVALID;
      sd.op = OP_EXPAND;
      sd.key_mode = KEY_CONST;
      sd.kstart = 0;
      sd.pid = uint32_t(p.subject);
      sd.seg = st.seg_index(1, PREDICATE_ID, sd.dir);
      bind(p.object, 0);
      cols = 1;
      sd.out_cols = 1;
      plan.steps.push_back(sd);
      continue;
    }
    if (p.predicate < 0 || p.predicate >= (1 << NBITS_IDX)) return plan.status = WK_E_UNKNOWN_PATTERN;
    sd.pid = uint32_t(p.predicate);
    const int ss = stat(p.subject), so = stat(p.object);
    if (err) return plan.status = err;
    switch ((ss << 4) | so) {
      case (CONST_VAR << 4) | CONST_VAR:
        return plan.status = WK_E_UNKNOWN_PATTERN;
      case (CONST_VAR << 4) | KNOWN_VAR:  // const_to_known (sparql.hpp:139-188)
        sd.op = OP_FILTER_KNOWN;
        sd.key_mode = KEY_CONST;
        sd.kstart = uint32_t(p.subject);
        sd.ec = col_of(p.object);
        sd.seg = seg_of_key(sd.kstart, sd.pid, sd.dir);
        break;
      case (CONST_VAR << 4) | UNKNOWN_VAR:  // const_to_unknown (sparql.hpp:241-288)
        if (cols != 0) return plan.status = WK_E_FIRST_PATTERN;
        sd.op = OP_EXPAND;
        sd.key_mode = KEY_CONST;
        sd.kstart = uint32_t(p.subject);
        sd.seg = seg_of_key(sd.kstart, sd.pid, sd.dir);
        bind(p.object, cols);
        break;
      case (KNOWN_VAR << 4) | CONST_VAR:  // known_to_const (sparql.hpp:487-552)
        sd.op = OP_FILTER_CONST;
        sd.key_mode = KEY_FROM_ROW;
        sd.sc = col_of(p.subject);
        sd.kend = uint32_t(p.object);
        sd.seg = st.seg_index(0, sd.pid, sd.dir);
        break;
      case (KNOWN_VAR << 4) | KNOWN_VAR:  // known_to_known (sparql.hpp:419-479)
        sd.op = OP_FILTER_KNOWN;
        sd.key_mode = KEY_FROM_ROW;
        sd.sc = col_of(p.subject);
        sd.ec = col_of(p.object);
        sd.seg = st.seg_index(0, sd.pid, sd.dir);
        break;
      case (KNOWN_VAR << 4) | UNKNOWN_VAR:  // known_to_unknown (sparql.hpp:298-410)
        sd.op = OP_EXPAND;
        sd.sc = col_of(p.subject);
        if (sd.pid == TYPE_ID && sd.dir == DIR_IN) {
          sd.key_mode = KEY_TYPE_INDEX;  // get_index(cur, IN), sparql.hpp:342-343
          sd.seg = st.seg_index(1, PREDICATE_ID, DIR_IN);
        } else {
          sd.key_mode = KEY_FROM_ROW;
          sd.seg = st.seg_index(0, sd.pid, sd.dir);
        }
        bind(p.object, cols);
        break;
      default:  // UNKNOWN start: incorrect plan
        return plan.status = WK_E_UNKNOWN_SUB;
    }
    if (sd.op == OP_EXPAND) cols++;
    sd.out_cols = cols;
    plan.steps.push_back(sd);
  }
  plan.cols = cols;
  return WK_OK;
}

// ----------------------------------------------------------- segment cache

namespace {

int evict(wk_dev* dev, int v) {
  SegSlot& s = dev->seg[size_t(v)];
  // pushed last-first, so the next segment loaded takes them in order again
  for (auto it = s.kblk.rbegin(); it != s.kblk.rend(); ++it) dev->free_k.push_back(*it);
  for (auto it = s.vblk.rbegin(); it != s.vblk.rend(); ++it) dev->free_v.push_back(*it);
  s.kblk.clear();
  s.vblk.clear();
  s.resident = false;
  dev->h_segs[v].resident = 0;
  dev->h_segs[v].kphys = kNoPhys;
  dev->h_segs[v].vphys = kNoPhys;
  dev->copy_pending = true;
  WK_HIP(hipMemcpyAsync(dev->d_segs + v, dev->h_segs + v, sizeof(DevSeg), hipMemcpyHostToDevice,
                        dev->copy_stream));
  dev->stats.segments_evicted++;
  return WK_OK;
}

int load_segment(wk_dev* dev, int si, const std::vector<int>& keep) {
  SegSlot& s = dev->seg[size_t(si)];
  if (s.resident) {
    s.last_use = ++dev->tick;
    return WK_OK;
  }
  if (s.nkb > dev->nkblk || s.nvb > dev->nvblk) return WK_E_SEG_TOO_LARGE;
  while (dev->free_k.size() < s.nkb || dev->free_v.size() < s.nvb) {
    int victim = -1;
    uint64_t best = ~uint64_t(0);
    for (size_t i = 0; i < dev->seg.size(); i++) {
      const SegSlot& c = dev->seg[i];
      if (!c.resident || c.pins > 0 || int(i) == si) continue;
      if (std::find(keep.begin(), keep.end(), int(i)) != keep.end()) continue;
      if (c.last_use < best) {
        best = c.last_use;
        victim = int(i);
      }
    }
    if (victim < 0) return WK_E_NOMEM_BUDGET;
    int rc = evict(dev, victim);
    if (rc) return rc;
  }
  // blocks being reused may still be read by queued kernels
  dev->copy_pending = true;
  WK_HIP(hipEventRecord(dev->compute_done, dev->stream));
  WK_HIP(hipStreamWaitEvent(dev->copy_stream, dev->compute_done, 0));
  const SegMeta& m = dev->st->segs[size_t(si)];
  DevSeg& hs = dev->h_segs[si];
  const uint64_t kb = uint64_t(1) << dev->kshift, vb = uint64_t(1) << dev->vshift;
  for (uint32_t j = 0; j < s.nkb; j++) {
    const uint32_t blk = dev->free_k.back();
    dev->free_k.pop_back();
    s.kblk.push_back(blk);
    dev->h_ktab[hs.kb_base + j] = blk;
    const uint64_t b0 = uint64_t(j) * kb;
    const uint64_t n = std::min<uint64_t>(kb, m.total_buckets() - b0);
    WK_HIP(hipMemcpyAsync(dev->karena + (uint64_t(blk) << dev->kshift) * ASSOCIATIVITY,
                          dev->st->slots + (m.bucket_start + b0) * ASSOCIATIVITY,
                          n * ASSOCIATIVITY * sizeof(Slot), hipMemcpyHostToDevice, dev->copy_stream));
    dev->stats.h2d_bytes += n * ASSOCIATIVITY * sizeof(Slot);
  }
  for (uint32_t j = 0; j < s.nvb; j++) {
    const uint32_t blk = dev->free_v.back();
    dev->free_v.pop_back();
    s.vblk.push_back(blk);
    dev->h_vtab[hs.vb_base + j] = blk;
    const uint64_t e0 = uint64_t(j) * vb;
    const uint64_t n = std::min<uint64_t>(vb, m.num_edges - e0);
    WK_HIP(hipMemcpyAsync(dev->varena + (uint64_t(blk) << dev->vshift), dev->st->values + m.edge_start + e0,
                          n * sizeof(uint32_t), hipMemcpyHostToDevice, dev->copy_stream));
    dev->stats.h2d_bytes += n * sizeof(uint32_t);
  }
  if (s.nkb)
    WK_HIP(hipMemcpyAsync(dev->d_ktab + hs.kb_base, dev->h_ktab + hs.kb_base, s.nkb * sizeof(uint32_t),
                          hipMemcpyHostToDevice, dev->copy_stream));
  if (s.nvb)
    WK_HIP(hipMemcpyAsync(dev->d_vtab + hs.vb_base, dev->h_vtab + hs.vb_base, s.nvb * sizeof(uint32_t),
                          hipMemcpyHostToDevice, dev->copy_stream));
  auto first_if_consecutive = [](const std::vector<uint32_t>& blk) {
    for (size_t j = 1; j < blk.size(); j++)
      if (blk[j] != blk[0] + j) return kNoPhys;
    return blk.empty() ? 0u : blk[0];
  };
  hs.kphys = first_if_consecutive(s.kblk);
  hs.vphys = first_if_consecutive(s.vblk);
  hs.resident = 1;
  WK_HIP(hipMemcpyAsync(dev->d_segs + si, dev->h_segs + si, sizeof(DevSeg), hipMemcpyHostToDevice,
                        dev->copy_stream));
  s.resident = true;
  s.last_use = ++dev->tick;
  dev->stats.segments_loaded++;
  return WK_OK;
}

}  // namespace

int ensure_segments(wk_dev* dev, const std::vector<int>& segs) {
  for (int s : segs)
    if (s >= 0) dev->seg[size_t(s)].pins++;
  for (int s : segs) {
    if (s < 0) continue;
    int rc = load_segment(dev, s, segs);
    if (rc) {
      unpin_segments(dev, segs);
      return rc;
    }
  }
  // the compute stream waits for the copy stream only when something was
  // queued there since it last waited (a load, an eviction or a prefetch)
  if (dev->copy_pending) {
    WK_HIP(hipEventRecord(dev->copy_done, dev->copy_stream));
    WK_HIP(hipStreamWaitEvent(dev->stream, dev->copy_done, 0));
    dev->copy_pending = false;
  }
  return WK_OK;
}

void unpin_segments(wk_dev* dev, const std::vector<int>& segs) {
  for (int s : segs)
    if (s >= 0 && dev->seg[size_t(s)].pins > 0) dev->seg[size_t(s)].pins--;
}

int prefetch_segment(wk_dev* dev, int seg, const std::vector<int>& in_use) {
  if (seg < 0 || dev->seg[size_t(seg)].resident) return WK_OK;
  // a failed prefetch is not an error: the segment is loaded on demand
  (void)load_segment(dev, seg, in_use);
  return WK_OK;
}

// ------------------------------------------------------------ step arrays

int reserve_steps(wk_dev* dev, uint32_t steps, uint32_t nq) {
  if (steps <= dev->steps_cap && nq <= dev->nq_cap) return WK_OK;
  const uint32_t sc = std::max(steps, dev->steps_cap), qc = std::max(nq, dev->nq_cap);
  WK_HIP(hipStreamSynchronize(dev->stream));
  for (auto& g : dev->graphs) (void)hipGraphExecDestroy(g.second);
  dev->graphs.clear();
  if (dev->d_desc) (void)hipFree(dev->d_desc);
  if (dev->d_rs) (void)hipFree(dev->d_rs);
  if (dev->d_state) (void)hipFree(dev->d_state);
  if (dev->h_desc) (void)hipHostFree(dev->h_desc);
  if (dev->h_rs) (void)hipHostFree(dev->h_rs);
  if (dev->h_state) (void)hipHostFree(dev->h_state);
  if (dev->d_qres) (void)hipFree(dev->d_qres);
  if (dev->h_qres) (void)hipHostFree(dev->h_qres);
  if (dev->d_qagg) (void)hipFree(dev->d_qagg);
  dev->d_qagg = nullptr;
  dev->d_qres = nullptr;
  dev->h_qres = nullptr;
  dev->d_desc = nullptr;
  dev->d_rs = nullptr;
  dev->d_state = nullptr;
  dev->h_desc = nullptr;
  dev->h_rs = nullptr;
  dev->h_state = nullptr;
  dev->steps_cap = dev->nq_cap = 0;
  const size_t nd = size_t(sc + 1) * qc, nr = size_t(sc + 1) * (qc + 1), ns = size_t(sc + 1);
  WK_HIP(hipMalloc(&dev->d_desc, nd * sizeof(StepDesc)));
  WK_HIP(hipMalloc(&dev->d_rs, nr * sizeof(uint64_t)));
  WK_HIP(hipMalloc(&dev->d_state, ns * sizeof(StepState)));
  WK_HIP(hipHostMalloc(&dev->h_desc, nd * sizeof(StepDesc), hipHostMallocDefault));
  WK_HIP(hipHostMalloc(&dev->h_rs, nr * sizeof(uint64_t), hipHostMallocDefault));
  WK_HIP(hipHostMalloc(&dev->h_state, ns * sizeof(StepState), hipHostMallocDefault));
  WK_HIP(hipMalloc(&dev->d_qres, size_t(qc) * 2 * sizeof(uint64_t)));
  WK_HIP(hipHostMalloc(&dev->h_qres, size_t(qc) * 2 * sizeof(uint64_t), hipHostMallocDefault));
  WK_HIP(hipMalloc(&dev->d_qagg, (size_t(qc) + kBlock - 1) / kBlock * 3 * sizeof(uint64_t)));
  dev->steps_cap = sc;
  dev->nq_cap = qc;
  return WK_OK;
}

StepArgs make_args(wk_dev* dev, uint32_t step, uint32_t nq, const uint32_t* tin, uint32_t* tout, uint64_t tcap,
                   bool has_next) {
  StepArgs a{};
  a.karena = dev->karena;
  a.varena = dev->varena;
  a.ktab = dev->d_ktab;
  a.vtab = dev->d_vtab;
  a.segs = dev->d_segs;
  a.kshift = dev->kshift;
  a.vshift = dev->vshift;
  a.tin = tin;
  a.tout = tout;
  a.arena = dev->arena;
  a.tcap = tcap;
  a.arena_cap = dev->arena_cap;
  a.desc = dev->d_desc + size_t(step) * dev->nq_cap;
  a.next = has_next ? dev->d_desc + size_t(step + 1) * dev->nq_cap : nullptr;
  a.row_start = dev->d_rs + size_t(step) * (dev->nq_cap + 1);
  a.next_row_start = has_next ? dev->d_rs + size_t(step + 1) * (dev->nq_cap + 1) : nullptr;
  a.nq = nq;
  a.cnt = dev->cnt;
  a.eoff = dev->eoff;
  a.pfx = dev->pfx;
  a.rcap = dev->rcap;
  a.tile_sum = dev->tile_sum;
  a.tile_cap = dev->tile_cap;
  a.tile_row = dev->tile_row;
  a.otile_cap = dev->otile_cap;
  a.st = dev->d_state + step;
  a.next_st = has_next ? dev->d_state + step + 1 : nullptr;
  a.qres = nullptr;
  a.qagg = dev->d_qagg;
  return a;
}

// ----------------------------------------------------------------- timing

void timed_begin(wk_dev* dev, hipEvent_t* ev) {
  *ev = nullptr;
  if (!dev->opts.profile) return;
  if (dev->event_pool.empty()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    dev->event_pool.push_back(e);
  }
  *ev = dev->event_pool.back();
  dev->event_pool.pop_back();
  (void)hipEventRecord(*ev, dev->stream);
}

void timed_end(wk_dev* dev, hipEvent_t a, const char* name, double bytes) {
  if (!a) return;
  hipEvent_t b;
  if (dev->event_pool.empty()) {
    if (hipEventCreate(&b) != hipSuccess) return;
  } else {
    b = dev->event_pool.back();
    dev->event_pool.pop_back();
  }
  (void)hipEventRecord(b, dev->stream);
  dev->pending.push_back(PendingTiming{name, a, b, bytes});
}

double host_now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

double host_phase(wk_dev* dev, const char* name, double t0) {
  const double t = host_now();
  if (dev->opts.profile) {
    KernelStat& k = dev->kstats[name];
    k.launches++;
    k.ms += (t - t0) * 1e3;
  }
  return t;
}

int flush_timings(wk_dev* dev) {
  if (dev->pending.empty()) return WK_OK;
  WK_HIP(hipStreamSynchronize(dev->stream));
  for (auto& p : dev->pending) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    KernelStat& k = dev->kstats[p.name];
    k.launches++;
    k.ms += ms;
    k.bytes += p.bytes;
    dev->event_pool.push_back(p.a);
    dev->event_pool.push_back(p.b);
  }
  dev->pending.clear();
  return WK_OK;
}

// ------------------------------------------------------------ single step

int run_single_step(wk_dev* dev, StepDesc sd, const uint32_t* tin, uint32_t** tout, uint64_t tcap,
                    bool alloc_out, uint64_t* out_rows, int prefetch_seg) {
  *out_rows = 0;
  if (sd.in_rows == 0) {
    if (alloc_out) *tout = nullptr;
    return WK_OK;
  }
  if (sd.in_rows > dev->rcap) return WK_E_NOMEM_BUDGET;
  int rc = reserve_steps(dev, 1, 1);
  if (rc) return rc;
  std::vector<int> segs{sd.seg};
  rc = ensure_segments(dev, segs);
  if (rc) return rc;
  // the next pattern's segment streams in on the copy stream meanwhile
  // (GPUEngineCuda pipeline, src/gpu/gpu_engine_cuda.hpp:143-151)
  if (dev->opts.enable_pipeline && prefetch_seg >= 0) prefetch_segment(dev, prefetch_seg, segs);
  sd.final_step = 0;
  sd.in_off = 0;
  dev->h_desc[0] = sd;
  dev->h_rs[0] = 0;
  dev->h_rs[1] = sd.in_rows;
  dev->h_state[0] = StepState{};
  auto fail = [&](int code) {
    unpin_segments(dev, segs);
    return code;
  };
  if (hip_failed(hipMemcpyAsync(dev->d_desc, dev->h_desc, sizeof(StepDesc), hipMemcpyHostToDevice, dev->stream)) ||
      hip_failed(hipMemcpyAsync(dev->d_rs, dev->h_rs, 2 * sizeof(uint64_t), hipMemcpyHostToDevice, dev->stream)) ||
      hip_failed(hipMemcpyAsync(dev->d_state, dev->h_state, sizeof(StepState), hipMemcpyHostToDevice, dev->stream)))
    return fail(WK_E_HIP);
  StepArgs a = make_args(dev, 0, 1, tin, alloc_out ? nullptr : *tout, alloc_out ? ~uint64_t(0) : tcap, false);
  const double R = double(sd.in_rows), C = double(sd.in_cols);
  hipEvent_t ev;
  // per-kernel HIP-event timing (opts.profile) with the algorithmic bytes of
  // each launch: key column 4 B + slot 16 B + cnt/eoff 8 B per row for the
  // probe, cnt 4 B + prefix 8 B per row for the row scan
  const double ntiles = std::ceil(R / double(kTile));
  timed_begin(dev, &ev);
  launch_probe(a, dev->opts.probe_width, sd.in_rows, dev->stream);
  timed_end(dev, ev, "k_probe", 28 * R);
  timed_begin(dev, &ev);
  launch_scan_tiles(a, dev->stream);
  timed_end(dev, ev, "k_scan_tiles", 16 * ntiles);
  timed_begin(dev, &ev);
  launch_scan_rows(a, sd.in_rows, dev->stream);
  timed_end(dev, ev, "k_scan_rows", 12 * R);
  timed_begin(dev, &ev);
  launch_scan_queries(a, dev->stream);
  timed_end(dev, ev, "k_scan_queries", 96);
  if (hip_failed(hipMemcpyAsync(dev->h_state, dev->d_state, sizeof(StepState), hipMemcpyDeviceToHost, dev->stream)) ||
      hip_failed(hipStreamSynchronize(dev->stream)))
    return fail(WK_E_HIP);
  const StepState stt = dev->h_state[0];
  if (stt.status) return fail(WK_E_NOMEM_BUDGET);
  const uint64_t total = stt.total_out_rows;
  if (total > dev->otile_cap * uint64_t(kOutTile)) return fail(WK_E_NOMEM_BUDGET);
  if (alloc_out) {
    uint32_t* p = nullptr;
    if (total) {
      if (hipMalloc(&p, total * sd.out_cols * sizeof(uint32_t)) != hipSuccess) return fail(WK_E_NOMEM_BUDGET);
    }
    *tout = p;
    a.tout = p;
  }
  if (total) {
    const double E = double(total), Co = double(sd.out_cols);
    const double bytes = sd.op == OP_EXPAND ? 4 * R * C + 4 * E + 4 * E * Co : 4 * R * C + 4 * E * Co;
    if (dev->opts.rows_per_wave <= 0) {
      timed_begin(dev, &ev);
      launch_tile_rows(a, total, dev->stream);
      timed_end(dev, ev, "k_tile_rows", 8 * std::ceil(E / kOutTile));
    }
    timed_begin(dev, &ev);
    launch_gather_only(a, dev->opts.rows_per_wave, total, dev->stream);
    timed_end(dev, ev, dev->opts.rows_per_wave <= 0 ? "k_gather" : "k_gather_rows", bytes);
  }
  unpin_segments(dev, segs);
  *out_rows = total;
  return WK_OK;
}

int run_query_device(wk_dev* dev, const Plan& plan, uint64_t* rows_out, int* buf_out) {
  *rows_out = 0;
  *buf_out = -1;
  if (plan.status) return plan.status;
  uint64_t rows = 1;  // the start pattern reads one virtual row
  int cur = -1;
  for (size_t k = 0; k < plan.steps.size(); k++) {
    StepDesc sd = plan.steps[k];
    sd.in_rows = (k == 0) ? 1 : rows;
    if (sd.in_rows == 0) break;  // later patterns of an empty table stay empty
    const int out = (cur == 0) ? 1 : 0;
    uint32_t* tout = dev->tbuf[out];
    const int next_seg = (k + 1 < plan.steps.size()) ? plan.steps[k + 1].seg : -1;
    int rc = run_single_step(dev, sd, cur >= 0 ? dev->tbuf[cur] : nullptr, &tout, dev->tcap, false, &rows,
                             next_seg);
    if (rc) return rc;
    cur = out;
  }
  if (plan.steps.empty()) rows = 0;
  *rows_out = rows;
  *buf_out = cur;
  dev->stats.queries_run++;
  return WK_OK;
}

}  // namespace wk
