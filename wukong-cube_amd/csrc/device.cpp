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

int compile_query(const Store& st, const wk_query& q, Plan& plan, const int32_t* type_slot) {
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
        if (type_slot && sd.pid == TYPE_ID && sd.dir == DIR_OUT && sd.kend < (1u << NBITS_IDX) &&
            type_slot[sd.kend] >= 0) {
          // T in the types of ?x  <=>  ?x in the instances of T
          sd.op = OP_FILTER_TYPE;
          sd.ec = uint32_t(type_slot[sd.kend]);
          sd.seg = -1;  // no segment is read
        }
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
  dev->h_segs[v].koff = dev->h_segs[v].voff = kNotResident;
  all_lanes_copy_pending(dev);  // every lane waits for it
  WK_HIP(hipMemcpyAsync(dev->d_segs + v, dev->h_segs + v, sizeof(DevSeg), hipMemcpyHostToDevice,
                        dev->copy_stream));
  dev->stats.segments_evicted++;
  return WK_OK;
}

// A host-to-device copy of store bytes on the copy stream.  From a pinned
// store it is one DMA; from a pageable one, the bytes go through the ring of
// pinned staging buffers: host threads copy a piece into the next buffer
// (once its previous transfer is done) and the DMA goes from there, so
// filling one buffer overlaps the transfer of the one before.
int copy_in(wk_dev* dev, void* dst, const void* src, uint64_t n) {
  if (dev->st->pinned || !dev->stage[0]) {
    WK_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, dev->copy_stream));
    return WK_OK;
  }
  uint8_t* d = static_cast<uint8_t*>(dst);
  const uint8_t* s = static_cast<const uint8_t*>(src);
  while (n) {
    const int k = dev->stage_next;
    dev->stage_next = (k + 1) % kStageBufs;
    if (dev->stage_used[k]) WK_HIP(hipEventSynchronize(dev->stage_ev[k]));
    const uint64_t m = std::min<uint64_t>(n, dev->stage_bytes);
    uint8_t* buf = dev->stage[k];
    constexpr uint64_t kPiece = uint64_t(1) << 20;
#pragma omp parallel for schedule(static) if (m >= 4 * kPiece)
    for (uint64_t o = 0; o < m; o += kPiece) std::memcpy(buf + o, s + o, std::min(kPiece, m - o));
    WK_HIP(hipMemcpyAsync(d, buf, m, hipMemcpyHostToDevice, dev->copy_stream));
    WK_HIP(hipEventRecord(dev->stage_ev[k], dev->copy_stream));
    dev->stage_used[k] = 1;
    d += m;
    s += m;
    n -= m;
  }
  return WK_OK;
}

// Device address of byte `off` of a pinned host allocation at `base`.
int mapped_addr(const void* base, uint64_t off, intptr_t* out) {
  void* d = nullptr;
  WK_HIP(hipHostGetDevicePointer(&d, const_cast<void*>(base), 0));
  *out = reinterpret_cast<intptr_t>(d) + intptr_t(off);
  return WK_OK;
}

// A segment larger than the whole gstore budget is read in place from pinned
// host memory: the store's own when it is pinned, else a pinned copy made
// once (kept until the device is destroyed).  Its buckets and edges then
// cross the host link on every read; its direct key index, when it has one,
// is built once into HBM, so dense-keyed segments read only edges there.
int map_segment(wk_dev* dev, int si) {
  SegSlot& s = dev->seg[size_t(si)];
  const SegMeta& m = dev->st->segs[size_t(si)];
  const uint64_t kbytes = m.total_buckets() * ASSOCIATIVITY * sizeof(Slot);
  const uint64_t vbytes = m.num_edges * sizeof(uint32_t);
  const void* kbase = dev->st->slots;
  const void* vbase = dev->st->values;
  uint64_t koff = m.bucket_start * ASSOCIATIVITY * sizeof(Slot), voff = m.edge_start * sizeof(uint32_t);
  if (!dev->st->pinned) {
    void* kp = nullptr;
    void* vp = nullptr;
    if (hipHostMalloc(&kp, std::max<uint64_t>(kbytes, 1), hipHostMallocDefault) != hipSuccess)
      return WK_E_SEG_TOO_LARGE;
    dev->mapped_host.push_back(kp);
    if (hipHostMalloc(&vp, std::max<uint64_t>(vbytes, 1), hipHostMallocDefault) != hipSuccess)
      return WK_E_SEG_TOO_LARGE;
    dev->mapped_host.push_back(vp);
    const uint8_t* ks = reinterpret_cast<const uint8_t*>(dev->st->slots) + koff;
    const uint8_t* vs = reinterpret_cast<const uint8_t*>(dev->st->values) + voff;
    constexpr uint64_t kPiece = uint64_t(1) << 20;
#pragma omp parallel for schedule(static)
    for (uint64_t o = 0; o < kbytes; o += kPiece)
      std::memcpy(static_cast<uint8_t*>(kp) + o, ks + o, std::min(kPiece, kbytes - o));
#pragma omp parallel for schedule(static)
    for (uint64_t o = 0; o < vbytes; o += kPiece)
      std::memcpy(static_cast<uint8_t*>(vp) + o, vs + o, std::min(kPiece, vbytes - o));
    kbase = kp;
    vbase = vp;
    koff = voff = 0;
  }
  intptr_t kd = 0, vd = 0;
  int rc = mapped_addr(kbase, koff, &kd);
  if (!rc) rc = mapped_addr(vbase, voff, &vd);
  if (rc) return rc;
  // offsets from the arenas, in buckets and edges (both sides are aligned to them)
  const intptr_t dk = kd - reinterpret_cast<intptr_t>(dev->karena);
  const intptr_t dv = vd - reinterpret_cast<intptr_t>(dev->varena);
  constexpr intptr_t kBucketBytes = ASSOCIATIVITY * sizeof(Slot);
  if (dk % kBucketBytes != 0 || dv % intptr_t(sizeof(uint32_t)) != 0) return WK_E_SEG_TOO_LARGE;
  DevSeg& hs = dev->h_segs[si];
  hs.koff = dk / kBucketBytes;
  hs.voff = dv / intptr_t(sizeof(uint32_t));
  all_lanes_copy_pending(dev);  // every lane waits for it
  if (dev->dense[size_t(si)] && !dev->dense_built[size_t(si)]) {
    hs.dvmin = m.vmin;
    hs.dvn = m.vmax - m.vmin + 1;
    WK_HIP(hipMemsetAsync(dev->dense[size_t(si)], 0, uint64_t(hs.dvn) * sizeof(uint64_t), dev->copy_stream));
    launch_dense_build(dev->karena, dev->d_ktab, dev->kshift, hs, dev->dense[size_t(si)], dev->copy_stream);
    hs.dense = dev->dense[size_t(si)];
    dev->dense_built[size_t(si)] = 1;
  }
  WK_HIP(hipMemcpyAsync(dev->d_segs + si, dev->h_segs + si, sizeof(DevSeg), hipMemcpyHostToDevice,
                        dev->copy_stream));
  s.resident = s.mapped = true;
  s.last_use = ++dev->tick;
  dev->stats.mapped_segments++;
  dev->stats.mapped_bytes += kbytes + vbytes;
  return WK_OK;
}

int load_segment(wk_dev* dev, int si, const std::vector<int>& keep) {
  SegSlot& s = dev->seg[size_t(si)];
  if (s.resident) {
    s.last_use = ++dev->tick;
    return WK_OK;
  }
  if (s.nkb > dev->nkblk || s.nvb > dev->nvblk) return map_segment(dev, si);
  while (dev->free_k.size() < s.nkb || dev->free_v.size() < s.nvb) {
    int victim = -1;
    uint64_t best = ~uint64_t(0);
    for (size_t i = 0; i < dev->seg.size(); i++) {
      const SegSlot& c = dev->seg[i];
      if (!c.resident || c.mapped || c.pins > 0 || int(i) == si) continue;
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
  all_lanes_copy_pending(dev);  // every lane waits for it
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
    int rc = copy_in(dev, dev->karena + (uint64_t(blk) << dev->kshift) * ASSOCIATIVITY,
                     dev->st->slots + (m.bucket_start + b0) * ASSOCIATIVITY, n * ASSOCIATIVITY * sizeof(Slot));
    if (rc) return rc;
    dev->stats.h2d_bytes += n * ASSOCIATIVITY * sizeof(Slot);
  }
  for (uint32_t j = 0; j < s.nvb; j++) {
    const uint32_t blk = dev->free_v.back();
    dev->free_v.pop_back();
    s.vblk.push_back(blk);
    dev->h_vtab[hs.vb_base + j] = blk;
    const uint64_t e0 = uint64_t(j) * vb;
    const uint64_t n = std::min<uint64_t>(vb, m.num_edges - e0);
    int rc = copy_in(dev, dev->varena + (uint64_t(blk) << dev->vshift), dev->st->values + m.edge_start + e0,
                     n * sizeof(uint32_t));
    if (rc) return rc;
    dev->stats.h2d_bytes += n * sizeof(uint32_t);
  }
  if (s.nkb)
    WK_HIP(hipMemcpyAsync(dev->d_ktab + hs.kb_base, dev->h_ktab + hs.kb_base, s.nkb * sizeof(uint32_t),
                          hipMemcpyHostToDevice, dev->copy_stream));
  if (s.nvb)
    WK_HIP(hipMemcpyAsync(dev->d_vtab + hs.vb_base, dev->h_vtab + hs.vb_base, s.nvb * sizeof(uint32_t),
                          hipMemcpyHostToDevice, dev->copy_stream));
  // consecutive blocks: addressed from the first one's offset
  auto offset_if_consecutive = [](const std::vector<uint32_t>& blk, uint32_t shift) {
    for (size_t j = 1; j < blk.size(); j++)
      if (blk[j] != blk[0] + j) return kBlockTable;
    return blk.empty() ? int64_t(0) : int64_t(blk[0]) << shift;
  };
  hs.koff = offset_if_consecutive(s.kblk, dev->kshift);
  hs.voff = offset_if_consecutive(s.vblk, dev->vshift);
  if (dev->dense[si] && !dev->dense_built[si]) {
    // the direct key index from the buckets just copied (copy stream, so the
    // compute stream sees it once it waits for the copies)
    hs.dvmin = m.vmin;
    hs.dvn = m.vmax - m.vmin + 1;
    WK_HIP(hipMemsetAsync(dev->dense[si], 0, uint64_t(hs.dvn) * sizeof(uint64_t), dev->copy_stream));
    launch_dense_build(dev->karena, dev->d_ktab, dev->kshift, hs, dev->dense[si], dev->copy_stream);
    hs.dense = dev->dense[si];
    dev->dense_built[si] = 1;
  }
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
  if (dev->d_ck) (void)hipFree(dev->d_ck);
  if (dev->h_ck) (void)hipHostFree(dev->h_ck);
  dev->d_ck = nullptr;
  dev->h_ck = nullptr;
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
  WK_HIP(hipMalloc(&dev->d_ck, size_t(qc) * sizeof(uint64_t)));
  WK_HIP(hipHostMalloc(&dev->h_ck, size_t(qc) * sizeof(uint64_t), hipHostMallocDefault));
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
  a.tbits = dev->d_tbits;
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
  a.eoff = dev->eoff;
  a.lpfx = dev->pfx;
  a.rcap = dev->rcap;
  a.tile_sum = dev->tile_sum;
  a.tile_cap = dev->tile_cap;
  a.tile_row = dev->tile_row;
  a.otile_cap = dev->otile_cap;
  a.st = dev->d_state + step;
  a.next_st = has_next ? dev->d_state + step + 1 : nullptr;
  a.qres = nullptr;
  a.qagg = dev->d_qagg;
  a.win_lo = 0;
  a.win_hi = ~uint64_t(0);
  return a;
}

// ----------------------------------------------------------------- timing

namespace {

hipEvent_t take_event(wk_dev* dev) {
  if (!dev->event_pool.empty()) {
    hipEvent_t e = dev->event_pool.back();
    dev->event_pool.pop_back();
    return e;
  }
  // timing only: no system-scope fence when the event completes, which
  // would add an L2 write-back to the measured kernel
  hipEvent_t e = nullptr;
  return hipEventCreateWithFlags(&e, hipEventDisableSystemFence) == hipSuccess ? e : nullptr;
}

}  // namespace

// The launches between timed_begin and timed_end take their start and end
// times from their own dispatch packets (arm_launch_timer).
void timed_begin(wk_dev* dev, hipEvent_t* ev) {
  *ev = nullptr;
  if (!dev->opts.profile) return;
  hipEvent_t a = take_event(dev), b = a ? take_event(dev) : nullptr;
  if (!b) {
    if (a) dev->event_pool.push_back(a);
    return;
  }
  *ev = a;
  dev->timing_stop = b;
  arm_launch_timer(a, b);
}

void timed_end(wk_dev* dev, hipEvent_t a, const char* name, double bytes) {
  if (!a) return;
  hipEvent_t b = dev->timing_stop;
  dev->timing_stop = nullptr;
  if (disarm_launch_timer()) {
    dev->pending.push_back(PendingTiming{name, a, b, bytes});
  } else {
    // nothing was launched
    dev->event_pool.push_back(a);
    dev->event_pool.push_back(b);
  }
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
  for (auto& p : dev->pending) {
    WK_HIP(hipEventSynchronize(p.b));  // on whichever lane's stream it was taken
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

int probe_scan(wk_dev* dev, const StepDesc& sd0, const uint32_t* tin, uint64_t R, uint64_t* total) {
  *total = 0;
  if (R == 0) return WK_OK;
  if (R > dev->rcap) return WK_E_NOMEM_BUDGET;
  int rc = reserve_steps(dev, 1, 1);
  if (rc) return rc;
  StepDesc sd = sd0;
  sd.final_step = 0;
  sd.in_off = 0;
  sd.in_rows = R;
  dev->h_desc[0] = sd;
  dev->h_rs[0] = 0;
  dev->h_rs[1] = R;
  dev->h_state[0] = StepState{};
  if (hip_failed(hipMemcpyAsync(dev->d_desc, dev->h_desc, sizeof(StepDesc), hipMemcpyHostToDevice, dev->stream)) ||
      hip_failed(hipMemcpyAsync(dev->d_rs, dev->h_rs, 2 * sizeof(uint64_t), hipMemcpyHostToDevice, dev->stream)) ||
      hip_failed(hipMemcpyAsync(dev->d_state, dev->h_state, sizeof(StepState), hipMemcpyHostToDevice, dev->stream)))
    return WK_E_HIP;
  // no output buffer yet: the host sizes it (or its windows) from the total
  StepArgs a = make_args(dev, 0, 1, tin, nullptr, ~uint64_t(0), false);
  const double Rd = double(R);
  hipEvent_t ev;
  // per-kernel HIP-event timing (opts.profile) with the algorithmic bytes of
  // each launch: key column 4 B + slot 16 B + prefix 8 B and edge offset 4 B
  // per row for the probe
  const double ntiles = std::ceil(Rd / double(kTile));
  // a key read from a direct index is one 8-B word, so lanes sharing a key
  // would only repeat it: one lane per key
  const bool direct = sd.seg >= 0 && dev->h_segs[sd.seg].dense && sd.key_mode != KEY_TYPE_INDEX;
  timed_begin(dev, &ev);
  launch_probe(a, direct ? 1 : dev->opts.probe_width, R, dev->stream);
  // direct key index: 8-B ptr word instead of the 16-B slot; a type bitmap:
  // one 4-B word
  timed_end(dev, ev, "k_probe", (sd.op == OP_FILTER_TYPE ? 20 : direct ? 24 : 32) * Rd);
  timed_begin(dev, &ev);
  launch_scan_small(a, dev->stream);  // tile scan and query scan of the one query
  timed_end(dev, ev, "k_scan_small", 16 * ntiles + 96);
  if (hip_failed(hipMemcpyAsync(dev->h_state, dev->d_state, sizeof(StepState), hipMemcpyDeviceToHost, dev->stream)) ||
      hip_failed(hipStreamSynchronize(dev->stream)))
    return WK_E_HIP;
  if (dev->h_state[0].status) return WK_E_NOMEM_BUDGET;
  *total = dev->h_state[0].total_out_rows;
  return WK_OK;
}

int gather_window(wk_dev* dev, const StepDesc& sd, const uint32_t* tin, uint64_t R, uint32_t* dst, uint64_t j0,
                  uint64_t j1) {
  if (j1 <= j0) return WK_OK;
  StepArgs a = make_args(dev, 0, 1, tin, dst, ~uint64_t(0), false);
  a.win_lo = j0;
  a.win_hi = j1;
  const double Rd = double(R), C = double(sd.in_cols), Co = double(sd.out_cols);
  const uint64_t total = dev->h_state[0].total_out_rows;
  const double E = double(j1 - j0);
  // input rows read in proportion to the window's share of the step
  const double Rw = total ? Rd * E / double(total) : 0;
  const double bytes = sd.op == OP_EXPAND ? 4 * Rw * C + 4 * E + 4 * E * Co : 4 * Rw * C + 4 * E * Co;
  hipEvent_t ev;
  if (dev->opts.rows_per_wave <= 0) {
    timed_begin(dev, &ev);
    launch_tile_rows(a, j1 - j0, dev->stream);
    timed_end(dev, ev, "k_tile_rows", 8 * std::ceil(E / kOutTile));
  }
  timed_begin(dev, &ev);
  launch_gather_only(a, dev->opts.rows_per_wave, j1 - j0, dev->stream);
  timed_end(dev, ev, dev->opts.rows_per_wave <= 0 ? "k_gather" : "k_gather_rows", bytes);
  return WK_OK;
}

namespace {

// Depth-first execution of one query with one host sync per pattern.
// Intermediate tables live in the pool (the two table buffers, one
// allocation of 2 * tcap elements); final rows go to a separate sink.  A
// step whose output does not fit the room left in the pool runs in output
// windows (chunks): each chunk is gathered, then carried through the
// remaining steps, before the next chunk is re-probed from its input rows.
// [0, pinned) of the pool holds the inputs of chunked steps above the
// current one, which must outlive it.
struct QueryExec {
  wk_dev* dev;
  const Plan& plan;
  uint32_t* pool;
  uint64_t pool_cap;
  uint32_t* sink;
  uint64_t sink_cap;
  uint64_t sink_used = 0;  // elements
  uint64_t rows = 0;

  QueryExec(wk_dev* d, const Plan& p, uint32_t* dst, uint64_t cap)
      : dev(d), plan(p), pool(d->tbuf[0]), pool_cap(2 * d->tcap), sink(dst), sink_cap(cap) {}

  // Steps k.. over R rows at `in`, which occupy pool elements [in_lo, in_hi)
  // (empty for the first step's virtual row).
  int run(uint32_t k, const uint32_t* in, uint64_t R, uint64_t in_lo, uint64_t in_hi, uint64_t pinned) {
    if (R == 0) return WK_OK;
    const StepDesc& sd = plan.steps[k];
    const uint32_t ic = sd.in_cols;
    if (R > dev->rcap) {
      // slices of at most rcap rows, each through the remaining steps; the
      // whole input stays live meanwhile
      for (uint64_t r0 = 0; r0 < R; r0 += dev->rcap) {
        const int rc = run(k, in + r0 * ic, std::min(dev->rcap, R - r0), in_lo, in_hi, std::max(pinned, in_hi));
        if (rc) return rc;
      }
      return WK_OK;
    }
    const bool fin = k + 1 == plan.steps.size();
    const uint32_t oc = sd.out_cols;
    std::vector<int> segs{sd.seg};
    int rc = ensure_segments(dev, segs);
    if (rc) return rc;
    // the next pattern's segment streams in on the copy stream meanwhile
    // (GPUEngineCuda pipeline, src/gpu/gpu_engine_cuda.hpp:143-151)
    if (dev->opts.enable_pipeline && !fin) prefetch_segment(dev, plan.steps[k + 1].seg, segs);
    uint64_t T = 0;
    rc = probe_scan(dev, sd, in, R, &T);
    if (rc || T == 0) {
      unpin_segments(dev, segs);
      return rc;
    }
    if (fin) {
      if (T * oc > sink_cap - sink_used) {
        unpin_segments(dev, segs);
        return WK_E_NOMEM_BUDGET;  // the result itself does not fit its buffer
      }
      rc = gather_window(dev, sd, in, R, sink + sink_used, 0, T);
      unpin_segments(dev, segs);
      sink_used += T * oc;
      rows += T;
      return rc;
    }
    const uint64_t need = T * oc;
    const uint64_t live = std::max(pinned, in_hi);
    if (in_lo >= pinned + need && in_hi > in_lo) {
      // into the dead space below the input (ping-pong)
      rc = gather_window(dev, sd, in, R, pool + pinned, 0, T);
      unpin_segments(dev, segs);
      if (rc) return rc;
      return run(k + 1, pool + pinned, T, pinned, pinned + need, pinned);
    }
    if (need <= (pool_cap - live) / 2) {
      // after the input, leaving at least as much room for the steps below
      rc = gather_window(dev, sd, in, R, pool + live, 0, T);
      unpin_segments(dev, segs);
      if (rc) return rc;
      return run(k + 1, pool + live, T, live, live + need, pinned);
    }
    // chunks of C output rows after the (now pinned) input
    const uint64_t C = ((pool_cap - live) / 2) / oc;
    if (C == 0) {
      unpin_segments(dev, segs);
      return WK_E_NOMEM_BUDGET;
    }
    if (dev->opts.profile) dev->kstats["query:chunked_steps"].launches++;
    uint32_t* const cbuf = pool + live;
    const uint64_t nch = (T + C - 1) / C;
    bool fresh = true;  // the scratch holds the probe of all R rows
    for (uint64_t g0 = 0; g0 < nch; g0 += kWinGroup) {
      const uint32_t n = uint32_t(std::min<uint64_t>(kWinGroup, nch - g0));
      if (!fresh) {
        // deeper steps reused the scratch: probe all R rows again
        rc = ensure_segments(dev, segs);
        if (rc) return rc;
        rc = probe_scan(dev, sd, in, R, &T);
        if (rc) {
          unpin_segments(dev, segs);
          return rc;
        }
        fresh = true;
      }
      // each chunk's first and last input row, and the prefix at the first
      launch_window_rows(dev->tile_sum, dev->pfx, R, g0 * C, C, T, n, dev->d_win, dev->stream);
      if (hip_failed(hipMemcpyAsync(dev->h_win, dev->d_win, size_t(n) * 3 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                    dev->stream)) ||
          hip_failed(hipStreamSynchronize(dev->stream))) {
        unpin_segments(dev, segs);
        return WK_E_HIP;
      }
      std::vector<uint64_t> win(dev->h_win, dev->h_win + size_t(n) * 3);
      for (uint32_t i = 0; i < n; i++) {
        const uint64_t j0 = (g0 + i) * C, j1 = std::min(T, j0 + C);
        if (fresh) {
          rc = gather_window(dev, sd, in, R, cbuf, j0, j1);
          fresh = false;
        } else {
          // re-probe the chunk's input rows [ra, rb]: their output rows are
          // the chunk's, shifted by the prefix before row ra
          const uint64_t ra = win[3 * i], rb = win[3 * i + 1], pa = win[3 * i + 2];
          rc = ensure_segments(dev, segs);
          if (rc) return rc;
          uint64_t Tc = 0;
          rc = probe_scan(dev, sd, in + ra * ic, rb - ra + 1, &Tc);
          if (!rc) rc = gather_window(dev, sd, in + ra * ic, rb - ra + 1, cbuf, j0 - pa, j1 - pa);
        }
        unpin_segments(dev, segs);
        if (rc) return rc;
        rc = run(k + 1, cbuf, j1 - j0, live, live + (j1 - j0) * oc, live);
        if (rc) return rc;
      }
    }
    return WK_OK;
  }
};

}  // namespace

int run_query_to(wk_dev* dev, const Plan& plan, uint32_t* dst, uint64_t cap, uint64_t* rows_out) {
  *rows_out = 0;
  if (plan.status) return plan.status;
  if (plan.steps.empty()) return WK_OK;
  QueryExec ex(dev, plan, dst, cap);
  // the start pattern reads one virtual row
  const int rc = ex.run(0, nullptr, 1, 0, 0, 0);
  if (rc) return rc;
  *rows_out = ex.rows;
  dev->stats.queries_run++;
  return WK_OK;
}

}  // namespace wk
