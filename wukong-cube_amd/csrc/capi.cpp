// extern "C" entry points of include/wk_gpu.h (store, device, tables,
// per-pattern calls, per-query and batched execution).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "device.hpp"

using namespace wk;

struct wk_store {
  Store* s;
};

namespace {

uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

void free_dev(wk_dev* d) {
  if (!d) return;
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  for (LaneState& l : d->alt)
    if (l.stream) (void)hipStreamSynchronize(l.stream);
  if (d->copy_stream) (void)hipStreamSynchronize(d->copy_stream);
  for (auto& p : d->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : d->event_pool) (void)hipEventDestroy(e);
  free_lane(*d);
  for (LaneState& l : d->alt) free_lane(l);
  void* dev_ptrs[] = {d->karena, d->varena, d->d_segs, d->d_ktab, d->d_vtab};
  for (void* p : dev_ptrs)
    if (p) (void)hipFree(p);
  void* host_ptrs[] = {d->h_segs, d->h_ktab, d->h_vtab};
  for (void* p : host_ptrs)
    if (p) (void)hipHostFree(p);
  for (uint64_t* p : d->dense)
    if (p) (void)hipFree(p);
  for (uint32_t* p : d->tbits_words) (void)hipFree(p);
  if (d->d_tbits) (void)hipFree(d->d_tbits);
  for (int k = 0; k < kStageBufs; k++) {
    if (d->stage[k]) (void)hipHostFree(d->stage[k]);
    if (d->stage_ev[k]) (void)hipEventDestroy(d->stage_ev[k]);
  }
  for (void* p : d->mapped_host) (void)hipHostFree(p);
  if (d->copy_done) (void)hipEventDestroy(d->copy_done);
  if (d->compute_done) (void)hipEventDestroy(d->compute_done);
  if (d->copy_stream) (void)hipStreamDestroy(d->copy_stream);
  delete d;
}

}  // namespace

namespace wk {

void free_lane(LaneState& l) {
  for (auto& g : l.graphs) (void)hipGraphExecDestroy(g.second);
  l.graphs.clear();
  // tbuf[1] lies inside tbuf[0]'s allocation (one pool of two tables)
  void* dev_ptrs[] = {l.tbuf[0], l.arena,   l.eoff,   l.pfx,    l.tile_sum, l.tile_row, l.d_desc,
                      l.d_rs,    l.d_state, l.d_qres, l.d_qagg, l.d_ck,     l.d_win};
  for (void* p : dev_ptrs)
    if (p) (void)hipFree(p);
  void* host_ptrs[] = {l.h_desc, l.h_rs, l.h_state, l.h_qres, l.h_ck, l.h_win};
  for (void* p : host_ptrs)
    if (p) (void)hipHostFree(p);
  if (l.stream) (void)hipStreamDestroy(l.stream);
  l = LaneState{};
}

int alloc_lane(wk_dev* d, uint64_t budget) {
  // history tables + scratch: 2 ping-pong tables + arena (4 B each per
  // element) and per-row edge offset and prefix (12 B), rows <= elements
  const uint64_t tb = budget ? budget : (uint64_t(8) << 30);
  d->tcap = std::max<uint64_t>(tb / 24, 1 << 16);
  d->arena_cap = d->tcap;
  d->rcap = d->tcap;
  d->tile_cap = d->rcap / kTile + 2;
  d->otile_cap = (d->tcap + d->arena_cap) / kOutTile + 2;
  // the two ping-pong tables are one pool, which the per-query path uses as
  // a stack of intermediate tables
  WK_HIP(hipMalloc(&d->tbuf[0], 2 * d->tcap * sizeof(uint32_t)));
  d->tbuf[1] = d->tbuf[0] + d->tcap;
  WK_HIP(hipMalloc(&d->d_win, size_t(kWinGroup) * 3 * sizeof(uint64_t)));
  WK_HIP(hipHostMalloc(&d->h_win, size_t(kWinGroup) * 3 * sizeof(uint64_t), hipHostMallocDefault));
  WK_HIP(hipMalloc(&d->arena, d->arena_cap * sizeof(uint32_t)));
  WK_HIP(hipMalloc(&d->eoff, d->rcap * sizeof(uint32_t)));
  WK_HIP(hipMalloc(&d->pfx, (d->rcap + 1) * sizeof(uint64_t)));
  WK_HIP(hipMalloc(&d->tile_sum, (d->tile_cap + 1) * sizeof(uint64_t)));
  WK_HIP(hipMalloc(&d->tile_row, (d->otile_cap + 1) * sizeof(uint64_t)));
  return reserve_steps(d, 8, 64);
}

}  // namespace wk

namespace {

// Direct key indexes for the normal segments whose key vids are dense enough
// (at least one key per 16 vids of their range: 8 B per vid, so at most
// 128 B per key, one bucket line), those with the most keys first, within a
// quarter of the HBM still free and, with a gstore budget, a quarter of it.
// Off with opts.dense_index < 0.
int alloc_dense(wk_dev* d) {
  const Store* st = d->st;
  d->dense.assign(st->segs.size(), nullptr);
  d->dense_built.assign(st->segs.size(), 0);
  d->type_slot.assign(size_t(1) << NBITS_IDX, -1);
  if (d->opts.dense_index < 0) return WK_OK;
  size_t free_b = 0, total_b = 0;
  WK_HIP(hipMemGetInfo(&free_b, &total_b));
  uint64_t room = free_b / 4;
  if (d->opts.gstore_budget) room = std::min<uint64_t>(room, d->opts.gstore_budget / 4);
  // instance bitmaps of the types first (a bit per vid of each type's range)
  const size_t nt = st->type_ids.size();
  std::vector<std::vector<uint32_t>> words(nt);
  std::vector<TypeBits> tb(nt);
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t k = 0; k < nt; k++) {
    uint64_t n = 0;
    const uint32_t* inst = st->get_index(st->type_ids[k], DIR_IN, n);
    if (!n) continue;
    tb[k].vmin = inst[0];
    tb[k].nbits = inst[n - 1] - inst[0] + 1;  // the list is ascending
    words[k].assign((uint64_t(tb[k].nbits) + 31) / 32, 0u);
    for (uint64_t i = 0; i < n; i++) {
      const uint32_t b = inst[i] - inst[0];
      words[k][b >> 5] |= 1u << (b & 31);
    }
  }
  std::vector<TypeBits> slots;
  for (size_t k = 0; k < nt; k++) {
    const uint64_t bytes = words[k].size() * sizeof(uint32_t);
    if (!bytes || bytes > room) continue;
    uint32_t* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      break;
    }
    WK_HIP(hipMemcpy(p, words[k].data(), bytes, hipMemcpyHostToDevice));
    room -= bytes;
    d->tbits_words.push_back(p);
    tb[k].words = p;
    d->type_slot[st->type_ids[k]] = int32_t(slots.size());
    slots.push_back(tb[k]);
    d->stats.dense_bytes += bytes;
  }
  if (!slots.empty()) {
    WK_HIP(hipMalloc(&d->d_tbits, slots.size() * sizeof(TypeBits)));
    WK_HIP(hipMemcpy(d->d_tbits, slots.data(), slots.size() * sizeof(TypeBits), hipMemcpyHostToDevice));
  }
  std::vector<std::pair<uint64_t, size_t>> cand;  // (keys, segment)
  for (size_t i = 0; i < st->segs.size(); i++) {
    const SegMeta& m = st->segs[i];
    if (m.id.index != 0 || m.num_keys < 64) continue;
    const uint64_t span = uint64_t(m.vmax) - m.vmin + 1;
    if (span <= 16 * m.num_keys) cand.emplace_back(m.num_keys, i);
  }
  std::sort(cand.rbegin(), cand.rend());
  for (auto& c : cand) {
    const SegMeta& m = st->segs[c.second];
    const uint64_t bytes = (uint64_t(m.vmax) - m.vmin + 1) * sizeof(uint64_t);
    if (bytes > room) continue;
    uint64_t* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      break;
    }
    room -= bytes;
    d->dense[c.second] = p;
    d->stats.dense_segments++;
    d->stats.dense_bytes += bytes;
  }
  return WK_OK;
}

int create_dev(const Store* st, const wk_dev_opts* o, wk_dev* d) {
  d->st = st;
  d->opts = *o;
  if (d->opts.probe_width == 0) d->opts.probe_width = 2;  // fastest measured on the probe at 2,560 universities
  if (d->opts.probe_width != 1 && d->opts.probe_width != 2 && d->opts.probe_width != 4 && d->opts.probe_width != 8)
    return WK_E_INVALID_ARG;
  if (d->opts.batch_lanes < 0 || d->opts.batch_lanes > kMaxLanes) return WK_E_INVALID_ARG;
  if (d->opts.rows_per_wave != 1 && d->opts.rows_per_wave != 4 && d->opts.rows_per_wave != 16 &&
      d->opts.rows_per_wave != 64)
    d->opts.rows_per_wave = 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return WK_E_NO_DEVICE;
  }
  if (o->device < 0 || o->device >= n) return WK_E_INVALID_ARG;
  d->device = o->device;
  WK_HIP(hipSetDevice(d->device));
  init_launch_grids();
  // the no-allocation entry points' fixed scratch on this device
  int src = noalloc_init(d->device);
  if (src) return src;
  WK_HIP(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
  WK_HIP(hipStreamCreateWithFlags(&d->copy_stream, hipStreamNonBlocking));
  WK_HIP(hipEventCreateWithFlags(&d->copy_done, hipEventDisableTiming));
  WK_HIP(hipEventCreateWithFlags(&d->compute_done, hipEventDisableTiming));
  d->kshift = o->key_block_shift ? o->key_block_shift : 14;
  d->vshift = o->value_block_shift ? o->value_block_shift : 19;
  if (d->kshift < 4 || d->kshift > 24 || d->vshift < 8 || d->vshift > 28) return WK_E_INVALID_ARG;

  // segment -> block tables
  const size_t ns = st->segs.size();
  d->seg.resize(ns);
  WK_HIP(hipHostMalloc(&d->h_segs, std::max<size_t>(ns, 1) * sizeof(DevSeg), hipHostMallocDefault));
  uint64_t kt = 0, vt = 0;
  for (size_t i = 0; i < ns; i++) {
    const SegMeta& m = st->segs[i];
    if (m.num_edges >= (uint64_t(1) << 32)) return WK_E_OVERFLOW;
    SegSlot& s = d->seg[i];
    s.nkb = uint32_t(ceil_div(m.total_buckets(), uint64_t(1) << d->kshift));
    s.nvb = uint32_t(ceil_div(m.num_edges, uint64_t(1) << d->vshift));
    if (m.total_buckets() >= (uint64_t(1) << 32)) return WK_E_OVERFLOW;
    DevSeg& h = d->h_segs[i];
    h = DevSeg{};
    h.num_buckets = uint32_t(m.num_buckets);
    h.nb_magic = bucket_magic(m.num_buckets);
    h.total_buckets = uint32_t(m.total_buckets());
    h.edge_start = m.edge_start;
    h.kb_base = uint32_t(kt);
    h.vb_base = uint32_t(vt);
    h.koff = h.voff = kNotResident;
    kt += s.nkb;
    vt += s.nvb;
  }
  d->ktab_n = kt;
  d->vtab_n = vt;
  WK_HIP(hipHostMalloc(&d->h_ktab, std::max<uint64_t>(kt, 1) * sizeof(uint32_t), hipHostMallocDefault));
  WK_HIP(hipHostMalloc(&d->h_vtab, std::max<uint64_t>(vt, 1) * sizeof(uint32_t), hipHostMallocDefault));
  std::memset(d->h_ktab, 0, std::max<uint64_t>(kt, 1) * sizeof(uint32_t));
  std::memset(d->h_vtab, 0, std::max<uint64_t>(vt, 1) * sizeof(uint32_t));
  WK_HIP(hipMalloc(&d->d_segs, std::max<size_t>(ns, 1) * sizeof(DevSeg)));
  WK_HIP(hipMalloc(&d->d_ktab, std::max<uint64_t>(kt, 1) * sizeof(uint32_t)));
  WK_HIP(hipMalloc(&d->d_vtab, std::max<uint64_t>(vt, 1) * sizeof(uint32_t)));
  WK_HIP(hipMemcpy(d->d_segs, d->h_segs, std::max<size_t>(ns, 1) * sizeof(DevSeg), hipMemcpyHostToDevice));
  WK_HIP(hipMemcpy(d->d_ktab, d->h_ktab, std::max<uint64_t>(kt, 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
  WK_HIP(hipMemcpy(d->d_vtab, d->h_vtab, std::max<uint64_t>(vt, 1) * sizeof(uint32_t), hipMemcpyHostToDevice));

  // gstore cache: enough blocks for the whole store, or the budget split in
  // proportion to key and value bytes (GPUCache ctor, gpu_cache.hpp:428-470)
  const uint64_t kbytes = (uint64_t(1) << d->kshift) * ASSOCIATIVITY * sizeof(Slot);
  const uint64_t vbytes = (uint64_t(1) << d->vshift) * sizeof(uint32_t);
  const uint64_t need = kt * kbytes + vt * vbytes;
  uint64_t nk = kt, nv = vt;
  if (o->gstore_budget && o->gstore_budget < need) {
    const double kshare = double(kt * kbytes) / double(std::max<uint64_t>(need, 1));
    nk = uint64_t(double(o->gstore_budget) * kshare) / kbytes;
    nv = (o->gstore_budget - nk * kbytes) / vbytes;
  }
  nk = std::max<uint64_t>(nk, 1);
  nv = std::max<uint64_t>(nv, 1);
  d->nkblk = uint32_t(nk);
  d->nvblk = uint32_t(nv);
  d->swapping = nk < kt || nv < vt;
  WK_HIP(hipMalloc(&d->karena, nk * kbytes));
  WK_HIP(hipMalloc(&d->varena, nv * vbytes));
  if (!st->pinned) {
    // swaps from a pageable store go through pinned staging buffers of a
    // block each (the larger of a key and a value block)
    d->stage_bytes = std::max(kbytes, vbytes);
    for (int k = 0; k < kStageBufs; k++) {
      WK_HIP(hipHostMalloc(&d->stage[k], d->stage_bytes, hipHostMallocDefault));
      WK_HIP(hipEventCreateWithFlags(&d->stage_ev[k], hipEventDisableTiming));
    }
  }
  for (uint64_t i = nk; i-- > 0;) d->free_k.push_back(uint32_t(i));
  for (uint64_t i = nv; i-- > 0;) d->free_v.push_back(uint32_t(i));

  int rc = alloc_lane(d, o->table_budget);
  if (rc) return rc;
  rc = alloc_dense(d);
  if (rc) return rc;
  d->stats.key_blocks_total = nk;
  d->stats.value_blocks_total = nv;
  d->stats.table_capacity_elems = d->tcap;
  return WK_OK;
}

}  // namespace

namespace {

// Plans of a batch into the current lane's plan storage (kept from call to
// call), with each query's column count and var->column map written out.
void compile_batch(wk_dev* dev, const wk_query* qs, uint32_t nq, const wk_batch_result* out) {
  const double t0 = host_now();
  std::vector<Plan>& plans = dev->plan_cache;
  if (plans.size() < nq) plans.resize(nq);
#pragma omp parallel for schedule(static) num_threads(batch_threads()) if (nq >= 1024)
  for (uint32_t i = 0; i < nq; i++) {
    compile_query(*dev->st, qs[i], plans[i], dev->type_slot.data());
    out->cols[i] = plans[i].cols;
    if (out->v2c) {
      // the query's own variables only (a full row of WK_MAX_VARS entries
      // per query was a megabyte of host writes per 4,096-query batch)
      const int32_t nv = std::min<int32_t>(std::max<int32_t>(qs[i].nvars, 0), WK_MAX_VARS);
      int32_t* m = out->v2c + size_t(i) * WK_MAX_VARS;
      for (int32_t v = 0; v < nv; v++) m[v] = plans[i].status == WK_OK ? plans[i].v2c[size_t(v)] : -1;
    }
  }
  host_phase(dev, "host:batch_compile", t0);
}

bool valid_result(const wk_batch_result* out) {
  return out && out->rows && out->cols && out->status && out->offset;
}

// Completes the current lane's submitted batch: waits for its steps and
// fills the caller's result arrays.
int finish_lane(wk_dev* dev) {
  if (!dev->inflight) return WK_OK;
  dev->inflight = false;
  const wk_batch_result& o = dev->b_out;
  const int rc = finish_batch(dev, dev->plan_cache, 0, dev->b_nq, dev->b_graph, o.rows, o.offset, o.status,
                              dev->launch);
  dev->batch_used = rc ? 0 : dev->launch.arena_end;
  dev->fetchable = rc == WK_OK;
  dev->done_rc = rc;
  return rc;
}

// Lane 0 is the device's own, lane i > 0 is alt[i - 1] (swapped in around
// the call).
template <typename F>
int on_lane(wk_dev* dev, int lane, F f) {
  if (lane == 0) return f();
  swap_lanes(dev, lane);
  const int rc = f();
  swap_lanes(dev, lane);
  return rc;
}

LaneState& lane_state(wk_dev* dev, int lane) {
  return lane == 0 ? static_cast<LaneState&>(*dev) : dev->alt[lane - 1];
}

int batch_lanes(const wk_dev* dev) {
  const int n = dev->opts.batch_lanes;
  return n <= 0 ? 2 : std::min(n, kMaxLanes);
}

}  // namespace

namespace wk {

int ensure_alt_lane(wk_dev* dev, int lane) {
  if (dev->alt[lane - 1].stream) return WK_OK;
  swap_lanes(dev, lane);
  int rc = WK_OK;
  if (hip_failed(hipStreamCreateWithFlags(&dev->stream, hipStreamNonBlocking))) rc = WK_E_HIP;
  if (!rc) rc = alloc_lane(dev, dev->opts.table_budget);
  if (rc) free_lane(*dev);
  swap_lanes(dev, lane);
  return rc;
}

int drain_batches(wk_dev* dev, int lanes_used) {
  int rc = WK_OK;
  for (int lane = 0; lane < kMaxLanes; lane++) {
    if (!lane_state(dev, lane).inflight) continue;
    const int r = on_lane(dev, lane, [&]() -> int { return finish_lane(dev); });
    if (!rc) rc = r;
  }
  // the call that drains reuses these lanes' arenas
  for (int lane = 0; lane < lanes_used && lane < kMaxLanes; lane++) lane_state(dev, lane).fetchable = false;
  return rc;
}

}  // namespace wk

extern "C" {

// ------------------------------------------------------------------ store

int wk_store_build(const uint32_t* pso, uint64_t n_pso, const uint32_t* pos, uint64_t n_pos,
                   const wk_store_opts* opts, wk_store** out) {
  if (!out) return WK_E_INVALID_ARG;
  *out = nullptr;
  BuildOpts bo;
  if (opts) {
    bo.pinned = opts->pinned != 0;
    if (opts->load_factor_pct > 0) bo.load_factor_pct = opts->load_factor_pct;
  }
  Store* s = nullptr;
  int rc;
  try {
    rc = build_store(pso, n_pso, pos, n_pos, bo, &s);
  } catch (const std::bad_alloc&) {
    return WK_E_NOMEM_HOST;
  }
  if (rc) return rc;
  *out = new (std::nothrow) wk_store{s};
  if (!*out) {
    delete s;
    return WK_E_NOMEM_HOST;
  }
  return WK_OK;
}

void wk_store_destroy(wk_store* st) {
  if (!st) return;
  delete st->s;
  delete st;
}

int wk_store_info_get(const wk_store* st, wk_store_info* info) {
  if (!st || !info) return WK_E_INVALID_ARG;
  info->num_triples_out = st->s->n_out;
  info->num_triples_in = st->s->n_in;
  info->num_keys = st->s->total_keys;
  info->num_buckets = st->s->total_buckets;
  info->num_values = st->s->total_values;
  info->num_segments = uint32_t(st->s->segs.size());
  info->pinned = st->s->pinned ? 1 : 0;
  return WK_OK;
}

int wk_store_segment(const wk_store* st, uint32_t i, wk_seg_info* seg) {
  if (!st || !seg || i >= st->s->segs.size()) return WK_E_INVALID_ARG;
  const SegMeta& m = st->s->segs[i];
  seg->index = m.id.index;
  seg->pid = m.id.pid;
  seg->dir = m.id.dir;
  seg->pad = 0;
  seg->num_keys = m.num_keys;
  seg->num_buckets = m.num_buckets;
  seg->num_ext_buckets = m.num_ext;
  seg->bucket_start = m.bucket_start;
  seg->num_edges = m.num_edges;
  seg->edge_start = m.edge_start;
  return WK_OK;
}

int wk_store_raw(const wk_store* st, const void** slots, uint64_t* num_buckets, const uint32_t** values,
                 uint64_t* num_values) {
  if (!st) return WK_E_INVALID_ARG;
  if (slots) *slots = st->s->slots;
  if (num_buckets) *num_buckets = st->s->total_buckets;
  if (values) *values = st->s->values;
  if (num_values) *num_values = st->s->total_values;
  return WK_OK;
}

int wk_store_get_triples(const wk_store* st, uint32_t vid, uint32_t pid, int32_t dir, const uint32_t** edges,
                         uint64_t* size) {
  if (!st || !edges || !size) return WK_E_INVALID_ARG;
  if (dir != WK_IN && dir != WK_OUT) return WK_E_INVALID_ARG;
  *edges = st->s->get_triples(vid, pid, dir, *size);
  return WK_OK;
}

int wk_store_get_index(const wk_store* st, uint32_t tpid, int32_t dir, const uint32_t** edges, uint64_t* size) {
  if (!st || !edges || !size) return WK_E_INVALID_ARG;
  if (dir != WK_IN && dir != WK_OUT) return WK_E_INVALID_ARG;
  *edges = st->s->get_index(tpid, dir, *size);
  return WK_OK;
}

// ----------------------------------------------------------------- device

int wk_dev_create(const wk_store* st, const wk_dev_opts* opts, wk_dev** out) {
  if (!st || !out) return WK_E_INVALID_ARG;
  *out = nullptr;
  wk_dev_opts o{};
  if (opts) o = *opts;
  wk_dev* d = new (std::nothrow) wk_dev();
  if (!d) return WK_E_NOMEM_HOST;
  int rc;
  try {
    rc = create_dev(st->s, &o, d);
  } catch (const std::bad_alloc&) {
    rc = WK_E_NOMEM_HOST;
  }
  if (rc) {
    free_dev(d);
    return rc;
  }
  *out = d;
  return WK_OK;
}

void wk_dev_destroy(wk_dev* dev) { free_dev(dev); }

int wk_dev_stats_get(wk_dev* dev, wk_dev_stats* s) {
  if (!dev || !s) return WK_E_INVALID_ARG;
  *s = dev->stats;
  s->key_blocks_free = dev->free_k.size();
  s->value_blocks_free = dev->free_v.size();
  return WK_OK;
}

int wk_dev_kernel_stats(wk_dev* dev, wk_kernel_stat* out, uint32_t cap, uint32_t* n) {
  if (!dev || !n) return WK_E_INVALID_ARG;
  int rc = flush_timings(dev);
  if (rc) return rc;
  uint32_t i = 0;
  for (auto& e : dev->kstats) {
    if (out && i < cap) {
      std::memset(&out[i], 0, sizeof(wk_kernel_stat));
      std::strncpy(out[i].name, e.first.c_str(), sizeof(out[i].name) - 1);
      out[i].launches = e.second.launches;
      out[i].total_ms = e.second.ms;
      out[i].alg_bytes = e.second.bytes;
    }
    i++;
  }
  *n = i;
  return WK_OK;
}

int wk_dev_reset_kernel_stats(wk_dev* dev) {
  if (!dev) return WK_E_INVALID_ARG;
  int rc = flush_timings(dev);
  dev->kstats.clear();
  return rc;
}

int wk_dev_set_profile(wk_dev* dev, int32_t on) {
  if (!dev) return WK_E_INVALID_ARG;
  int rc = flush_timings(dev);
  dev->opts.profile = on ? 1 : 0;
  return rc;
}

int wk_dev_set_tuning(wk_dev* dev, int32_t probe_width, int32_t rows_per_wave) {
  if (!dev) return WK_E_INVALID_ARG;
  if (int drc = drain_batches(dev, 0)) return drc;
  if (probe_width != 1 && probe_width != 2 && probe_width != 4 && probe_width != 8) return WK_E_INVALID_ARG;
  if (rows_per_wave != 0 && rows_per_wave != 1 && rows_per_wave != 4 && rows_per_wave != 16 &&
      rows_per_wave != 64)
    return WK_E_INVALID_ARG;
  if (dev->opts.probe_width != probe_width || dev->opts.rows_per_wave != rows_per_wave) {
    (void)hipStreamSynchronize(dev->stream);
    for (auto& g : dev->graphs) (void)hipGraphExecDestroy(g.second);
    dev->graphs.clear();
    for (LaneState& l : dev->alt) {
      if (l.stream) (void)hipStreamSynchronize(l.stream);
      for (auto& g : l.graphs) (void)hipGraphExecDestroy(g.second);
      l.graphs.clear();
    }
  }
  dev->opts.probe_width = probe_width;
  dev->opts.rows_per_wave = rows_per_wave;
  return WK_OK;
}

int wk_dev_sync(wk_dev* dev) {
  if (!dev) return WK_E_INVALID_ARG;
  if (int drc = drain_batches(dev, 0)) return drc;
  WK_HIP(hipStreamSynchronize(dev->copy_stream));
  WK_HIP(hipStreamSynchronize(dev->stream));
  return WK_OK;
}

int wk_dev_load_all(wk_dev* dev) {
  if (!dev) return WK_E_INVALID_ARG;
  if (int drc = drain_batches(dev, 0)) return drc;
  std::vector<int> none;
  for (size_t i = 0; i < dev->seg.size(); i++) {
    std::vector<int> one{int(i)};
    int rc = ensure_segments(dev, one);
    unpin_segments(dev, one);
    if (rc == WK_E_NOMEM_BUDGET || rc == WK_E_SEG_TOO_LARGE) break;
    if (rc) return rc;
  }
  return wk_dev_sync(dev);
}

int wk_seg_ensure(wk_dev* dev, uint32_t index, uint32_t pid, int32_t dir) {
  if (!dev || (dir != WK_IN && dir != WK_OUT)) return WK_E_INVALID_ARG;
  if (int drc = drain_batches(dev, 0)) return drc;
  const int s = dev->st->seg_index(index, pid, uint32_t(dir));
  if (s < 0) return WK_E_INVALID_ARG;
  std::vector<int> one{s};
  int rc = ensure_segments(dev, one);
  unpin_segments(dev, one);
  return rc;
}

int wk_seg_resident(wk_dev* dev, uint32_t index, uint32_t pid, int32_t dir, int32_t* resident) {
  if (!dev || !resident) return WK_E_INVALID_ARG;
  const int s = dev->st->seg_index(index, pid, uint32_t(dir));
  *resident = (s >= 0 && dev->seg[size_t(s)].resident) ? 1 : 0;
  return WK_OK;
}

// ----------------------------------------------------------------- tables

int wk_table_upload(wk_dev* dev, const uint32_t* host, uint64_t rows, uint32_t cols, wk_table** out) {
  if (!dev || !out || (rows * cols && !host)) return WK_E_INVALID_ARG;
  *out = nullptr;
  wk_table* t = new (std::nothrow) wk_table();
  if (!t) return WK_E_NOMEM_HOST;
  t->rows = rows;
  t->cols = cols;
  if (rows * cols) {
    if (hipMalloc(&t->d, rows * cols * sizeof(uint32_t)) != hipSuccess) {
      delete t;
      return WK_E_NOMEM_BUDGET;
    }
    if (hip_failed(hipMemcpy(t->d, host, rows * cols * sizeof(uint32_t), hipMemcpyHostToDevice))) {
      (void)hipFree(t->d);
      delete t;
      return WK_E_HIP;
    }
  }
  *out = t;
  return WK_OK;
}

int wk_table_shape(const wk_table* t, uint64_t* rows, uint32_t* cols) {
  if (!t) return WK_E_INVALID_ARG;
  if (rows) *rows = t->rows;
  if (cols) *cols = t->cols;
  return WK_OK;
}

int wk_table_download(wk_dev* dev, const wk_table* t, uint32_t* dst, uint64_t cap_elems) {
  if (!dev || !t) return WK_E_INVALID_ARG;
  const uint64_t n = t->rows * t->cols;
  if (n > cap_elems || (n && !dst)) return WK_E_INVALID_ARG;
  WK_HIP(hipStreamSynchronize(dev->stream));
  if (n) WK_HIP(hipMemcpy(dst, t->d, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return WK_OK;
}

int wk_table_free(wk_dev* dev, wk_table* t) {
  if (!t) return WK_OK;
  if (dev) (void)hipStreamSynchronize(dev->stream);
  if (t->d) (void)hipFree(t->d);
  delete t;
  return WK_OK;
}

int wk_table_partition(wk_dev* dev, const wk_table* in, uint32_t col, uint32_t nparts, wk_table** out,
                       uint64_t* sizes, uint64_t* heads) {
  if (!dev || !in || !out || !sizes || !heads) return WK_E_INVALID_ARG;
  if (int drc = drain_batches(dev)) return drc;
  *out = nullptr;
  if (nparts == 0 || nparts > kMaxParts || (in->rows && col >= in->cols)) return WK_E_INVALID_ARG;
  const uint64_t rows = in->rows;
  // scratch: the step prefix array (rcap + 1 words) holds the histogram,
  // the chunk totals and the heads
  if (partition_scratch_words(rows, nparts) > dev->rcap + 1) return WK_E_NOMEM_BUDGET;
  wk_table* t = new (std::nothrow) wk_table();
  if (!t) return WK_E_NOMEM_HOST;
  t->rows = rows;
  t->cols = in->cols;
  for (uint32_t p = 0; p < nparts; p++) sizes[p] = heads[p] = 0;
  if (rows * in->cols) {
    if (hipMalloc(&t->d, rows * in->cols * sizeof(uint32_t)) != hipSuccess) {
      delete t;
      return WK_E_NOMEM_BUDGET;
    }
    const uint64_t* dh =
        launch_partition(in->d, rows, in->cols, col, OwnerFn{nparts, 0, 0}, dev->pfx, t->d, dev->stream);
    std::vector<uint64_t> h(size_t(nparts) + 1);
    if (hip_failed(hipMemcpyAsync(h.data(), dh, h.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, dev->stream)) ||
        hip_failed(hipStreamSynchronize(dev->stream))) {
      (void)hipFree(t->d);
      delete t;
      return WK_E_HIP;
    }
    for (uint32_t p = 0; p < nparts; p++) {
      heads[p] = h[p];
      sizes[p] = h[p + 1] - h[p];
    }
  }
  *out = t;
  return WK_OK;
}

// --------------------------------------------------------------- patterns

// One pattern over a whole device table into a new table of its exact size:
// slices of at most rcap input rows, each probed for its output size, then
// probed again (unless it is the only one) and gathered at its offset.
static int pattern_step(wk_dev* dev, StepDesc sd, const wk_table* in, wk_table** out) {
  if (!dev || !out) return WK_E_INVALID_ARG;
  if (int drc = drain_batches(dev)) return drc;
  *out = nullptr;
  const uint64_t R = in ? in->rows : sd.in_rows;
  const uint32_t* tin = in ? in->d : nullptr;
  std::unique_ptr<wk_table> t(new (std::nothrow) wk_table());
  if (!t) return WK_E_NOMEM_HOST;
  t->cols = sd.out_cols;
  std::vector<int> segs{sd.seg};
  int rc = ensure_segments(dev, segs);
  if (rc) return rc;
  struct Unpin {
    wk_dev* d;
    std::vector<int>& s;
    ~Unpin() { unpin_segments(d, s); }
  } unpin{dev, segs};
  const uint64_t nsl = R ? (R + dev->rcap - 1) / dev->rcap : 0;
  std::vector<uint64_t> sl_rows(nsl, 0);
  uint64_t total = 0;
  for (uint64_t i = 0; i < nsl; i++) {
    const uint64_t r0 = i * dev->rcap, n = std::min(dev->rcap, R - r0);
    rc = probe_scan(dev, sd, tin ? tin + r0 * sd.in_cols : nullptr, n, &sl_rows[i]);
    if (rc) return rc;
    total += sl_rows[i];
  }
  if (total) {
    if (hipMalloc(&t->d, total * sd.out_cols * sizeof(uint32_t)) != hipSuccess) {
      (void)hipGetLastError();
      return WK_E_NOMEM_BUDGET;
    }
    uint64_t at = 0;
    for (uint64_t i = 0; i < nsl; i++) {
      if (!sl_rows[i]) continue;
      const uint64_t r0 = i * dev->rcap, n = std::min(dev->rcap, R - r0);
      const uint32_t* src = tin ? tin + r0 * sd.in_cols : nullptr;
      uint64_t T = sl_rows[i];
      if (nsl > 1) rc = probe_scan(dev, sd, src, n, &T);
      if (!rc) rc = gather_window(dev, sd, src, n, t->d + at * sd.out_cols, 0, T);
      if (rc) {
        (void)hipStreamSynchronize(dev->stream);
        (void)hipFree(t->d);
        return rc;
      }
      at += T;
    }
  }
  t->rows = total;
  *out = t.release();
  return WK_OK;
}

int wk_start_i2u(wk_dev* dev, uint32_t tpid, int32_t dir, wk_table** out) {
  if (!dev || (dir != WK_IN && dir != WK_OUT) || tpid >= (1u << NBITS_IDX)) return WK_E_INVALID_ARG;
  StepDesc sd{};
  sd.op = OP_EXPAND;
  sd.key_mode = KEY_CONST;
  sd.kstart = 0;
  sd.pid = tpid;
  sd.dir = uint32_t(dir);
  sd.seg = dev->st->seg_index(1, PREDICATE_ID, uint32_t(dir));
  sd.in_rows = 1;
  sd.out_cols = 1;
  return pattern_step(dev, sd, nullptr, out);
}

int wk_start_c2u(wk_dev* dev, uint32_t vid, uint32_t pid, int32_t dir, wk_table** out) {
  if (!dev || (dir != WK_IN && dir != WK_OUT) || pid >= (1u << NBITS_IDX)) return WK_E_INVALID_ARG;
  StepDesc sd{};
  sd.op = OP_EXPAND;
  sd.key_mode = KEY_CONST;
  sd.kstart = vid;
  sd.pid = pid;
  sd.dir = uint32_t(dir);
  const SegId s = segid_of_key(make_key(vid, pid, uint64_t(dir)));
  sd.seg = dev->st->seg_index(s.index, s.pid, s.dir);
  sd.in_rows = 1;
  sd.out_cols = 1;
  return pattern_step(dev, sd, nullptr, out);
}

static int known_step(wk_dev* dev, const wk_table* in, uint32_t start_col, uint32_t pid, int32_t dir,
                      uint32_t op, uint32_t end, wk_table** out) {
  if (!dev || !in || !out || (dir != WK_IN && dir != WK_OUT) || pid >= (1u << NBITS_IDX)) return WK_E_INVALID_ARG;
  if (start_col >= in->cols || (op == OP_FILTER_KNOWN && end >= in->cols)) return WK_E_INVALID_ARG;
  StepDesc sd{};
  sd.op = op;
  sd.pid = pid;
  sd.dir = uint32_t(dir);
  sd.sc = start_col;
  sd.in_cols = in->cols;
  sd.in_rows = in->rows;
  if (op == OP_EXPAND && pid == TYPE_ID && dir == WK_IN) {
    sd.key_mode = KEY_TYPE_INDEX;
    sd.seg = dev->st->seg_index(1, PREDICATE_ID, DIR_IN);
  } else {
    sd.key_mode = KEY_FROM_ROW;
    sd.seg = dev->st->seg_index(0, pid, uint32_t(dir));
  }
  if (op == OP_FILTER_KNOWN) sd.ec = end;
  if (op == OP_FILTER_CONST) sd.kend = end;
  if (op == OP_FILTER_CONST && pid == TYPE_ID && dir == WK_OUT && end < (1u << NBITS_IDX) &&
      dev->type_slot[end] >= 0) {
    sd.op = OP_FILTER_TYPE;  // T in the types of ?x  <=>  ?x in the instances of T
    sd.ec = uint32_t(dev->type_slot[end]);
    sd.seg = -1;
  }
  sd.out_cols = op == OP_EXPAND ? in->cols + 1 : in->cols;
  return pattern_step(dev, sd, in, out);
}

int wk_expand_k2u(wk_dev* dev, const wk_table* in, uint32_t start_col, uint32_t pid, int32_t dir, wk_table** out) {
  return known_step(dev, in, start_col, pid, dir, OP_EXPAND, 0, out);
}

int wk_expand_k2k(wk_dev* dev, const wk_table* in, uint32_t start_col, uint32_t pid, int32_t dir,
                  uint32_t end_col, wk_table** out) {
  return known_step(dev, in, start_col, pid, dir, OP_FILTER_KNOWN, end_col, out);
}

int wk_expand_k2c(wk_dev* dev, const wk_table* in, uint32_t start_col, uint32_t pid, int32_t dir, uint32_t end,
                  wk_table** out) {
  return known_step(dev, in, start_col, pid, dir, OP_FILTER_CONST, end, out);
}

int wk_expand_c2k(wk_dev* dev, const wk_table* in, uint32_t vid, uint32_t pid, int32_t dir, uint32_t end_col,
                  wk_table** out) {
  if (!dev || !in || !out || (dir != WK_IN && dir != WK_OUT) || pid >= (1u << NBITS_IDX)) return WK_E_INVALID_ARG;
  if (end_col >= in->cols) return WK_E_INVALID_ARG;
  StepDesc sd{};
  sd.op = OP_FILTER_KNOWN;
  sd.key_mode = KEY_CONST;
  sd.kstart = vid;
  sd.pid = pid;
  sd.dir = uint32_t(dir);
  sd.ec = end_col;
  sd.in_cols = in->cols;
  sd.out_cols = in->cols;
  sd.in_rows = in->rows;
  const SegId s = segid_of_key(make_key(vid, pid, uint64_t(dir)));
  sd.seg = dev->st->seg_index(s.index, s.pid, s.dir);
  return pattern_step(dev, sd, in, out);
}

// ------------------------------------------------------------------ query

int wk_query_run(wk_dev* dev, const wk_query* q, uint64_t* rows, uint32_t* cols, int32_t* v2c) {
  if (!dev || !q) return WK_E_INVALID_ARG;
  if (int drc = drain_batches(dev)) return drc;
  dev->last_ptr = nullptr;
  dev->last_rows = 0;
  dev->last_cols = 0;
  Plan plan;
  compile_query(*dev->st, *q, plan, dev->type_slot.data());
  if (rows) *rows = 0;
  if (cols) *cols = 0;
  if (plan.status) return plan.status;
  uint64_t r = 0;
  // the result goes to the result arena, in place of the last batch's
  dev->batch_used = 0;
  int rc = run_one(dev, plan, &r);
  if (rc) return rc;
  dev->last_ptr = dev->arena;
  dev->last_rows = r;
  dev->last_cols = plan.cols;
  if (rows) *rows = dev->last_rows;
  if (cols) *cols = plan.cols;
  if (v2c)
    for (int32_t i = 0; i < q->nvars; i++) v2c[i] = plan.v2c[size_t(i)];
  return WK_OK;
}

int wk_result_download(wk_dev* dev, uint32_t* dst, uint64_t cap_elems) {
  if (!dev) return WK_E_INVALID_ARG;
  const uint64_t n = dev->last_rows * dev->last_cols;
  if (n > cap_elems || (n && !dst)) return WK_E_INVALID_ARG;
  if (n == 0) return WK_OK;
  if (!dev->last_ptr) return WK_E_STATE;
  WK_HIP(hipStreamSynchronize(dev->stream));
  WK_HIP(hipMemcpy(dst, dev->last_ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return WK_OK;
}

// ------------------------------------------------------------------ batch


int wk_batch_run(wk_dev* dev, const wk_query* qs, uint32_t nq, int32_t use_graph, wk_batch_result* out) {
  if (!dev || (!qs && nq) || !valid_result(out)) return WK_E_INVALID_ARG;
  int rc = drain_batches(dev);
  if (rc) return rc;
  compile_batch(dev, qs, nq, out);
  uint64_t end = 0;
  rc = run_batch(dev, dev->plan_cache, 0, nq, 0, use_graph != 0, out->rows, out->offset, out->status, &end);
  dev->batch_used = rc ? 0 : end;
  return rc;
}

int wk_batch_submit(wk_dev* dev, const wk_query* qs, uint32_t nq, int32_t use_graph, wk_batch_result* out,
                    uint32_t* ticket) {
  if (!dev || (!qs && nq) || !valid_result(out) || !ticket) return WK_E_INVALID_ARG;
  *ticket = 0;
  // the preferred lane if it is free, else the next free one; all busy:
  // wait first
  const int nl = batch_lanes(dev);
  int lane = -1;
  for (int i = 0; i < nl && lane < 0; i++) {
    const int c = (dev->next_lane + i) % nl;
    const LaneState& l = lane_state(dev, c);
    // a lane whose last batch finished (completed by another call) but was
    // not waited for yet keeps that ticket until its wait
    if (!l.inflight && l.awaited) lane = c;
  }
  if (lane < 0) return WK_E_STATE;
  if (lane > 0) {
    const int rc = ensure_alt_lane(dev, lane);
    if (rc) return rc;
  }
  if (dev->swapping) {
    // under a gstore budget smaller than the store, the other lanes' pinned
    // segments could leave this batch no room: finish theirs first
    for (int o = 0; o < kMaxLanes; o++) {
      if (o == lane || !lane_state(dev, o).inflight) continue;
      const int rc = on_lane(dev, o, [&]() -> int { return finish_lane(dev); });
      if (rc) return rc;
    }
  }
  const uint32_t t = dev->next_ticket++;
  if (dev->next_ticket == 0) dev->next_ticket = 1;
  const int rc = on_lane(dev, lane, [&]() -> int {
    compile_batch(dev, qs, nq, out);
    dev->ticket = t;
    dev->awaited = false;
    dev->done_rc = WK_OK;
    dev->fetchable = false;
    dev->batch_used = 0;
    dev->b_out = *out;
    dev->b_nq = nq;
    dev->b_graph = use_graph != 0;
    const int r = begin_batch(dev, dev->plan_cache, 0, nq, 0, dev->b_graph, out->rows, out->offset, out->status,
                              dev->launch);
    if (r) {
      dev->ticket = 0;
      dev->awaited = true;
      return r;
    }
    // a batch that already completed (no steps, or split for lack of
    // gstore room) is finished by its wait like any other
    dev->inflight = true;
    return WK_OK;
  });
  if (rc) return rc;
  dev->next_lane = (lane + 1) % nl;
  *ticket = t;
  return WK_OK;
}

int wk_batch_wait(wk_dev* dev, uint32_t ticket) {
  if (!dev || ticket == 0) return WK_E_INVALID_ARG;
  for (int lane = 0; lane < kMaxLanes; lane++) {
    LaneState& l = lane_state(dev, lane);
    if (l.ticket != ticket) continue;
    l.awaited = true;
    if (!l.inflight) return l.done_rc;  // already finished, by a wait or by another call
    return on_lane(dev, lane, [&]() -> int { return finish_lane(dev); });
  }
  return WK_E_STATE;
}

int wk_batch_fetch(wk_dev* dev, uint32_t ticket, uint32_t* dst, uint64_t cap_elems, uint64_t* used_elems) {
  if (!dev || ticket == 0) return WK_E_INVALID_ARG;
  for (int lane = 0; lane < kMaxLanes; lane++) {
    LaneState& l = lane_state(dev, lane);
    if (l.ticket != ticket) continue;
    if (l.inflight || !l.fetchable) return WK_E_STATE;  // not waited for, or its arena reused since
    if (used_elems) *used_elems = l.batch_used;
    if (!dst) return WK_OK;
    if (cap_elems < l.batch_used) return WK_E_INVALID_ARG;
    if (l.batch_used) WK_HIP(hipMemcpy(dst, l.arena, l.batch_used * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return WK_OK;
  }
  return WK_E_STATE;
}

int wk_dev_autotune(wk_dev* dev, const wk_query* qs, uint32_t nq, int32_t use_graph, int32_t* probe_width,
                    int32_t* rows_per_wave, double* batch_ms) {
  if (!dev || (!qs && nq)) return WK_E_INVALID_ARG;
  if (int drc = drain_batches(dev)) return drc;
  std::vector<Plan>& plans = dev->plan_cache;
  if (plans.size() < nq) plans.resize(nq);
#pragma omp parallel for schedule(static) num_threads(batch_threads()) if (nq >= 1024)
  for (uint32_t i = 0; i < nq; i++) compile_query(*dev->st, qs[i], plans[i], dev->type_slot.data());
  std::vector<uint64_t> rows(nq), offs(nq);
  std::vector<int32_t> status(nq);
  static const int kWidths[] = {1, 2, 4, 8};
  static const int kRpw[] = {0, 1, 4, 16, 64};
  int best_pw = dev->opts.probe_width, best_rpw = dev->opts.rows_per_wave;
  double best = -1;
  for (int pw : kWidths) {
    for (int rpw : kRpw) {
      int rc = wk_dev_set_tuning(dev, pw, rpw);
      if (rc) return rc;
      uint64_t end = 0;
      // the first run builds the graph and warms the caches; the setting's
      // time is the median of the three runs after it
      double t[4];
      for (int rep = 0; rep < 4; rep++) {
        const auto t0 = std::chrono::steady_clock::now();
        rc = run_batch(dev, plans, 0, nq, 0, use_graph != 0, rows.data(), offs.data(), status.data(), &end);
        if (rc) return rc;
        t[rep] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      }
      std::sort(t + 1, t + 4);
      const double ms = t[2];
      if (best < 0 || ms < best) {
        best = ms;
        best_pw = pw;
        best_rpw = rpw;
      }
    }
  }
  int rc = wk_dev_set_tuning(dev, best_pw, best_rpw);
  if (rc) return rc;
  dev->batch_used = 0;
  if (probe_width) *probe_width = best_pw;
  if (rows_per_wave) *rows_per_wave = best_rpw;
  if (batch_ms) *batch_ms = best;
  return WK_OK;
}

int wk_batch_traffic(wk_dev* dev, int32_t enable, double* alg_bytes, double* edges, uint64_t* batches) {
  if (!dev) return WK_E_INVALID_ARG;
  if (alg_bytes) *alg_bytes = dev->traffic_bytes;
  if (edges) *edges = dev->traffic_edges;
  if (batches) *batches = dev->traffic_batches;
  dev->traffic_bytes = dev->traffic_edges = 0;
  dev->traffic_batches = 0;
  dev->traffic_on = enable != 0;
  return WK_OK;
}

int wk_batch_download(wk_dev* dev, uint32_t* dst, uint64_t cap_elems, uint64_t* used_elems) {
  if (!dev) return WK_E_INVALID_ARG;
  if (int drc = drain_batches(dev, 0)) return drc;
  if (used_elems) *used_elems = dev->batch_used;
  if (!dst) return WK_OK;
  if (cap_elems < dev->batch_used) return WK_E_INVALID_ARG;
  if (dev->batch_used)
    WK_HIP(hipMemcpy(dst, dev->arena, dev->batch_used * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return WK_OK;
}

// ------------------------------------------------------------------- misc

uint64_t wk_table_checksum(const uint32_t* table, uint64_t rows, uint32_t cols) {
  uint64_t s = 0;
  for (uint64_t r = 0; r < rows; r++) s += row_hash(table + r * cols, cols);
  return s;
}

const char* wk_strerror(int code) {
  switch (code) {
    case WK_OK: return "ok";
    case WK_E_UNKNOWN: return "something wrong happened";
    case WK_E_SYNTAX: return "query syntax error";
    case WK_E_UNKNOWN_PATTERN: return "unsupported triple pattern";
    case WK_E_ATTR_DISABLE: return "attribute support is disabled";
    case WK_E_NO_REQUIRED_VAR: return "no required variables";
    case WK_E_UNSUPPORT_UNION: return "unsupported UNION on attribute results";
    case WK_E_OBJ_ERROR: return "object should not be an index";
    case WK_E_VERTEX_INVALID: return "subject or object is not valid";
    case WK_E_UNKNOWN_SUB: return "triple pattern should not start from unknown subject";
    case WK_E_SETTING: return "setting error";
    case WK_E_FIRST_PATTERN: return "const_X_X or index_X_X must be the first pattern";
    case WK_E_UNKNOWN_FILTER: return "unsupported filter type";
    case WK_E_FILE_NOT_FOUND: return "file not found";
    case WK_E_INVALID_ARG: return "invalid argument";
    case WK_E_INVALID_TRIPLE: return "triple violates the id conventions";
    case WK_E_NOMEM_HOST: return "host allocation failed";
    case WK_E_NOMEM_BUDGET: return "device budget exceeded";
    case WK_E_HIP: return "HIP runtime error";
    case WK_E_NO_DEVICE: return "no usable GPU";
    case WK_E_OVERFLOW: return "count overflows its field";
    case WK_E_SEG_TOO_LARGE: return "segment exceeds the gstore budget";
    case WK_E_STATE: return "call out of order";
    case WK_E_TOO_MANY_VARS: return "too many variables";
    case WK_E_WIRE: return "malformed or unsupported query archive";
    case WK_E_SCRATCH_LIMIT: return "input exceeds the fixed scratch area";
    default: return "unknown status";
  }
}

int wk_last_hip_error(void) { return g_last_hip_error; }

int wk_device_count(int* n) {
  if (!n) return WK_E_INVALID_ARG;
  *n = 0;
  if (hipGetDeviceCount(n) != hipSuccess) {
    (void)hipGetLastError();
    *n = 0;
  }
  return WK_OK;
}

}  // extern "C"
