// CDNA4 (gfx950) kernels of the Wukong GPU query path.
//
// Replaces the reference's per-pattern CUDA pipeline (src/gpu/gpu_hash.cu):
// key generation, slot probe and edge lookup (:96-326) are fused into one
// probe pass; the thrust inclusive_scan (:589-598) becomes a project-written
// three-phase wave64/LDS scan with 64-bit offsets; the thread-per-row gathers
// (:525-587,764-836) become an output-balanced gather (each output row finds
// its source row by a bounded binary search over the 64-bit prefix) or a
// row-centric gather with 1..64 lanes per row.  Membership tests of k2k/k2c
// (:328-447) use binary search: edge lists are sorted and unique
// (loader/base_loader.hpp:481-498), so the result equals the linear scan's.
// All kernels handle many queries at once through per-query descriptors.
#include <hip/hip_ext.h>

#include <mutex>

#include "kernels.hpp"

namespace wk {

namespace {

// Armed by the library's profiling mode around a group of launches (see
// arm_launch_timer).  Launches then take their times from their own dispatch
// packets: host launch latency, which a pair of stream events around a launch
// to an idle GPU would count, is left out.
struct LaunchTimer {
  hipEvent_t start = nullptr, stop = nullptr;
  bool used = false;
};
thread_local LaunchTimer t_timer;

}  // namespace

void arm_launch_timer(hipEvent_t start, hipEvent_t stop) { t_timer = LaunchTimer{start, stop, false}; }

bool disarm_launch_timer() {
  const bool used = t_timer.used;
  t_timer = LaunchTimer{};
  return used;
}

// Every kernel launch of the library: the first launch of an armed group
// records its start into `start`, each records its end into `stop` (the last
// one's stays).
#define WK_LAUNCH(K, G, B, S, ST, ...)                                                        \
  do {                                                                                        \
    if (t_timer.stop) {                                                                       \
      hipExtLaunchKernelGGL(K, G, B, S, ST, t_timer.used ? nullptr : t_timer.start, t_timer.stop, \
                            0u, __VA_ARGS__);                                                 \
      t_timer.used = true;                                                                    \
    } else {                                                                                  \
      hipLaunchKernelGGL(K, G, B, S, ST, __VA_ARGS__);                                        \
    }                                                                                         \
  } while (0)
namespace {

constexpr int kWave = 64;
constexpr int kPersistentGrid = 2048;

// A segment whose buckets (edges) are contiguous, in consecutive arena blocks
// or in pinned host memory, is addressed directly from its offset; otherwise
// through its block table, one more dependent load.  Both choose an index
// relative to the arena, not a pointer: the two ways meet on an integer, so
// no pointer phi reaches the load.  A host-resident segment's offset is
// negative or beyond the arena; the index arithmetic is modulo 2^64.
__device__ __forceinline__ const Slot* bucket_addr(const StepArgs& a, const DevSeg& s, uint64_t b) {
  uint64_t i;
  if (s.koff > kBlockTable) {
    i = uint64_t(s.koff) + b;
  } else {
    const uint64_t blk = a.ktab[s.kb_base + (b >> a.kshift)];
    i = (blk << a.kshift) | (b & ((uint64_t(1) << a.kshift) - 1));
  }
  return a.karena + i * ASSOCIATIVITY;
}

__device__ __forceinline__ uint32_t edge_at(const StepArgs& a, const DevSeg& s, uint64_t e) {
  uint64_t i;
  if (s.voff > kBlockTable) {
    i = uint64_t(s.voff) + e;
  } else {
    const uint64_t blk = a.vtab[s.vb_base + (e >> a.vshift)];
    i = (blk << a.vshift) | (e & ((uint64_t(1) << a.vshift) - 1));
  }
  return a.varena[i];
}

// Output rows of the step before input row g (g <= R), once k_scan_tiles has
// run: its tile's prefix plus its prefix within the tile.
__device__ __forceinline__ uint64_t pfx_at(const StepArgs& a, uint64_t g) {
  return a.tile_sum[g >> kTileShift] + a.lpfx[g];
}


// Largest q in [ql, qh) with rs[q] <= g, given rs[ql] <= g.
__device__ __forceinline__ uint32_t query_in(const uint64_t* rs, uint32_t ql, uint32_t qh, uint64_t g) {
  while (qh - ql > 1) {
    const uint32_t mid = (ql + qh) >> 1;
    if (rs[mid] <= g) ql = mid;
    else qh = mid;
  }
  return ql;
}

// Largest q in [0, nq) with rs[q] <= g (queries with no rows are skipped).
__device__ __forceinline__ uint32_t find_query(const uint64_t* rs, uint32_t nq, uint64_t g) {
  uint32_t lo = 0, hi = nq;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (rs[mid] <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

// find_query for two values g0 <= g1 at once, by all NT threads of a block:
// each round reads NT evenly spaced entries of the current interval together
// (one memory latency) and keeps the stretch between the two samples around
// g, so up to NT queries take one round instead of log2(nq) dependent loads.
// Every thread of the block must call it (it synchronizes the block).
template <int NT>
__device__ __forceinline__ void block_find_queries(const uint64_t* rs, uint32_t nq, uint64_t g0, uint64_t g1,
                                                   uint32_t& q0, uint32_t& q1) {
  __shared__ uint32_t s_at[2][2];
  uint32_t lo[2] = {0, 0}, len[2] = {nq, nq};
  const uint64_t g[2] = {g0, g1};
  for (int round = 0; len[0] > 1 || len[1] > 1; ++round) {
    uint32_t st[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      st[k] = (len[k] + NT - 1) / NT;
      const uint64_t o = uint64_t(threadIdx.x) * st[k];
      // the one sample <= g whose successor sample is > g or past the end
      if (len[k] > 1 && o < len[k] && rs[lo[k] + o] <= g[k] &&
          (o + st[k] >= len[k] || rs[lo[k] + o + st[k]] > g[k]))
        s_at[round & 1][k] = threadIdx.x;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (len[k] > 1) {
        const uint32_t o = s_at[round & 1][k] * st[k];
        lo[k] += o;
        len[k] = (st[k] < len[k] - o) ? st[k] : len[k] - o;
      }
    }
  }
  q0 = lo[0];
  q1 = lo[1];
}

// Largest g in [0, n) with pfx_at(a, g) <= x, for two values x0 <= x1 at
// once, by all NT threads of a block (pfx_at(a, 0) = 0 <= x0):
// block_find_queries over a 64-bit index range.  Every thread of the block
// must call it.
template <int NT>
__device__ __forceinline__ void block_find_rows(const StepArgs& a, uint64_t n, uint64_t x0, uint64_t x1,
                                                uint64_t& g0, uint64_t& g1) {
  __shared__ uint32_t s_at[2][2];
  uint64_t lo[2] = {0, 0}, len[2] = {n, n};
  const uint64_t x[2] = {x0, x1};
  for (int round = 0; len[0] > 1 || len[1] > 1; ++round) {
    uint64_t st[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      st[k] = (len[k] + NT - 1) / NT;
      const uint64_t o = uint64_t(threadIdx.x) * st[k];
      if (len[k] > 1 && o < len[k] && pfx_at(a, lo[k] + o) <= x[k] &&
          (o + st[k] >= len[k] || pfx_at(a, lo[k] + o + st[k]) > x[k]))
        s_at[round & 1][k] = threadIdx.x;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (len[k] > 1) {
        const uint64_t o = uint64_t(s_at[round & 1][k]) * st[k];
        lo[k] += o;
        len[k] = (st[k] < len[k] - o) ? st[k] : len[k] - o;
      }
    }
  }
  g0 = lo[0];
  g1 = lo[1];
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T x) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const T y = __shfl_up(x, o, kWave);
    if (lane >= o) x += y;
  }
  return x;
}

// Exclusive scan over the block; every thread must call it.
template <int NT>
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t* total) {
  constexpr int NW = NT / kWave;
  __shared__ uint64_t wsum[NW];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const uint64_t x = wave_incl_scan<unsigned long long>(v);
  if (lane == kWave - 1) wsum[w] = x;
  __syncthreads();
  if (w == 0) {
    uint64_t s = lane < NW ? wsum[lane] : 0;
    s = wave_incl_scan<unsigned long long>(s);
    if (lane < NW) wsum[lane] = s;
  }
  __syncthreads();
  const uint64_t wbase = w ? wsum[w - 1] : 0;
  if (total) *total = wsum[NW - 1];
  __syncthreads();
  return wbase + x - v;
}

// Maximum over the block; every thread must call it.
template <int NT>
__device__ __forceinline__ uint64_t block_max(uint64_t v) {
  constexpr int NW = NT / kWave;
  __shared__ uint64_t wmax[NW];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const uint64_t y = __shfl_xor((unsigned long long)v, o, kWave);
    v = v > y ? v : y;
  }
  if (lane == 0) wmax[w] = v;
  __syncthreads();
  uint64_t m = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) m = m > wmax[i] ? m : wmax[i];
  __syncthreads();
  return m;
}

template <int NT>
__device__ __forceinline__ uint64_t block_sum(uint64_t v) {
  uint64_t t;
  block_excl_scan<NT>(v, &t);
  return t;
}

// Cluster-chained probe of one 128-byte bucket per hop by W cooperating
// lanes (KVStore::get_slot_local, core/store/kvstore.hpp:292-314): slots 0..6
// are compared with the key, slot 7 links to the next bucket or is empty.
// One bucket of the cluster chain in two halves, so that the loads of
// several buckets can be issued before any of them is compared: load_bucket reads the
// lane's 8/W whole 16-B slots (key and ptr in one load each); scan_bucket
// compares slots 0..6 with the key and takes slot 7 as the link to the next
// bucket (0: end of chain).  With W > 1 the group's result is combined, so
// hit, ptr and link are the same in all W lanes.
template <int W>
__device__ __forceinline__ void load_bucket(const StepArgs& a, const DevSeg& s, uint64_t b,
                                            uint4 (&v)[ASSOCIATIVITY / W]) {
  const int gl = threadIdx.x & (W - 1);
  const uint4* bk = reinterpret_cast<const uint4*>(bucket_addr(a, s, b));
#pragma unroll
  for (int k = 0; k < ASSOCIATIVITY / W; ++k) v[k] = bk[gl + k * W];
}

template <int W>
__device__ __forceinline__ void scan_bucket(const uint4 (&v)[ASSOCIATIVITY / W], uint64_t key, bool& hit,
                                            uint64_t& ptr, uint64_t& link) {
  const int lane = threadIdx.x & (kWave - 1);
  const int gl = lane & (W - 1);
  const int gbase = lane & ~(W - 1);
  bool h = false;
  uint64_t hptr = 0, l = 0;
#pragma unroll
  for (int k = 0; k < ASSOCIATIVITY / W; ++k) {
    const int si = gl + k * W;
    const uint64_t sk = uint64_t(v[k].x) | (uint64_t(v[k].y) << 32);
    const uint64_t sp = uint64_t(v[k].z) | (uint64_t(v[k].w) << 32);
    const bool m = si < ASSOCIATIVITY - 1 && sk == key;
    h = h || m;
    hptr = m ? sp : hptr;
    l = (si == ASSOCIATIVITY - 1) ? sk : l;
  }
  if constexpr (W == 1) {
    hit = h;
    ptr = hptr;
    link = l;
  } else {
    const uint64_t m = (__ballot(h) >> gbase) & ((uint64_t(1) << W) - 1);
    hit = m != 0;
    ptr = __shfl(hptr, gbase + (m ? __ffsll((long long)m) - 1 : 0), kWave);
    link = __shfl(l, gbase + ((ASSOCIATIVITY - 1) & (W - 1)), kWave);
  }
}

template <int W>
__device__ __forceinline__ void probe_hop(const StepArgs& a, const DevSeg& s, uint64_t key, uint64_t b, bool& hit,
                                          uint64_t& ptr, uint64_t& link) {
  uint4 v[ASSOCIATIVITY / W];
  load_bucket<W>(a, s, b, v);
  scan_bucket<W>(v, key, hit, ptr, link);
}

// The rest of a cluster chain after a first bucket that missed with `link`.
template <int W>
__device__ __forceinline__ bool probe_chain(const StepArgs& a, const DevSeg& s, uint64_t key, uint64_t link,
                                            uint64_t& ptr) {
  for (uint64_t hop = 1; hop < s.total_buckets && link != 0; ++hop) {
    const uint64_t b = key_vid(link);
    if (b >= s.total_buckets) return false;  // malformed link: treat as absent
    bool hit;
    probe_hop<W>(a, s, key, b, hit, ptr, link);
    if (hit) return true;
  }
  return false;
}

// Cluster-chained probe of one 128-byte bucket per hop by W cooperating
// lanes (KVStore::get_slot_local, core/store/kvstore.hpp:292-314).
template <int W>
__device__ __forceinline__ bool probe(const StepArgs& a, const DevSeg& s, uint64_t key, uint64_t& ptr) {
  bool hit;
  uint64_t link;
  probe_hop<W>(a, s, key, bucket_of(hash_u64(key), s.num_buckets, s.nb_magic), hit, ptr, link);
  if (hit) return true;
  return probe_chain<W>(a, s, key, link, ptr);
}

// Membership of x in the sorted, duplicate-free list of n edges read by
// at(i): binary search down to <= 4 entries, which are then read together
// (independent loads, one memory latency).  A wider multi-pivot search costs
// the probe kernels enough registers to lower their occupancy, and the lists
// probed here are mostly short.
template <typename At>
__device__ __forceinline__ bool search_sorted(At at, uint64_t n, uint32_t x) {
  uint64_t lo = 0, hi = n;
  while (hi - lo > 4) {
    const uint64_t mid = (lo + hi) >> 1;
    const uint32_t v = at(mid);
    if (v < x) lo = mid + 1;
    else if (v > x) hi = mid;
    else return true;
  }
  bool found = false;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (lo + k < hi) found = found || at(lo + k) == x;
  return found;
}

// Membership of x in the edge list [off, off+n) of segment s: directly in
// the arena when the segment's value blocks are consecutive, else through
// the block table.  Each way keeps its addressing to itself, so only the
// boolean answer leaves the branch.
__device__ __forceinline__ bool contains(const StepArgs& a, const DevSeg& s, uint64_t off, uint64_t n,
                                         uint32_t x) {
  if (s.voff > kBlockTable) {
    const uint32_t* e = a.varena + (uint64_t(s.voff) + off);
    return search_sorted([e](uint64_t i) { return e[i]; }, n, x);
  }
  const uint64_t mask = (uint64_t(1) << a.vshift) - 1;
  return search_sorted(
      [&a, &s, off, mask](uint64_t i) {
        const uint64_t g = off + i;
        return a.varena[(uint64_t(a.vtab[s.vb_base + (g >> a.vshift)]) << a.vshift) | (g & mask)];
      },
      n, x);
}

__device__ __forceinline__ uint64_t step_rows(const StepArgs& a) {
  const uint64_t r = a.row_start[a.nq];
  return r < a.rcap ? r : a.rcap;
}

// The StepDesc fields the probe needs, read as three 16-byte words so that
// every later choice between them is a choice between register values.
// hipcc (ROCm 7.2) miscompiles the form `op == FILTER_KNOWN ? row[d.ec] :
// d.kend` inside the probe: structurizecfg turns the selected pointer into
// poison on the known_to_const path, which then loads from a stale register.
struct ProbeDesc {
  uint32_t op, pid, dir, key_mode, sc, ec, kstart, kend, in_cols;
  int32_t seg;
};
static_assert(offsetof(StepDesc, key_mode) == 16 && offsetof(StepDesc, kend) == 32 &&
                  offsetof(StepDesc, in_cols) == 36 && sizeof(StepDesc) % 16 == 0,
              "StepDesc word layout");

__device__ __forceinline__ ProbeDesc read_probe_desc(const StepDesc* d) {
  const uint4* w = reinterpret_cast<const uint4*>(d);
  const uint4 w0 = w[0], w1 = w[1], w2 = w[2];
  ProbeDesc p;
  p.op = w0.x;
  p.seg = int32_t(w0.y);
  p.pid = w0.z;
  p.dir = w0.w;
  p.key_mode = w1.x;
  p.sc = w1.y;
  p.ec = w1.z;
  p.kstart = w1.w;
  p.kend = w2.x;
  p.in_cols = w2.y;
  return p;
}

// ---- probe: per input row, edge count (expand) or keep flag (filters) -----

// OP_FILTER_TYPE: whether row[sc] is an instance of the step's type.
__device__ __forceinline__ uint32_t type_bit(const StepArgs& a, const ProbeDesc& d, const uint32_t* row) {
  const TypeBits t = a.tbits[d.ec];
  const uint32_t i = row[d.sc] - t.vmin;
  return i < t.nbits ? (t.words[i >> 5] >> (i & 31)) & 1u : 0u;
}

// One input row: key from the row (or the constant / type index), bucket
// probe, then the edge count (expand) or the membership flag (filters).  The
// W lanes of a group call it together with the same row.
template <int W>
__device__ __forceinline__ void probe_row(const StepArgs& a, const ProbeDesc& d, const DevSeg& s,
                                          const uint32_t* row, uint32_t& c, uint32_t& eo) {
  // the key's vertex: a column of the row, or the constant start
  const uint32_t kv = (d.key_mode == KEY_CONST) ? d.kstart : row[d.sc];
  // KEY_TYPE_INDEX reads [0|t|IN]: only type/predicate ids fit the 17-bit pid field
  const bool tix = d.key_mode == KEY_TYPE_INDEX;
  const uint64_t key = tix ? make_key(0, kv, DIR_IN) : make_key(kv, d.pid, d.dir);
  const bool valid = seg_resident(s) && (!tix || kv < (1u << NBITS_IDX));
  uint64_t ptr = 0;
  bool hit;
  if (s.dense && !tix) {
    // the key's slot word straight from the direct index (one 8-B read)
    const uint32_t i = kv - s.dvmin;
    if (valid && i < s.dvn) ptr = s.dense[i];
    hit = ptr != 0;
  } else {
    hit = valid && probe<W>(a, s, key, ptr);
  }
  if (hit) {
    const uint64_t sz = ptr_size(ptr);
    const uint64_t off = ptr_off(ptr) - s.edge_start;
    if (d.op == OP_EXPAND) {
      c = uint32_t(sz);
      eo = uint32_t(off);
    } else {
      uint32_t target = d.kend;
      if (d.op == OP_FILTER_KNOWN) target = row[d.ec];
      c = contains(a, s, off, sz, target) ? 1u : 0u;
    }
  }
}

// Exclusive scan of the per-tile edge counts by one block of NT threads;
// sets the step's total output rows.  Entry nt of the tile prefixes and
// lpfx[R] are set so that pfx_at(a, R) is the total too.
template <int NT>
__device__ __forceinline__ void scan_tiles_block(const StepArgs& a) {
  const uint64_t R = step_rows(a);
  const uint64_t nt = (R + kTile - 1) / kTile;  // R <= rcap, so nt < tile_cap
  const uint64_t per = (nt + NT - 1) / NT;
  uint64_t lo = uint64_t(threadIdx.x) * per;
  if (lo > nt) lo = nt;
  const uint64_t hi = (lo + per < nt) ? lo + per : nt;
  uint64_t s = 0;
  for (uint64_t i = lo; i < hi; ++i) s += a.tile_sum[i];
  uint64_t total = 0;
  uint64_t ex = block_excl_scan<NT>(s, &total);
  for (uint64_t i = lo; i < hi; ++i) {
    const uint64_t v = a.tile_sum[i];
    a.tile_sum[i] = ex;
    // row R inside the last tile: the tile's rows before it are all of them
    if (i == nt - 1 && (R & (kTile - 1))) a.lpfx[R] = v;
    ex += v;
  }
  if (threadIdx.x == 0) {
    a.st->total_out_rows = total;
    a.st->tiles_marked = 0;
    a.tile_sum[nt] = total;
    if (!(R & (kTile - 1))) a.lpfx[R] = 0;  // row R starts tile nt
  }
}

// Two rows of one query at once: both keys, then both first buckets are
// loaded before either is compared, so each lane keeps two dependent load
// chains in flight.  Chains past the first bucket and the membership tests
// of filters follow row by row.  `s` must be resident.
template <int W>
__device__ __forceinline__ void finish_row(const StepArgs& a, const ProbeDesc& d, const DevSeg& s,
                                           const uint32_t* row, bool hit, uint64_t ptr, uint32_t& c, uint32_t& eo) {
  if (!hit) return;
  const uint64_t sz = ptr_size(ptr);
  const uint64_t off = ptr_off(ptr) - s.edge_start;
  if (d.op == OP_EXPAND) {
    c = uint32_t(sz);
    eo = uint32_t(off);
  } else {
    uint32_t target = d.kend;
    if (d.op == OP_FILTER_KNOWN) target = row[d.ec];
    c = contains(a, s, off, sz, target) ? 1u : 0u;
  }
}

template <int W>
__device__ __forceinline__ void probe_row2(const StepArgs& a, const ProbeDesc& d, const DevSeg& s,
                                           const uint32_t* row0, const uint32_t* row1, uint32_t& c0, uint32_t& eo0,
                                           uint32_t& c1, uint32_t& eo1) {
  const bool konst = d.key_mode == KEY_CONST, tix = d.key_mode == KEY_TYPE_INDEX;
  const uint32_t kv0 = konst ? d.kstart : row0[d.sc];
  const uint32_t kv1 = konst ? d.kstart : row1[d.sc];
  if (s.dense && !tix) {
    // both keys' slot words from the direct index, loaded together at
    // in-bounds indices and masked after
    const uint32_t i0 = kv0 - s.dvmin, i1 = kv1 - s.dvmin;
    const bool in0 = i0 < s.dvn, in1 = i1 < s.dvn;
    const uint64_t w0 = s.dense[in0 ? i0 : 0], w1 = s.dense[in1 ? i1 : 0];
    const uint64_t p0 = in0 ? w0 : 0, p1 = in1 ? w1 : 0;
    finish_row<W>(a, d, s, row0, p0 != 0, p0, c0, eo0);
    finish_row<W>(a, d, s, row1, p1 != 0, p1, c1, eo1);
    return;
  }
  const uint64_t key0 = tix ? make_key(0, kv0, DIR_IN) : make_key(kv0, d.pid, d.dir);
  const uint64_t key1 = tix ? make_key(0, kv1, DIR_IN) : make_key(kv1, d.pid, d.dir);
  uint4 v0[ASSOCIATIVITY / W], v1[ASSOCIATIVITY / W];
  load_bucket<W>(a, s, bucket_of(hash_u64(key0), s.num_buckets, s.nb_magic), v0);
  load_bucket<W>(a, s, bucket_of(hash_u64(key1), s.num_buckets, s.nb_magic), v1);
  bool h0, h1;
  uint64_t p0, p1, l0, l1;
  scan_bucket<W>(v0, key0, h0, p0, l0);
  scan_bucket<W>(v1, key1, h1, p1, l1);
  if (!h0 && l0 != 0) h0 = probe_chain<W>(a, s, key0, l0, p0);
  if (!h1 && l1 != 0) h1 = probe_chain<W>(a, s, key1, l1, p1);
  // KEY_TYPE_INDEX reads [0|t|IN]: only type/predicate ids fit the 17-bit pid field
  finish_row<W>(a, d, s, row0, h0 && (!tix || kv0 < (1u << NBITS_IDX)), p0, c0, eo0);
  finish_row<W>(a, d, s, row1, h1 && (!tix || kv1 < (1u << NBITS_IDX)), p1, c1, eo1);
}

// The rows of tile [tlo, thi), all of one query (descriptor d, segment s,
// input rows from tbase, first row r0): counts into s_cnt, edge offsets into
// eoff.  The W lanes of a group share a row.
template <int W, int NT>
__device__ __forceinline__ void probe_tile_one(const StepArgs& a, const ProbeDesc& d, const DevSeg& s,
                                               const uint32_t* tbase, uint64_t r0, uint64_t tlo, uint64_t thi,
                                               uint32_t* s_cnt) {
  constexpr int G = NT / W;
  constexpr int kPasses = kTile / G;
  const int grp = threadIdx.x / W, gl = threadIdx.x % W;
  const bool live = d.op != OP_SKIP && d.seg >= 0;
  const bool tb = d.op == OP_FILTER_TYPE;
  if constexpr (kPasses % 2 == 0) {
    // rows g and g + G of each pair of passes together
    const bool res = live && seg_resident(s);
    for (int pass = 0; pass < kPasses; pass += 2) {
      const uint64_t g = tlo + uint64_t(pass) * G + grp;
      if (g < thi) {
        const uint64_t g1 = g + G;
        const bool has1 = g1 < thi;
        const uint64_t gr1 = has1 ? g1 : g;  // a lone last row probes itself twice
        uint32_t c = 0, eo = 0, c1 = 0, eo1 = 0;
        if (tb) {
          c = type_bit(a, d, tbase + (g - r0) * d.in_cols);
          c1 = type_bit(a, d, tbase + (gr1 - r0) * d.in_cols);
        } else if (res) {
          probe_row2<W>(a, d, s, tbase + (g - r0) * d.in_cols, tbase + (gr1 - r0) * d.in_cols, c, eo, c1, eo1);
        }
        if (gl == 0) {
          s_cnt[g - tlo] = c;
          a.eoff[g] = eo;
          if (has1) {
            s_cnt[g1 - tlo] = c1;
            a.eoff[g1] = eo1;
          }
        }
      }
    }
  } else {
    for (int pass = 0; pass < kPasses; ++pass) {
      const uint64_t g = tlo + uint64_t(pass) * G + grp;
      if (g < thi) {
        uint32_t c = 0, eo = 0;
        if (tb) c = type_bit(a, d, tbase + (g - r0) * d.in_cols);
        else if (live) probe_row<W>(a, d, s, tbase + (g - r0) * d.in_cols, c, eo);
        if (gl == 0) {
          s_cnt[g - tlo] = c;
          a.eoff[g] = eo;
        }
      }
    }
  }
}

// The end of a tile: the prefix of its counts within it (kTile / NT
// consecutive rows per thread) to lpfx, and its sum to tile_sum[t].
template <int NT>
__device__ __forceinline__ void probe_tile_end(const StepArgs& a, uint64_t t, uint64_t tlo, uint64_t thi,
                                               const uint32_t* s_cnt, uint64_t* s_pfx) {
  constexpr int kPer = kTile / NT;
  // the tile's prefix of counts, kPer consecutive rows per thread, and its
  // sum
  __syncthreads();
  const uint32_t nrow = uint32_t(thi - tlo);
  // this thread's index, opaque to the compiler, so that the addresses below
  // are formed here and not kept in registers across the whole tile loop
  uint32_t tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  uint32_t cc[kPer];  // a row's count is below 2^28
  uint64_t cs = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const uint32_t i = tid * kPer + k;
    cc[k] = i < nrow ? s_cnt[i] : 0u;
    cs += cc[k];
  }
  uint64_t tsum = 0;
  uint64_t ex = block_excl_scan<NT>(cs, &tsum);  // ends with a barrier: s_cnt is free again
  if constexpr (kPer == 1) {
    if (tid < nrow) a.lpfx[tlo + tid] = ex;
  } else {
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      s_pfx[tid * kPer + k] = ex;
      ex += cc[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const uint32_t i = k * NT + tid;
      if (i < nrow) a.lpfx[tlo + i] = s_pfx[i];
    }
    __syncthreads();  // before the next tile's prefixes overwrite s_pfx
  }
  if (threadIdx.x == 0) a.tile_sum[t] = tsum;
}

// One workgroup of NT threads per tile of kTile rows, in kTile / (NT / W)
// passes.  A tile inside one query reads the descriptor and segment once,
// into scalar registers; a tile spanning several queries stages their row
// starts, descriptors and input offsets in LDS and each row finds its query
// there.  NT = 512 (two rows per lane at W = 1, two workgroups per CU) serves
// batches; NT = 1024 a batched step of one row per query, whose tiles span
// as many queries as rows; NT = 256 serves one query, whose descriptor and
// segment each workgroup reads once for all its tiles.  (Capping the
// registers for one more wave per SIMD made the large query's probe 30%
// slower: fewer loads stay in flight per wave.)
template <int W, int NT>
__global__ __launch_bounds__(NT) void k_probe(StepArgs a) {
  static_assert(kTile % NT == 0 && NT % W == 0, "probe tile shape");
  if (a.st->status) return;
  // the queries of a tile that spans several (up to a tile's rows), staged by
  // the batch probes only (a one-query launch has no such tile)
  constexpr int QS = NT >= 512 ? kTile : 1;
  __shared__ uint64_t s_rs[QS];
  __shared__ ProbeDesc s_pd[QS];
  __shared__ uint64_t s_in[QS];
  // the tile's counts, scanned at its end; with several rows per thread the
  // prefixes go through s_pfx, so that their stores are coalesced
  constexpr int kPer = kTile / NT;
  __shared__ uint32_t s_cnt[kTile];
  __shared__ uint64_t s_pfx[kPer > 1 ? kTile : 1];
  const uint64_t R = step_rows(a);
  const uint64_t ntiles = (R + kTile - 1) / kTile;
  constexpr int G = NT / W;
  constexpr int kPasses = kTile / G;
  const int grp = threadIdx.x / W, gl = threadIdx.x % W;
  if constexpr (NT == 256) {
    // one query (launch_probe uses NT = 256 for one-query launches only):
    // its descriptor, segment and row base are read from memory once per
    // workgroup, into LDS, and from there for each tile
    __shared__ ProbeDesc s_d1[1];
    __shared__ DevSeg s_s1[1];
    __shared__ uint64_t s_x1[2];
    if (threadIdx.x == 0) {
      const ProbeDesc d = read_probe_desc(a.desc);
      s_d1[0] = d;
      s_s1[0] = a.segs[d.op != OP_SKIP && d.seg >= 0 ? d.seg : 0];
      s_x1[0] = a.desc[0].in_off;
      s_x1[1] = a.row_start[0];
    }
    __syncthreads();
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
      const uint64_t tlo = t * kTile, thi = (tlo + kTile < R) ? tlo + kTile : R;
      // an index of 0 the compiler cannot see through, so that the LDS reads
      // happen per tile and their values are not held across the loop
      uint32_t z = 0;
      asm volatile("" : "+s"(z));
      const ProbeDesc d = s_d1[z];
      probe_tile_one<W, NT>(a, d, s_s1[z], a.tin + s_x1[z], s_x1[z + 1], tlo, thi, s_cnt);
      probe_tile_end<NT>(a, t, tlo, thi, s_cnt, s_pfx);
    }
  } else {
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
      const uint64_t tlo = t * kTile, thi = (tlo + kTile < R) ? tlo + kTile : R;
      uint32_t q0, q1;
      block_find_queries<NT>(a.row_start, a.nq, tlo, thi - 1, q0, q1);
      if (q0 == q1) {
        // one query owns the tile: descriptor, segment and row base once, in
        // scalar registers
        const uint32_t q = __builtin_amdgcn_readfirstlane(q0);
        const ProbeDesc d = read_probe_desc(a.desc + q);
        const DevSeg s = a.segs[d.op != OP_SKIP && d.seg >= 0 ? d.seg : 0];
        probe_tile_one<W, NT>(a, d, s, a.tin + a.desc[q].in_off, a.row_start[q], tlo, thi, s_cnt);
      } else {
        const uint32_t nqt = q1 - q0 + 1;
        const bool staged = QS > 1 && nqt <= uint32_t(QS);
        if (staged) {
          // the tile's queries: row start, probe descriptor and input offset,
          // so that a row's chain starts from LDS instead of dependent global
          // loads (its segment, shared by many queries, is read from cache)
          for (uint32_t i = threadIdx.x; i < nqt; i += NT) {
            const uint32_t q = q0 + i;
            s_rs[i] = a.row_start[q];
            s_pd[i] = read_probe_desc(a.desc + q);
            s_in[i] = a.desc[q].in_off;
          }
        }
        __syncthreads();
        for (int pass = 0; pass < kPasses; ++pass) {
          const uint64_t g = tlo + uint64_t(pass) * G + grp;
          if (g < thi) {
            uint32_t c = 0, eo = 0;
            // each branch probes by itself, so only values leave LDS or memory
            if (staged) {
              uint32_t l = 0, h = nqt;
              while (h - l > 1) {
                const uint32_t m = (l + h) >> 1;
                if (s_rs[m] <= g) l = m;
                else h = m;
              }
              const ProbeDesc d = s_pd[l];
              if (d.op == OP_FILTER_TYPE)
                c = type_bit(a, d, a.tin + s_in[l] + (g - s_rs[l]) * d.in_cols);
              else if (d.op != OP_SKIP && d.seg >= 0)
                probe_row<W>(a, d, a.segs[d.seg], a.tin + s_in[l] + (g - s_rs[l]) * d.in_cols, c, eo);
            } else {
              const uint32_t q = query_in(a.row_start, q0, q1 + 1, g);
              const ProbeDesc d = read_probe_desc(a.desc + q);
              if (d.op == OP_FILTER_TYPE)
                c = type_bit(a, d, a.tin + a.desc[q].in_off + (g - a.row_start[q]) * d.in_cols);
              else if (d.op != OP_SKIP && d.seg >= 0)
                probe_row<W>(a, d, a.segs[d.seg], a.tin + a.desc[q].in_off + (g - a.row_start[q]) * d.in_cols, c,
                             eo);
            }
            if (gl == 0) {
              s_cnt[g - tlo] = c;
              a.eoff[g] = eo;
            }
          }
        }
      }
      probe_tile_end<NT>(a, t, tlo, thi, s_cnt, s_pfx);
    }
  }
}

// ---- scan -----------------------------------------------------------------

__global__ __launch_bounds__(1024) void k_scan_tiles(StepArgs a) {
  if (a.st->status) return;
  scan_tiles_block<1024>(a);
}

// End of the output rows the gather writes in this launch (the step's total,
// or the host's window end), and the number of output tiles from win_lo.
__device__ __forceinline__ uint64_t out_end(const StepArgs& a) {
  const uint64_t t = a.st->total_out_rows;
  return t < a.win_hi ? t : a.win_hi;
}

// Output tiles of the launch.  The tile -> row map covers the first
// otile_cap of them (enough for every row that fits the tables and the
// arena); tiles past it, which only a step with dropped final tables has,
// find their rows themselves.
__device__ __forceinline__ uint64_t out_tiles(const StepArgs& a) {
  const uint64_t e = out_end(a);
  return e > a.win_lo ? (e - a.win_lo + kOutTile - 1) / kOutTile : 0;
}

__device__ __forceinline__ uint64_t mapped_tiles(const StepArgs& a) {
  const uint64_t n = out_tiles(a);
  return n < a.otile_cap ? n : a.otile_cap;
}

// The largest g in [0, R) with pfx[g] <= j, found by one wavefront (j is
// the same in every lane): each round samples 64 evenly spaced rows of the
// remaining range, one per lane, and keeps the span after the last sample
// <= j, so R rows take ceil(log64 R) rounds of one dependent load each
// (4 for 16M rows) instead of log2 R.  pfx[0] = 0 <= j, and the samples <= j
// are a prefix of the lanes, as pfx does not decrease.
__device__ __forceinline__ uint64_t wave_find_row(const StepArgs& a, uint64_t R, uint64_t j) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  uint64_t lo = 0, len = R;
  while (len > 1) {
    const uint64_t step = (len + kWave - 1) / kWave;
    const uint64_t o = uint64_t(lane) * step;
    const bool le = o < len && pfx_at(a, lo + o) <= j;
    const uint64_t m = __ballot(le);
    const uint64_t at = uint64_t(63 - __builtin_clzll(m)) * step;  // lane 0's sample is lo itself: m != 0
    lo += at;
    len = len - at < step ? len - at : step;
  }
  return lo;
}

// Output tile -> source row of its first output row: the largest g with
// pfx[g] <= t * kOutTile (a zero-count row can never be that largest g, as
// the row after it has the same prefix), for the nt mapped tiles, by the
// threads of blocks [b0, gridDim.x).  Up to two tiles per wavefront of those
// blocks, each tile takes a wavefront (a few dependent loads); more tiles take
// a thread each, by binary search (fewer loads in all).  Either way a single
// row with millions of edges costs nothing extra.
__device__ __forceinline__ void mark_tiles(const StepArgs& a, uint64_t R, uint64_t nt, uint32_t b0) {
  constexpr uint32_t kWaves = kBlock / kWave;
  const uint64_t nb = gridDim.x - b0, b = blockIdx.x - b0;
  if (nt <= 2 * nb * kWaves) {
    for (uint64_t t = b * kWaves + threadIdx.x / kWave; t < nt; t += nb * kWaves) {
      const uint64_t g = wave_find_row(a, R, a.win_lo + t * kOutTile);
      if ((threadIdx.x & (kWave - 1)) == 0) a.tile_row[t] = g;
    }
    return;
  }
  for (uint64_t t = b * kBlock + threadIdx.x; t < nt; t += nb * kBlock) {
    const uint64_t j = a.win_lo + t * kOutTile;
    uint64_t lo = 0, hi = R;
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) >> 1;
      if (pfx_at(a, mid) <= j) lo = mid;
      else hi = mid;
    }
    a.tile_row[t] = lo;
  }
}

// A step whose few output tiles k_scan_small mapped for the whole output
// (no window) returns at once.
__global__ __launch_bounds__(kBlock) void k_tile_rows(StepArgs a) {
  if (a.st->status || (a.st->tiles_marked && a.win_lo == 0 && a.win_hi == ~uint64_t(0))) return;
  mark_tiles(a, step_rows(a), mapped_tiles(a), 0);
}

// Elements of the result arena left for this step's final tables.  A final
// table larger than all of it can never fit: it takes no arena space in the
// query scan, so it cannot push the tables after it out of the arena.
__device__ __forceinline__ uint64_t arena_room(const StepArgs& a) {
  const uint64_t base = a.st->arena_base;
  return a.arena_cap > base ? a.arena_cap - base : 0;
}

// Status of a step once its output sizes are known: 1 when its intermediate
// tables overflow the table capacity.
__device__ __forceinline__ uint32_t step_status(const StepArgs& a, uint64_t carry_t) {
  return carry_t > a.tcap ? 1u : 0u;
}

// The query scan of one step by one 1,024-thread block: every query's output
// rows and offsets, the next step's row space and descriptors, the step's
// status.  Every thread of the block must call it.
__device__ __forceinline__ void scan_queries_block(const StepArgs& a) {
  const uint64_t R = step_rows(a);
  if (a.nq == 1) {
    // one query (every single query): its offsets are the bases themselves,
    // written by one thread without the block scans
    if (threadIdx.x != 0) return;
    uint64_t r0 = a.row_start[0], r1 = a.row_start[1];
    r0 = r0 < R ? r0 : R;
    r1 = r1 < R ? r1 : R;
    const uint64_t rows = pfx_at(a, r1) - pfx_at(a, r0);
    StepDesc& d = a.desc[0];
    const uint32_t oc = d.out_cols;
    const bool fin = d.final_step != 0;
    const uint64_t el = rows * oc;
    const uint64_t tbase = a.st->tout_base, abase = a.st->arena_base;
    d.out_rows = rows;
    if (a.ck && fin) a.ck[0] = 0;
    if (a.next) {
      a.next[0].in_rows = fin ? 0 : rows;
      a.next[0].in_off = tbase;
      a.next_row_start[0] = 0;
    }
    uint64_t aend = abase, tt = 0, tn = 0;
    if (fin) {
      const bool fits = abase + el <= a.arena_cap;
      d.out_off = fits ? abase : kNoSpace;
      if (!fits) d.out_rows = 0;
      if (a.qres) {
        a.qres[0] = fits ? rows : 0;
        a.qres[1] = fits ? abase : kNoSpace;
      }
      if (fits) aend = abase + el;
    } else {
      d.out_off = tbase;
      tt = el;
      tn = rows;
    }
    const uint32_t status = step_status(a, tbase + tt);
    a.st->out_elems = tt;
    a.st->status = status;
    if (a.next) {
      a.next_row_start[1] = tn;
      if (a.next_st) {
        a.next_st->arena_base = aend;
        a.next_st->tout_base = 0;
        a.next_st->status = status;
      }
    }
    return;
  }
  // each thread owns a contiguous run of queries: three block scans in all
  const uint32_t per = (a.nq + 1023) / 1024;
  const uint32_t qlo = threadIdx.x * per < a.nq ? threadIdx.x * per : a.nq;
  const uint32_t qhi = qlo + per < a.nq ? qlo + per : a.nq;
  auto rows_of = [&](uint32_t q) {
    uint64_t r0 = a.row_start[q], r1 = a.row_start[q + 1];
    r0 = r0 < R ? r0 : R;
    r1 = r1 < R ? r1 : R;
    return pfx_at(a, r1) - pfx_at(a, r0);
  };
  // up to kPer queries per thread (nq <= 8192): their row counts are loaded
  // together and kept in registers for the second pass
  constexpr uint32_t kPer = 8;
  const bool cached = per <= kPer;
  uint64_t crows[kPer];
  uint32_t ccols[kPer];  // out_cols, or 0x80000000 | out_cols for a final step
  const uint64_t room = arena_room(a);
  uint64_t et = 0, ea = 0, nr = 0;
  if (cached) {
    // every load is unconditional at a clamped, in-bounds index, so they all
    // issue back to back; entries past the thread's run are masked after
    uint64_t rs[kPer + 1];
#pragma unroll
    for (uint32_t i = 0; i <= kPer; ++i) {
      const uint32_t qi = qlo + i < qhi ? qlo + i : qhi;  // row_start has nq + 1 entries
      const uint64_t r = a.row_start[qi];
      rs[i] = r < R ? r : R;
    }
    uint64_t pf[kPer + 1];
#pragma unroll
    for (uint32_t i = 0; i <= kPer; ++i) pf[i] = pfx_at(a, rs[i]);
    uint32_t dc[kPer], df[kPer];
#pragma unroll
    for (uint32_t i = 0; i < kPer; ++i) {
      const uint32_t qi = qlo + i < a.nq ? qlo + i : a.nq - 1;
      dc[i] = a.desc[qi].out_cols;
      df[i] = a.desc[qi].final_step;
    }
#pragma unroll
    for (uint32_t i = 0; i < kPer; ++i) {
      const bool in = qlo + i < qhi;
      const uint64_t rows = in ? pf[i + 1] - pf[i] : 0;
      crows[i] = rows;
      ccols[i] = in ? (dc[i] | (df[i] ? 0x80000000u : 0u)) : 0u;
      if (df[i]) {
        const uint64_t el = rows * dc[i];
        ea += el <= room ? el : 0;
      } else {
        et += rows * dc[i];
        nr += rows;
      }
    }
  } else {
    for (uint32_t q = qlo; q < qhi; ++q) {
      const uint64_t rows = rows_of(q);
      const StepDesc& d = a.desc[q];
      if (d.final_step) {
        const uint64_t el = rows * d.out_cols;
        ea += el <= room ? el : 0;
      } else {
        et += rows * d.out_cols;
        nr += rows;
      }
    }
  }
  uint64_t tt = 0, tn = 0;
  const uint64_t tbase = a.st->tout_base, abase = a.st->arena_base;
  uint64_t xt = block_excl_scan<1024>(et, &tt) + tbase;
  uint64_t xa = block_excl_scan<1024>(ea, nullptr) + abase;
  uint64_t xn = block_excl_scan<1024>(nr, &tn);
  uint64_t aend = abase;  // end of this thread's last final table that fits
  for (uint32_t q = qlo; q < qhi; ++q) {
    uint64_t rows = 0;
    uint32_t oc = 0;
    bool fin = false;
    if (cached) {
#pragma unroll
      for (uint32_t i = 0; i < kPer; ++i)
        if (q - qlo == i) {
          rows = crows[i];
          oc = ccols[i] & 0x7FFFFFFFu;
          fin = (ccols[i] >> 31) != 0;
        }
    } else {
      rows = rows_of(q);
      oc = a.desc[q].out_cols;
      fin = a.desc[q].final_step != 0;
    }
    StepDesc& d = a.desc[q];
    const uint64_t el = rows * oc;
    d.out_rows = rows;
    if (a.ck && fin) a.ck[q] = 0;  // the gather adds the final rows' hashes
    if (a.next) {
      a.next[q].in_rows = fin ? 0 : rows;
      a.next[q].in_off = xt;
      a.next_row_start[q] = xn;
    }
    if (fin) {
      // a final table that does not fit the arena is dropped alone; one
      // larger than the whole room takes no arena space
      const bool fits = xa + el <= a.arena_cap;
      d.out_off = fits ? xa : kNoSpace;
      if (!fits) d.out_rows = 0;
      if (a.qres) {
        a.qres[2 * uint64_t(q)] = fits ? rows : 0;
        a.qres[2 * uint64_t(q) + 1] = fits ? xa : kNoSpace;
      }
      if (fits) aend = xa + el;
      xa += el <= room ? el : 0;
    } else {
      d.out_off = xt;
      xt += el;
      xn += rows;
    }
  }
  // the tables that fit are a prefix of the step's final tables in query
  // order: the next step's tables follow the last of them, not the dropped
  // ones, which take no room
  const uint64_t carry_a = block_max<1024>(aend);
  if (threadIdx.x == 0) {
    const uint64_t carry_t = tbase + tt;
    const uint32_t status = step_status(a, carry_t);
    a.st->out_elems = tt;
    a.st->status = status;
    if (a.next) {
      a.next_row_start[a.nq] = tn;
      if (a.next_st) {
        a.next_st->arena_base = carry_a;
        a.next_st->tout_base = 0;
        a.next_st->status = status;
      }
    }
  }
}

__global__ __launch_bounds__(1024) void k_scan_queries(StepArgs a) {
  if (a.st->status) {
    // an overflowed step skips everything after it: pass the status on so
    // the next step's kernels also return before reading its (unwritten) rows
    if (threadIdx.x == 0 && a.next_st) a.next_st->status = a.st->status;
    return;
  }
  scan_queries_block(a);
}

// The tile scan and the query scan of a step in one 1,024-thread block (steps
// of at most kBlock queries: every single query, small serving iterations):
// the query scan reads the tile prefixes the same block wrote before the
// barrier between them.
__global__ __launch_bounds__(1024) void k_scan_small(StepArgs a) {
  if (a.st->status) {
    // an overflowed step skips everything after it: pass the status on so
    // the next step's kernels also return before reading its (unwritten) rows
    if (threadIdx.x == 0 && a.next_st) a.next_st->status = a.st->status;
    return;
  }
  scan_tiles_block<1024>(a);
  __syncthreads();
  scan_queries_block(a);
  // a step of at most kSmallTiles output tiles, gathered whole: its output
  // tile map too, two tiles per wavefront, instead of a k_tile_rows launch
  constexpr uint64_t kSmallTiles = 2 * (1024 / kWave);
  if (a.win_lo != 0 || a.win_hi != ~uint64_t(0)) return;
  __syncthreads();
  if (a.st->status) return;
  const uint64_t nto = mapped_tiles(a);
  if (nto > kSmallTiles) return;
  const uint64_t R = step_rows(a);
  for (uint64_t t = threadIdx.x / kWave; t < nto; t += 1024 / kWave) {
    const uint64_t g = wave_find_row(a, R, t * kOutTile);
    if ((threadIdx.x & (kWave - 1)) == 0) a.tile_row[t] = g;
  }
  if (threadIdx.x == 0) a.st->tiles_marked = 1;
}

// Multi-block form of k_scan_queries for large batches.  One block scanning
// thousands of queries is bound by its own scattered stores (each query's
// descriptor and next-step descriptor are separate cache lines), so the
// queries are spread over ceil(nq / kBlock) blocks in two launches.
//
// First launch: blocks [0, nqb) sum the non-final output elements, final
// output elements and next-step rows of their kBlock queries into qagg; the
// blocks after them mark output tiles (the work of k_tile_rows), which needs
// only the prefix computed before this launch.
__global__ __launch_bounds__(kBlock) void k_query_sums(StepArgs a, uint32_t nqb) {
  if (a.st->status) return;
  const uint64_t R = step_rows(a);
  if (blockIdx.x < nqb) {
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    uint64_t et = 0, ea = 0, nr = 0;
    if (q < a.nq) {
      uint64_t r0 = a.row_start[q], r1 = a.row_start[q + 1];
      r0 = r0 < R ? r0 : R;
      r1 = r1 < R ? r1 : R;
      const uint64_t rows = pfx_at(a, r1) - pfx_at(a, r0);
      const uint32_t oc = a.desc[q].out_cols;
      if (a.desc[q].final_step) {
        ea = rows * oc;
        if (ea > arena_room(a)) ea = 0;
      } else {
        et = rows * oc;
        nr = rows;
      }
    }
    et = block_sum<kBlock>(et);
    ea = block_sum<kBlock>(ea);
    nr = block_sum<kBlock>(nr);
    if (threadIdx.x == 0) {
      a.qagg[3 * uint64_t(blockIdx.x)] = et;
      a.qagg[3 * uint64_t(blockIdx.x) + 1] = ea;
      a.qagg[3 * uint64_t(blockIdx.x) + 2] = nr;
    }
    return;
  }
  mark_tiles(a, R, mapped_tiles(a), nqb);
}

// Second launch: each block adds the sums of the blocks before it (and the
// totals of all of them) to its own block scans, then writes its queries'
// offsets exactly as k_scan_queries does; the last block writes the step's
// totals and status.
__global__ __launch_bounds__(kBlock) void k_scan_queries_mb(StepArgs a, uint32_t nqb) {
  if (a.st->status) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.next_st) a.next_st->status = a.st->status;
    return;
  }
  __shared__ uint64_t s_red[6][kBlock / kWave];
  const uint64_t R = step_rows(a);
  uint64_t v[6] = {0, 0, 0, 0, 0, 0};  // prefix t/a/n, then totals t/a/n
  for (uint32_t i = threadIdx.x; i < nqb; i += kBlock) {
    const uint64_t x = a.qagg[3 * uint64_t(i)], y = a.qagg[3 * uint64_t(i) + 1], z = a.qagg[3 * uint64_t(i) + 2];
    if (i < blockIdx.x) {
      v[0] += x;
      v[1] += y;
      v[2] += z;
    }
    v[3] += x;
    v[4] += y;
    v[5] += z;
  }
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const uint64_t x = wave_incl_scan<unsigned long long>(v[k]);
    if (lane == kWave - 1) s_red[k][w] = x;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    uint64_t t = 0;
#pragma unroll
    for (int i = 0; i < kBlock / kWave; ++i) t += s_red[k][i];
    v[k] = t;
  }
  const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
  uint64_t rows = 0;
  uint32_t oc = 0;
  bool fin = false;
  if (q < a.nq) {
    uint64_t r0 = a.row_start[q], r1 = a.row_start[q + 1];
    r0 = r0 < R ? r0 : R;
    r1 = r1 < R ? r1 : R;
    rows = pfx_at(a, r1) - pfx_at(a, r0);
    oc = a.desc[q].out_cols;
    fin = a.desc[q].final_step != 0;
  }
  const uint64_t el = rows * oc;
  const uint64_t tbase = a.st->tout_base, abase = a.st->arena_base;
  const uint64_t xt = block_excl_scan<kBlock>(fin ? 0 : el, nullptr) + v[0] + tbase;
  const uint64_t share = (fin && el <= arena_room(a)) ? el : 0;  // as k_query_sums counted it
  const uint64_t xa = block_excl_scan<kBlock>(share, nullptr) + v[1] + abase;
  const uint64_t xn = block_excl_scan<kBlock>(fin ? 0 : rows, nullptr) + v[2];
  if (q < a.nq) {
    StepDesc& d = a.desc[q];
    if (a.ck && fin) a.ck[q] = 0;  // the gather adds the final rows' hashes
    // a final table that does not fit the arena is dropped alone
    const bool fits = !fin || xa + el <= a.arena_cap;
    d.out_rows = fits ? rows : 0;
    d.out_off = fin ? (fits ? xa : kNoSpace) : xt;
    if (a.next) {
      a.next[q].in_rows = fin ? 0 : rows;
      a.next[q].in_off = xt;
      a.next_row_start[q] = xn;
    }
    if (fin && a.qres) {
      a.qres[2 * uint64_t(q)] = fits ? rows : 0;
      a.qres[2 * uint64_t(q) + 1] = fits ? xa : kNoSpace;
    }
  }
  // the next step's arena base: the end of the last final table that fits
  // (a prefix of the step's final tables), the largest over all blocks; the
  // host uploads the later steps' bases as 0
  const uint64_t bend = block_max<kBlock>((q < a.nq && fin && xa + el <= a.arena_cap) ? xa + el : abase);
  if (threadIdx.x == 0 && a.next_st) atomicMax((unsigned long long*)&a.next_st->arena_base, (unsigned long long)bend);
  if (blockIdx.x == nqb - 1 && threadIdx.x == 0) {
    const uint64_t tt = v[3], tn = v[5];
    const uint64_t carry_t = tbase + tt;
    const uint32_t status = step_status(a, carry_t);
    a.st->out_elems = tt;
    a.st->status = status;
    if (a.next) {
      a.next_row_start[a.nq] = tn;
      if (a.next_st) {
        a.next_st->tout_base = 0;
        a.next_st->status = status;
      }
    }
  }
}

// ---- gather ---------------------------------------------------------------

// Writes output row j of query q: input row g (whose first output is row
// pg of the step) and, for an expansion, edge eo + (j - pg) of its list.
// r0 is the query's first input row and p0 = pfx[r0] its first output row.
// With CK, returns the row's hash when the step keeps checksums of final
// tables (a.ck) and this is one, else 0; the caller adds it to the query's
// sum.  Without CK no hashing code is generated at all.
template <bool CK>
__device__ __forceinline__ uint64_t emit_row(const StepArgs& a, uint32_t q, uint64_t r0, uint64_t p0,
                                             uint64_t g, uint64_t j, uint64_t pg, uint32_t eo) {
  const StepDesc& d = a.desc[q];
  if (d.out_off == kNoSpace) return 0;  // final table dropped: no room in the arena
  const uint64_t orow = j - p0 - a.win_lo;
  uint32_t* dst = (d.final_step ? a.arena : a.tout) + d.out_off + orow * d.out_cols;
  const uint32_t* src = a.tin + d.in_off + (g - r0) * d.in_cols;
  const uint32_t ic = d.in_cols;
  const bool ck = CK && a.ck && d.final_step;
  uint64_t h = row_hash_init(d.out_cols);
  for (uint32_t c = 0; c < ic; ++c) {
    const uint32_t x = src[c];
    dst[c] = x;
    if (ck) h = row_hash_step(h, x, c);
  }
  if (d.op == OP_EXPAND) {
    const DevSeg s = a.segs[d.seg];
    const uint32_t x = edge_at(a, s, uint64_t(eo) + (j - pg));
    dst[ic] = x;
    if (ck) h = row_hash_step(h, x, ic);
  }
  return ck ? h : 0;
}

template <bool CK>
__device__ __forceinline__ uint64_t emit_row(const StepArgs& a, uint32_t q, uint64_t g, uint64_t j, uint64_t pg,
                                             uint32_t eo) {
  const uint64_t r0 = a.row_start[q];
  return emit_row<CK>(a, q, r0, pfx_at(a, r0), g, j, pg, eo);
}

template <bool CK>
__device__ __forceinline__ uint64_t emit_row(const StepArgs& a, uint32_t q, uint64_t g, uint64_t j) {
  return emit_row<CK>(a, q, g, j, pfx_at(a, g), a.eoff[g]);
}

__device__ __forceinline__ void add_checksum(const StepArgs& a, uint32_t q, uint64_t h) {
  if (h) atomicAdd(&a.ck[q], (unsigned long long)h);
}

// Output-balanced: a block writes one tile of kOutTile consecutive output
// rows, kOutTile / kBlock per thread in coalesced order.  The tile's source
// rows [lo, hi) come from the tile -> row map.  When they fit, their prefix
// and edge offsets are staged in LDS and each output row finds its source row
// by a binary search there (a tile of a heavy row stages one entry).  When
// they do not (filters that keep a fraction of their rows), every S-th row's
// prefix is staged, S = ceil(rows / kOutTile): the search in LDS brackets the
// source row within S rows, and log2(S) loads of the global prefix finish it.
constexpr int kStage = kOutTile + 1;
constexpr int kRegCols = 8;  // input columns a gather thread holds in registers
constexpr int kQStage = kBlock;  // queries of a mixed tile staged in LDS

// Last row g in [g0, g1) whose prefix is <= j, given the prefix p of g0 is;
// returns g and sets p to its prefix.
__device__ __forceinline__ uint64_t refine_row(const StepArgs& a, uint64_t g0, uint64_t g1, uint64_t j,
                                               uint64_t& p) {
  while (g1 - g0 > 1) {
    const uint64_t m = (g0 + g1) >> 1;
    const uint64_t v = pfx_at(a, m);
    if (v <= j) {
      g0 = m;
      p = v;
    } else {
      g1 = m;
    }
  }
  return g0;
}

// Rank of this thread's flag among the set flags of its kBlock-thread block,
// in thread order, and their number.  Every thread of the block must call it.
__device__ __forceinline__ uint32_t block_flag_rank(bool f, uint32_t& total) {
  __shared__ uint32_t s_wc[kBlock / kWave];
  const uint64_t m = __ballot(f);
  const uint32_t lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const uint32_t below = uint32_t(__popcll(m & ((uint64_t(1) << lane) - 1)));
  if (lane == 0) s_wc[w] = uint32_t(__popcll(m));
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kBlock / kWave; ++i) {
    const uint32_t c = s_wc[i];
    off += uint32_t(i) < w ? c : 0u;
    tot += c;
  }
  __syncthreads();
  total = tot;
  return off + below;
}

// What a mixed tile stages in LDS of each of its queries (up to kQStage):
// first input row, first output row, the checksum share, and the descriptor
// and segment fields a row's copy needs, so that a row reads no descriptor or
// segment from memory.
struct QStage {
  uint32_t qid[kQStage];       // the query
  uint64_t r0[kQStage];
  uint64_t p0[kQStage];
  unsigned long long ck[kQStage];
  uint64_t out_off[kQStage];  // element offset of its table
  uint64_t in_off[kQStage];
  int64_t voff[kQStage];      // the expanded segment's edge addressing (DevSeg)
  uint32_t vb_base[kQStage];
  uint32_t meta[kQStage];     // in_cols | out_cols << 8 | expand << 16 | final << 17 | no room << 18
};

__device__ __forceinline__ void stage_query(const StepArgs& a, QStage& qs, uint32_t i, uint32_t q, uint64_t R) {
  const StepDesc& d = a.desc[q];
  const uint64_t r0 = a.row_start[q];
  const bool expand = d.op == OP_EXPAND;
  qs.qid[i] = q;
  qs.r0[i] = r0;
  qs.ck[i] = 0;
  qs.out_off[i] = d.out_off;
  qs.in_off[i] = d.in_off;
  qs.meta[i] = d.in_cols | (d.out_cols << 8) | (expand ? 1u << 16 : 0u) | (d.final_step ? 1u << 17 : 0u) |
               (d.out_off == kNoSpace ? 1u << 18 : 0u);
  const DevSeg& s = a.segs[expand ? d.seg : 0];
  qs.voff[i] = s.voff;
  qs.vb_base[i] = s.vb_base;
  qs.p0[i] = pfx_at(a, r0 < R ? r0 : R);
}

// emit_row for query i of the tile's staged queries.
template <bool CK>
__device__ __forceinline__ uint64_t emit_row_staged(const StepArgs& a, const QStage& qs, uint32_t i, uint64_t g,
                                                    uint64_t j, uint64_t pg, uint32_t eo) {
  const uint32_t m = qs.meta[i];
  if (m & (1u << 18)) return 0;  // final table dropped: no room in the arena
  const uint32_t ic = m & 0xffu, oc = (m >> 8) & 0xffu;
  const bool expand = (m >> 16) & 1u, fin = (m >> 17) & 1u;
  const uint64_t orow = j - qs.p0[i] - a.win_lo;
  uint32_t* dst = (fin ? a.arena : a.tout) + qs.out_off[i] + orow * oc;
  const uint32_t* src = a.tin + qs.in_off[i] + (g - qs.r0[i]) * ic;
  const bool ck = CK && a.ck && fin;
  uint64_t h = row_hash_init(oc);
  for (uint32_t c = 0; c < ic; ++c) {
    const uint32_t x = src[c];
    dst[c] = x;
    if (ck) h = row_hash_step(h, x, c);
  }
  if (expand) {
    DevSeg sg;  // the fields edge_at reads
    sg.voff = qs.voff[i];
    sg.vb_base = qs.vb_base[i];
    const uint32_t x = edge_at(a, sg, uint64_t(eo) + (j - pg));
    dst[ic] = x;
    if (ck) h = row_hash_step(h, x, ic);
  }
  return ck ? h : 0;
}

// The rows of a tile that spans several queries: each output row finds its
// source row (in the staged rows, or bracketed by the staged samples and
// finished in the global prefix) and its query (in LDS when the tile's
// queries are staged there, else in the row starts).  One instantiation per
// kind of staging, so that every source of a row's values is read in its own
// code and only values meet.
template <bool CK, bool STAGED>
__device__ __forceinline__ void gather_mixed(const StepArgs& a, uint64_t j0, uint64_t total, uint64_t lo,
                                             uint64_t hi, uint64_t S, uint32_t ns, uint32_t qlo, uint32_t qhi,
                                             uint32_t nqs, const int32_t* s_rel, const uint32_t* s_eoff,
                                             QStage& qs) {
  const bool qlds = nqs > 0;
  // this thread's index, opaque to the compiler (fewer registers held)
  uint32_t tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
#pragma unroll
  for (int k = 0; k < kOutTile / kBlock; ++k) {
    const uint64_t j = j0 + uint64_t(k) * kBlock + tid;
    if (j >= total) continue;
    const int32_t jl = int32_t(j - j0);
    uint32_t l = 0, h = ns;
    while (h - l > 1) {
      const uint32_t m = (l + h) >> 1;
      if (s_rel[m] <= jl) l = m;
      else h = m;
    }
    // the source row g, its first output row pg and its edge offset eo
    uint64_t pg = uint64_t(int64_t(j0) + s_rel[l]);
    uint64_t g;
    uint32_t eo;
    if constexpr (STAGED) {
      g = lo + l;
      eo = s_eoff[l];
    } else {
      const uint64_t gl = lo + uint64_t(l) * S;
      g = refine_row(a, gl, gl + S < hi ? gl + S : hi, j, pg);
      eo = a.eoff[g];
    }
    if (qlds) {
      uint32_t ql = 0, qh = nqs;
      while (qh - ql > 1) {
        const uint32_t m = (ql + qh) >> 1;
        if (qs.r0[m] <= g) ql = m;
        else qh = m;
      }
      // checksum shares gather in LDS, one global add per query and tile
      const uint64_t rh = emit_row_staged<CK>(a, qs, ql, g, j, pg, eo);
      if (CK && rh) atomicAdd(&qs.ck[ql], (unsigned long long)rh);
    } else {
      const uint32_t q = query_in(a.row_start, qlo, qhi, g);
      const uint64_t rh = emit_row<CK>(a, q, g, j, pg, eo);
      if (CK) add_checksum(a, q, rh);
    }
  }
  if (CK && a.ck && qlds) {
    __syncthreads();
    if (threadIdx.x < nqs && qs.ck[threadIdx.x]) atomicAdd(&a.ck[qs.qid[threadIdx.x]], qs.ck[threadIdx.x]);
  }
}

// What the gather needs of a query that owns a whole output tile.
struct GatherQuery {
  uint32_t q;
  bool nospace, expand, ck;
  uint32_t ic, oc;
  uint64_t r0, p0;
  uint32_t* dbase;
  const uint32_t* sbase;
  int64_t voff;      // the expanded segment's edge addressing (DevSeg)
  uint32_t vb_base;

  __device__ __forceinline__ void load(const StepArgs& a, uint32_t qi) {
    const StepDesc& d = a.desc[qi];
    q = qi;
    nospace = d.out_off == kNoSpace;
    expand = d.op == OP_EXPAND;
    ck = a.ck && d.final_step;
    ic = d.in_cols;
    oc = d.out_cols;
    r0 = a.row_start[qi];
    p0 = pfx_at(a, r0);
    dbase = (d.final_step ? a.arena : a.tout) + d.out_off;
    sbase = a.tin + d.in_off;
    const DevSeg& s = a.segs[expand ? d.seg : 0];
    voff = s.voff;
    vb_base = s.vb_base;
  }
};

template <bool CK>
__global__ __launch_bounds__(kBlock) void k_gather(StepArgs a) {
  if (a.st->status) return;
  // the tile's source rows (or every S-th of them): their first output row
  // relative to the tile's first, pfx - j0, in (-2^28, kOutTile] for every
  // row holding one of the tile's output rows (a row has fewer than 2^28
  // edges), clamped above
  __shared__ int32_t s_rel[kStage];
  __shared__ uint32_t s_eoff[kStage];
  __shared__ QStage s_qs;  // the queries of a tile that spans several
  const uint64_t R = step_rows(a);
  const uint64_t total = out_end(a);
  const uint64_t nt = out_tiles(a);
  // the last query that owned a whole tile of this workgroup: its descriptor
  // fields, row start, first output row and segment, kept in scalar
  // registers while the next tiles belong to it too (a one-query step, or a
  // query larger than a tile)
  GatherQuery gq;
  gq.q = ~0u;
  for (uint64_t t = blockIdx.x; t < nt; t += gridDim.x) {
    const uint64_t j0 = a.win_lo + t * kOutTile;
    // this thread's index, opaque to the compiler, so that the values derived
    // from it are formed per tile and not held in registers across the loop
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    uint64_t lo, hi;
    if (t + 1 < a.otile_cap) {
      lo = a.tile_row[t];
      hi = (t + 1 < nt) ? a.tile_row[t + 1] + 1 : R;
    } else {
      // past the tile map (dropped final tables before this tile): the rows
      // holding this tile's first row and the next tile's
      uint64_t r1;
      block_find_rows<kBlock>(a, R, j0, j0 + kOutTile, lo, r1);
      hi = (t + 1 < nt) ? r1 + 1 : R;
    }
    const uint64_t nr = hi - lo;
    const bool staged = nr <= uint64_t(kStage);
    // staged: every row (S = 1); else every S-th row, ns <= kOutTile samples
    const uint64_t S = staged ? 1 : (nr + kOutTile - 1) / kOutTile;
    const uint32_t ns = uint32_t(staged ? nr : (nr + S - 1) / S);
    // unconditional loads at clamped indices (rows lo..hi-1 exist), so all
    // of them are in flight together; a tile of at most kBlock source rows
    // (a heavy expansion's) takes one load of each per thread
    if (staged && nr <= uint64_t(kBlock)) {
      const uint64_t gi = lo + (tid < uint32_t(nr) ? tid : uint32_t(nr) - 1);
      const uint64_t pv = pfx_at(a, gi);
      const uint32_t ev = a.eoff[gi];
      if (tid < uint32_t(nr)) {
        const int64_t r = int64_t(pv - j0);
        s_rel[tid] = int32_t(r < (int64_t(1) << 30) ? r : (int64_t(1) << 30));
        s_eoff[tid] = ev;
      }
    } else if (staged) {
#pragma unroll
      for (int k = 0; k < (kStage + kBlock - 1) / kBlock; ++k) {
        const uint32_t i = tid + k * kBlock;
        const uint64_t gi = lo + (i < uint32_t(nr) ? i : uint32_t(nr) - 1);
        const uint64_t pv = pfx_at(a, gi);
        const uint32_t ev = a.eoff[gi];
        if (i < uint32_t(nr)) {
          const int64_t r = int64_t(pv - j0);
          s_rel[i] = int32_t(r < (int64_t(1) << 30) ? r : (int64_t(1) << 30));
          s_eoff[i] = ev;
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < (kOutTile + kBlock - 1) / kBlock; ++k) {
        const uint32_t i = tid + k * kBlock;
        const uint64_t pv = pfx_at(a, lo + uint64_t(i < ns ? i : ns - 1) * S);
        if (i < ns) {
          const int64_t r = int64_t(pv - j0);
          s_rel[i] = int32_t(r < (int64_t(1) << 30) ? r : (int64_t(1) << 30));
        }
      }
    }
    // the queries of rows lo and hi-1 bound every row's query in this tile
    uint32_t qlo, qlast;
    block_find_queries<kBlock>(a.row_start, a.nq, lo, hi - 1, qlo, qlast);
    const uint32_t qhi = qlast + 1;
    // the queries of a tile that spans several, staged in LDS when they fit:
    // all of them, or those with rows when more lie between (queries that
    // skip this step have none; not in the checksum variant, whose registers
    // it would cost a wave).  The one query of a tile whose source rows are
    // too many for LDS is staged too, for the sampled search.
    uint32_t nqs = 0;
    if (qhi - qlo > 1 || !staged) {
      const bool all = CK || qhi - qlo <= uint32_t(kQStage);  // then one pass (kQStage = kBlock)
      for (uint32_t q0 = qlo; q0 < qhi && nqs <= uint32_t(kQStage); q0 += kBlock) {
        const uint32_t q = q0 + tid;
        bool has;
        uint32_t at, cnt;
        if (all) {
          has = q < qhi;
          at = tid;
          cnt = qhi - qlo;
        } else {
          has = q < qhi && a.row_start[q + 1] > a.row_start[q];
          at = nqs + block_flag_rank(has, cnt);
        }
        if (has && at < uint32_t(kQStage)) stage_query(a, s_qs, at, q, R);
        nqs += cnt;
      }
      if (nqs > uint32_t(kQStage)) nqs = 0;  // too many: rows find their queries in memory
    }
    __syncthreads();
    if (staged && qhi - qlo == 1) {
      // one query owns the whole tile: its descriptor, row start and segment
      // come from scalar registers, loaded when the workgroup's query changes
      const uint32_t q = __builtin_amdgcn_readfirstlane(qlo);
      if (q != gq.q) gq.load(a, q);
    }
    if (staged && qhi - qlo == 1 && gq.nospace) {
      // the tile's one query lost its final table (no room in the arena)
    } else if (staged && qhi - qlo == 1) {
      const uint32_t q = gq.q;
      const uint64_t r0 = gq.r0, p0 = gq.p0;
      const uint32_t ic = gq.ic, oc = gq.oc;
      const bool expand = gq.expand;
      uint32_t* const dbase = gq.dbase;
      const uint32_t* const sbase = gq.sbase;
      DevSeg sg;  // the fields edge_at reads
      sg.voff = gq.voff;
      sg.vb_base = gq.vb_base;
      const bool ck = CK && gq.ck;
      uint64_t hsum = 0;  // row hashes of this thread's final rows
      constexpr int kR = kOutTile / kBlock;
      // each output row's source row in the tile, by independent searches
      uint32_t ls[kR];
      bool act[kR];
#pragma unroll
      for (int k = 0; k < kR; ++k) {
        const int32_t jl = k * kBlock + tid;
        act[k] = j0 + uint64_t(jl) < total;
        uint32_t l = 0, h = uint32_t(nr);
        while (h - l > 1) {
          const uint32_t m = (l + h) >> 1;
          if (s_rel[m] <= jl) l = m;
          else h = m;
        }
        ls[k] = l;
      }
      if (ic <= kRegCols) {
        // every row's words and edge are loaded before anything is stored:
        // a store may alias a later load, so loads after it would each wait
        // for the one before
        uint32_t v[kR][kRegCols];
        uint32_t e[kR];
#pragma unroll
        for (int k = 0; k < kR; ++k) {
          const uint32_t* src = sbase + (lo + ls[k] - r0) * ic;  // an input row of the tile
#pragma unroll
          for (int c = 0; c < kRegCols; ++c) v[k][c] = uint32_t(c) < ic ? src[c] : 0u;
          const int32_t jl = k * kBlock + tid;
          e[k] = (expand && act[k]) ? edge_at(a, sg, uint64_t(s_eoff[ls[k]]) + uint64_t(int64_t(jl) - s_rel[ls[k]]))
                                    : 0u;
        }
        // two-column rows at an 8-byte aligned table go out as one 8-byte
        // store each (rows are 8 bytes apart, so the table's alignment decides;
        // not in the checksum variant, whose registers it would cost a wave)
        const bool pair = !CK && oc == 2 && ic + (expand ? 1u : 0u) == 2 && (reinterpret_cast<uintptr_t>(dbase) & 7) == 0;
#pragma unroll
        for (int k = 0; k < kR; ++k) {
          if (!act[k]) continue;
          const uint32_t jl = k * kBlock + tid;
          uint32_t* dst = dbase + (j0 + jl - p0 - a.win_lo) * oc;
          uint64_t rh = row_hash_init(oc);
          if (pair) {
            const uint32_t w1 = expand ? e[k] : v[k][1];
            *reinterpret_cast<uint2*>(dst) = make_uint2(v[k][0], w1);
            if (ck) rh = row_hash_step(row_hash_step(rh, v[k][0], 0), w1, 1);
          } else {
#pragma unroll
            for (int c = 0; c < kRegCols; ++c) {
              if (uint32_t(c) < ic) {
                dst[c] = v[k][c];
                if (ck) rh = row_hash_step(rh, v[k][c], c);
              }
            }
            if (expand) {
              dst[ic] = e[k];
              if (ck) rh = row_hash_step(rh, e[k], ic);
            }
          }
          if (ck) hsum += rh;
        }
      } else {
#pragma unroll
        for (int k = 0; k < kR; ++k) {
          if (!act[k]) continue;
          const uint64_t j = j0 + uint64_t(k) * kBlock + tid;
          const uint32_t l = ls[k];
          const uint64_t g = lo + l;
          uint32_t* dst = dbase + (j - p0 - a.win_lo) * oc;
          const uint32_t* src = sbase + (g - r0) * ic;
          uint64_t rh = row_hash_init(oc);
          for (uint32_t c = 0; c < ic; ++c) {
            const uint32_t x = src[c];
            dst[c] = x;
            if (ck) rh = row_hash_step(rh, x, c);
          }
          if (expand) {
            const uint32_t x = edge_at(a, sg, uint64_t(s_eoff[l]) + uint64_t(int64_t(j - j0) - s_rel[l]));
            dst[ic] = x;
            if (ck) rh = row_hash_step(rh, x, ic);
          }
          if (ck) hsum += rh;
        }
      }
      if (ck) {
        hsum = block_sum<kBlock>(hsum);
        if (tid == 0 && hsum) atomicAdd(&a.ck[q], (unsigned long long)hsum);
      }
    } else if (staged) {
      gather_mixed<CK, true>(a, j0, total, lo, hi, 1, ns, qlo, qhi, nqs, s_rel, s_eoff, s_qs);
    } else {
      gather_mixed<CK, false>(a, j0, total, lo, hi, S, ns, qlo, qhi, nqs, s_rel, s_eoff, s_qs);
    }
    __syncthreads();
  }
}

// Row-centric: LPR lanes share one input row (64/LPR rows per wavefront).
template <int LPR>
__global__ __launch_bounds__(kBlock) void k_gather_rows(StepArgs a) {
  if (a.st->status) return;
  const uint64_t R = step_rows(a);
  constexpr int G = kBlock / LPR;
  const int grp = threadIdx.x / LPR, gl = threadIdx.x % LPR;
  for (uint64_t g = uint64_t(blockIdx.x) * G + grp; g < R; g += uint64_t(gridDim.x) * G) {
    const uint64_t p = pfx_at(a, g);
    const uint64_t c = pfx_at(a, g + 1) - p;
    if (!c) continue;
    const uint32_t q = find_query(a.row_start, a.nq, g);
    uint64_t hsum = 0;
    const uint64_t e = out_end(a);
    for (uint32_t k = gl; k < c; k += LPR)
      if (p + k >= a.win_lo && p + k < e) hsum += emit_row<true>(a, q, g, p + k);
    if (a.ck) add_checksum(a, q, hsum);
  }
}

// ---- fork-join dispatch: rows grouped by owner = row[col] % nparts --------
//
// The split of a history table among its next-hop owners before a remote
// step (CPU: SPARQLEngine::generate_sub_query, sparql.hpp:741-793, rows
// appended in order; GPU: gpu_shuffle_result_buf / gpu_split_result_buf,
// gpu_hash.cu:601-762, positions by atomics).  This keeps the CPU order: a
// row's place is its owner's base, plus the owner's rows in earlier tiles,
// plus its rank among the owner's rows earlier in its own tile.

__device__ __forceinline__ uint32_t owner_of(OwnerFn f, uint32_t x) { return f.mod ? x % f.mod : (x >> f.shift) & f.mask; }

__device__ __forceinline__ uint32_t row_owner(const uint32_t* in, uint32_t cols, uint32_t col, OwnerFn f, uint64_t r) {
  return owner_of(f, in[r * cols + col]);
}

// Per-tile owner counts, owner-major: hist[p * ntiles + t].
__global__ __launch_bounds__(kBlock) void k_part_hist(const uint32_t* in, uint64_t rows, uint32_t cols,
                                                      uint32_t col, OwnerFn f, uint64_t* hist) {
  const uint32_t nparts = owner_count(f);
  __shared__ uint32_t s_h[kMaxParts];
  const uint64_t ntiles = (rows + kTile - 1) / kTile;
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    for (uint32_t p = threadIdx.x; p < nparts; p += kBlock) s_h[p] = 0;
    __syncthreads();
    for (uint32_t k = 0; k < kTile / kBlock; ++k) {
      const uint64_t r = t * kTile + uint64_t(k) * kBlock + threadIdx.x;
      if (r < rows) atomicAdd(&s_h[row_owner(in, cols, col, f, r)], 1u);
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < nparts; p += kBlock) hist[uint64_t(p) * ntiles + t] = s_h[p];
    __syncthreads();
  }
}

// Local exclusive scan of the histogram in chunks of kPartChunk entries
// (8 consecutive entries per thread); chunk totals go to csum[chunk].
constexpr uint32_t kPartChunk = kBlock * 8;

__global__ __launch_bounds__(kBlock) void k_part_scan_local(uint64_t* v, uint64_t n, uint64_t* csum) {
  const uint64_t nch = (n + kPartChunk - 1) / kPartChunk;
  for (uint64_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const uint64_t lo = c * kPartChunk + uint64_t(threadIdx.x) * 8;
    uint64_t x[8], s = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      x[k] = lo + k < n ? v[lo + k] : 0;
      s += x[k];
    }
    uint64_t total = 0;
    uint64_t ex = block_excl_scan<kBlock>(s, &total);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (lo + k < n) v[lo + k] = ex;
      ex += x[k];
    }
    if (threadIdx.x == 0) csum[c] = total;
  }
}

// Exclusive scan of the chunk totals by one 1,024-thread block, then each
// owner's first output row: heads[p] = global position of (owner p, tile 0);
// heads[nparts] = the row count.
__global__ __launch_bounds__(1024) void k_part_scan_chunks(const uint64_t* hist, uint64_t* csum, uint64_t nch,
                                                           uint64_t ntiles, uint32_t nparts, uint64_t* heads) {
  const uint64_t per = (nch + 1023) / 1024;
  uint64_t lo = uint64_t(threadIdx.x) * per;
  if (lo > nch) lo = nch;
  const uint64_t hi = (lo + per < nch) ? lo + per : nch;
  uint64_t s = 0;
  for (uint64_t i = lo; i < hi; ++i) s += csum[i];
  uint64_t total = 0;
  uint64_t ex = block_excl_scan<1024>(s, &total);
  for (uint64_t i = lo; i < hi; ++i) {
    const uint64_t x = csum[i];
    csum[i] = ex;
    ex += x;
  }
  __syncthreads();  // the block's csum writes are visible to the block
  if (threadIdx.x < nparts) {
    const uint64_t i = uint64_t(threadIdx.x) * ntiles;
    heads[threadIdx.x] = hist[i] + csum[i / kPartChunk];
  }
  if (threadIdx.x == 0) heads[nparts] = total;
}

// Scatter: wave w of a block takes rows [w * 256, w * 256 + 256) of the
// tile, 64 per pass.  A row's rank among the wave's rows of its owner in one
// pass comes from one ballot per owner bit (the lanes whose owner equals
// this lane's); wave-private counters in LDS carry it across passes, and one
// barrier per tile adds the rows of the owner in earlier waves.  So a row's
// place is the owner's base, plus the owner's rows in earlier tiles, plus
// its rank among the tile's rows of its owner: the input order.
//
// MAXC > 0 (tables of at most MAXC columns): a lane loads all its rows of
// the tile into registers at once, takes the owners from them, and writes
// each row to its place.  MAXC = 0 (wider tables): a lane reads only the
// owner words; the rows' places go to LDS, and the block then copies the
// tile word by word, consecutive lanes on consecutive words.
template <int MAXC>
__global__ __launch_bounds__(kBlock) void k_part_scatter(const uint32_t* __restrict__ in, uint64_t rows,
                                                         uint32_t cols, uint32_t col, OwnerFn f,
                                                         const uint64_t* __restrict__ hist,
                                                         const uint64_t* __restrict__ csum,
                                                         uint32_t* __restrict__ out) {
  constexpr int kWaves = kBlock / kWave;
  constexpr int kPass = kTile / kBlock;         // passes of 64 rows per wave
  constexpr int RC = MAXC ? MAXC : 1;           // row words held in registers
  __shared__ uint32_t s_wcnt[kWaves][kMaxParts];  // per wave: its rows of owner p so far
  __shared__ uint64_t s_base[kMaxParts];          // output row of the tile's first row of owner p
  __shared__ uint64_t s_dst[MAXC ? 1 : kTile];    // wide tables: output row of each tile row
  const uint64_t ntiles = (rows + kTile - 1) / kTile;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const uint64_t lt = (uint64_t(1) << lane) - 1;  // lanes below this one
  const uint32_t nparts = owner_count(f);
  const int bits = nparts > 1 ? 32 - __clz(int(nparts - 1)) : 0;
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    for (uint32_t p = lane; p < nparts; p += kWave) s_wcnt[w][p] = 0;
    for (uint32_t p = threadIdx.x; p < nparts; p += kBlock) {
      const uint64_t i = uint64_t(p) * ntiles + t;
      s_base[p] = hist[i] + csum[i / kPartChunk];
    }
    const uint32_t l0 = uint32_t(w) * (kPass * kWave) + lane;  // this lane's first row in the tile
    const uint64_t r0 = t * kTile + l0;
    uint32_t v[kPass][RC], own[kPass], rk[kPass];
#pragma unroll
    for (int k = 0; k < kPass; ++k) {
      const uint64_t r = r0 + uint64_t(k) * kWave;
      if (MAXC) {
#pragma unroll
        for (int c = 0; c < RC; ++c) v[k][c] = (r < rows && c < int(cols)) ? in[r * cols + c] : 0;
      } else {
        v[k][0] = r < rows ? in[r * cols + col] : 0;
      }
    }
#pragma unroll
    for (int k = 0; k < kPass; ++k) {
      uint32_t key = v[k][0];
#pragma unroll
      for (int c = 1; c < RC; ++c)
        if (c == int(col)) key = v[k][c];
      own[k] = owner_of(f, key);
    }
#pragma unroll
    for (int k = 0; k < kPass; ++k) {
      const bool valid = r0 + uint64_t(k) * kWave < rows;
      const uint32_t my = own[k];
      uint64_t m = __ballot(valid);
      for (int b = 0; b < bits; ++b) {
        const bool on = (my >> b) & 1;
        const uint64_t bb = __ballot(on);
        m &= on ? bb : ~bb;
      }
      const uint32_t rank = __popcll(m & lt);
      rk[k] = valid ? s_wcnt[w][my] + rank : 0;
      __builtin_amdgcn_wave_barrier();  // every lane has read before the counts move on
      if (valid && rank == 0) s_wcnt[w][my] = rk[k] + __popcll(m);  // the owner's first lane
      __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();  // every wave's owner counts of this tile are in s_wcnt
#pragma unroll
    for (int k = 0; k < kPass; ++k) {
      const uint64_t r = r0 + uint64_t(k) * kWave;
      if (r < rows) {
        for (int v2 = 0; v2 < w; ++v2) rk[k] += s_wcnt[v2][own[k]];
        const uint64_t d = s_base[own[k]] + rk[k];
        if (MAXC) {
#pragma unroll
          for (int c = 0; c < RC; ++c)
            if (c < int(cols)) out[d * cols + c] = v[k][c];
        } else {
          s_dst[l0 + k * kWave] = d;
        }
      }
    }
    if (!MAXC) {
      __syncthreads();
      const uint64_t tr = t * kTile;
      const uint32_t nrow = uint32_t(rows - tr < uint64_t(kTile) ? rows - tr : uint64_t(kTile));
      const uint32_t nword = nrow * cols;
      const uint32_t* src = in + tr * cols;
      for (uint32_t e0 = threadIdx.x; e0 < nword; e0 += 4 * kBlock) {
        uint32_t x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t e = e0 + u * kBlock;
          x[u] = e < nword ? src[e] : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t e = e0 + u * kBlock;
          if (e < nword) {
            const uint32_t j = e / cols;
            out[s_dst[j] * cols + (e - j * cols)] = x[u];
          }
        }
      }
    }
    __syncthreads();  // s_wcnt, s_base and s_dst are free for the next tile
  }
}

// Per-table checksums (sum of row hashes) of non-empty tables whose rows are
// numbered consecutively; a block's 256 rows span at most 256 tables.
__global__ __launch_bounds__(kBlock) void k_checksum(const uint32_t* base, const CkDesc* d, uint32_t n,
                                                     uint64_t total, unsigned long long* sums) {
  __shared__ unsigned long long acc[kBlock];
  for (uint64_t t0 = uint64_t(blockIdx.x) * kBlock; t0 < total; t0 += uint64_t(gridDim.x) * kBlock) {
    acc[threadIdx.x] = 0;
    __syncthreads();
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (d[mid].row0 <= t0) lo = mid;
      else hi = mid;
    }
    const uint32_t qa = lo;
    const uint64_t r = t0 + threadIdx.x;
    if (r < total) {
      uint32_t a = qa, b = n;
      while (b - a > 1) {
        const uint32_t mid = (a + b) >> 1;
        if (d[mid].row0 <= r) a = mid;
        else b = mid;
      }
      const CkDesc& c = d[a];
      const uint64_t h = row_hash(base + c.off + (r - c.row0) * c.cols, c.cols);
      atomicAdd(&acc[a - qa], (unsigned long long)h);
    }
    __syncthreads();
    if (qa + threadIdx.x < n && acc[threadIdx.x]) atomicAdd(&sums[qa + threadIdx.x], acc[threadIdx.x]);
    __syncthreads();
  }
}

__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t v) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// One thread per bucket of the segment: every key in slots 0..6 writes its
// ptr word at its vid's place in the direct index.
__global__ __launch_bounds__(kBlock) void k_dense_build(const Slot* karena, const uint32_t* ktab, uint32_t kshift,
                                                        DevSeg s, uint64_t* dense) {
  const uint64_t mask = (uint64_t(1) << kshift) - 1;
  for (uint64_t b = uint64_t(blockIdx.x) * kBlock + threadIdx.x; b < s.total_buckets;
       b += uint64_t(gridDim.x) * kBlock) {
    const uint64_t i = s.koff > kBlockTable ? uint64_t(s.koff) + b
                                            : (uint64_t(ktab[s.kb_base + (b >> kshift)]) << kshift) | (b & mask);
    const Slot* bk = karena + i * ASSOCIATIVITY;
    for (int k = 0; k < ASSOCIATIVITY - 1; ++k) {
      const Slot sl = bk[k];
      if (sl.key == 0) continue;
      const uint64_t v = key_vid(sl.key) - s.dvmin;
      if (v < s.dvn) dense[v] = sl.ptr;
    }
  }
}

// Prefix of row g from the tile prefixes and the prefixes within tiles.
__device__ __forceinline__ uint64_t prefix_of(const uint64_t* tile_sum, const uint64_t* lpfx, uint64_t g) {
  return tile_sum[g >> kTileShift] + lpfx[g];
}

// Largest g < R whose prefix is <= j (row 0's is 0 <= j).
__device__ __forceinline__ uint64_t row_of_output(const uint64_t* tile_sum, const uint64_t* lpfx, uint64_t R,
                                                  uint64_t j) {
  uint64_t lo = 0, hi = R;
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    if (prefix_of(tile_sum, lpfx, mid) <= j) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kBlock) void k_window_rows(const uint64_t* tile_sum, const uint64_t* lpfx, uint64_t R,
                                                        uint64_t j0, uint64_t C, uint64_t T, uint32_t n,
                                                        uint64_t* out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint64_t j = j0 + uint64_t(i) * C;
  const uint64_t je = (j + C < T ? j + C : T) - 1;
  const uint64_t ra = row_of_output(tile_sum, lpfx, R, j);
  out[3 * uint64_t(i)] = ra;
  out[3 * uint64_t(i) + 1] = row_of_output(tile_sum, lpfx, R, je);
  out[3 * uint64_t(i) + 2] = prefix_of(tile_sum, lpfx, ra);
}

inline unsigned grid_for(uint64_t units, unsigned cap) {
  if (units == 0) return 1;
  return unsigned(units < cap ? units : cap);
}

// Grids of the launches whose row counts live on the device (batched steps
// and serving iterations): the workgroups resident on the whole device at
// once, whose workgroups then stride over the tiles.  A larger grid only adds
// rounds of workgroups that find no tile: the batch probe's 1,024-thread
// workgroup holds 120 KB of LDS, one per CU, so 2,048 of them took 8 rounds.
struct LaunchGrids {
  unsigned probe[4];       // k_probe<W, 512>, W = 1, 2, 4, 8
  unsigned probe1k[4];     // k_probe<W, 1024>, steps of one row per query
  unsigned probe1q[4];     // k_probe<W, 256>, the one-query probe
  unsigned tile_rows;
  unsigned gather[2];      // k_gather<false>, k_gather<true>
  unsigned gather_rows[4]; // k_gather_rows<1, 4, 16, 64>
  unsigned query_sums;
};
LaunchGrids g_grids = {{2 * kPersistentGrid, 2 * kPersistentGrid, 2 * kPersistentGrid, 2 * kPersistentGrid},
                       {kPersistentGrid, kPersistentGrid, kPersistentGrid, kPersistentGrid},
                       {4 * kPersistentGrid, 4 * kPersistentGrid, 4 * kPersistentGrid, 4 * kPersistentGrid},
                       kPersistentGrid,
                       {4 * kPersistentGrid, 4 * kPersistentGrid},
                       {2 * kPersistentGrid, 2 * kPersistentGrid, 2 * kPersistentGrid, 2 * kPersistentGrid},
                       kPersistentGrid};
std::once_flag g_grids_once;

unsigned resident_grid(const void* kernel, int block, int cus, unsigned fallback) {
  int per = 0;
  if (cus <= 0 || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, 0) != hipSuccess || per <= 0) {
    (void)hipGetLastError();
    return fallback;
  }
  return unsigned(cus) * unsigned(per);
}

}  // namespace

void init_launch_grids() {
  std::call_once(g_grids_once, [] {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
      (void)hipGetLastError();
      return;
    }
    auto rg = [cus](const void* k, int block, unsigned fallback) { return resident_grid(k, block, cus, fallback); };
    LaunchGrids g = g_grids;
    g.probe[0] = rg(reinterpret_cast<const void*>(k_probe<1, 512>), 512, g.probe[0]);
    g.probe1k[0] = rg(reinterpret_cast<const void*>(k_probe<1, 1024>), 1024, g.probe1k[0]);
    g.probe[1] = rg(reinterpret_cast<const void*>(k_probe<2, 512>), 512, g.probe[1]);
    g.probe1k[1] = rg(reinterpret_cast<const void*>(k_probe<2, 1024>), 1024, g.probe1k[1]);
    g.probe[2] = rg(reinterpret_cast<const void*>(k_probe<4, 512>), 512, g.probe[2]);
    g.probe1k[2] = rg(reinterpret_cast<const void*>(k_probe<4, 1024>), 1024, g.probe1k[2]);
    g.probe[3] = rg(reinterpret_cast<const void*>(k_probe<8, 512>), 512, g.probe[3]);
    g.probe1k[3] = rg(reinterpret_cast<const void*>(k_probe<8, 1024>), 1024, g.probe1k[3]);
    g.probe1q[0] = rg(reinterpret_cast<const void*>(k_probe<1, 256>), 256, g.probe1q[0]);
    g.probe1q[1] = rg(reinterpret_cast<const void*>(k_probe<2, 256>), 256, g.probe1q[1]);
    g.probe1q[2] = rg(reinterpret_cast<const void*>(k_probe<4, 256>), 256, g.probe1q[2]);
    g.probe1q[3] = rg(reinterpret_cast<const void*>(k_probe<8, 256>), 256, g.probe1q[3]);
    g.tile_rows = rg(reinterpret_cast<const void*>(k_tile_rows), kBlock, g.tile_rows);
    g.gather[0] = rg(reinterpret_cast<const void*>(k_gather<false>), kBlock, g.gather[0]);
    g.gather[1] = rg(reinterpret_cast<const void*>(k_gather<true>), kBlock, g.gather[1]);
    g.gather_rows[0] = rg(reinterpret_cast<const void*>(k_gather_rows<1>), kBlock, g.gather_rows[0]);
    g.gather_rows[1] = rg(reinterpret_cast<const void*>(k_gather_rows<4>), kBlock, g.gather_rows[1]);
    g.gather_rows[2] = rg(reinterpret_cast<const void*>(k_gather_rows<16>), kBlock, g.gather_rows[2]);
    g.gather_rows[3] = rg(reinterpret_cast<const void*>(k_gather_rows<64>), kBlock, g.gather_rows[3]);
    g.query_sums = rg(reinterpret_cast<const void*>(k_query_sums), kBlock, g.query_sums);
    g_grids = g;
  });
}

// Workgroup size of a probe launch: 256 threads for one query; for several,
// 512 (two rows per lane at W = 1, two workgroups per CU), or 1,024 for a
// step of one row per query, whose tiles each span a tile's worth of queries.
template <int W>
void probe_launch(const StepArgs& a, unsigned grid, int nt, hipStream_t s) {
  if (nt == 256)
    WK_LAUNCH((k_probe<W, 256>), dim3(grid), dim3(256), 0, s, a);
  else if (nt == 1024)
    WK_LAUNCH((k_probe<W, 1024>), dim3(grid), dim3(1024), 0, s, a);
  else
    WK_LAUNCH((k_probe<W, 512>), dim3(grid), dim3(512), 0, s, a);
}

void launch_probe(const StepArgs& a, int w, uint64_t rows_hint, hipStream_t s, bool row_per_query) {
  const int wi = w == 8 ? 3 : w == 4 ? 2 : w == 2 ? 1 : 0;
  const int nt = a.nq == 1 ? 256 : row_per_query ? 1024 : 512;
  // the workgroups resident at once stride over the tiles (fewer when the
  // host knows the step has fewer tiles)
  const unsigned resident = nt == 256 ? g_grids.probe1q[wi] : nt == 1024 ? g_grids.probe1k[wi] : g_grids.probe[wi];
  const unsigned grid = rows_hint ? grid_for((rows_hint + kTile - 1) / kTile, resident) : resident;
  // w lanes share one key: each reads 8/w slots of the 128-B bucket, so a
  // wave's loads of one slot index cover 64/w buckets, each line coalesced
  switch (w) {
    case 8: probe_launch<8>(a, grid, nt, s); break;
    case 4: probe_launch<4>(a, grid, nt, s); break;
    case 2: probe_launch<2>(a, grid, nt, s); break;
    default: probe_launch<1>(a, grid, nt, s); break;
  }
}

void launch_scan_tiles(const StepArgs& a, hipStream_t s) {
  WK_LAUNCH(k_scan_tiles, dim3(1), dim3(1024), 0, s, a);
}


void launch_scan_queries(const StepArgs& a, hipStream_t s) {
  WK_LAUNCH(k_scan_queries, dim3(1), dim3(1024), 0, s, a);
}

void launch_scan_small(const StepArgs& a, hipStream_t s) {
  WK_LAUNCH(k_scan_small, dim3(1), dim3(1024), 0, s, a);
}

void launch_scan_queries_mb(const StepArgs& a, hipStream_t s) {
  const uint32_t nqb = (a.nq + kBlock - 1) / kBlock;
  // blocks past nqb do k_tile_rows' work
  const unsigned extra = g_grids.query_sums / 2 > 64 ? g_grids.query_sums / 2 : 64;
  WK_LAUNCH(k_query_sums, dim3(nqb + extra), dim3(kBlock), 0, s, a, nqb);
  WK_LAUNCH(k_scan_queries_mb, dim3(nqb), dim3(kBlock), 0, s, a, nqb);
}

void launch_expand_tail(const StepArgs& a, int rows_per_wave, hipStream_t s) {
  if (a.nq > uint32_t(kBlock) && a.qagg) {
    launch_scan_tiles(a, s);
    launch_scan_queries_mb(a, s);  // output tiles are marked by its first launch
    launch_gather_only(a, rows_per_wave, 0, s);
  } else {
    launch_scan_small(a, s);
    launch_gather(a, rows_per_wave, 0, s);
  }
}

void launch_scan(const StepArgs& a, hipStream_t s) { launch_scan_small(a, s); }

void launch_tile_rows(const StepArgs& a, uint64_t out_hint, hipStream_t s) {
  const uint64_t tiles = (out_hint + kOutTile - 1) / kOutTile;
  constexpr unsigned kWaves = kBlock / kWave;  // one output tile per wavefront
  const unsigned grid = out_hint ? grid_for((tiles + kWaves - 1) / kWaves, g_grids.tile_rows) : g_grids.tile_rows;
  WK_LAUNCH(k_tile_rows, dim3(grid), dim3(kBlock), 0, s, a);
}

void launch_gather_only(const StepArgs& a, int rows_per_wave, uint64_t out_hint, hipStream_t s) {
  if (rows_per_wave <= 0) {
    // the workgroups resident at once stride over the tiles (fewer when the
    // host knows the step has fewer tiles)
    const unsigned resident = g_grids.gather[a.ck ? 1 : 0];
    const unsigned grid = out_hint ? grid_for((out_hint + kOutTile - 1) / kOutTile, resident) : resident;
    // the checksum variant only when final tables are checksummed (serving)
    if (a.ck)
      WK_LAUNCH(k_gather<true>, dim3(grid), dim3(kBlock), 0, s, a);
    else
      WK_LAUNCH(k_gather<false>, dim3(grid), dim3(kBlock), 0, s, a);
    return;
  }
  // rows per wavefront -> lanes per row
  switch (rows_per_wave) {
    case 64: WK_LAUNCH(k_gather_rows<1>, dim3(g_grids.gather_rows[0]), dim3(kBlock), 0, s, a); break;
    case 16: WK_LAUNCH(k_gather_rows<4>, dim3(g_grids.gather_rows[1]), dim3(kBlock), 0, s, a); break;
    case 4: WK_LAUNCH(k_gather_rows<16>, dim3(g_grids.gather_rows[2]), dim3(kBlock), 0, s, a); break;
    default: WK_LAUNCH(k_gather_rows<64>, dim3(g_grids.gather_rows[3]), dim3(kBlock), 0, s, a); break;
  }
}

void launch_gather(const StepArgs& a, int rows_per_wave, uint64_t out_hint, hipStream_t s) {
  if (rows_per_wave <= 0) launch_tile_rows(a, out_hint, s);
  launch_gather_only(a, rows_per_wave, out_hint, s);
}

void launch_checksum(const uint32_t* base, const CkDesc* d, uint32_t n, uint64_t total_rows, uint64_t* sums,
                     hipStream_t s) {
  if (n == 0 || total_rows == 0) return;
  WK_LAUNCH(k_checksum, dim3(grid_for((total_rows + kBlock - 1) / kBlock, 8192)), dim3(kBlock), 0, s,
                     base, d, n, total_rows, reinterpret_cast<unsigned long long*>(sums));
}


uint64_t partition_scratch_words(uint64_t rows, uint32_t nparts) {
  const uint64_t n = uint64_t(nparts) * ((rows + kTile - 1) / kTile);
  return n + (n + kPartChunk - 1) / kPartChunk + nparts + 1;
}

uint64_t* launch_partition(const uint32_t* in, uint64_t rows, uint32_t cols, uint32_t col, OwnerFn f,
                           uint64_t* scratch, uint32_t* out, hipStream_t s) {
  const uint32_t nparts = owner_count(f);
  const uint64_t ntiles = (rows + kTile - 1) / kTile;
  const uint64_t n = uint64_t(nparts) * ntiles, nch = (n + kPartChunk - 1) / kPartChunk;
  uint64_t* hist = scratch;
  uint64_t* csum = hist + n;
  uint64_t* heads = csum + nch;
  const unsigned grid = grid_for(ntiles, 8192);
  WK_LAUNCH(k_part_hist, dim3(grid), dim3(kBlock), 0, s, in, rows, cols, col, f, hist);
  WK_LAUNCH(k_part_scan_local, dim3(grid_for(nch, 8192)), dim3(kBlock), 0, s, hist, n, csum);
  WK_LAUNCH(k_part_scan_chunks, dim3(1), dim3(1024), 0, s, (const uint64_t*)hist, csum, nch, ntiles,
                     nparts, heads);
  const uint64_t* h = hist;
  const uint64_t* c = csum;
  if (cols <= 2)
    WK_LAUNCH(k_part_scatter<2>, dim3(grid), dim3(kBlock), 0, s, in, rows, cols, col, f, h, c, out);
  else if (cols <= 4)
    WK_LAUNCH(k_part_scatter<4>, dim3(grid), dim3(kBlock), 0, s, in, rows, cols, col, f, h, c, out);
  else if (cols <= 8)
    WK_LAUNCH(k_part_scatter<8>, dim3(grid), dim3(kBlock), 0, s, in, rows, cols, col, f, h, c, out);
  else
    WK_LAUNCH(k_part_scatter<0>, dim3(grid), dim3(kBlock), 0, s, in, rows, cols, col, f, h, c, out);
  return heads;
}

// ------------------------------------------------------------ final_process
// SPARQLEngine::final_process (core/engine/sparql.hpp:1419-1546) on the
// device.  Rows are sorted as a permutation: (key, row) pairs go through a
// stable LSD radix sort, one 32-bit key column at a time, 6 bits per stable
// split.  The pairs' second word is the row index into the input table.

__device__ __forceinline__ uint32_t sort_key(uint32_t x, uint32_t mode) {
  return mode == kKeySigned ? x ^ 0x80000000u : mode == kKeyDesc ? ~x : x;
}

// pairs[i] = (key of row idx_i at column c, idx_i), idx_i = prev ? prev[2i + 1]
// : i.  prev may equal pairs: thread i reads and writes only pair i.
__global__ __launch_bounds__(kBlock) void k_fp_keys(const uint32_t* table, uint32_t cols, uint32_t c, uint32_t mode,
                                                    const uint32_t* prev, uint64_t n, uint32_t* pairs) {
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
    const uint32_t r = prev ? prev[2 * i + 1] : uint32_t(i);
    pairs[2 * i] = sort_key(table[uint64_t(r) * cols + c], mode);
    pairs[2 * i + 1] = r;
  }
}

// DISTINCT over rows in sorted order: flag 0 keeps a row (the first, or one
// that differs from its predecessor on a required column), 1 drops it.
// Equality on the required columns is transitive, so comparing with the
// predecessor is comparing with the last kept row, as the reference does.
__global__ __launch_bounds__(kBlock) void k_fp_dedup(const uint32_t* table, uint32_t cols, ColList req,
                                                     const uint32_t* sorted, uint64_t n, uint32_t* pairs) {
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
    const uint32_t r = sorted[2 * i + 1];
    uint32_t drop = 0;
    if (i > 0) {
      const uint32_t q = sorted[2 * i - 1];
      drop = 1;
      for (uint32_t j = 0; j < req.n; ++j)
        if (table[uint64_t(r) * cols + req.col[j]] != table[uint64_t(q) * cols + req.col[j]]) drop = 0;
    }
    pairs[2 * i] = drop;
    pairs[2 * i + 1] = r;
  }
}

// out[i][j] = table[idx_(first + i)][req.col[j]], consecutive threads on
// consecutive output words; idx_k = sorted ? sorted[2k + 1] : k.
__global__ __launch_bounds__(kBlock) void k_fp_project(const uint32_t* table, uint32_t cols, ColList req,
                                                       const uint32_t* sorted, uint64_t first, uint64_t n,
                                                       uint32_t* out) {
  const uint64_t total = n * req.n;
  for (uint64_t e = uint64_t(blockIdx.x) * kBlock + threadIdx.x; e < total; e += uint64_t(gridDim.x) * kBlock) {
    const uint64_t i = e / req.n;
    const uint32_t j = uint32_t(e - i * req.n);
    const uint64_t k = first + i;
    const uint64_t r = sorted ? sorted[2 * k + 1] : k;
    out[e] = table[r * cols + req.col[j]];
  }
}

void launch_fp_keys(const uint32_t* table, uint32_t cols, uint32_t c, uint32_t mode, const uint32_t* prev, uint64_t n,
                    uint32_t* pairs, hipStream_t s) {
  if (n == 0) return;
  WK_LAUNCH(k_fp_keys, dim3(grid_for((n + kBlock - 1) / kBlock, 8192)), dim3(kBlock), 0, s, table, cols, c,
                     mode, prev, n, pairs);
}

void launch_fp_dedup(const uint32_t* table, uint32_t cols, const ColList& req, const uint32_t* sorted, uint64_t n,
                     uint32_t* pairs, hipStream_t s) {
  if (n == 0) return;
  WK_LAUNCH(k_fp_dedup, dim3(grid_for((n + kBlock - 1) / kBlock, 8192)), dim3(kBlock), 0, s, table, cols,
                     req, sorted, n, pairs);
}

void launch_fp_project(const uint32_t* table, uint32_t cols, const ColList& req, const uint32_t* sorted,
                       uint64_t first, uint64_t n, uint32_t* out, hipStream_t s) {
  if (n == 0) return;
  WK_LAUNCH(k_fp_project, dim3(grid_for((n * req.n + kBlock - 1) / kBlock, 8192)), dim3(kBlock), 0, s,
                     table, cols, req, sorted, first, n, out);
}

void launch_dense_build(const Slot* karena, const uint32_t* ktab, uint32_t kshift, const DevSeg& s, uint64_t* dense,
                        hipStream_t st) {
  WK_LAUNCH(k_dense_build, dim3(grid_for((uint64_t(s.total_buckets) + kBlock - 1) / kBlock, 8192)),
                     dim3(kBlock), 0, st, karena, ktab, kshift, s, dense);
}

void launch_window_rows(const uint64_t* tile_sum, const uint64_t* lpfx, uint64_t R, uint64_t j0, uint64_t C,
                        uint64_t T, uint32_t n, uint64_t* out, hipStream_t s) {
  if (n == 0) return;
  WK_LAUNCH(k_window_rows, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, tile_sum, lpfx, R, j0, C, T, n,
            out);
}

void launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s) {
  WK_LAUNCH(k_fill_u64, dim3(grid_for((n + 255) / 256, 4096)), dim3(256), 0, s, p, n, v);
}

}  // namespace wk
