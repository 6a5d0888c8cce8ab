// Host gstore builder: sorted, de-duplicated triples -> per-segment
// cluster-chained hash buckets plus contiguous edge lists.
//
// Semantics restated from the reference's static gstore:
//   * pso/pos sorted and de-duplicated (loader/base_loader.hpp:90-103,481-498),
//     so every key's edge list is ascending and unique;
//   * one OUT key [s|p|OUT] per (s,p) and one IN key [o|p|IN] per (o,p); IN
//     keys skip triples whose object is a type/predicate id
//     (segment_rdf_dgraph.hpp:513-590, the skip at :581-584);
//   * index keys (vid = 0): [0|p|IN] = subjects having p, [0|p|OUT] = objects
//     of p, [0|t|IN] = instances of type t (collect_idx_info :951-994,
//     insert_idx :284-347);
//   * main bucket = hash_u64(key) % num_buckets of the segment; slots 0..6 hold
//     keys, slot 7 links to the next bucket of the cluster (kvstore.hpp:292-314,
//     insert_key_to_seg segment_rdf_dgraph.hpp:203-271);
//   * #buckets = #keys * 100 / (8 * load_factor) (segment_rdf_dgraph.hpp:138).
// Layout choices of this build (results do not depend on them): extended
// buckets are sized exactly and placed right after the segment's main
// buckets; the link in slot 7 stores the segment-local bucket number in the
// vid field; keys sharing a bucket are placed in ascending key order, so the
// build is deterministic for any thread count.
#include "store.hpp"

#include <hip/hip_runtime_api.h>
#include <omp.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "sort_util.hpp"
#include "wk_gpu.h"

namespace wk {

namespace {

constexpr uint64_t TPID_LIMIT = uint64_t(1) << NBITS_IDX;

// Triples of one direction grouped by predicate: run of predicate p is
// packed[off[p] .. off[p+1]), each word (hi << 32 | lo), sorted, unique.
struct PredRuns {
  std::vector<uint64_t> packed;
  std::vector<uint64_t> off;  // size P+1
  uint32_t num_p = 0;
  const uint64_t* run(uint32_t p, uint64_t& n) const {
    if (p >= num_p) {
      n = 0;
      return nullptr;
    }
    n = off[p + 1] - off[p];
    return packed.data() + off[p];
  }
};

int validate(const uint32_t* t, uint64_t n, uint32_t& maxp) {
  int bad = 0;
  uint32_t mp = 0;
#pragma omp parallel for reduction(| : bad) reduction(max : mp)
  for (uint64_t i = 0; i < n; i++) {
    const uint32_t s = t[3 * i], p = t[3 * i + 1], o = t[3 * i + 2];
    // vid 0/1 would alias index keys; predicate 0 is the index class; type
    // triples must point at a type id (collect_idx_info asserts otherwise).
    if (p == PREDICATE_ID || p >= TPID_LIMIT || s < 2 || o < 2 || (p == TYPE_ID && !is_tpid(o)))
      bad = 1;
    mp = std::max(mp, p);
  }
  maxp = mp;
  return bad ? WK_E_INVALID_TRIPLE : WK_OK;
}

int group_sort(const uint32_t* t, uint64_t n, bool by_object, uint32_t num_p, PredRuns& r) {
  r.num_p = num_p;
  const int nt = omp_get_max_threads();
  std::vector<uint64_t> hist(size_t(nt) * num_p, 0);
  r.packed.resize(n);
  std::vector<uint64_t> tmp(n);
#pragma omp parallel num_threads(nt)
  {
    const int tid = omp_get_thread_num();
    const uint64_t lo = n * tid / nt, hi = n * (tid + 1) / nt;
    uint64_t* h = &hist[size_t(tid) * num_p];
    for (uint64_t i = lo; i < hi; i++) h[t[3 * i + 1]]++;
#pragma omp barrier
#pragma omp single
    {
      uint64_t run = 0;
      for (uint32_t p = 0; p < num_p; p++)
        for (int k = 0; k < nt; k++) {
          uint64_t c = hist[size_t(k) * num_p + p];
          hist[size_t(k) * num_p + p] = run;
          run += c;
        }
    }
    for (uint64_t i = lo; i < hi; i++) {
      const uint64_t s = t[3 * i], p = t[3 * i + 1], o = t[3 * i + 2];
      r.packed[h[p]++] = by_object ? ((o << 32) | s) : ((s << 32) | o);
    }
  }
  // run starts: after the scatter, thread nt-1's cursor for p is the end of p
  std::vector<uint64_t> start(num_p + 1, 0), count(num_p, 0);
  for (uint32_t p = 0; p < num_p; p++) {
    const uint64_t end = hist[size_t(nt - 1) * num_p + p];
    start[p + 1] = end;
  }
  for (uint32_t p = 0; p < num_p; p++) {
    uint64_t c = start[p + 1] - start[p];
    if (c == 0) continue;
    radix_sort_u64(r.packed.data() + start[p], tmp.data() + start[p], c);
    count[p] = dedup_sorted(r.packed.data() + start[p], c);
  }
  r.off.assign(num_p + 1, 0);
  uint64_t w = 0;
  for (uint32_t p = 0; p < num_p; p++) {
    r.off[p] = w;
    if (count[p] && w != start[p])
      std::memmove(r.packed.data() + w, r.packed.data() + start[p], count[p] * sizeof(uint64_t));
    w += count[p];
  }
  r.off[num_p] = w;
  r.packed.resize(w);
  r.packed.shrink_to_fit();
  return WK_OK;
}

// Distinct `hi` values of a sorted run with their multiplicities.
void distinct_hi(const uint64_t* a, uint64_t n, std::vector<uint32_t>& his,
                 std::vector<uint32_t>& sizes) {
  his.clear();
  sizes.clear();
  if (n == 0) return;
  const int nt = std::max(1, std::min(omp_get_max_threads(), int(n >> 16) + 1));
  std::vector<std::vector<uint64_t>> starts(nt);
#pragma omp parallel num_threads(nt)
  {
    const int tid = omp_get_thread_num();
    const uint64_t lo = n * tid / nt, hi = n * (tid + 1) / nt;
    auto& v = starts[tid];
    for (uint64_t i = lo; i < hi; i++)
      if (i == 0 || (a[i] >> 32) != (a[i - 1] >> 32)) v.push_back(i);
  }
  std::vector<uint64_t> all;
  for (auto& v : starts) all.insert(all.end(), v.begin(), v.end());
  his.resize(all.size());
  sizes.resize(all.size());
#pragma omp parallel for if (all.size() > 65536)
  for (uint64_t k = 0; k < all.size(); k++) {
    const uint64_t e = (k + 1 < all.size()) ? all[k + 1] : n;
    his[k] = uint32_t(a[all[k]] >> 32);
    sizes[k] = uint32_t(e - all[k]);
  }
}

// One segment under construction.
struct SegBuild {
  SegMeta meta;
  std::vector<uint64_t> keys;     // key words, in value order
  std::vector<uint32_t> sizes;    // edge-list length per key
  // values: either the lo halves of a packed run, or explicit lists
  const uint64_t* src = nullptr;
  uint64_t src_n = 0;
  std::vector<uint32_t> vals;
  // bucketing: (bucket << 32 | key index) sorted; run starts; ext prefix
  std::vector<uint64_t> order;
  std::vector<uint64_t> run_start;  // index into order, size runs+1
  std::vector<uint64_t> run_ext;    // ext bucket prefix per run
};

void bucketize(SegBuild& sb, int lf) {
  const uint64_t nk = sb.keys.size();
  if (nk && sb.meta.id.index == 0) {
    // keys are built in ascending vid order
    sb.meta.vmin = uint32_t(key_vid(sb.keys.front()));
    sb.meta.vmax = uint32_t(key_vid(sb.keys.back()));
  }
  uint64_t nb = (nk * 100 + uint64_t(ASSOCIATIVITY) * lf - 1) / (uint64_t(ASSOCIATIVITY) * lf);
  if (nb == 0) nb = 1;
  sb.meta.num_keys = nk;
  sb.meta.num_buckets = nb;
  sb.order.resize(nk);
#pragma omp parallel for if (nk > 65536)
  for (uint64_t k = 0; k < nk; k++) sb.order[k] = ((hash_u64(sb.keys[k]) % nb) << 32) | k;
  {
    std::vector<uint64_t> tmp(nk);
    radix_sort_u64(sb.order.data(), tmp.data(), nk);
  }
  sb.run_start.clear();
  sb.run_ext.clear();
  uint64_t ext = 0;
  for (uint64_t i = 0; i < nk;) {
    uint64_t j = i + 1;
    while (j < nk && (sb.order[j] >> 32) == (sb.order[i] >> 32)) j++;
    sb.run_start.push_back(i);
    sb.run_ext.push_back(ext);
    const uint64_t c = j - i;
    if (c > ASSOCIATIVITY - 1) ext += (c - (ASSOCIATIVITY - 1) + ASSOCIATIVITY - 2) / (ASSOCIATIVITY - 1);
    i = j;
  }
  sb.run_start.push_back(nk);
  sb.meta.num_ext = ext;
}

int fill_segment(SegBuild& sb, Slot* slots, uint32_t* values) {
  const SegMeta& m = sb.meta;
  const uint64_t nk = sb.keys.size();
  // local value offsets (exclusive scan of sizes, keys are in value order)
  std::vector<uint64_t> loff(nk);
  uint64_t run = 0;
  for (uint64_t k = 0; k < nk; k++) {
    loff[k] = run;
    run += sb.sizes[k];
  }
  if (run != m.num_edges) return WK_E_UNKNOWN;
  if (m.edge_start + m.num_edges >= (uint64_t(1) << NBITS_PTR)) return WK_E_OVERFLOW;
  int bad = 0;
  const uint64_t nruns = sb.run_ext.size();
#pragma omp parallel for schedule(dynamic, 1024) reduction(| : bad)
  for (uint64_t r = 0; r < nruns; r++) {
    const uint64_t i0 = sb.run_start[r], i1 = sb.run_start[r + 1];
    uint64_t bucket = sb.order[i0] >> 32;
    uint64_t next_ext = m.num_buckets + sb.run_ext[r];
    int slot = 0;
    for (uint64_t i = i0; i < i1; i++) {
      const uint64_t k = sb.order[i] & 0xffffffffu;
      if (slot == ASSOCIATIVITY - 1) {
        slots[(m.bucket_start + bucket) * ASSOCIATIVITY + slot].key = make_key(next_ext, 0, 0);
        bucket = next_ext++;
        slot = 0;
      }
      if (sb.sizes[k] >= (uint64_t(1) << NBITS_SIZE)) bad = 1;
      Slot& s = slots[(m.bucket_start + bucket) * ASSOCIATIVITY + slot];
      s.key = sb.keys[k];
      s.ptr = make_ptr(sb.sizes[k], m.edge_start + loff[k]);
      slot++;
    }
  }
  if (bad) return WK_E_OVERFLOW;
  uint32_t* dst = values + m.edge_start;
  if (sb.src) {
#pragma omp parallel for if (sb.src_n > 65536)
    for (uint64_t i = 0; i < sb.src_n; i++) dst[i] = uint32_t(sb.src[i]);
  } else {
    std::memcpy(dst, sb.vals.data(), sb.vals.size() * sizeof(uint32_t));
  }
  return WK_OK;
}

void* host_alloc(size_t bytes, bool want_pinned, bool& pinned) {
  void* p = nullptr;
  pinned = false;
  if (bytes == 0) bytes = 64;
  if (want_pinned && hipHostMalloc(&p, bytes, hipHostMallocDefault) == hipSuccess) {
    pinned = true;
    return p;
  }
  (void)hipGetLastError();
  if (posix_memalign(&p, 4096, bytes) != 0) return nullptr;
  return p;
}

void host_free(void* p, bool pinned) {
  if (!p) return;
  if (pinned)
    (void)hipHostFree(p);
  else
    std::free(p);
}

}  // namespace

Store::~Store() {
  host_free(slots, pinned);
  host_free(values, pinned);
}

const Slot* Store::find_slot(uint64_t key) const {
  const SegId sid = segid_of_key(key);
  const int si = seg_index(sid.index, sid.pid, sid.dir);
  if (si < 0) return nullptr;
  const SegMeta& m = segs[si];
  uint64_t b = hash_u64(key) % m.num_buckets;
  while (true) {
    const Slot* bk = slots + (m.bucket_start + b) * ASSOCIATIVITY;
    for (int i = 0; i < ASSOCIATIVITY - 1; i++)
      if (bk[i].key == key) return &bk[i];
    const uint64_t link = bk[ASSOCIATIVITY - 1].key;
    if (link == 0) return nullptr;
    b = key_vid(link);
  }
}

const uint32_t* Store::get_triples(uint32_t vid, uint32_t pid, int dir, uint64_t& sz) const {
  sz = 0;
  if (pid >= TPID_LIMIT || (dir != DIR_IN && dir != DIR_OUT)) return nullptr;
  const Slot* s = find_slot(make_key(vid, pid, dir));
  if (!s) return nullptr;
  sz = ptr_size(s->ptr);
  return values + ptr_off(s->ptr);
}

const uint32_t* Store::get_index(uint32_t tpid, int dir, uint64_t& sz) const {
  return get_triples(0, tpid, dir, sz);
}

int build_store(const uint32_t* pso, uint64_t n_pso, const uint32_t* pos, uint64_t n_pos,
                const BuildOpts& opts, Store** out) {
  *out = nullptr;
  if ((n_pso && !pso) || (n_pos && !pos)) return WK_E_INVALID_ARG;
  const int lf = (opts.load_factor_pct > 0 && opts.load_factor_pct < 100) ? opts.load_factor_pct : 55;
  uint32_t maxp_o = 0, maxp_i = 0;
  int rc = validate(pso, n_pso, maxp_o);
  if (rc) return rc;
  rc = validate(pos, n_pos, maxp_i);
  if (rc) return rc;
  const uint32_t num_p = std::max(maxp_o, maxp_i) + 1;

  PredRuns outr, inr;
  group_sort(pso, n_pso, false, num_p, outr);
  group_sort(pos, n_pos, true, num_p, inr);

  std::unique_ptr<Store> st(new Store());
  st->n_out = outr.packed.size();
  st->n_in = inr.packed.size();

  std::vector<std::unique_ptr<SegBuild>> builds;
  std::vector<uint32_t> his, sizes;
  // index lists: [0|p|IN] subjects of p, [0|p|OUT] objects of p
  std::vector<std::pair<uint32_t, std::vector<uint32_t>>> pidx_in, pidx_out;
  for (uint32_t p = 1; p < num_p; p++) {
    uint64_t n = 0;
    const uint64_t* a = outr.run(p, n);
    if (n) {
      distinct_hi(a, n, his, sizes);
      auto sb = std::make_unique<SegBuild>();
      sb->meta.id = SegId{0, p, DIR_OUT};
      sb->keys.resize(his.size());
      for (size_t k = 0; k < his.size(); k++) sb->keys[k] = make_key(his[k], p, DIR_OUT);
      sb->sizes = sizes;
      sb->src = a;
      sb->src_n = n;
      sb->meta.num_edges = n;
      if (p != TYPE_ID) pidx_in.emplace_back(p, his);
      builds.push_back(std::move(sb));
    }
    a = inr.run(p, n);
    if (n) {
      // skip objects that are type/predicate ids (a prefix: the run is sorted by object)
      const uint64_t* first = std::lower_bound(a, a + n, TPID_LIMIT << 32);
      const uint64_t m = n - uint64_t(first - a);
      if (m) {
        distinct_hi(first, m, his, sizes);
        auto sb = std::make_unique<SegBuild>();
        sb->meta.id = SegId{0, p, DIR_IN};
        sb->keys.resize(his.size());
        for (size_t k = 0; k < his.size(); k++) sb->keys[k] = make_key(his[k], p, DIR_IN);
        sb->sizes = sizes;
        sb->src = first;
        sb->src_n = m;
        sb->meta.num_edges = m;
        pidx_out.emplace_back(p, his);
        builds.push_back(std::move(sb));
      }
    }
  }
  // type index from the OUT type triples: (s << 32 | t) -> (t << 32 | s)
  std::vector<std::pair<uint32_t, std::vector<uint32_t>>> tidx;
  {
    uint64_t n = 0;
    const uint64_t* a = outr.run(TYPE_ID, n);
    if (n) {
      std::vector<uint64_t> ts(n), tmp(n);
#pragma omp parallel for if (n > 65536)
      for (uint64_t i = 0; i < n; i++) ts[i] = ((a[i] & 0xffffffffu) << 32) | (a[i] >> 32);
      radix_sort_u64(ts.data(), tmp.data(), n);
      distinct_hi(ts.data(), n, his, sizes);
      uint64_t o = 0;
      for (size_t k = 0; k < his.size(); k++) {
        std::vector<uint32_t> v(sizes[k]);
        for (uint32_t j = 0; j < sizes[k]; j++) v[j] = uint32_t(ts[o + j]);
        o += sizes[k];
        st->type_ids.push_back(his[k]);
        tidx.emplace_back(his[k], std::move(v));
      }
    }
  }
  auto make_index_seg = [&](int dir,
                            std::vector<std::pair<uint32_t, std::vector<uint32_t>>>& a,
                            std::vector<std::pair<uint32_t, std::vector<uint32_t>>>* b) -> int {
    auto sb = std::make_unique<SegBuild>();
    sb->meta.id = SegId{1, PREDICATE_ID, uint32_t(dir)};
    std::vector<uint32_t> ids;
    for (auto* lst : {&a, b}) {
      if (!lst) continue;
      for (auto& e : *lst) {
        sb->keys.push_back(make_key(0, e.first, dir));
        sb->sizes.push_back(uint32_t(e.second.size()));
        sb->vals.insert(sb->vals.end(), e.second.begin(), e.second.end());
        ids.push_back(e.first);
      }
    }
    std::sort(ids.begin(), ids.end());
    if (std::adjacent_find(ids.begin(), ids.end()) != ids.end()) return WK_E_INVALID_TRIPLE;
    sb->meta.num_edges = sb->vals.size();
    if (!sb->keys.empty()) builds.push_back(std::move(sb));
    return WK_OK;
  };
  rc = make_index_seg(DIR_IN, pidx_in, &tidx);
  if (rc) return rc;
  rc = make_index_seg(DIR_OUT, pidx_out, nullptr);
  if (rc) return rc;

  // bucketize, then lay out segments back to back
  uint64_t nbk = 0, nval = 0, nkeys = 0;
  for (auto& sb : builds) {
    bucketize(*sb, lf);
    sb->meta.bucket_start = nbk;
    sb->meta.edge_start = nval;
    nbk += sb->meta.total_buckets();
    nval += sb->meta.num_edges;
    nkeys += sb->meta.num_keys;
  }
  bool pin_s = false, pin_v = false;
  st->slots = static_cast<Slot*>(host_alloc(nbk * sizeof(Slot) * ASSOCIATIVITY, opts.pinned, pin_s));
  st->values = static_cast<uint32_t*>(host_alloc(nval * sizeof(uint32_t), opts.pinned && pin_s, pin_v));
  st->pinned = pin_s;
  if (pin_s != pin_v) {
    // keep one allocation policy for both arrays
    host_free(st->slots, pin_s);
    host_free(st->values, pin_v);
    st->slots = static_cast<Slot*>(host_alloc(nbk * sizeof(Slot) * ASSOCIATIVITY, false, pin_s));
    st->values = static_cast<uint32_t*>(host_alloc(nval * sizeof(uint32_t), false, pin_v));
    st->pinned = false;
  }
  if (!st->slots || !st->values) return WK_E_NOMEM_HOST;
  const uint64_t nslots = nbk * ASSOCIATIVITY;
#pragma omp parallel for if (nslots > 65536)
  for (uint64_t i = 0; i < nslots; i++) st->slots[i] = Slot{0, 0};
  st->total_buckets = nbk;
  st->total_values = nval;
  st->total_keys = nkeys;
  for (auto& sb : builds) {
    rc = fill_segment(*sb, st->slots, st->values);
    if (rc) return rc;
    const SegId& id = sb->meta.id;
    st->seg_lookup[seg_code(id.index, id.pid, id.dir)] = int(st->segs.size());
    if (id.index == 0 && id.dir < 2) {
      const size_t i = 2 * size_t(id.pid) + id.dir;
      if (st->seg_flat.size() <= i) st->seg_flat.resize(i + 1, -1);
      st->seg_flat[i] = int32_t(st->segs.size());
    }
    st->segs.push_back(sb->meta);
    sb.reset();
  }
  *out = st.release();
  return WK_OK;
}

}  // namespace wk
