// Host gstore: the static, segment-organised key/value store that the device
// cache mirrors (core/store/segment_rdf_dgraph.hpp, core/store/kvstore.hpp).
#pragma once

#include <cstdint>
#include <unordered_map>
#include <vector>

#include "wk_common.hpp"

namespace wk {

// rdf_seg_meta_t (core/store/segment_meta.hpp:87-154), restated for this
// layout: a segment's main buckets are followed by its extended buckets, so
// one contiguous bucket range [bucket_start, bucket_start+num_buckets+num_ext).
struct SegMeta {
  SegId id;
  uint64_t num_keys = 0;
  uint64_t num_buckets = 0;  // hash space (main buckets)
  uint64_t num_ext = 0;      // chained buckets, local ids num_buckets..+num_ext-1
  uint64_t bucket_start = 0; // global bucket offset
  uint64_t num_edges = 0;
  uint64_t edge_start = 0;   // global edge offset
  uint32_t vmin = 0, vmax = 0;  // smallest and largest key vid (normal segments)
  uint64_t total_buckets() const { return num_buckets + num_ext; }
};

inline uint64_t seg_code(uint32_t index, uint32_t pid, uint32_t dir) {
  return (uint64_t(index) << 40) | (uint64_t(pid) << 1) | dir;
}

struct Store {
  std::vector<SegMeta> segs;
  std::unordered_map<uint64_t, int> seg_lookup;
  std::vector<int32_t> seg_flat;  // normal segment of (pid, dir) at 2 * pid + dir, -1 if none
  Slot* slots = nullptr;       // 8 per bucket
  uint64_t total_buckets = 0;
  uint32_t* values = nullptr;
  uint64_t total_values = 0;
  uint64_t total_keys = 0;
  uint64_t n_out = 0, n_in = 0;
  bool pinned = false;
  std::vector<uint32_t> type_ids;  // ids with instances [0|t|IN], ascending

  ~Store();
  int seg_index(uint32_t index, uint32_t pid, uint32_t dir) const {
    if (index == 0 && dir < 2) {
      const std::size_t i = 2 * std::size_t(pid) + dir;
      return i < seg_flat.size() ? seg_flat[i] : -1;
    }
    auto it = seg_lookup.find(seg_code(index, pid, dir));
    return it == seg_lookup.end() ? -1 : it->second;
  }
  // KVStore::get_slot_local restated on this layout; returns nullptr if absent.
  const Slot* find_slot(uint64_t key) const;
  // DGraph::get_triples / get_index.
  const uint32_t* get_triples(uint32_t vid, uint32_t pid, int dir, uint64_t& sz) const;
  const uint32_t* get_index(uint32_t tpid, int dir, uint64_t& sz) const;
};

struct BuildOpts {
  bool pinned = false;
  int load_factor_pct = 55;
};

// Returns a WK_* status; on success *out owns the store.
int build_store(const uint32_t* pso, uint64_t n_pso, const uint32_t* pos, uint64_t n_pos,
                const BuildOpts& opts, Store** out);

}  // namespace wk
