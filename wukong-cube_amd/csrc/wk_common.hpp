// Shared host/device definitions of the gstore layout used by the Wukong GPU
// query path: keys, pointers, slots, buckets, segment ids and the probe hash.
//
// Layout follows Wukong's static gstore:
//   key  = 64-bit word  vid:46 | pid:17 | dir:1     (core/store/vertex.hpp:44-117)
//          raw word == (vid << 18) | (pid << 1) | dir on a little-endian
//          bitfield ABI, which is also the pre-hash value at vertex.hpp:107-112.
//   ptr  = 64-bit word  type:2 | off:34 | size:28   (core/store/vertex.hpp:133-173)
//   slot = key + ptr (16 B); bucket = 8 slots (128 B), slot 7 is the chain link
//          to the next bucket of the cluster      (core/store/kvstore.hpp:292-314)
//   edge = uint32 vertex id                          (core/store/vertex.hpp:182)
#pragma once

#include <cstdint>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define WK_HD __host__ __device__ __forceinline__
#else
#define WK_HD inline
#endif

namespace wk {

using sid_t = uint32_t;   // core/common/type.hpp:39
using ssid_t = int32_t;   // core/common/type.hpp:40

enum : int { DIR_IN = 0, DIR_OUT = 1 };          // core/common/type.hpp:50
enum : uint32_t { PREDICATE_ID = 0, TYPE_ID = 1 }; // core/store/vertex.hpp:52-53
enum : int { NBITS_DIR = 1, NBITS_IDX = 17, NBITS_VID = 64 - NBITS_IDX - NBITS_DIR };
enum : int { NBITS_SIZE = 28, NBITS_PTR = 34, NBITS_TYPE = 2 };
enum : int { ASSOCIATIVITY = 8 };                 // core/store/kvstore.hpp:83
constexpr uint32_t BLANK_ID = 0xFFFFFFFFu;        // core/common/type.hpp:41

// 2 <= id < 2^17 are index ids (predicates and types); >= 2^17 are vertices.
WK_HD bool is_tpid(int64_t id) { return id > 1 && id < (int64_t(1) << NBITS_IDX); }
WK_HD bool is_vid(int64_t id) { return id >= (int64_t(1) << NBITS_IDX); }

WK_HD uint64_t make_key(uint64_t vid, uint64_t pid, uint64_t dir) {
  return (vid << (NBITS_IDX + NBITS_DIR)) | (pid << NBITS_DIR) | dir;
}
WK_HD uint64_t key_vid(uint64_t k) { return k >> (NBITS_IDX + NBITS_DIR); }
WK_HD uint64_t key_pid(uint64_t k) { return (k >> NBITS_DIR) & ((uint64_t(1) << NBITS_IDX) - 1); }
WK_HD uint64_t key_dir(uint64_t k) { return k & 1; }

WK_HD uint64_t make_ptr(uint64_t size, uint64_t off, uint64_t type = 0) {
  return size | (off << NBITS_SIZE) | (type << (NBITS_SIZE + NBITS_PTR));
}
WK_HD uint64_t ptr_size(uint64_t p) { return p & ((uint64_t(1) << NBITS_SIZE) - 1); }
WK_HD uint64_t ptr_off(uint64_t p) { return (p >> NBITS_SIZE) & ((uint64_t(1) << NBITS_PTR) - 1); }

// Thomas Wang's 64-bit integer mix, the probe hash of the gstore
// (utils/math.hpp:59-68; same function as the GPU copy gpu_hash.cu:35-54).
WK_HD uint64_t hash_u64(uint64_t key) {
  key = (~key) + (key << 21);
  key = key ^ (key >> 24);
  key = (key + (key << 3)) + (key << 8);
  key = key ^ (key >> 14);
  key = (key + (key << 2)) + (key << 4);
  key = key ^ (key >> 28);
  key = key + (key << 31);
  return key;
}

// High 64 bits of the 128-bit product a*b.
WK_HD uint64_t mulhi_u64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return uint64_t((unsigned __int128)a * b >> 64);
#endif
}

// Reciprocal for bucket_of: floor(2^64 / d) for d >= 2, 2^64 - 1 for d = 1.
WK_HD uint64_t bucket_magic(uint64_t d) {
  if (d <= 1) return ~uint64_t(0);
  const uint64_t m = ~uint64_t(0) / d;
  return (~uint64_t(0) % d == d - 1) ? m + 1 : m;  // d divides 2^64
}

// h mod d (d >= 1) without a 64-bit division, with m = bucket_magic(d)
// (Barrett reduction).  m > 2^64/d - 1 gives h*m/2^64 > h/d - 1, so the
// quotient estimate is q or q-1 and one conditional subtraction completes it.
WK_HD uint64_t bucket_of(uint64_t h, uint64_t d, uint64_t m) {
  const uint64_t r = h - mulhi_u64(h, m) * d;
  return r >= d ? r - d : r;
}

// Hash of one history-table row; a table's checksum is the sum (mod 2^64) of
// its row hashes, so it does not depend on row order.
WK_HD uint64_t row_hash_init(uint32_t cols) { return 0x9E3779B97F4A7C15ull ^ cols; }
WK_HD uint64_t row_hash_step(uint64_t h, uint32_t x, uint32_t c) {
  return hash_u64(h ^ (uint64_t(x) + (uint64_t(c) << 32)));
}
WK_HD uint64_t row_hash(const uint32_t* row, uint32_t cols) {
  uint64_t h = row_hash_init(cols);
  for (uint32_t c = 0; c < cols; c++) h = row_hash_step(h, row[c], c);
  return h;
}

struct Slot {
  uint64_t key;
  uint64_t ptr;
};
static_assert(sizeof(Slot) == 16, "slot is 16 bytes");

// Segment identity: one segment per (predicate, direction) of normal keys and
// one per direction for the index keys (vid == 0), core/store/segment_meta.hpp:159-190.
struct SegId {
  uint32_t index;  // 1 = index segment, 0 = normal
  uint32_t pid;    // PREDICATE_ID for index segments
  uint32_t dir;
};

WK_HD SegId segid_of_key(uint64_t key) {
  SegId s;
  s.dir = uint32_t(key_dir(key));
  if (key_vid(key) == 0) {
    s.index = 1;
    s.pid = PREDICATE_ID;
  } else {
    s.index = 0;
    s.pid = uint32_t(key_pid(key));
  }
  return s;
}

}  // namespace wk
