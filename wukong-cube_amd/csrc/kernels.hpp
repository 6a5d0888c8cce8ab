// Structures shared by the device kernels and the host orchestration, and the
// host-side launch wrappers of the expansion kernels.
//
// One pattern step of one or many queries runs as:
//   probe        -> per input row: edge count (k2u) or keep flag (k2k/k2c),
//                   kept as the output rows before each row within its
//                   kTile-row tile, and each tile's sum
//   scan_tiles   -> exclusive scan of the per-tile sums; a row's output
//                   offset is its tile's prefix plus its prefix in the tile
//   scan_queries -> per-query output rows/offsets, next step's row space,
//                   output-tile -> row map
//   gather       -> write the next history table(s)
// All sizes live in device memory, so a multi-step batch can be captured in a
// hipGraph with no host round trip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "wk_common.hpp"

namespace wk {

// OP_FILTER_TYPE: known_to_const `?x rdf:type T` answered from T's instance
// bitmap (StepDesc.ec holds the bitmap's slot) instead of x's type list
enum : uint32_t { OP_EXPAND = 0, OP_FILTER_KNOWN = 1, OP_FILTER_CONST = 2, OP_SKIP = 3, OP_FILTER_TYPE = 4 };
enum : uint32_t { KEY_FROM_ROW = 0, KEY_CONST = 1, KEY_TYPE_INDEX = 2 };

// Segment descriptor as seen by the kernels (device array, one per segment).
struct DevSeg {
  uint64_t nb_magic;      // bucket_magic(num_buckets)
  uint64_t edge_start;    // global edge offset of the segment (ptr.off - edge_start = local)
  const uint64_t* dense;  // direct key index of a segment whose key vids are dense: the ptr
                          // word of key vid dvmin + i at dense[i], 0 for no key; or null
  int64_t koff;           // where bucket 0 is, in buckets from the key arena, when the
                          // segment's buckets are contiguous (consecutive key blocks, or a
                          // segment read in place from pinned host memory); kBlockTable when
                          // they are found through the key-block table; kNotResident
  int64_t voff;           // the same for edge 0, in edges from the value arena
  uint32_t num_buckets;   // main buckets: hash modulus
  uint32_t total_buckets; // main + extended: bound on chain links
  uint32_t kb_base;       // first entry in the key-block table
  uint32_t vb_base;       // first entry in the value-block table
  uint32_t dvmin;         // first vid the direct index covers
  uint32_t dvn;           // vids it covers (>= 1 when dense is set)
};
static_assert(sizeof(DevSeg) == 64, "DevSeg is one 64-byte line");
constexpr int64_t kNotResident = INT64_MIN;
constexpr int64_t kBlockTable = INT64_MIN + 1;
__host__ __device__ inline bool seg_resident(const DevSeg& s) { return s.koff != kNotResident; }
// StepDesc.out_off (and the arena offset in qres) of a final table that did
// not fit the result arena: the query reports WK_E_NOMEM_BUDGET, the others
// of the step go on.
constexpr uint64_t kNoSpace = ~uint64_t(0);

// Instances of one type as a bitmap over vids [vmin, vmin + nbits): bit
// i of words is set when vmin + i has the type, i.e. is in the type index
// [0|t|IN] (core/store/segment_rdf_dgraph.hpp:951-994), which holds exactly
// the subjects of the type's rdf:type triples.
struct TypeBits {
  const uint32_t* words;
  uint32_t vmin;
  uint32_t nbits;
};

// One query's work in one pattern step.
struct StepDesc {
  uint32_t op;
  int32_t seg;           // segment index, -1: key cannot exist (no rows)
  uint32_t pid, dir;
  uint32_t key_mode;
  uint32_t sc, ec;       // start / end column
  uint32_t kstart;       // constant start vid (KEY_CONST)
  uint32_t kend;         // constant end (k2c)
  uint32_t in_cols, out_cols;
  uint32_t final_step;   // 1: write into the result arena
  uint64_t in_off;       // element offset of the input table
  uint64_t in_rows;
  uint64_t out_rows;     // set by scan_queries
  uint64_t out_off;      // set by scan_queries (element offset in its destination)
};

// Per-step shared state (device).
struct StepState {
  uint64_t total_out_rows;  // output rows over all queries (set by scan_tiles)
  uint64_t tout_base;       // first element of the ping-pong out table to use
  uint64_t out_elems;       // elements written to the out table (set by scan_queries)
  uint64_t arena_base;      // first free element of the result arena
  uint32_t status;          // 0 ok, 1 table overflow
  uint32_t tiles_marked;    // 1: k_scan_small wrote the output-tile map of the whole step
};

// Pointers and capacities used by every step kernel.
struct StepArgs {
  const Slot* karena;
  const uint32_t* varena;
  const uint32_t* ktab;
  const uint32_t* vtab;
  const DevSeg* segs;
  const TypeBits* tbits;   // type bitmaps by slot (OP_FILTER_TYPE)
  uint32_t kshift, vshift;
  const uint32_t* tin;     // input table buffer
  uint32_t* tout;          // ping-pong output buffer
  uint32_t* arena;         // result arena (final steps)
  uint64_t tcap;           // elements of tout
  uint64_t arena_cap;      // elements of arena
  StepDesc* desc;          // [nq] this step
  StepDesc* next;          // [nq] next step (rows/offsets filled by scan_queries) or null
  uint64_t* row_start;     // [nq+1] flattened row starts of this step
  uint64_t* next_row_start;// [nq+1] of the next step or null
  uint32_t nq;
  uint32_t* eoff;          // [rcap]
  uint64_t* lpfx;          // [rcap+1] output rows before each row within its kTile-row tile; entry R:
                           // the rows of the tile holding row R before it
  uint64_t rcap;
  uint64_t* tile_sum;      // [tile_cap] output rows of each tile, then (k_scan_tiles) before each tile;
                           // entry ceil(R / kTile): the step's total
  uint64_t tile_cap;
  uint64_t* tile_row;      // [otile_cap+1] first source row of each output tile
  uint64_t otile_cap;
  StepState* st;
  StepState* next_st;      // or null
  uint64_t* qres;          // [nq][2] rows and arena offset of each final table, or null
  uint64_t* qagg;          // [ceil(nq / kBlock)][3] per-block query sums (multi-block query scan)
  unsigned long long* ck;  // [nq] sums of the row hashes of final tables (zeroed before), or null
  // Output window of the gather: it writes the step's output rows
  // [win_lo, min(win_hi, total)), row j at (j - win_lo) past its table's
  // start.  The batched and served paths use the whole step (0, ~0); the
  // per-query path runs a step too large for its room window by window.
  uint64_t win_lo;
  uint64_t win_hi;
};

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kRowsPerThread = 4;
constexpr int kTile = kBlock * kRowsPerThread;  // rows per probe/scan tile
constexpr int kTileShift = 10;
static_assert((1 << kTileShift) == kTile, "kTile is 2^kTileShift rows");
constexpr int kOutTile = 4 * kBlock;            // output rows per gather tile (4 per thread)

// Launch wrappers (all on `stream`; grids sized from host hints, kernels
// re-read the device sizes and stride).  `rows_hint`/`out_hint` = 0 means
// unknown on the host: a fixed persistent grid is used.
// row_per_query: a batched step whose input is one row per query (its first).
void launch_probe(const StepArgs& a, int probe_width, uint64_t rows_hint, hipStream_t s, bool row_per_query = false);
// scan = scan_tiles + scan_queries, in one launch (launch_scan_small) for
// steps of at most kBlock queries.
// The scans (multi-block query scan for more than kBlock queries, which also
// marks output tiles) and the gather of one step with no host size hints: the
// batched and served paths.
void launch_expand_tail(const StepArgs& a, int rows_per_wave, hipStream_t s);
void launch_scan_queries_mb(const StepArgs& a, hipStream_t s);
void launch_scan(const StepArgs& a, hipStream_t s);
void launch_scan_tiles(const StepArgs& a, hipStream_t s);
void launch_scan_queries(const StepArgs& a, hipStream_t s);
void launch_scan_small(const StepArgs& a, hipStream_t s);
// gather = tile_rows (output-balanced mapping only) + the gather kernel
void launch_gather(const StepArgs& a, int rows_per_wave, uint64_t out_hint, hipStream_t s);
void launch_tile_rows(const StepArgs& a, uint64_t out_hint, hipStream_t s);
void launch_gather_only(const StepArgs& a, int rows_per_wave, uint64_t out_hint, hipStream_t s);
// Fork-join dispatch: out = in's rows grouped by owner row[col] % nparts
// (nparts <= kMaxParts), in row order within each owner.  hist needs
// nparts * ceil(rows / kTile) entries; on return hist[p * ntiles] is owner
// p's first row in out and *total the row count.
constexpr uint32_t kMaxParts = 64;
// Owner of a row, from one of its words x: x % mod when mod > 0, else the
// digit (x >> shift) & mask.
struct OwnerFn {
  uint32_t mod, shift, mask;
};
__host__ __device__ inline uint32_t owner_count(OwnerFn f) { return f.mod ? f.mod : f.mask + 1; }

// Scratch words launch_partition needs: per-tile owner histogram, chunk
// totals, and the nparts + 1 heads.
uint64_t partition_scratch_words(uint64_t rows, uint32_t nparts);
// Groups rows by owner f(in[r * cols + col]) (at most kMaxParts owners) into
// out, keeping input order within each owner.  Returns the device address,
// inside scratch, of nparts + 1 words: each owner's first output row, then
// the row count.
uint64_t* launch_partition(const uint32_t* in, uint64_t rows, uint32_t cols, uint32_t col, OwnerFn f,
                           uint64_t* scratch, uint32_t* out, hipStream_t s);

// final_process: key transforms, so that unsigned order on the key is the
// wanted order on the word.
enum : uint32_t { kKeySigned = 0, kKeyAsc = 1, kKeyDesc = 2 };
constexpr uint32_t kSortDigitBits = 6;  // bits per stable split (kMaxParts owners)
struct ColList {
  uint32_t n;
  uint32_t col[64];
};
void launch_fp_keys(const uint32_t* table, uint32_t cols, uint32_t c, uint32_t mode, const uint32_t* prev, uint64_t n,
                    uint32_t* pairs, hipStream_t s);
void launch_fp_dedup(const uint32_t* table, uint32_t cols, const ColList& req, const uint32_t* sorted, uint64_t n,
                     uint32_t* pairs, hipStream_t s);
void launch_fp_project(const uint32_t* table, uint32_t cols, const ColList& req, const uint32_t* sorted,
                       uint64_t first, uint64_t n, uint32_t* out, hipStream_t s);
void launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s);

// Kernel timing for profiling: between arm_launch_timer and
// disarm_launch_timer, launches made on this thread take their times from
// their own dispatch packets (hipExtLaunchKernelGGL): `start` from the first
// launch, `stop` from each (the last one's stays).  disarm_launch_timer
// returns whether any launch took them.
void arm_launch_timer(hipEvent_t start, hipEvent_t stop);
// Sizes the grids of launches whose row counts live on the device from the
// current device's CU count and each kernel's occupancy (once per process,
// before any graph capture).
void init_launch_grids();
bool disarm_launch_timer();

// Direct key index of a resident segment s (its buckets at karena through
// ktab, s.dense/dvmin/dvn set): dense[vid - dvmin] = ptr of every key in the
// segment's buckets.  dense must be zeroed before.
void launch_dense_build(const Slot* karena, const uint32_t* ktab, uint32_t kshift, const DevSeg& s, uint64_t* dense,
                        hipStream_t st);

// Chunk boundaries of a step run in output windows of C rows: for chunk i
// (output rows [j, min(T, j + C)) with j = j0 + i * C), out[3i] and
// out[3i + 1] are the input rows holding its first and last output row and
// out[3i + 2] the prefix at the first.  The step's R + 1 prefixes are
// tile_sum[g / kTile] + lpfx[g], as k_scan_tiles leaves them.
constexpr uint32_t kWinGroup = 4096;
void launch_window_rows(const uint64_t* tile_sum, const uint64_t* lpfx, uint64_t R, uint64_t j0, uint64_t C,
                        uint64_t T, uint32_t n, uint64_t* out, hipStream_t s);

// One non-empty table for the checksum kernel: `rows` rows of `cols`
// elements at element offset `off`; rows of all tables numbered from `row0`.
struct CkDesc {
  uint64_t off;
  uint64_t rows;
  uint64_t row0;
  uint32_t cols;
  uint32_t pad;
};
// sums[i] += checksum of table i (sums must be zeroed by the caller).
void launch_checksum(const uint32_t* base, const CkDesc* d, uint32_t n, uint64_t total_rows, uint64_t* sums,
                     hipStream_t s);
// sums[f] += checksum of the final table desc[idx[f]] in the arena.


}  // namespace wk
