// Device handle: HBM gstore cache, history tables, scratch and the per-step
// orchestration shared by per-query, batched and served execution.
#pragma once

#include <hip/hip_runtime.h>
#include <omp.h>

#include <algorithm>
#include <map>
#include <utility>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "store.hpp"
#include "wk_gpu.h"

namespace wk {

extern thread_local int g_last_hip_error;

// Threads of the host loops that every batch runs (plan compile, descriptor
// fill).  OpenMP workers spin between parallel regions, and a batch every few
// hundred microseconds keeps a full team spinning all the time: with as many
// threads as the process's CPU quota allows, the quota runs out within its
// period and the whole process is stopped for milliseconds.  Four threads
// fewer (at most 12) leave the runtime's and the caller's threads room.
inline int batch_threads() { return std::max(1, std::min(12, omp_get_max_threads() - 4)); }

// True when `e` is a failure; the code is kept for wk_last_hip_error().
inline bool hip_failed(hipError_t e) {
  if (e == hipSuccess) return false;
  g_last_hip_error = int(e);
  return true;
}

#define WK_HIP(x)                            \
  do {                                       \
    hipError_t e_ = (x);                     \
    if (e_ != hipSuccess) {                  \
      ::wk::g_last_hip_error = int(e_);      \
      return WK_E_HIP;                       \
    }                                        \
  } while (0)

// A compiled query: one StepDesc per pattern (static fields only), the final
// var->column map and column count, or an error status
// (SPARQLEngine::execute_one_pattern dispatch, core/engine/sparql.hpp:932-1055).
struct Plan {
  std::vector<StepDesc> steps;
  std::vector<int32_t> v2c;
  uint32_t cols = 0;
  int status = WK_OK;
};

// type_slot (may be null): for each id below 2^17, the slot of its instance
// bitmap, or -1; a `?x rdf:type T` known_to_const step of a type with a
// bitmap becomes OP_FILTER_TYPE.
int compile_query(const Store& st, const wk_query& q, Plan& plan, const int32_t* type_slot = nullptr);

struct SegSlot {
  uint32_t nkb = 0, nvb = 0;      // blocks needed
  std::vector<uint32_t> kblk, vblk;
  bool resident = false;
  bool mapped = false;            // read in place from pinned host memory (never evicted)
  uint64_t last_use = 0;
  int pins = 0;
};

constexpr int kStageBufs = 4;  // pinned staging buffers for swaps from pageable memory
constexpr int kMaxLanes = 4;   // lanes of a device: its own and up to three others

struct KernelStat {
  uint64_t launches = 0;
  double ms = 0;
  double bytes = 0;
};

// What begin_batch left on the lane's stream for finish_batch: the pattern
// steps launched and the segments pinned for them.  `running` is false when
// nothing is left to wait for (no steps, or the batch already completed by
// itself, split or alone).
struct BatchLaunch {
  bool running = false;
  uint32_t S = 0;
  uint64_t arena_base = 0;
  uint64_t arena_end = 0;
  std::vector<int> segs;
};

struct PendingTiming {
  std::string name;
  hipEvent_t a, b;
  double bytes;
};

}  // namespace wk

struct wk_table {
  uint32_t* d = nullptr;
  uint64_t rows = 0;
  uint32_t cols = 0;
};

namespace wk {

// What one lane of work owns: a stream, its history tables and result
// arena, per-row scratch and step descriptors.  The device has a main lane
// and, for serving, a second one that runs heavy queries concurrently on its
// own stream; the gstore, its cache and the direct indexes are shared.
struct LaneState {
  hipStream_t stream = nullptr;
  bool copy_pending = false;  // copy-stream work not yet waited for by this lane's stream

  // tables and scratch
  uint32_t* tbuf[2] = {nullptr, nullptr};
  uint64_t tcap = 0;
  uint32_t* arena = nullptr;
  uint64_t arena_cap = 0;
  uint32_t* eoff = nullptr;
  uint64_t* pfx = nullptr;  // a step's prefixes within tiles (StepArgs.lpfx); scratch of the split
  uint64_t rcap = 0;
  uint64_t* tile_sum = nullptr;
  uint64_t tile_cap = 0;
  uint64_t* tile_row = nullptr;
  uint64_t otile_cap = 0;

  // step descriptors: device arrays sized for nsteps_cap x nq_cap
  StepDesc* d_desc = nullptr;
  uint64_t* d_rs = nullptr;
  StepState* d_state = nullptr;
  StepDesc* h_desc = nullptr;   // pinned staging
  uint64_t* h_rs = nullptr;
  StepState* h_state = nullptr;
  uint64_t* d_qres = nullptr;       // [nq_cap][2] final rows / arena offset per query
  uint64_t* h_qres = nullptr;       // pinned staging
  uint64_t* d_qagg = nullptr;       // [ceil(nq_cap / 256)][3] query-scan block sums
  uint64_t* d_ck = nullptr;         // [nq_cap] row-hash sums of final tables
  uint64_t* h_ck = nullptr;         // pinned staging
  uint32_t nq_cap = 0, steps_cap = 0;

  // hipGraph cache for batches: key (steps, nq)
  std::map<std::pair<uint32_t, uint32_t>, hipGraphExec_t> graphs;

  // chunk boundaries of a step run in output windows: [kWinGroup][3]
  uint64_t* d_win = nullptr;
  uint64_t* h_win = nullptr;  // pinned mirror

  // compiled plans of the lane's last batch (storage reused by the next)
  std::vector<Plan> plan_cache;
  // the plan of a query run by itself (run_one)
  std::vector<Plan> one_plan;
  // elements of the arena holding the last batch's results
  uint64_t batch_used = 0;

  // A batch submitted on this lane (wk_batch_submit) whose GPU work may
  // still run: its launch state and the caller's result arrays, which
  // finish_batch fills.  `ticket` stays the lane's last submitted ticket
  // until its next submit, whatever else runs on the lane in between, so
  // that its wait still answers with the status it finished with; the lane
  // takes no new submit before that wait.
  bool inflight = false;
  bool fetchable = false;  // the arena still holds that batch's tables
  bool awaited = true;     // the caller has waited for the lane's last ticket
  int done_rc = 0;         // what the lane's last batch finished with
  uint32_t ticket = 0;
  uint32_t b_nq = 0;
  bool b_graph = false;
  BatchLaunch launch;
  wk_batch_result b_out{};
};

}  // namespace wk

struct wk_dev : wk::LaneState {
  const wk::Store* st = nullptr;
  int device = 0;
  hipStream_t copy_stream = nullptr;
  wk_dev_opts opts{};
  // the other lanes: one is swapped with the device's own fields while its
  // work is issued.  alt[0] carries the serving loop's heavy queries; the
  // pipelined batches rotate over the device's own lane and the first
  // opts.batch_lanes - 1 of these.
  wk::LaneState alt[wk::kMaxLanes - 1];

  // gstore cache
  uint32_t kshift = 14, vshift = 19;
  uint32_t nkblk = 0, nvblk = 0;
  bool swapping = false;  // the gstore budget holds fewer blocks than the store
  wk::Slot* karena = nullptr;
  uint32_t* varena = nullptr;
  std::vector<uint32_t> free_k, free_v;
  std::vector<wk::SegSlot> seg;
  // direct key indexes (null where a segment has none); built on a
  // segment's first load and kept across evictions (they hold ptr words, not
  // arena places)
  std::vector<uint64_t*> dense;
  std::vector<uint8_t> dense_built;
  // instance bitmaps of the types (OP_FILTER_TYPE): slot per id, device
  // descriptors by slot, and their word arrays
  std::vector<int32_t> type_slot;
  wk::TypeBits* d_tbits = nullptr;
  std::vector<uint32_t*> tbits_words;
  wk::DevSeg* h_segs = nullptr;   // pinned mirror
  wk::DevSeg* d_segs = nullptr;
  uint32_t* h_ktab = nullptr;     // pinned mirror
  uint32_t* h_vtab = nullptr;
  uint32_t* d_ktab = nullptr;
  uint32_t* d_vtab = nullptr;
  uint64_t ktab_n = 0, vtab_n = 0;
  uint64_t tick = 0;
  hipEvent_t copy_done = nullptr, compute_done = nullptr;
  // pinned staging ring for swaps from a pageable host store
  uint8_t* stage[wk::kStageBufs] = {};
  hipEvent_t stage_ev[wk::kStageBufs] = {};
  uint8_t stage_used[wk::kStageBufs] = {};
  int stage_next = 0;
  uint64_t stage_bytes = 0;
  // pinned host copies of segments larger than the gstore budget, made when
  // the store itself is pageable; they are read in place by the kernels
  std::vector<void*> mapped_host;

  // last per-query result (in the result arena)
  const uint32_t* last_ptr = nullptr;
  uint64_t last_rows = 0;
  uint32_t last_cols = 0;
  // tickets of pipelined batches (wk_batch_submit), and the lane the next
  // one prefers (0: this lane, 1: the other)
  uint32_t next_ticket = 1;
  int next_lane = 0;
  // batched-step traffic accounting (wk_batch_traffic)
  bool traffic_on = false;
  double traffic_bytes = 0, traffic_edges = 0;
  uint64_t traffic_batches = 0;
  // last agent reply bundle, and the agent's query ids
  std::vector<uint8_t> agent_reply;
  int32_t agent_qid = 0;

  // served results kept for checking
  std::vector<std::vector<uint32_t>> served_rows;
  std::vector<uint32_t> served_cols;

  wk_dev_stats stats{};
  std::map<std::string, wk::KernelStat> kstats;
  std::vector<wk::PendingTiming> pending;
  std::vector<hipEvent_t> event_pool;
  hipEvent_t timing_stop = nullptr;  // the open timed_begin's end event
};

namespace wk {

// Exchange the device's own lane fields with dev->alt (the lanes' streams,
// tables, scratch and descriptors).
inline void swap_lanes(wk_dev* dev, int lane = 1) { std::swap(static_cast<LaneState&>(*dev), dev->alt[lane - 1]); }

// Marks every lane as having copy-stream work to wait for.
inline void all_lanes_copy_pending(wk_dev* dev) {
  dev->copy_pending = true;
  for (LaneState& l : dev->alt) l.copy_pending = true;
}

// Tables, scratch and step descriptors of the lane now in the device's own
// fields, for a table budget of `budget` bytes (0: 8 GiB); its stream exists.
int alloc_lane(wk_dev* dev, uint64_t budget);
void free_lane(wk::LaneState& l);

// Make `segs` resident (pinned against eviction while in use by the launch
// sequence that follows); the compute stream waits for their copies.
int ensure_segments(wk_dev* dev, const std::vector<int>& segs);
void unpin_segments(wk_dev* dev, const std::vector<int>& segs);
int prefetch_segment(wk_dev* dev, int seg, const std::vector<int>& in_use);

// Device address of the no-allocation entry points' fixed scratch on
// `device` (the current device), looked up once per device.
int noalloc_init(int device);

// Capacity management for step descriptors.
int reserve_steps(wk_dev* dev, uint32_t steps, uint32_t nq);

StepArgs make_args(wk_dev* dev, uint32_t step, uint32_t nq, const uint32_t* tin, uint32_t* tout,
                   uint64_t tcap, bool has_next);

// Timing helpers (active when opts.profile).
void timed_begin(wk_dev* dev, hipEvent_t* ev);
void timed_end(wk_dev* dev, hipEvent_t ev, const char* name, double bytes);
// Host-side phase timing (active when opts.profile): adds the milliseconds
// since `t0` (steady clock, seconds) to the stat `name` and returns now.
double host_phase(wk_dev* dev, const char* name, double t0);
double host_now();
int flush_timings(wk_dev* dev);

// One step of one query over R (<= rcap) input rows at `tin`, step segment
// resident: probe and scans, then the step's output rows come back to the
// host (one sync) in *total.  The per-row prefix stays in the scratch for
// gather_window until the next step reuses it.
int probe_scan(wk_dev* dev, const StepDesc& sd, const uint32_t* tin, uint64_t R, uint64_t* total);
// Output rows [j0, j1) of the step probe_scan ran last, written to `dst`
// (row j at dst + (j - j0) * out_cols), in as many launches as the output
// tile map needs.  No sync.
int gather_window(wk_dev* dev, const StepDesc& sd, const uint32_t* tin, uint64_t R, uint32_t* dst, uint64_t j0,
                  uint64_t j1);

// Per-query execution, one host sync per pattern, depth first: the final
// table goes to `dst` (capacity `cap` elements), intermediate tables to the
// table pool; a step whose output does not fit the pool's room runs in
// output chunks, each carried through the remaining steps.  WK_E_NOMEM_BUDGET
// only when the result does not fit `dst` or one output row does not fit the
// pool.
int run_query_to(wk_dev* dev, const Plan& plan, uint32_t* dst, uint64_t cap, uint64_t* rows);

// Lane `lane`'s (1 .. kMaxLanes - 1) stream, tables and scratch (same table
// budget), made on first use.
int ensure_alt_lane(wk_dev* dev, int lane = 1);
// Completes every batch submitted with wk_batch_submit and not yet waited
// for (their result arrays are filled); every entry point that uses a lane
// calls it first.  The arenas of lanes [0, lanes_used), which the caller is
// about to reuse, can no longer be fetched by ticket.
int drain_batches(wk_dev* dev, int lanes_used = 1);

// Batched execution of queries [q0, q1) of `plans`, in two halves:
// begin_batch fills and uploads the step descriptors, makes the segments
// resident and launches every pattern step on the lane's stream without
// waiting; finish_batch waits for them, releases the segments and fills
// rows/offsets/status per query, rerunning the batch in halves when a step
// overflowed.  Results land in the arena from `arena_base`.
int begin_batch(wk_dev* dev, const std::vector<Plan>& plans, uint32_t q0, uint32_t q1, uint64_t arena_base,
                bool use_graph, uint64_t* rows, uint64_t* offsets, int32_t* status, BatchLaunch& bl);
int finish_batch(wk_dev* dev, const std::vector<Plan>& plans, uint32_t q0, uint32_t q1, bool use_graph,
                 uint64_t* rows, uint64_t* offsets, int32_t* status, BatchLaunch& bl);
// begin_batch then finish_batch.
int run_batch(wk_dev* dev, const std::vector<Plan>& plans, uint32_t q0, uint32_t q1, uint64_t arena_base,
              bool use_graph, uint64_t* rows, uint64_t* offsets, int32_t* status, uint64_t* arena_end);
// Queries [q0, q1) as two batches, one after the other.
int split_batch(wk_dev* dev, const std::vector<Plan>& plans, uint32_t q0, uint32_t q1, uint64_t arena_base,
                bool use_graph, uint64_t* rows, uint64_t* offsets, int32_t* status, uint64_t* arena_end);
// One query by itself, its result at the start of the lane's result arena
// and its rows in *rows; returns the query's status (or WK_E_HIP).  Its steps
// run from device-side sizes with one host round trip: the batch path with
// one query, whose overflowing step falls back to the chunked executor.  With
// per-kernel profiling on, the step-by-step path runs instead, whose launches
// carry the host's row counts and algorithmic bytes.
int run_one(wk_dev* dev, const Plan& plan, uint64_t* rows);
// One query of a batch by itself (per-query path, chunked as needed), its
// result into the arena at arena_base.
int run_alone(wk_dev* dev, const Plan& plan, uint64_t arena_base, uint64_t* rows, uint64_t* offset, int32_t* status,
              uint64_t* arena_end);

}  // namespace wk
