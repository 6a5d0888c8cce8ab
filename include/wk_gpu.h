/*
 * wk_gpu.h -- C ABI of the MI355X-native Wukong GPU query path.
 *
 * This is the drop-in boundary that replaces the reference's GPU extension
 * (src/gpu/).  Every entry point takes plain pointers and sizes; no C++ or
 * HIP types cross it.  A handle is used from one host thread at a time;
 * independent handles are independent.  All calls return a status code
 * (0 = WK_OK); nothing aborts.  Tables are row-major uint32 `[rows][cols]`,
 * the reference's history-table layout (core/sparql/query.hpp:392-397).
 *
 * Reference interface each group replaces:
 *   store   : SegmentRDFGraph::init_gstore / get_triples / get_index
 *             (core/store/segment_rdf_dgraph.hpp:1114-1129,1182-1269)
 *   device  : GPUMem + GPUCache + GPUStreamPool set-up in main()
 *             (src/wukong.cpp:182-190,280-295; src/gpu/gpu_mem.hpp:38-139;
 *             src/gpu/gpu_cache.hpp:428-470)
 *   segment : GPUCache::load_segment / prefetch_segment
 *             (src/gpu/gpu_cache.hpp:551-559)
 *   table   : GPUEngine::load_result_buf, GPUEngineCuda::load_result_buf
 *             and the final thrust::copy D2H
 *             (src/gpu/gpu_engine.hpp:248-263; src/gpu/gpu_engine_cuda.hpp:90-106,190-193)
 *   pattern : GPUEngineCuda::known_to_unknown / known_to_known / known_to_const
 *             (src/gpu/gpu_engine_cuda.hpp:113-198,200-282,284-363) and the
 *             CPU-side starts GPUEngine::index_to_unknown / const_to_unknown
 *             (src/gpu/gpu_engine.hpp:65-125), which run on the device here
 *   query   : GPUEngine::execute_one_pattern looped by
 *             GPUAgent::execute_sparql_query (src/gpu/gpu_engine.hpp:265-338;
 *             src/gpu/gpu_agent.hpp:173-246)
 *   wire    : Bundle(const SPARQLQuery&) / Bundle::get_sparql_query and the
 *             GPU agent's request-to-reply handling
 *             (src/core/common/bundle.hpp:64-70,105-115;
 *             src/gpu/gpu_agent.hpp:115-125,173-244)
 */
#ifndef WK_GPU_H
#define WK_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  0..13 keep the numbering of core/common/errors.hpp:30-45 so
 * a reply's status_code means the same thing on both engines. */
enum {
  WK_OK = 0,
  WK_E_UNKNOWN = 1,
  WK_E_SYNTAX = 2,
  WK_E_UNKNOWN_PATTERN = 3,
  WK_E_ATTR_DISABLE = 4,
  WK_E_NO_REQUIRED_VAR = 5,
  WK_E_UNSUPPORT_UNION = 6,
  WK_E_OBJ_ERROR = 7,
  WK_E_VERTEX_INVALID = 8,
  WK_E_UNKNOWN_SUB = 9,
  WK_E_SETTING = 10,
  WK_E_FIRST_PATTERN = 11,
  WK_E_UNKNOWN_FILTER = 12,
  WK_E_FILE_NOT_FOUND = 13,
  /* codes of this library (no counterpart in the reference, which aborts) */
  WK_E_INVALID_ARG = 100,     /* null handle, bad column, bad direction ...  */
  WK_E_INVALID_TRIPLE = 101,  /* triple violates the id conventions          */
  WK_E_NOMEM_HOST = 102,      /* host allocation failed                      */
  WK_E_NOMEM_BUDGET = 103,    /* device budget (gstore or tables) exceeded   */
  WK_E_HIP = 104,             /* a HIP runtime call failed; see wk_last_hip_error */
  WK_E_NO_DEVICE = 105,       /* no usable GPU                               */
  WK_E_OVERFLOW = 106,        /* a count exceeds the 64-bit/28-bit fields    */
  WK_E_SEG_TOO_LARGE = 107,   /* a segment exceeds the gstore budget and no
                                 pinned host copy of it could be made         */
  WK_E_STATE = 108,           /* call out of order (e.g. no result yet)      */
  WK_E_TOO_MANY_VARS = 109,   /* more variables than WK_MAX_VARS             */
  WK_E_WIRE = 110,            /* malformed archive, or a form not read here  */
  WK_E_SCRATCH_LIMIT = 111    /* input exceeds a fixed scratch area          */
};

#define WK_IN 0
#define WK_OUT 1
#define WK_PREDICATE_ID 0
#define WK_TYPE_ID 1
#define WK_MAX_VARS 64

typedef struct wk_store wk_store;
typedef struct wk_dev wk_dev;
typedef struct wk_table wk_table;

/* A triple pattern, SPARQLQuery::Pattern (core/sparql/query.hpp:145-170):
 * negative ids are variables -1, -2, ...; direction is WK_IN or WK_OUT. */
typedef struct {
  int32_t subject;
  int32_t predicate;
  int32_t object;
  int32_t direction;
} wk_pattern;

/* A BASIC pattern group already in plan order (after the .fmt rewrite of
 * Planner::set_direction, optimizer/planner.hpp:667-699). */
typedef struct {
  const wk_pattern* patterns;
  uint32_t npatterns;
  int32_t nvars;
} wk_query;

/* ------------------------------------------------------------------ store */

typedef struct {
  int32_t pinned;          /* 1: allocate slots/values in pinned host memory */
  int32_t load_factor_pct; /* #keys/#slots target; <=0 -> 55 (global.hpp:127) */
} wk_store_opts;

typedef struct {
  uint64_t num_triples_out; /* distinct (s,p,o) after the pso sort+dedup */
  uint64_t num_triples_in;  /* distinct (s,p,o) after the pos sort+dedup */
  uint64_t num_keys;
  uint64_t num_buckets;     /* main + extended, all segments */
  uint64_t num_values;      /* edges, all segments */
  uint32_t num_segments;
  int32_t pinned;
} wk_store_info;

typedef struct {
  uint32_t index;           /* 1 = index segment */
  uint32_t pid;
  uint32_t dir;
  uint32_t pad;
  uint64_t num_keys;
  uint64_t num_buckets;     /* main buckets: hash space of the segment */
  uint64_t num_ext_buckets; /* chained buckets after the main ones */
  uint64_t bucket_start;    /* global bucket offset of bucket 0 */
  uint64_t num_edges;
  uint64_t edge_start;      /* global edge offset of the segment's values */
} wk_seg_info;

/* SegmentRDFGraph::init_gstore(triple_pso, triple_pos, ...)
 * (core/store/segment_rdf_dgraph.hpp:1182-1269).  `pso` and `pos` are flat
 * arrays of (s,p,o) uint32 triples; they may be the same array. */
int wk_store_build(const uint32_t* pso, uint64_t n_pso, const uint32_t* pos, uint64_t n_pos,
                   const wk_store_opts* opts, wk_store** out);
/* BaseLoader::load's normal triples (loader/base_loader.hpp:536-576) from a
 * POSIX directory (loader/posix_loader.hpp:49-69): every "id_*" file, read
 * as whitespace-separated "s p o" ids until the first read that fails (an
 * incomplete last triple is dropped).  Server `server` of `nservers` keeps
 * the triples with s % nservers == server as out-edges and those with
 * o % nservers == server as in-edges (PARTITION, core/store/vertex.hpp:59),
 * then builds its store as wk_store_build does.  `triples_read` (may be
 * NULL) receives the number of triples read from all files.
 * WK_E_FILE_NOT_FOUND: the directory or a file cannot be read. */
int wk_store_load_dir(const char* dir, uint32_t server, uint32_t nservers, const wk_store_opts* opts,
                      wk_store** out, uint64_t* triples_read);
void wk_store_destroy(wk_store* st);
int wk_store_info_get(const wk_store* st, wk_store_info* info);
int wk_store_segment(const wk_store* st, uint32_t i, wk_seg_info* seg);
/* Raw arrays: `slots` = 16-byte {key,ptr} pairs, 8 per bucket. */
int wk_store_raw(const wk_store* st, const void** slots, uint64_t* num_buckets,
                 const uint32_t** values, uint64_t* num_values);
/* DGraph::get_triples / get_index (segment_rdf_dgraph.hpp:1114-1129). */
int wk_store_get_triples(const wk_store* st, uint32_t vid, uint32_t pid, int32_t dir,
                         const uint32_t** edges, uint64_t* size);
int wk_store_get_index(const wk_store* st, uint32_t tpid, int32_t dir,
                       const uint32_t** edges, uint64_t* size);

/* ----------------------------------------------------------------- device */

typedef struct {
  int32_t device;             /* HIP device ordinal */
  uint64_t gstore_budget;     /* bytes of HBM for the gstore cache; 0 -> whole store */
  uint64_t table_budget;      /* bytes for history tables + scratch; 0 -> 8 GiB */
  uint32_t key_block_shift;   /* log2 buckets per key block; 0 -> 14 (2 MiB) */
  uint32_t value_block_shift; /* log2 edges per value block; 0 -> 19 (2 MiB) */
  int32_t enable_pipeline;    /* prefetch next pattern's segment (global.hpp:136) */
  int32_t probe_width;        /* lanes sharing one key's bucket probe: 1,2,4,8 (0 -> 2) */
  int32_t rows_per_wave;      /* gather mapping: 1,4,16,64 or 0 = balanced search */
  int32_t profile;            /* 1: time every kernel with HIP events */
  int32_t dense_index;        /* direct key indexes of segments whose key vids are dense
                                 (at least one key per 16 vids of the segment's range; one 8-B
                                 ptr word per vid, probed instead of the buckets): 0 = build
                                 them within a quarter of free HBM (and of gstore_budget when
                                 set), most keys first; -1 = never */
  int32_t batch_lanes;        /* lanes pipelined batches rotate over (wk_batch_submit), each
                                 with its own stream, tables and arena: 1..4 (0 -> 2) */
} wk_dev_opts;

typedef struct {
  uint64_t h2d_bytes;         /* gstore bytes swapped in */
  uint64_t segments_loaded;
  uint64_t segments_evicted;
  uint64_t key_blocks_total, key_blocks_free;
  uint64_t value_blocks_total, value_blocks_free;
  uint64_t table_capacity_elems;
  uint64_t queries_run;
  uint64_t dense_segments;    /* segments with a direct key index */
  uint64_t dense_bytes;
  uint64_t mapped_segments;   /* segments larger than the gstore budget, read in place
                                 from pinned host memory */
  uint64_t mapped_bytes;      /* their bytes (copied to pinned memory first when the store
                                 is pageable) */
} wk_dev_stats;

/* Per-kernel timing (opts.profile = 1): name, launches, total ms on the
 * library's stream, algorithmic bytes moved. */
typedef struct {
  char name[32];
  uint64_t launches;
  double total_ms;
  double alg_bytes;
} wk_kernel_stat;

int wk_dev_create(const wk_store* st, const wk_dev_opts* opts, wk_dev** out);
void wk_dev_destroy(wk_dev* dev);
int wk_dev_stats_get(wk_dev* dev, wk_dev_stats* stats);
int wk_dev_kernel_stats(wk_dev* dev, wk_kernel_stat* out, uint32_t cap, uint32_t* n);
int wk_dev_reset_kernel_stats(wk_dev* dev);
int wk_dev_set_tuning(wk_dev* dev, int32_t probe_width, int32_t rows_per_wave);
/* Turn per-kernel and host-phase timing (opts.profile) on or off; timings
 * already taken are kept.  No counterpart in the reference: it lets one
 * handle time single queries and then serve without event overhead. */
int wk_dev_set_profile(wk_dev* dev, int32_t on);
int wk_dev_sync(wk_dev* dev);
/* Load every segment that fits the budget (warm cache). */
int wk_dev_load_all(wk_dev* dev);
/* GPUCache::load_segment of (index, pid, dir). */
int wk_seg_ensure(wk_dev* dev, uint32_t index, uint32_t pid, int32_t dir);
int wk_seg_resident(wk_dev* dev, uint32_t index, uint32_t pid, int32_t dir, int32_t* resident);

/* ----------------------------------------------------------------- tables */

int wk_table_upload(wk_dev* dev, const uint32_t* host, uint64_t rows, uint32_t cols, wk_table** out);
int wk_table_shape(const wk_table* t, uint64_t* rows, uint32_t* cols);
int wk_table_download(wk_dev* dev, const wk_table* t, uint32_t* dst, uint64_t cap_elems);
int wk_table_free(wk_dev* dev, wk_table* t);
/* Fork-join dispatch of a history table among nparts owners (1..64): out
 * holds in's rows grouped by owner row[col] % nparts, in row order within
 * each owner, owner p at rows [heads[p], heads[p] + sizes[p]).  Replaces
 * SPARQLEngine::generate_sub_query's split (core/engine/sparql.hpp:741-793,
 * PARTITION = vid % num_servers, core/store/vertex.hpp:59) and the GPU
 * engine's gpu_shuffle_result_buf / gpu_split_result_buf
 * (src/gpu/gpu_hash.cu:601-762; GPUEngine::generate_sub_query,
 * src/gpu/gpu_engine.hpp:340-393), whose atomic positions lose the order.
 * WK_E_INVALID_ARG: nparts outside 1..64, or col >= cols of a non-empty
 * table.  WK_E_NOMEM_BUDGET: the owner x tile histogram does not fit the
 * device's row scratch, or the output table cannot be allocated. */
int wk_table_partition(wk_dev* dev, const wk_table* in, uint32_t col, uint32_t nparts, wk_table** out,
                       uint64_t* sizes, uint64_t* heads);
/* wk_table_partition with no allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid it).  `in` and `out` are the
 * caller's device buffers of rows x cols uint32 (out must not overlap in);
 * the owner x tile histogram lives in a fixed 8 MiB device array per GPU,
 * set up by wk_dev_create, so the call makes no device or host allocation.
 * nparts * ceil(rows / 1024) must stay below about 10^6 (64 owners over 16M
 * rows, 8 owners over 128M rows), else WK_E_SCRATCH_LIMIT.  Calls that
 * overlap on one GPU share that array and race.  Returns once the split is
 * done (sizes and heads as wk_table_partition).  WK_E_INVALID_ARG: nparts
 * outside 1..64, col >= cols, or null buffers for a non-empty table. */
int wk_table_partition_noalloc(wk_dev* dev, const uint32_t* in, uint64_t rows, uint32_t cols, uint32_t col,
                               uint32_t nparts, uint32_t* out, uint64_t* sizes, uint64_t* heads);
/* SPARQLEngine::final_process (core/engine/sparql.hpp:1419-1546) on a device
 * table: DISTINCT (rows sorted on all columns as signed words, a row dropped
 * when it equals its predecessor on every required variable), ORDER BY
 * (order_vars, order_desc[i] != 0 for DESC; ids compared as unsigned, ties
 * in full-row order), OFFSET, LIMIT (< 0: none), then projection onto
 * `required`.  v2c[v] is the column of variable -(v + 1), -1 for none.  The
 * reference's GPU agent skips this step (gpu/gpu_agent.hpp:220-224); its CPU
 * engine orders by the ids' strings, which this library does not hold.  An
 * empty table comes back empty with its columns, unchecked.
 * WK_E_NO_REQUIRED_VAR: nreq == 0.  WK_E_VERTEX_INVALID: a variable has no
 * column.  WK_E_TOO_MANY_VARS: nreq or norder > WK_MAX_VARS. */
int wk_table_final_process(wk_dev* dev, const wk_table* in, const int32_t* v2c, int32_t nvars,
                           const int32_t* required, uint32_t nreq, int32_t distinct, const int32_t* order_vars,
                           const int32_t* order_desc, uint32_t norder, int64_t offset, int64_t limit,
                           wk_table** out);

/* --------------------------------------------------------------- patterns */

/* index_to_unknown: ?X from the index list [0|tpid|dir] (1 column). */
int wk_start_i2u(wk_dev* dev, uint32_t tpid, int32_t dir, wk_table** out);
/* const_to_unknown: ?X from [vid|pid|dir] (1 column). */
int wk_start_c2u(wk_dev* dev, uint32_t vid, uint32_t pid, int32_t dir, wk_table** out);
/* known_to_unknown: append column of neighbours of row[start_col]. */
int wk_expand_k2u(wk_dev* dev, const wk_table* in, uint32_t start_col, uint32_t pid, int32_t dir,
                  wk_table** out);
/* known_to_known: keep rows with row[end_col] in neighbours of row[start_col]. */
int wk_expand_k2k(wk_dev* dev, const wk_table* in, uint32_t start_col, uint32_t pid, int32_t dir,
                  uint32_t end_col, wk_table** out);
/* known_to_const: keep rows with `end` in neighbours of row[start_col]. */
int wk_expand_k2c(wk_dev* dev, const wk_table* in, uint32_t start_col, uint32_t pid, int32_t dir,
                  uint32_t end, wk_table** out);

/* const_to_known: keep rows whose row[end_col] is in the edges of
 * [vid|pid|dir] (core/engine/sparql.hpp:139-188); with vid = 0 the edges are
 * the index list [0|pid|dir], the index_to_known of a request that arrives
 * with a table (sparql.hpp:83-137).  Each row is kept at most once. */
int wk_expand_c2k(wk_dev* dev, const wk_table* in, uint32_t vid, uint32_t pid, int32_t dir, uint32_t end_col,
                  wk_table** out);

/* ------------------------------------------------------------------ query */

/* Run one query (per-query GPU execution; the reference's loop of one
 * execute_one_pattern per pattern).  Its pattern steps run from sizes kept on
 * the device, with one host round trip for the query; a step whose tables
 * exceed the capacity runs in chunks, depth-first, with a host sync per
 * chunk.  With opts.profile on, every step syncs, so that each launch is
 * timed with its row counts.  On success the history table stays on the
 * device as the handle's last result: shape and var->column map are returned
 * here, rows are fetched with wk_result_download. */
int wk_query_run(wk_dev* dev, const wk_query* q, uint64_t* rows, uint32_t* cols, int32_t* v2c);
int wk_result_download(wk_dev* dev, uint32_t* dst, uint64_t cap_elems);

/* ------------------------------------------------------------------ batch */

typedef struct {
  uint64_t* rows;     /* [nq] result rows */
  uint32_t* cols;     /* [nq] result columns */
  int32_t* status;    /* [nq] per-query status code */
  uint64_t* offset;   /* [nq] element offset of the table in the batch arena */
  int32_t* v2c;       /* [nq * WK_MAX_VARS] var->col maps (-1 = unbound): row i holds query
                         i's map in its first nvars entries; the rest are not written */
} wk_batch_result;

/* Run nq queries together: each pattern step of all queries is one set of
 * launches over a concatenated table with per-query offsets, captured in a
 * hipGraph when opts allow.  Result tables stay in the batch arena, in query
 * order; a query whose result table does not fit the rest of the arena
 * reports WK_E_NOMEM_BUDGET and the others go on.  A step whose intermediate
 * tables overflow is rerun in halves, down to single queries. */
int wk_batch_run(wk_dev* dev, const wk_query* qs, uint32_t nq, int32_t use_graph,
                 wk_batch_result* out);
int wk_batch_download(wk_dev* dev, uint32_t* dst, uint64_t cap_elems, uint64_t* used_elems);
/* Pipelined batches (no counterpart in the reference, whose agent runs one
 * query at a time): wk_batch_submit compiles, uploads and launches a batch
 * as wk_batch_run does and returns without waiting for the GPU, so that the
 * caller can prepare and submit the next batch while this one runs.  Up to
 * opts.batch_lanes batches (2 by default) can be in flight at once, each on
 * its own lane (stream, history tables, scratch and result arena, allocated
 * with the handle's table budget on a lane's first use).  A lane takes a new
 * batch only once the ticket of its last one has been waited for, so one
 * more submit returns WK_E_STATE.
 * `out` must stay valid until the batch is waited for: cols and v2c are
 * written by the submit, rows, status and offset by wk_batch_wait, or by any
 * other call on the handle that uses a lane, which first completes the
 * batches in flight.  The wait of a batch completed that way returns the
 * status it completed with.  wk_batch_fetch copies a waited-for ticket's result
 * arena (offsets as in `out`) to the host until its lane starts another
 * batch or another call uses the lanes (WK_E_STATE after that).  Under a
 * gstore budget smaller than the store, a submit first completes the batches
 * on the other lanes, so that their pinned segments cannot starve this one. */
int wk_batch_submit(wk_dev* dev, const wk_query* qs, uint32_t nq, int32_t use_graph, wk_batch_result* out,
                    uint32_t* ticket);
int wk_batch_wait(wk_dev* dev, uint32_t ticket);
int wk_batch_fetch(wk_dev* dev, uint32_t ticket, uint32_t* dst, uint64_t cap_elems, uint64_t* used_elems);
/* Measurement (no counterpart in the reference).  With enable != 0, each
 * later wk_batch_run reads its pattern steps' descriptors back once it has
 * completed and adds their algorithmic bytes to a running total, by the
 * per-pattern formulas over the device-side row counts (R input rows of C
 * columns, E output rows of Co columns): an index or constant start
 * 16 + 8E; known_to_unknown 4RC + 16R + 4E + 4E*Co; known_to_known and
 * known_to_const 4RC + 16R + 4R + 4E*Co (one edge word per membership
 * test).  `edges` adds E for expansions and R for membership tests.  Returns
 * the totals since the last call and resets them; batches counts the
 * batched launches (sub-batches included) they cover. */
int wk_batch_traffic(wk_dev* dev, int32_t enable, double* alg_bytes, double* edges, uint64_t* batches);
/* Autotuner (no counterpart in the reference): times the batch of nq
 * queries under every probe width {1,2,4,8} x gather mapping {0,1,4,16,64}
 * (one warm-up run, then the median of three) and keeps the fastest on the
 * handle.  Outputs may be null. */
int wk_dev_autotune(wk_dev* dev, const wk_query* qs, uint32_t nq, int32_t use_graph, int32_t* probe_width,
                    int32_t* rows_per_wave, double* batch_ms);

/* ---------------------------------------------------------------- serving */

typedef struct {
  uint32_t max_inflight;     /* queries admitted at once */
  uint64_t footprint_budget; /* bytes of projected history tables admitted */
  int32_t per_query;         /* 1: FIFO, one query at a time (per-query baseline) */
  int32_t keep_results;      /* 1: keep each result's rows (for checking) */
  uint64_t heavy_footprint;  /* queries whose projected history tables reach this many bytes
                                run as batches on a second lane (its own stream, tables and
                                scratch, table_budget again), concurrently with the light
                                queries' iterations: projected at admission, and again from
                                the table each main-lane step made (the query then continues
                                on the lane from a copy of that table); 0 -> 1 MiB,
                                UINT64_MAX -> no second lane */
} wk_serve_opts;

typedef struct {
  double* latency_us;        /* [nq] completion - arrival */
  double* start_us;          /* [nq] admission - arrival */
  uint64_t* rows;            /* [nq] */
  int32_t* status;           /* [nq] */
  uint64_t* checksum;        /* [nq] order-independent hash of the rows */
  double wall_s;             /* first arrival -> last completion */
  uint64_t batches;          /* scheduler iterations that launched work */
} wk_serve_result;

/* Open-loop replay: query i arrives at arrival_s[i] seconds after the start
 * (non-decreasing).  A native scheduler admits arrived queries as slots and
 * footprint free up and advances every in-flight query by one pattern step
 * per iteration (continuous batching). */
int wk_serve_replay(wk_dev* dev, const wk_query* qs, const double* arrival_s, uint32_t nq,
                    const wk_serve_opts* opts, wk_serve_result* out);
/* Rows of query i kept by the last replay with keep_results = 1. */
int wk_serve_result_rows(wk_dev* dev, uint32_t i, uint32_t* dst, uint64_t cap_elems,
                         uint64_t* rows, uint32_t* cols);

/* Order-independent checksum of a host table (same function the serving
 * path uses): sum over rows of hash_u64 of a per-row fold. */
uint64_t wk_table_checksum(const uint32_t* table, uint64_t rows, uint32_t cols);

/* ------------------------------------------------------------------- wire */
/* SPARQLQuery as the reference's proxy and agents exchange it: a Bundle
 * string (Bundle::to_str / init, core/common/bundle.hpp:96-103,155-164) is
 * the 4-byte req_type SPARQL_QUERY = 0 (core/common/type.hpp:52-55) followed
 * by a boost binary_oarchive of the query (bundle.hpp:64-70; boost 1.67.0,
 * deps/deps.sh:5) with the member order of core/sparql/query.hpp:1086-1276,
 * every class object_serializable (query.hpp:1288-1296), TRDF_MODE off.
 * Only BASIC pattern groups cross: an archive that holds FILTERs, OPTIONAL
 * or UNION groups, or attribute results, is refused on decode with the
 * reference's status code for them.  Ids and enums as in the reference:
 * state SQ_PATTERN = 0 .. SQ_REPLY = 5, dev_type CPU = 0 / GPU = 1,
 * job_type FULL_JOB = 0 / SPLIT_JOB = 1 (query.hpp:79-101); v2c_map holds
 * Result's ext words, type << 16 | column, 0xffff for an unbound variable
 * (query.hpp:60-69). */
#define WK_SQ_PATTERN 0
#define WK_SQ_REPLY 5
#define WK_DEV_CPU 0
#define WK_DEV_GPU 1
#define WK_NO_RESULT 0xffff

typedef struct {
  int32_t subject, predicate, object, direction;
  int32_t pred_type;          /* a char in the archive; SID_t = 0 */
} wk_wire_pattern;

typedef struct {
  /* SPARQLQuery scalars, archive order (query.hpp:1210-1229) */
  int32_t q_type, qid, pqid, pg_type, state, dev_type, job_type, priority;
  int32_t mt_factor, mt_tid, pattern_step, local_var;
  int32_t corun_enabled, corun_step, fetch_step, union_done, optional_step;
  int32_t limit;
  uint32_t offset;
  int32_t distinct;
  /* PatternGroup (query.hpp:1113-1132): patterns and optional_new_vars (a
   * std::set, written ascending) */
  const wk_wire_pattern* patterns;
  uint32_t npatterns;
  const int32_t* new_vars;
  uint32_t nnew_vars;
  /* orders (Order: variable id, descending) */
  const int32_t* order_vars;
  const int32_t* order_desc;
  uint32_t norders;
  /* Result (query.hpp:1155-1181) */
  int32_t col_num, row_num, attr_col_num, status_code, blind, nvars;
  const int32_t* required_vars;
  uint32_t nrequired;
  const int32_t* v2c_map;
  uint32_t nv2c;
  const uint8_t* optional_matched;  /* vector<bool>, one byte per row */
  uint64_t noptional_matched;
  const uint32_t* table;            /* result_table; written when row_num > 0 */
  uint64_t ntable;
  uint64_t gpu_nelems;              /* Result::gpu (query.hpp:311-362), USE_GPU builds */
  int32_t gpu_col_num;
} wk_sparql_msg;

typedef struct {
  uint32_t library_version;  /* the archive's boost library version; 0 -> 16 (boost 1.67) */
  int32_t use_gpu;           /* 1: Result carries its GPUResult member (-DUSE_GPU builds) */
} wk_wire_opts;

/* Encode `m` as a Bundle string.  `len` receives the size; with dst NULL or
 * cap too small nothing is written and WK_E_INVALID_ARG is returned after
 * `len` is set (call once to size, once to write). */
int wk_wire_encode(const wk_sparql_msg* m, const wk_wire_opts* o, uint8_t* dst, uint64_t cap, uint64_t* len);
/* Decode a Bundle string.  On success *out points to a message whose arrays
 * the library owns until wk_wire_free.  WK_E_WIRE: malformed or truncated
 * bytes, another req_type, or a library version below 7.
 * WK_E_UNKNOWN_FILTER / WK_E_UNSUPPORT_UNION / WK_E_ATTR_DISABLE: FILTERs,
 * OPTIONAL or UNION groups, attribute results. */
int wk_wire_decode(const uint8_t* src, uint64_t len, const wk_wire_opts* o, wk_sparql_msg** out);
void wk_wire_free(wk_sparql_msg* m);

/* GPUAgent::execute_sparql_query (gpu/gpu_agent.hpp:173-244) on one server,
 * bundle in, bundle out.  The patterns from pattern_step run on the device:
 * from the first pattern (index or constant start), or, when the request
 * carries a history table, from that table uploaded as
 * GPUEngine::load_result_buf does (gpu/gpu_engine.hpp:248-263).  Then, unlike
 * the reference's GPU agent and as its CPU engine does, final_process
 * (core/engine/sparql.hpp:1419-1546, skipped when blind or empty), and
 * send_reply's fields (gpu_agent.hpp:115-125): state SQ_REPLY, FULL_JOB,
 * shrink() when blind (rows and columns kept, table and required_vars
 * dropped).  A qid of -1 gets the handle's next id.  A query this engine
 * cannot run (not dev_type GPU, a reply, unsupported pattern, ...) still
 * gets a reply, whose status_code says why, as the CPU engine reports
 * errors.  The reply is kept on the handle: `reply_len` gets its size and
 * wk_agent_reply copies it out.  Returns WK_E_WIRE for undecodable bytes
 * and device errors (WK_E_HIP, WK_E_NOMEM_BUDGET) as other calls do. */
int wk_agent_execute(wk_dev* dev, const uint8_t* req, uint64_t len, const wk_wire_opts* o, uint64_t* reply_len);
int wk_agent_reply(wk_dev* dev, uint8_t* dst, uint64_t cap);

/* ------------------------------------------------------------ generators */

/* Seeded university-schema generator: flat (s,p,o) triples, caller frees
 * with wk_free.  Ids follow the datagen conventions (datagen/generate_data.cpp:357-373):
 * predicates/types from 2, vertices from 2^17. */
int wk_gen_university(uint32_t seed, uint32_t n_univ, uint32_t** triples, uint64_t* n);
/* Id of a schema name ("rdf:type", "ub:memberOf", "ub:GraduateStudent"...); 0 if unknown. */
uint32_t wk_univ_id(const char* name);
/* Seeded power-law generator (Zipf predicates, Pareto out-degree). */
int wk_gen_powerlaw(uint32_t seed, uint64_t n_vertices, uint64_t n_triples, uint32_t n_preds,
                    uint32_t n_types, double zipf_s, double pareto_alpha, uint64_t max_degree,
                    uint32_t** triples, uint64_t* n);
void wk_free(void* p);

/* ------------------------------------------------------------------ misc */

const char* wk_strerror(int code);
int wk_last_hip_error(void);
int wk_device_count(int* n);

#ifdef __cplusplus
}
#endif

#endif /* WK_GPU_H */
