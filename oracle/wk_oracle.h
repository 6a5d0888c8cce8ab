/*
 * wk_oracle.h -- CPU oracle of the Wukong query path.  TEST INFRASTRUCTURE:
 * only tests/, the smoke check and the benchmark's CPU-baseline leg use it,
 * and only as the checker or as the timed CPU baseline.  The product path
 * never links it.
 *
 * It restates the reference CPU engine (src/core/engine/sparql.hpp) over an
 * independent store representation: sorted, de-duplicated (p,s,o) and
 * (p,o,s) arrays with binary search, instead of the hash store the device
 * path uses, so agreement between the two checks the device store too.
 */
#ifndef WK_ORACLE_H
#define WK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_store orc_store;

/* Thomas Wang's 64-bit hash and its inverse (utils/math.hpp:59-104). */
uint64_t orc_hash_u64(uint64_t key);
uint64_t orc_inverse_hash_u64(uint64_t key);

/* Build from flat (s,p,o) triples for the OUT side (pso) and the IN side
 * (pos), as SegmentRDFGraph::init_gstore receives them.  Returns 0 or 101
 * (invalid triple), 102 (out of memory). */
int orc_store_build(const uint32_t* pso, uint64_t n_pso, const uint32_t* pos, uint64_t n_pos,
                    orc_store** out);
void orc_store_free(orc_store* st);
uint64_t orc_store_num_out(const orc_store* st);
uint64_t orc_store_num_in(const orc_store* st);

/* DGraph::get_triples / get_index (segment_rdf_dgraph.hpp:1114-1129). */
const uint32_t* orc_get_triples(const orc_store* st, uint32_t vid, uint32_t pid, int dir, uint64_t* sz);
const uint32_t* orc_get_index(const orc_store* st, uint32_t tpid, int dir, uint64_t* sz);

/* Execute a planned BASIC pattern list (4 int32 per pattern: subject,
 * predicate, object, direction; negative = variable) the way
 * SPARQLEngine::execute_patterns does on one server with mt_factor 1.
 * On return *table is malloc'ed (free with orc_free), row-major rows x cols;
 * v2c[nvars] is the var->column map (-1 unbound).  Returns the status code
 * the engine would set (core/common/errors.hpp numbering) or 0. */
int orc_run_query(const orc_store* st, const int32_t* pats, uint32_t npat, int32_t nvars, uint32_t** table,
                  uint64_t* rows, uint32_t* cols, int32_t* v2c);

/* The same, from pattern `from` over a carried history table (in_rows x
 * in_cols, var->column map in_v2c[nvars]), as a request that arrives with a
 * table continues (GPUEngine::load_result_buf, gpu/gpu_engine.hpp:248-263);
 * an index pattern at step 0 whose object is bound is index_to_known. */
int orc_run_query_from(const orc_store* st, const int32_t* pats, uint32_t npat, int32_t nvars, uint32_t from,
                       const uint32_t* in, uint64_t in_rows, uint32_t in_cols, const int32_t* in_v2c,
                       uint32_t** table, uint64_t* rows, uint32_t* cols, int32_t* v2c);

/* orc_run_query with a query from an index split over nthreads threads the
 * way mt_factor splits it (sparql.hpp:214-224): thread t takes consecutive
 * part t of the index list (the last also the remainder) and runs the
 * remaining patterns on it; the parts' tables are concatenated in thread
 * order.  Other queries run as orc_run_query. */
int orc_run_query_mt(const orc_store* st, const int32_t* pats, uint32_t npat, int32_t nvars, int nthreads,
                     uint32_t** table, uint64_t* rows, uint32_t* cols, int32_t* v2c);

/* nq queries (query i: patterns [pat_off[i], pat_off[i+1]) of `pats`, nvars[i]
 * variables) on nthreads workers pulling from one shared queue (Engine::run,
 * core/engine/engine.hpp:132-233): row count, order-independent checksum
 * (orc_table_checksum) and status of each. */
int orc_run_stream(const orc_store* st, const int32_t* pats, const uint64_t* pat_off, const int32_t* nvars,
                   uint64_t nq, int nthreads, uint64_t* rows, uint64_t* checksums, int32_t* status);
uint64_t orc_table_checksum(const uint32_t* table, uint64_t rows, uint32_t cols);

/* Planner::set_direction (optimizer/planner.hpp:667-699): rewrite raw
 * patterns by a .fmt plan; dirs: 0 "<", 1 ">", 2 "<<", 3 ">>". */
void orc_set_direction(const int32_t* raw, const int32_t* orders, const int32_t* dirs, uint32_t n,
                       int32_t* out);

/* SPARQLEngine::final_process for a BASIC query (sparql.hpp:1419-1546):
 * project `required` variables, optional DISTINCT, ORDER BY (variables with
 * descending flags), OFFSET/LIMIT (limit < 0: none).  Output malloc'ed. */
int orc_final_process(const uint32_t* table, uint64_t rows, uint32_t cols, const int32_t* v2c, int32_t nvars,
                      const int32_t* required, uint32_t nreq, int distinct, const int32_t* order_vars,
                      const int32_t* order_desc, uint32_t norder, int64_t offset, int64_t limit,
                      uint32_t** out, uint64_t* out_rows);

void orc_free(void* p);

#ifdef __cplusplus
}
#endif

#endif
