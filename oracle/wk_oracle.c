/*
 * wk_oracle.c -- CPU oracle of the Wukong query path.  TEST INFRASTRUCTURE:
 * only tests/, the smoke check and the benchmark's CPU-baseline leg use it.
 *
 * Restates the reference CPU engine over sorted arrays:
 *   store   : loader sort+dedup (loader/base_loader.hpp:90-103,481-498);
 *             OUT keys per (s,p), IN keys per (o,p) skipping type/predicate
 *             objects (core/store/segment_rdf_dgraph.hpp:513-590);
 *             index lists [0|p|IN] subjects of p, [0|p|OUT] objects of p,
 *             [0|t|IN] instances of t (collect_idx_info :951-994,
 *             insert_idx :284-347); lookups get_triples/get_index (:1114-1129)
 *   engine  : index_to_unknown :197-234, const_to_unknown :241-288,
 *             known_to_unknown :298-410, known_to_known :419-479,
 *             known_to_const :487-552, const_to_known :139-188, dispatch
 *             execute_one_pattern :932-1055 (core/engine/sparql.hpp)
 *   planner : Planner::set_direction (optimizer/planner.hpp:667-699)
 *   final   : final_process (core/engine/sparql.hpp:1419-1546), ORDER BY on
 *             ids instead of strings (this path has no string server)
 *   hash    : math::hash_u64 / inverse_hash_u64 (utils/math.hpp:59-104)
 */
#include "wk_oracle.h"

#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <string.h>

#define NBITS_IDX 17
#define TPID_LIMIT (1u << NBITS_IDX)
#define PREDICATE_ID 0u
#define TYPE_ID 1u
#define DIR_IN 0
#define DIR_OUT 1

/* status codes, core/common/errors.hpp:30-45 */
enum { E_OK = 0, E_UNKNOWN_PATTERN = 3, E_OBJ_ERROR = 7, E_VERTEX_INVALID = 8, E_UNKNOWN_SUB = 9,
       E_FIRST_PATTERN = 11, E_INVALID_TRIPLE = 101, E_NOMEM = 102, E_TOO_MANY_VARS = 109 };

static int is_tpid(uint32_t id) { return id > 1 && id < TPID_LIMIT; }

uint64_t orc_hash_u64(uint64_t key) {
  key = (~key) + (key << 21);
  key = key ^ (key >> 24);
  key = (key + (key << 3)) + (key << 8);
  key = key ^ (key >> 14);
  key = (key + (key << 2)) + (key << 4);
  key = key ^ (key >> 28);
  key = key + (key << 31);
  return key;
}

uint64_t orc_inverse_hash_u64(uint64_t key) {
  uint64_t tmp;
  /* undo key += key << 31 */
  tmp = key - (key << 31);
  key = key - (tmp << 31);
  /* undo key ^= key >> 28 */
  tmp = key ^ key >> 28;
  key = key ^ tmp >> 28;
  /* undo key *= 21 */
  key *= 14933078535860113213ull;
  /* undo key ^= key >> 14 */
  tmp = key ^ key >> 14;
  tmp = key ^ tmp >> 14;
  tmp = key ^ tmp >> 14;
  key = key ^ tmp >> 14;
  /* undo key *= 265 */
  key *= 15244667743933553977ull;
  /* undo key ^= key >> 24 */
  tmp = key ^ key >> 24;
  key = key ^ tmp >> 24;
  /* undo key = ~key + (key << 21) */
  tmp = ~key;
  tmp = ~(key - (tmp << 21));
  tmp = ~(key - (tmp << 21));
  key = ~(key - (tmp << 21));
  return key;
}

/* ------------------------------------------------------------------ store */

typedef struct {
  uint32_t a, b, c; /* sort key (a,b,c) */
} trip;

typedef struct {
  uint64_t off, len;
} span;

struct orc_store {
  trip* pso; /* (p, s, o) */
  uint64_t n_pso;
  trip* pos; /* (p, o, s) */
  uint64_t n_pos;
  /* index lists, by id < 2^17 */
  span* pidx_in;  /* distinct subjects of predicate p */
  span* pidx_out; /* distinct non-index objects of predicate p */
  span* tidx;     /* instances of type t */
  uint32_t* ivals;
  uint64_t n_ivals;
};

/* LSD radix sort of triples by (a,b,c), 16-bit digits, stable passes.  Each
 * pass splits the array into one contiguous chunk per thread: per-chunk
 * digit histograms, then chunk t scatters after chunks 0..t-1 of the same
 * digit, which keeps every pass stable. */
static uint32_t trip_field(const trip* x, int field) { return field == 0 ? x->a : field == 1 ? x->b : x->c; }

static int sort_trips(trip* t, uint64_t n) {
  trip* tmp;
  uint64_t* cnt;
  int pass, nt = 1;
  if (n < 2) return 0;
#ifdef _OPENMP
  nt = n > (1u << 20) ? omp_get_max_threads() : 1;
#endif
  tmp = (trip*)malloc(n * sizeof(trip));
  cnt = (uint64_t*)malloc((size_t)nt * 65536 * sizeof(uint64_t));
  if (!tmp || !cnt) {
    free(tmp);
    free(cnt);
    return E_NOMEM;
  }
  for (pass = 0; pass < 6; pass++) {
    const int field = 2 - pass / 2, shift = (pass & 1) * 16;
    uint32_t orv = 0, andv = 0xffffffffu;
    uint64_t i, run = 0;
    int k;
    for (i = 0; i < n; i++) {
      const uint32_t v = trip_field(&t[i], field);
      orv |= v;
      andv &= v;
    }
    if ((((orv ^ andv) >> shift) & 0xffffu) == 0) continue; /* digit constant */
    memset(cnt, 0, (size_t)nt * 65536 * sizeof(uint64_t));
#pragma omp parallel num_threads(nt)
    {
      int tid = 0;
      uint64_t j, lo, hi;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      lo = n * (uint64_t)tid / (uint64_t)nt;
      hi = n * (uint64_t)(tid + 1) / (uint64_t)nt;
      for (j = lo; j < hi; j++) cnt[(size_t)tid * 65536 + ((trip_field(&t[j], field) >> shift) & 0xffffu)]++;
#pragma omp barrier
#pragma omp single
      {
        uint32_t d;
        for (d = 0; d < 65536; d++)
          for (k = 0; k < nt; k++) {
            const uint64_t c = cnt[(size_t)k * 65536 + d];
            cnt[(size_t)k * 65536 + d] = run;
            run += c;
          }
      }
      for (j = lo; j < hi; j++) {
        const uint32_t dg = (trip_field(&t[j], field) >> shift) & 0xffffu;
        tmp[cnt[(size_t)tid * 65536 + dg]++] = t[j];
      }
    }
    memcpy(t, tmp, n * sizeof(trip));
  }
  free(tmp);
  free(cnt);
  return 0;
}

static uint64_t dedup_trips(trip* t, uint64_t n) {
  uint64_t i, m = 0;
  for (i = 0; i < n; i++)
    if (m == 0 || t[i].a != t[m - 1].a || t[i].b != t[m - 1].b || t[i].c != t[m - 1].c) t[m++] = t[i];
  return m;
}

static int valid_triple(uint32_t s, uint32_t p, uint32_t o) {
  if (p == PREDICATE_ID || p >= TPID_LIMIT || s < 2 || o < 2) return 0;
  if (p == TYPE_ID && !is_tpid(o)) return 0;
  return 1;
}

/* first index in [lo,hi) with (a,b) >= (x,y) */
static uint64_t lower2(const trip* t, uint64_t lo, uint64_t hi, uint32_t x, uint32_t y) {
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (t[mid].a < x || (t[mid].a == x && t[mid].b < y)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

static int push_ival(orc_store* st, uint64_t* cap, uint32_t v) {
  if (st->n_ivals == *cap) {
    uint64_t nc = *cap ? *cap * 2 : 1024;
    uint32_t* p = (uint32_t*)realloc(st->ivals, nc * sizeof(uint32_t));
    if (!p) return E_NOMEM;
    st->ivals = p;
    *cap = nc;
  }
  st->ivals[st->n_ivals++] = v;
  return 0;
}

int orc_store_build(const uint32_t* pso, uint64_t n_pso, const uint32_t* pos, uint64_t n_pos, orc_store** out) {
  orc_store* st;
  uint64_t i, cap = 0;
  int rc;
  *out = NULL;
  for (i = 0; i < n_pso; i++)
    if (!valid_triple(pso[3 * i], pso[3 * i + 1], pso[3 * i + 2])) return E_INVALID_TRIPLE;
  for (i = 0; i < n_pos; i++)
    if (!valid_triple(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2])) return E_INVALID_TRIPLE;
  st = (orc_store*)calloc(1, sizeof(orc_store));
  if (!st) return E_NOMEM;
  st->pso = (trip*)malloc((n_pso ? n_pso : 1) * sizeof(trip));
  st->pos = (trip*)malloc((n_pos ? n_pos : 1) * sizeof(trip));
  st->pidx_in = (span*)calloc(TPID_LIMIT, sizeof(span));
  st->pidx_out = (span*)calloc(TPID_LIMIT, sizeof(span));
  st->tidx = (span*)calloc(TPID_LIMIT, sizeof(span));
  if (!st->pso || !st->pos || !st->pidx_in || !st->pidx_out || !st->tidx) {
    orc_store_free(st);
    return E_NOMEM;
  }
  for (i = 0; i < n_pso; i++) {
    st->pso[i].a = pso[3 * i + 1];
    st->pso[i].b = pso[3 * i];
    st->pso[i].c = pso[3 * i + 2];
  }
  for (i = 0; i < n_pos; i++) {
    st->pos[i].a = pos[3 * i + 1];
    st->pos[i].b = pos[3 * i + 2];
    st->pos[i].c = pos[3 * i];
  }
  if ((rc = sort_trips(st->pso, n_pso)) || (rc = sort_trips(st->pos, n_pos))) {
    orc_store_free(st);
    return rc;
  }
  st->n_pso = dedup_trips(st->pso, n_pso);
  st->n_pos = dedup_trips(st->pos, n_pos);

  /* [0|p|IN]: distinct subjects of each predicate except rdf:type */
  for (i = 0; i < st->n_pso; i++) {
    const trip* t = &st->pso[i];
    if (t->a == TYPE_ID) continue;
    if (i == 0 || t->a != st->pso[i - 1].a || t->b != st->pso[i - 1].b) {
      if (st->pidx_in[t->a].len == 0) st->pidx_in[t->a].off = st->n_ivals;
      if ((rc = push_ival(st, &cap, t->b))) {
        orc_store_free(st);
        return rc;
      }
      st->pidx_in[t->a].len++;
    }
  }
  /* [0|p|OUT]: distinct objects of each predicate that are not index ids */
  for (i = 0; i < st->n_pos; i++) {
    const trip* t = &st->pos[i];
    if (is_tpid(t->b)) continue;
    if (i == 0 || t->a != st->pos[i - 1].a || t->b != st->pos[i - 1].b) {
      if (st->pidx_out[t->a].len == 0) st->pidx_out[t->a].off = st->n_ivals;
      if ((rc = push_ival(st, &cap, t->b))) {
        orc_store_free(st);
        return rc;
      }
      st->pidx_out[t->a].len++;
    }
  }
  /* [0|t|IN]: instances of each type, from the OUT side's type triples */
  {
    uint64_t lo = lower2(st->pso, 0, st->n_pso, TYPE_ID, 0), hi = lower2(st->pso, 0, st->n_pso, TYPE_ID + 1, 0);
    uint64_t n = hi - lo, k;
    trip* tt = (trip*)malloc((n ? n : 1) * sizeof(trip));
    if (!tt) {
      orc_store_free(st);
      return E_NOMEM;
    }
    for (k = 0; k < n; k++) {
      tt[k].a = st->pso[lo + k].c; /* type */
      tt[k].b = st->pso[lo + k].b; /* instance */
      tt[k].c = 0;
    }
    sort_trips(tt, n);
    for (k = 0; k < n; k++) {
      if (st->pidx_in[tt[k].a].len) { /* a type id that is also a predicate with subjects */
        free(tt);
        orc_store_free(st);
        return E_INVALID_TRIPLE;
      }
    }
    for (k = 0; k < n; k++) {
      if (st->tidx[tt[k].a].len == 0) st->tidx[tt[k].a].off = st->n_ivals;
      if ((rc = push_ival(st, &cap, tt[k].b))) {
        free(tt);
        orc_store_free(st);
        return rc;
      }
      st->tidx[tt[k].a].len++;
    }
    free(tt);
  }
  *out = st;
  return 0;
}

void orc_store_free(orc_store* st) {
  if (!st) return;
  free(st->pso);
  free(st->pos);
  free(st->pidx_in);
  free(st->pidx_out);
  free(st->tidx);
  free(st->ivals);
  free(st);
}

uint64_t orc_store_num_out(const orc_store* st) { return st->n_pso; }
uint64_t orc_store_num_in(const orc_store* st) { return st->n_pos; }

/* The edge list of key [vid|pid|dir]; edge values are the third sort field,
 * so a run is returned through a strided copy-free view: we materialise it
 * into a per-call scratch buffer owned by the store user (see run_query). */
static uint64_t range_of(const orc_store* st, uint32_t vid, uint32_t pid, int dir, const trip** base) {
  const trip* t = dir == DIR_OUT ? st->pso : st->pos;
  const uint64_t n = dir == DIR_OUT ? st->n_pso : st->n_pos;
  uint64_t lo, hi;
  if (dir == DIR_IN && is_tpid(vid)) { /* IN keys skip type/predicate objects */
    *base = NULL;
    return 0;
  }
  lo = lower2(t, 0, n, pid, vid);
  hi = lo;
  while (hi < n && t[hi].a == pid && t[hi].b == vid) hi++;
  *base = t + lo;
  return hi - lo;
}

static const uint32_t* index_list(const orc_store* st, uint32_t tpid, int dir, uint64_t* sz) {
  const span* s;
  *sz = 0;
  if (tpid >= TPID_LIMIT) return NULL;
  if (dir == DIR_IN) {
    s = st->pidx_in[tpid].len ? &st->pidx_in[tpid] : &st->tidx[tpid];
  } else {
    s = &st->pidx_out[tpid];
  }
  *sz = s->len;
  return s->len ? st->ivals + s->off : NULL;
}

/* Lookup into a caller buffer (normal keys) or the index arrays. */
typedef struct {
  uint32_t* buf;
  uint64_t cap;
} scratch;

static const uint32_t* lookup(const orc_store* st, uint32_t vid, uint32_t pid, int dir, uint64_t* sz, scratch* sc) {
  const trip* base;
  uint64_t n, k;
  if (vid == 0) return index_list(st, pid, dir, sz);
  n = range_of(st, vid, pid, dir, &base);
  *sz = n;
  if (n == 0) return NULL;
  if (n > sc->cap) {
    uint32_t* p = (uint32_t*)realloc(sc->buf, n * sizeof(uint32_t));
    if (!p) {
      *sz = 0;
      return NULL;
    }
    sc->buf = p;
    sc->cap = n;
  }
  for (k = 0; k < n; k++) sc->buf[k] = base[k].c;
  return sc->buf;
}

static scratch g_api_scratch; /* for the single-threaded lookup API */

const uint32_t* orc_get_triples(const orc_store* st, uint32_t vid, uint32_t pid, int dir, uint64_t* sz) {
  *sz = 0;
  if (pid >= TPID_LIMIT || (dir != DIR_IN && dir != DIR_OUT)) return NULL;
  return lookup(st, vid, pid, dir, sz, &g_api_scratch);
}

const uint32_t* orc_get_index(const orc_store* st, uint32_t tpid, int dir, uint64_t* sz) {
  return orc_get_triples(st, 0, tpid, dir, sz);
}

/* ----------------------------------------------------------------- engine */

typedef struct {
  uint32_t* d;
  uint64_t n, cap;
} vec;

static int vpush(vec* v, uint32_t x) {
  if (v->n == v->cap) {
    uint64_t nc = v->cap ? v->cap * 2 : 1024;
    uint32_t* p = (uint32_t*)realloc(v->d, nc * sizeof(uint32_t));
    if (!p) return E_NOMEM;
    v->d = p;
    v->cap = nc;
  }
  v->d[v->n++] = x;
  return 0;
}

static int append_row(vec* v, const uint32_t* row, uint32_t cols) {
  uint32_t c;
  for (c = 0; c < cols; c++)
    if (vpush(v, row[c])) return E_NOMEM;
  return 0;
}

static int member(const uint32_t* e, uint64_t n, uint32_t x) {
  uint64_t k;
  for (k = 0; k < n; k++)
    if (e[k] == x) return 1;
  return 0;
}

enum { KNOWN_VAR = 0, UNKNOWN_VAR = 1, CONST_VAR = 2 };

/* Steps [from, npat) of a planned pattern list over the history table *res
 * of *col_io columns (empty for a query from its first pattern) with the
 * var->column map v2c, as execute_patterns loops execute_one_pattern
 * (sparql.hpp:932-1055,1107-1148).  *res holds the result on success. */
static int exec_steps(const orc_store* st, const int32_t* pats, uint32_t from, uint32_t npat, int32_t nvars,
                      vec* res_io, uint32_t* col_io, int32_t* v2c) {
  vec res = *res_io, upd = {0, 0, 0};
  uint32_t col_num = *col_io, step;
  int rc = 0;
  scratch sc = {0, 0};
#define STAT(x) ((x) >= 0 ? CONST_VAR : (v2c[-((x) + 1)] < 0 ? UNKNOWN_VAR : KNOWN_VAR))
#define COL(x) ((uint32_t)v2c[-((x) + 1)])
#define BAD_VAR(x) ((x) < 0 && -((int64_t)(x) + 1) >= nvars)
  for (step = from; step < npat && rc == 0; step++) {
    const int32_t s = pats[4 * step], p = pats[4 * step + 1], o = pats[4 * step + 2], d = pats[4 * step + 3];
    const uint64_t nrows = col_num ? res.n / col_num : 0;
    uint64_t r, sz, k;
    const uint32_t* e;
    if (d != DIR_IN && d != DIR_OUT) {
      rc = E_UNKNOWN_PATTERN;
      break;
    }
    if (BAD_VAR(s) || BAD_VAR(o)) {
      rc = E_VERTEX_INVALID;
      break;
    }
    upd.n = 0;
    if (step == 0 && s >= 0 && is_tpid((uint32_t)s)) {
      /* the first pattern from an index (start_from_index, query.hpp:816-838) */
      if (p != (int32_t)PREDICATE_ID && p != (int32_t)TYPE_ID) {
        rc = E_OBJ_ERROR;
        break;
      }
      if (o >= 0) {
        rc = E_VERTEX_INVALID;
        break;
      }
      e = index_list(st, (uint32_t)s, d, &sz);
      if (v2c[-(o + 1)] >= 0) {
        /* index_to_known (sparql.hpp:83-137): keep the rows of the carried
         * table whose ?o is in the index list, each once */
        const uint32_t c = COL(o);
        for (r = 0; r < nrows && !rc; r++)
          if (member(e, sz, res.d[r * col_num + c])) rc = append_row(&upd, res.d + r * col_num, col_num);
      } else {
        /* index_to_unknown (sparql.hpp:197-234) */
        if (col_num != 0) {
          rc = E_FIRST_PATTERN;
          break;
        }
        for (k = 0; k < sz && !rc; k++) rc = vpush(&upd, e[k]);
        v2c[-(o + 1)] = 0;
        col_num = 1;
      }
    } else if (p < 0 || p >= (int32_t)TPID_LIMIT) {
      rc = E_UNKNOWN_PATTERN;
      break;
    } else {
      const int ss = STAT(s), so = STAT(o);
      if (ss == CONST_VAR && so == CONST_VAR) {
        rc = E_UNKNOWN_PATTERN;
      } else if (ss == CONST_VAR && so == KNOWN_VAR) { /* const_to_known */
        const uint32_t c = COL(o);
        e = lookup(st, (uint32_t)s, (uint32_t)p, d, &sz, &sc);
        for (r = 0; r < nrows && !rc; r++)
          if (member(e, sz, res.d[r * col_num + c])) rc = append_row(&upd, res.d + r * col_num, col_num);
      } else if (ss == CONST_VAR && so == UNKNOWN_VAR) { /* const_to_unknown */
        if (col_num != 0) {
          rc = E_FIRST_PATTERN;
          break;
        }
        e = lookup(st, (uint32_t)s, (uint32_t)p, d, &sz, &sc);
        for (k = 0; k < sz && !rc; k++) rc = vpush(&upd, e[k]);
        v2c[-(o + 1)] = (int32_t)col_num;
        col_num++;
      } else if (ss == KNOWN_VAR && so == CONST_VAR) { /* known_to_const */
        const uint32_t c = COL(s);
        for (r = 0; r < nrows && !rc; r++) {
          e = lookup(st, res.d[r * col_num + c], (uint32_t)p, d, &sz, &sc);
          if (member(e, sz, (uint32_t)o)) rc = append_row(&upd, res.d + r * col_num, col_num);
        }
      } else if (ss == KNOWN_VAR && so == KNOWN_VAR) { /* known_to_known */
        const uint32_t cs = COL(s), ce = COL(o);
        for (r = 0; r < nrows && !rc; r++) {
          e = lookup(st, res.d[r * col_num + cs], (uint32_t)p, d, &sz, &sc);
          if (member(e, sz, res.d[r * col_num + ce])) rc = append_row(&upd, res.d + r * col_num, col_num);
        }
      } else if (ss == KNOWN_VAR && so == UNKNOWN_VAR) { /* known_to_unknown */
        const uint32_t c = COL(s);
        for (r = 0; r < nrows && !rc; r++) {
          const uint32_t cur = res.d[r * col_num + c];
          if (p == (int32_t)TYPE_ID && d == DIR_IN)
            e = index_list(st, cur, d, &sz); /* get_index(cur, IN) */
          else
            e = lookup(st, cur, (uint32_t)p, d, &sz, &sc);
          for (k = 0; k < sz && !rc; k++) {
            rc = append_row(&upd, res.d + r * col_num, col_num);
            if (!rc) rc = vpush(&upd, e[k]);
          }
        }
        v2c[-(o + 1)] = (int32_t)col_num;
        col_num++;
      } else {
        rc = E_UNKNOWN_SUB;
      }
    }
    if (rc) break;
    {
      vec t = res;
      res = upd;
      upd = t;
    }
  }
#undef STAT
#undef COL
#undef BAD_VAR
  free(upd.d);
  free(sc.buf);
  *res_io = res;
  *col_io = col_num;
  return rc;
}

static int finish_query(int rc, vec* res, uint32_t col_num, int32_t nvars, uint32_t** table, uint64_t* rows,
                        uint32_t* cols, int32_t* v2c) {
  int i;
  if (rc) {
    free(res->d);
    for (i = 0; i < nvars; i++) v2c[i] = -1;
    return rc;
  }
  *table = res->d;
  *rows = col_num ? res->n / col_num : 0;
  *cols = col_num;
  return 0;
}

int orc_run_query(const orc_store* st, const int32_t* pats, uint32_t npat, int32_t nvars, uint32_t** table,
                  uint64_t* rows, uint32_t* cols, int32_t* v2c) {
  vec res = {0, 0, 0};
  uint32_t col_num = 0;
  int i;
  *table = NULL;
  *rows = 0;
  *cols = 0;
  if (nvars < 0 || nvars > 64) return E_TOO_MANY_VARS;
  for (i = 0; i < nvars; i++) v2c[i] = -1;
  const int rc = exec_steps(st, pats, 0, npat, nvars, &res, &col_num, v2c);
  return finish_query(rc, &res, col_num, nvars, table, rows, cols, v2c);
}

int orc_run_query_from(const orc_store* st, const int32_t* pats, uint32_t npat, int32_t nvars, uint32_t from,
                       const uint32_t* in, uint64_t in_rows, uint32_t in_cols, const int32_t* in_v2c,
                       uint32_t** table, uint64_t* rows, uint32_t* cols, int32_t* v2c) {
  vec res = {0, 0, 0};
  uint64_t k;
  int i;
  *table = NULL;
  *rows = 0;
  *cols = 0;
  if (nvars < 0 || nvars > 64) return E_TOO_MANY_VARS;
  for (i = 0; i < nvars; i++) v2c[i] = in_v2c[i];
  for (k = 0; k < in_rows * in_cols; k++)
    if (vpush(&res, in[k])) {
      free(res.d);
      return E_NOMEM;
    }
  {
    const int rc = exec_steps(st, pats, from, npat, nvars, &res, &in_cols, v2c);
    return finish_query(rc, &res, in_cols, nvars, table, rows, cols, v2c);
  }
}

int orc_run_query_mt(const orc_store* st, const int32_t* pats, uint32_t npat, int32_t nvars, int nthreads,
                     uint32_t** table, uint64_t* rows, uint32_t* cols, int32_t* v2c) {
  uint64_t sz = 0, total = 0;
  vec* part;
  int32_t* pv;
  uint32_t ncol = 0;
  int t, rc = 0, i;
  const uint32_t* e;
  /* only a query from an index is split (dispatch, sparql.hpp:1058-1083) */
  if (nthreads <= 1 || npat == 0 || pats[0] < 0 || !is_tpid((uint32_t)pats[0]) || pats[2] >= 0 ||
      (pats[1] != (int32_t)PREDICATE_ID && pats[1] != (int32_t)TYPE_ID) || nvars < 1 || nvars > 64 ||
      -((int64_t)pats[2] + 1) >= nvars || (pats[3] != DIR_IN && pats[3] != DIR_OUT))
    return orc_run_query(st, pats, npat, nvars, table, rows, cols, v2c);
  *table = NULL;
  *rows = 0;
  *cols = 0;
  e = index_list(st, (uint32_t)pats[0], pats[3], &sz);
  part = (vec*)calloc((size_t)nthreads, sizeof(vec));
  pv = (int32_t*)malloc((size_t)nthreads * 64 * sizeof(int32_t));
  if (!part || !pv) {
    free(part);
    free(pv);
    return E_NOMEM;
  }
#pragma omp parallel for num_threads(nthreads) schedule(static, 1) reduction(| : rc)
  for (t = 0; t < nthreads; t++) {
    /* index_to_unknown with mt_factor = nthreads, mt_tid = t (sparql.hpp:214-224):
     * consecutive parts of length sz / nthreads, the rest to the last */
    const uint64_t len = sz / (uint64_t)nthreads;
    const uint64_t lo = (uint64_t)t * len, hi = (t == nthreads - 1) ? sz : lo + len;
    uint32_t c = 1;
    uint64_t k;
    int j, r = 0;
    int32_t* v = pv + (size_t)t * 64;
    for (j = 0; j < nvars; j++) v[j] = -1;
    v[-(pats[2] + 1)] = 0;
    for (k = lo; k < hi && !r; k++) r = vpush(&part[t], e[k]);
    if (!r) r = exec_steps(st, pats, 1, npat, nvars, &part[t], &c, v);
    if (t == 0) ncol = c;
    rc |= r;
  }
  for (i = 0; i < nvars; i++) v2c[i] = pv[i];
  if (!rc) {
    vec all = {0, 0, 0};
    for (t = 0; t < nthreads && !rc; t++) total += part[t].n;
    all.d = (uint32_t*)malloc((total ? total : 1) * sizeof(uint32_t));
    if (!all.d) rc = E_NOMEM;
    for (t = 0; t < nthreads && !rc; t++) {
      memcpy(all.d + all.n, part[t].d, part[t].n * sizeof(uint32_t));
      all.n += part[t].n;
    }
    for (t = 0; t < nthreads; t++) free(part[t].d);
    free(part);
    free(pv);
    return finish_query(rc, &all, ncol, nvars, table, rows, cols, v2c);
  }
  for (t = 0; t < nthreads; t++) free(part[t].d);
  free(part);
  free(pv);
  for (i = 0; i < nvars; i++) v2c[i] = -1;
  return rc;
}

uint64_t orc_table_checksum(const uint32_t* table, uint64_t rows, uint32_t cols) {
  /* sum over rows of a chained hash of (column, value) pairs: the same
   * order-independent fingerprint as the device library's wk_table_checksum */
  uint64_t s = 0, r;
  uint32_t c;
  for (r = 0; r < rows; r++) {
    uint64_t h = 0x9E3779B97F4A7C15ull ^ cols;
    for (c = 0; c < cols; c++) h = orc_hash_u64(h ^ ((uint64_t)table[r * cols + c] + ((uint64_t)c << 32)));
    s += h;
  }
  return s;
}

int orc_run_stream(const orc_store* st, const int32_t* pats, const uint64_t* pat_off, const int32_t* nvars,
                   uint64_t nq, int nthreads, uint64_t* rows, uint64_t* checksums, int32_t* status) {
  int64_t i;
  if (nthreads < 1) nthreads = 1;
  /* one worker per thread pulling the next query from a shared queue, as the
   * engine threads of Engine::run take requests (engine.hpp:132-233) */
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
  for (i = 0; i < (int64_t)nq; i++) {
    uint32_t* t = NULL;
    uint64_t r = 0;
    uint32_t c = 0;
    int32_t v[64];
    const int rc = orc_run_query(st, pats + 4 * pat_off[i], (uint32_t)(pat_off[i + 1] - pat_off[i]), nvars[i], &t,
                                 &r, &c, v);
    status[i] = rc;
    rows[i] = rc ? 0 : r;
    checksums[i] = rc ? 0 : orc_table_checksum(t, r, c);
    free(t);
  }
  return 0;
}

/* ---------------------------------------------------------------- planner */

void orc_set_direction(const int32_t* raw, const int32_t* orders, const int32_t* dirs, uint32_t n, int32_t* out) {
  uint32_t i;
  for (i = 0; i < n; i++) {
    const int32_t* pr = raw + 4 * (orders[i] - 1);
    int32_t s = pr[0], p = pr[1], o = pr[2], d = pr[3], t;
    switch (dirs[i]) {
      case 0: /* "<" */
        d = DIR_IN;
        t = s;
        s = o;
        o = t;
        break;
      case 1: /* ">" */
        d = DIR_OUT;
        break;
      case 2: /* "<<" */
        d = DIR_IN;
        o = s;
        s = p;
        p = PREDICATE_ID;
        break;
      case 3: /* ">>" */
        d = DIR_OUT;
        s = p;
        p = PREDICATE_ID;
        break;
    }
    out[4 * i] = s;
    out[4 * i + 1] = p;
    out[4 * i + 2] = o;
    out[4 * i + 3] = d;
  }
}

/* ------------------------------------------------------------ final stage */

typedef struct {
  const int32_t* cols;
  const int32_t* desc;
  uint32_t n, width;
} order_ctx;

static order_ctx g_ord;

/* ReduceCmp: lexicographic over all columns, signed (int*) comparison */
static int cmp_full(const void* x, const void* y) {
  const int32_t* a = (const int32_t*)x;
  const int32_t* b = (const int32_t*)y;
  uint32_t i;
  for (i = 0; i < g_ord.width; i++)
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  return 0;
}

/* ORDER BY ids; ties broken by the full row so the result is deterministic */
static int cmp_order(const void* x, const void* y) {
  const uint32_t* a = (const uint32_t*)x;
  const uint32_t* b = (const uint32_t*)y;
  uint32_t i;
  for (i = 0; i < g_ord.n; i++) {
    const uint32_t ca = a[g_ord.cols[i]], cb = b[g_ord.cols[i]];
    if (ca != cb) {
      int c = ca < cb ? -1 : 1;
      return g_ord.desc[i] ? -c : c;
    }
  }
  return cmp_full(x, y);
}

int orc_final_process(const uint32_t* table, uint64_t rows, uint32_t cols, const int32_t* v2c, int32_t nvars,
                      const int32_t* required, uint32_t nreq, int distinct, const int32_t* order_vars,
                      const int32_t* order_desc, uint32_t norder, int64_t offset, int64_t limit, uint32_t** out,
                      uint64_t* out_rows) {
  uint32_t* t;
  uint64_t n = rows, r, first, last;
  uint32_t i;
  int32_t ocols[64], rcols[64];
  *out = NULL;
  *out_rows = 0;
  if (rows == 0) return 0; /* nothing to do on an empty table */
  if (nreq == 0) return 5; /* NO_REQUIRED_VAR */
  if (nreq > 64 || norder > 64) return E_TOO_MANY_VARS;
  for (i = 0; i < nreq; i++) {
    if (required[i] >= 0 || -(required[i] + 1) >= nvars || v2c[-(required[i] + 1)] < 0) return E_VERTEX_INVALID;
    rcols[i] = v2c[-(required[i] + 1)];
  }
  for (i = 0; i < norder; i++) {
    if (order_vars[i] >= 0 || -(order_vars[i] + 1) >= nvars || v2c[-(order_vars[i] + 1)] < 0)
      return E_VERTEX_INVALID;
    ocols[i] = v2c[-(order_vars[i] + 1)];
  }
  t = (uint32_t*)malloc(rows * cols * sizeof(uint32_t));
  if (!t) return E_NOMEM;
  memcpy(t, table, rows * cols * sizeof(uint32_t));
  g_ord.width = cols;
  g_ord.cols = ocols;
  g_ord.desc = order_desc;
  g_ord.n = norder;
  if (distinct) {
    /* sort by all columns, then drop rows equal to their predecessor on the
     * required columns (std::unique over the sorted rows) */
    uint64_t p = 0;
    qsort(t, rows, cols * sizeof(uint32_t), cmp_full);
    for (r = 1; r < rows; r++) {
      int eq = 1;
      for (i = 0; i < nreq; i++)
        if (t[p * cols + rcols[i]] != t[r * cols + rcols[i]]) {
          eq = 0;
          break;
        }
      if (!eq) {
        p++;
        if (p != r) memcpy(t + p * cols, t + r * cols, cols * sizeof(uint32_t));
      }
    }
    n = p + 1;
  }
  if (norder) qsort(t, n, cols * sizeof(uint32_t), cmp_order);
  first = offset > 0 ? (uint64_t)offset : 0;
  if (first > n) first = n;
  last = n;
  if (limit >= 0 && first + (uint64_t)limit < last) last = first + (uint64_t)limit;
  {
    uint32_t* o = (uint32_t*)malloc(((last - first) * nreq + 1) * sizeof(uint32_t));
    if (!o) {
      free(t);
      return E_NOMEM;
    }
    for (r = first; r < last; r++)
      for (i = 0; i < nreq; i++) o[(r - first) * nreq + i] = t[r * cols + rcols[i]];
    *out = o;
    *out_rows = last - first;
  }
  free(t);
  return 0;
}

void orc_free(void* p) { free(p); }
