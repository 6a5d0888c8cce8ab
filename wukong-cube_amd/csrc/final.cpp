// wk_table_final_process: SPARQLEngine::final_process
// (core/engine/sparql.hpp:1419-1546) on a device history table.
//
// The reference copies the table into an array of row pointers and calls
// std::sort on it: ReduceCmp (all columns, signed) for DISTINCT, then
// Compare (the ORDER BY columns) for ORDER BY.  Here the rows never move
// until the final projection.  A permutation of row indices is sorted
// instead, as (key, row) pairs through a stable LSD radix sort, one 32-bit
// key column at a time.  Each 6-bit digit is one stable split by
// launch_partition.  ORDER BY compares ids, and ties keep the full-row order.
#include <algorithm>
#include <new>
#include <utility>
#include <vector>

#include "device.hpp"

using namespace wk;

namespace {

// The two pair buffers of one call, freed on every way out.
struct PairBufs {
  uint32_t* a = nullptr;
  uint32_t* b = nullptr;
  ~PairBufs() {
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
  }
};

// Stable sort of n pairs by their first word: one split per 6-bit digit,
// least significant first, alternating between cur and tmp.
void sort_pairs(wk_dev* dev, uint32_t*& cur, uint32_t*& tmp, uint64_t n) {
  for (uint32_t shift = 0; shift < 32; shift += kSortDigitBits) {
    const uint32_t bits = std::min<uint32_t>(kSortDigitBits, 32 - shift);
    launch_partition(cur, n, 2, 0, OwnerFn{0, shift, (1u << bits) - 1}, dev->pfx, tmp, dev->stream);
    std::swap(cur, tmp);
  }
}

}  // namespace

extern "C" int wk_table_final_process(wk_dev* dev, const wk_table* in, const int32_t* v2c, int32_t nvars,
                                      const int32_t* required, uint32_t nreq, int32_t distinct,
                                      const int32_t* order_vars, const int32_t* order_desc, uint32_t norder,
                                      int64_t offset, int64_t limit, wk_table** out) {
  if (!dev || !in || !out || nvars < 0 || (nvars && !v2c) || (nreq && !required) ||
      (norder && (!order_vars || !order_desc)))
    return WK_E_INVALID_ARG;
  if (int drc = drain_batches(dev)) return drc;
  *out = nullptr;
  const uint64_t rows = in->rows;
  const uint32_t cols = in->cols;
  if (rows == 0) {  // an empty result is left as it is (sparql.hpp:1420-1421)
    wk_table* t = new (std::nothrow) wk_table();
    if (!t) return WK_E_NOMEM_HOST;
    t->cols = cols;
    *out = t;
    return WK_OK;
  }
  if (nreq == 0) return WK_E_NO_REQUIRED_VAR;  // sparql.hpp:1507
  if (nreq > WK_MAX_VARS || norder > WK_MAX_VARS) return WK_E_TOO_MANY_VARS;
  // var2col of a variable, or -1 when it has no column in this table
  auto col_of = [&](int32_t v) -> int64_t {
    if (v >= 0 || -int64_t(v) - 1 >= nvars) return -1;
    const int32_t c = v2c[-int64_t(v) - 1];
    return (c >= 0 && uint32_t(c) < cols) ? c : -1;
  };
  ColList req{};
  req.n = nreq;
  for (uint32_t i = 0; i < nreq; i++) {
    const int64_t c = col_of(required[i]);
    if (c < 0) return WK_E_VERTEX_INVALID;
    req.col[i] = uint32_t(c);
  }
  std::vector<uint32_t> ocol(norder);
  for (uint32_t i = 0; i < norder; i++) {
    const int64_t c = col_of(order_vars[i]);
    if (c < 0) return WK_E_VERTEX_INVALID;
    ocol[i] = uint32_t(c);
  }
  if (rows > UINT32_MAX) return WK_E_OVERFLOW;  // the permutation holds 32-bit row indices

  const bool sort = distinct || norder;
  hipStream_t s = dev->stream;
  PairBufs pb;
  uint32_t* cur = nullptr;
  uint32_t* tmp = nullptr;
  uint64_t n = rows;
  if (sort) {
    if (partition_scratch_words(rows, kMaxParts) > dev->rcap + 1) return WK_E_NOMEM_BUDGET;
    if (hipMalloc(&pb.a, rows * 2 * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&pb.b, rows * 2 * sizeof(uint32_t)) != hipSuccess)
      return WK_E_NOMEM_BUDGET;
    cur = pb.a;
    tmp = pb.b;
    // full-row order, words compared as signed (ReduceCmp, sparql.hpp:1403-1417),
    // least significant column first
    for (uint32_t c = cols; c-- > 0;) {
      launch_fp_keys(in->d, cols, c, kKeySigned, c + 1 == cols ? nullptr : cur, n, cur, s);
      sort_pairs(dev, cur, tmp, n);
    }
    if (distinct) {
      // flag 0 = kept; one stable split puts the kept rows first, in order
      launch_fp_dedup(in->d, cols, req, cur, n, tmp, s);
      const uint64_t* heads = launch_partition(tmp, n, 2, 0, OwnerFn{0, 0, 1}, dev->pfx, cur, s);
      uint64_t kept = 0;
      if (hip_failed(hipMemcpyAsync(&kept, heads + 1, sizeof(kept), hipMemcpyDeviceToHost, s)) ||
          hip_failed(hipStreamSynchronize(s)))
        return WK_E_HIP;
      n = kept;
    }
    // ORDER BY, last key first; each sort is stable, so ties keep the
    // full-row order
    for (uint32_t o = norder; o-- > 0;) {
      launch_fp_keys(in->d, cols, ocol[o], order_desc[o] ? kKeyDesc : kKeyAsc, cur, n, cur, s);
      sort_pairs(dev, cur, tmp, n);
    }
  }
  // OFFSET, LIMIT (sparql.hpp:1488-1500)
  const uint64_t first = offset > 0 ? std::min<uint64_t>(uint64_t(offset), n) : 0;
  uint64_t last = n;
  if (limit >= 0 && first + uint64_t(limit) < last) last = first + uint64_t(limit);

  wk_table* t = new (std::nothrow) wk_table();
  if (!t) return WK_E_NOMEM_HOST;
  t->rows = last - first;
  t->cols = nreq;
  if (t->rows) {
    if (hipMalloc(&t->d, t->rows * nreq * sizeof(uint32_t)) != hipSuccess) {
      delete t;
      return WK_E_NOMEM_BUDGET;
    }
    // projection onto the required variables (sparql.hpp:1502-1526)
    launch_fp_project(in->d, cols, req, sort ? cur : nullptr, first, t->rows, t->d, s);
  }
  if (hip_failed(hipStreamSynchronize(s))) {
    if (t->d) (void)hipFree(t->d);
    delete t;
    return WK_E_HIP;
  }
  *out = t;
  return WK_OK;
}
