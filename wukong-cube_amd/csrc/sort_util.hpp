// Parallel sorting helpers for the host gstore builder.
//
// The loader sorts triples by (p,s,o) and (p,o,s) and de-duplicates them
// (loader/base_loader.hpp:90-103,481-498).  Here each predicate's run is
// packed into 64-bit words (hi:lo) and sorted with an OpenMP LSD radix sort
// that skips digit positions that are constant over the input.
#pragma once

#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace wk {

// Sort `n` u64 words in place; `tmp` must hold n words.
inline void radix_sort_u64(uint64_t* data, uint64_t* tmp, size_t n) {
  if (n < (size_t(1) << 16)) {
    std::sort(data, data + n);
    return;
  }
  constexpr int BITS = 11;
  constexpr int NB = 1 << BITS;
  const int nthreads = std::max(1, std::min(omp_get_max_threads(), int(n >> 15) + 1));
  uint64_t ormask = 0, andmask = ~uint64_t(0);
#pragma omp parallel for num_threads(nthreads) reduction(| : ormask) reduction(& : andmask)
  for (size_t i = 0; i < n; i++) {
    ormask |= data[i];
    andmask &= data[i];
  }
  const uint64_t varying = ormask ^ andmask;
  std::vector<size_t> hist(size_t(nthreads) * NB);
  uint64_t* src = data;
  uint64_t* dst = tmp;
  for (int shift = 0; shift < 64; shift += BITS) {
    if (((varying >> shift) & (NB - 1)) == 0) continue;
    std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nthreads)
    {
      const int t = omp_get_thread_num();
      const size_t lo = n * t / nthreads, hi = n * (t + 1) / nthreads;
      size_t* h = &hist[size_t(t) * NB];
      for (size_t i = lo; i < hi; i++) h[(src[i] >> shift) & (NB - 1)]++;
#pragma omp barrier
#pragma omp single
      {
        size_t run = 0;
        for (int d = 0; d < NB; d++)
          for (int tt = 0; tt < nthreads; tt++) {
            size_t c = hist[size_t(tt) * NB + d];
            hist[size_t(tt) * NB + d] = run;
            run += c;
          }
      }
      for (size_t i = lo; i < hi; i++) {
        const uint64_t v = src[i];
        dst[h[(v >> shift) & (NB - 1)]++] = v;
      }
    }
    std::swap(src, dst);
  }
  if (src != data) std::memcpy(data, src, n * sizeof(uint64_t));
}

// Remove adjacent duplicates of a sorted run; returns the new length.
inline size_t dedup_sorted(uint64_t* a, size_t n) {
  if (n <= 1) return n;
  size_t m = 1;
  for (size_t i = 1; i < n; i++)
    if (a[i] != a[m - 1]) a[m++] = a[i];
  return m;
}

}  // namespace wk
