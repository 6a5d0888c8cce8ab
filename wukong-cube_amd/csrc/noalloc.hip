// Fork-join split of a caller-owned device table with no allocation: the
// owner x tile histogram, its chunk totals and the owners' heads live in a
// fixed array of device memory that each GPU holds until the process ends.
//
// For callers that cannot allocate device memory (code under graph capture,
// tight serving loops, harnesses that forbid it).  Every call on one GPU
// shares that GPU's array: calls that overlap on one GPU race.  Nothing in
// the array needs to start at zero (every word a call reads, the same call
// wrote first).
#include <hip/hip_runtime.h>

#include <mutex>

#include "device.hpp"

namespace wk {
namespace {

// 8 MiB: owner x tile histograms of up to ~1M entries (64 owners over 16M
// rows, 8 owners over 128M rows).
constexpr uint64_t kScratchWords = uint64_t(1) << 20;
static __device__ __align__(256) uint64_t scratch[kScratchWords];

// scratch's device address on each GPU; the host-side name is only a stand-in
constexpr int kMaxDevices = 64;
uint64_t* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

}  // namespace

int noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return WK_E_INVALID_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return WK_OK;
  void* p = nullptr;
  // the caller made `device` current
  WK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch[device] = static_cast<uint64_t*>(p);
  return WK_OK;
}

}  // namespace wk

using namespace wk;

extern "C" int wk_table_partition_noalloc(wk_dev* dev, const uint32_t* in, uint64_t rows, uint32_t cols,
                                          uint32_t col, uint32_t nparts, uint32_t* out, uint64_t* sizes,
                                          uint64_t* heads) {
  if (!dev || !sizes || !heads || nparts == 0 || nparts > kMaxParts) return WK_E_INVALID_ARG;
  if (rows && (!in || !out || cols == 0 || col >= cols)) return WK_E_INVALID_ARG;
  if (partition_scratch_words(rows, nparts) > kScratchWords) return WK_E_SCRATCH_LIMIT;
  uint64_t* s = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev->device >= 0 && dev->device < kMaxDevices) s = g_scratch[dev->device];
  }
  if (!s) return WK_E_STATE;
  for (uint32_t p = 0; p < nparts; p++) sizes[p] = heads[p] = 0;
  if (rows == 0) return WK_OK;
  const uint64_t* dh = launch_partition(in, rows, cols, col, OwnerFn{nparts, 0, 0}, s, out, dev->stream);
  // the heads come back through the handle's pinned staging words
  if (hip_failed(hipMemcpyAsync(dev->h_win, dh, (size_t(nparts) + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                dev->stream)) ||
      hip_failed(hipStreamSynchronize(dev->stream)))
    return WK_E_HIP;
  for (uint32_t p = 0; p < nparts; p++) {
    heads[p] = dev->h_win[p];
    sizes[p] = dev->h_win[p + 1] - dev->h_win[p];
  }
  return WK_OK;
}
