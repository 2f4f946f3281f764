// kh_sort.hip -- device radix sort of the N-body launch's object order.
//
// The longest-first order of a batch with per-object final epochs sorts the
// objects by a 32-bit key, descending: the cost bucket of the counting sort
// (cost level and arc direction) in the high bits, the arc length below it,
// so that the lanes of a wavefront carry objects of like cost that also move
// through time at like rates (their records stay within a few of each other,
// which the wavefront windows of the main kernel rely on).
#include <hipcub/hipcub.hpp>

#include "kh_internal.h"

namespace kh {

static size_t cub_bytes(int64_t n) {
  size_t b = 0;
  (void)hipcub::DeviceRadixSort::SortPairsDescending(
      nullptr, b, (const uint32_t*)nullptr, (uint32_t*)nullptr, (const uint32_t*)nullptr,
      (uint32_t*)nullptr, (int)n, 0, 32);
  return b;
}

size_t order_sort_temp_bytes(int64_t n) { return 3 * order_sort_span(n) + cub_bytes(n) + 256; }

int32_t launch_order_sort(int64_t n, void* buf, size_t bytes, uint32_t* perm, hipStream_t s) {
  if (n <= 0) return 0;
  const size_t a = order_sort_span(n);
  char* q = static_cast<char*>(buf);
  const uint32_t* k_in = order_sort_keys(buf, n);
  const uint32_t* v_in = order_sort_vals(buf, n);
  uint32_t* k_out = reinterpret_cast<uint32_t*>(q + 2 * a);
  void* cub = q + 3 * a;
  size_t cb = bytes - 3 * a;
  return hipcub::DeviceRadixSort::SortPairsDescending(cub, cb, k_in, k_out, v_in, perm, (int)n, 0,
                                                      32, s) == hipSuccess
             ? 0
             : -1;
}

}  // namespace kh
