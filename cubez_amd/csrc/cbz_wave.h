// cbz_wave.h — wave64 primitives (CDNA: 64 lanes per wavefront).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cbz {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ int wave_sum_i32(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Inclusive prefix sum across the 64 lanes.
__device__ __forceinline__ int wave_incl_scan_i32(int v) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (l >= o) v += t;
  }
  return v;
}

__device__ __forceinline__ int64_t wave_incl_scan_i64(int64_t v) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int64_t t = __shfl_up(v, o, 64);
    if (l >= o) v += t;
  }
  return v;
}

// Argmax with the reference's tie rule: the larger key wins; equal keys go
// to the lower index.  A lane without a candidate carries idx = INT_MAX.
template <typename R>
__device__ __forceinline__ void wave_argmax(R& key, int& idx) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    R k2 = __shfl_xor(key, o, 64);
    int i2 = __shfl_xor(idx, o, 64);
    if (k2 > key || (k2 == key && i2 < idx)) {
      key = k2;
      idx = i2;
    }
  }
}

}  // namespace cbz
