// cbz_wave.h — wave64 primitives (CDNA: 64 lanes per wavefront).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>

namespace cbz {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ int wave_sum_i32(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
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

namespace cbz {

// DPP wave reductions (gfx9 DPP: row_shr within 16-lane rows, then
// row_bcast15 / row_bcast31 across rows).  The inclusive scan leaves the
// result in lane 63, which readlane broadcasts.  Lanes whose DPP source is
// outside the row read `ident`.
namespace dpp {
constexpr int kShr1 = 0x111, kShr2 = 0x112, kShr4 = 0x114, kShr8 = 0x118;
constexpr int kBcast15 = 0x142, kBcast31 = 0x143;

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int mov(int ident, int v) {
  return __builtin_amdgcn_update_dpp(ident, v, CTRL, ROW_MASK, 0xF, false);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double mov(double ident, double v) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(ident), __double2loint(v), CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(ident), __double2hiint(v), CTRL, ROW_MASK, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float mov(float ident, float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(ident), __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}
__device__ __forceinline__ int bcast63(int v) { return __builtin_amdgcn_readlane(v, 63); }
__device__ __forceinline__ double bcast63(double v) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                          __builtin_amdgcn_readlane(__double2loint(v), 63));
}
__device__ __forceinline__ float bcast63(float v) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
template <typename T, typename OP>
__device__ __forceinline__ T allreduce(T v, T ident, OP op) {
  v = op(v, mov<kShr1, 0xF>(ident, v));
  v = op(v, mov<kShr2, 0xF>(ident, v));
  v = op(v, mov<kShr4, 0xF>(ident, v));
  v = op(v, mov<kShr8, 0xF>(ident, v));
  v = op(v, mov<kBcast15, 0xA>(ident, v));
  v = op(v, mov<kBcast31, 0xC>(ident, v));
  return bcast63(v);
}
}  // namespace dpp

// Argmax with the reference's tie rule (larger key; equal keys -> lower
// index), as a max over keys followed by a min over the indices of the
// lanes holding that max.  Every lane's (key, idx) must already be its own
// lowest-index maximum; a lane without a candidate carries idx = INT_MAX.
template <typename R>
__device__ __forceinline__ void wave_argmax_dpp(R& key, int& idx) {
  const R m = dpp::allreduce(key, R(-__builtin_inf()), [](R x, R y) { return y > x ? y : x; });
  const int cand = key == m ? idx : INT_MAX;
  idx = dpp::allreduce(cand, INT_MAX, [](int x, int y) { return y < x ? y : x; });
  key = m;
}

}  // namespace cbz
