// k_util.hip — small utility kernels: an int32 fill (used instead of
// hipMemsetAsync so a captured step graph holds only kernel nodes), and the
// elementwise FP64 self-test used to check, on the device, that division,
// square root, product, sum and difference round exactly as IEEE 754
// binary64 (round to nearest even) — the property that makes the exact path
// reproduce the Go reference bit for bit.
#include "cbz_devmath.h"
#include "cbz_launch.h"

namespace cbz {

__global__ void __launch_bounds__(256) k_selftest(int op, int64_t n, const double* a, const double* b, double* out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = a[i], y = b[i];
  double r;
  switch (op) {
    case 0: r = x / y; break;
    case 1: r = dsqrt(x); break;
    case 2: r = x * y; break;
    case 3: r = x + y; break;
    default: r = x - y; break;
  }
  out[i] = r;
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Profiling counters of one step, accumulated on the device (so a profiled
// run need not stop the stream every step): acc[0] steps, acc[1] bodies
// integrated (the kAwakeSlots counters, then cleared), acc[2] contacts,
// acc[3] resolver iterations (the sum of iters[0 .. n), below 2^32 per step;
// the two sums share one 64-bit reduction).
__global__ void __launch_bounds__(1024) k_prof_accum(uint32_t* awake, const int64_t* total, const int64_t* iters,
                                                     int n, int64_t* acc) {
  __shared__ int64_t part[1024];
  int64_t s = 0;
  for (int i = threadIdx.x; i < n; i += 1024) s += iters[i];
  if (threadIdx.x < kAwakeSlots) {  // the awake counters, cleared for the next step
    s += int64_t(awake[threadIdx.x]) << 32;
    awake[threadIdx.x] = 0;
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    acc[0] += 1;
    acc[1] += part[0] >> 32;
    acc[2] += *total;
    acc[3] += part[0] & 0xFFFFFFFFll;
  }
}

hipError_t launch_prof_accum(uint32_t* awake, const int64_t* total, const int64_t* iters, int n, int64_t* acc,
                             hipStream_t st) {
  k_prof_accum<<<1, 1024, 0, st>>>(awake, total, iters, n, acc);
  return hipGetLastError();
}

hipError_t launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_fill_u8<<<int((n + 255) / 256), 256, 0, st>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_fill_i32<<<int((n + 255) / 256), 256, 0, st>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_selftest(int op, int64_t n, const double* a, const double* b, double* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_selftest<<<int((n + 255) / 256), 256, 0, st>>>(op, n, a, b, out);
  return hipGetLastError();
}

}  // namespace cbz
