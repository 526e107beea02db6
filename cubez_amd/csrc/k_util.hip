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
