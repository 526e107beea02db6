// k_broad.hip — sort-and-sweep broadphase in front of the explicit-list
// narrowphase, for worlds of thousands of bodies where the all-pairs space
// grows as n^2.
//
// Every body gets an axis-aligned box of half-width kBroadGrow * bounding
// radius around its collider centre.  The narrowphase's own far-pair test
// (k_narrow.hip, far_apart) rejects a body pair only when
// |t|^2 > k (rA + rB)^2 with k <= 3.003, so a pair it keeps has
// |t| <= sqrt(3.003) (rA + rB): its boxes overlap.  A body whose centre or
// radius is not finite overlaps everything (the reference still emits
// contacts for NaN states).  The kept ordered pairs (i, j) and (j, i) of
// every overlap, plus every body-plane pair, are sorted into the order of
// examples/cubedrop.go:47-64 (row i ascending, its planes first, then j
// ascending) and handed to the explicit-list narrowphase: its contacts are
// those of the all-pairs narrowphase, in the same order.  Both sorts are
// the stable radix sort of k_sort.hip.
#include "cbz_broad.h"
#include "cbz_launch.h"
#include "cbz_sort.h"
#include "cbz_wave.h"

namespace cbz {

// sqrt(3.003) = 1.73291..., rounded up with a margin for the rounding of
// the distance and radius sums
constexpr double kBroadGrow = 1.7330;

// float nearest to x in the direction of -inf / +inf (x finite)
__device__ __forceinline__ float f32_down(double x) {
  float f = float(x);
  if (double(f) > x) f = nextafterf(f, -INFINITY);
  return f;
}
__device__ __forceinline__ float f32_up(double x) {
  float f = float(x);
  if (double(f) < x) f = nextafterf(f, INFINITY);
  return f;
}
// order-preserving map of a float onto uint32 (-inf first)
__device__ __forceinline__ uint32_t f32_order(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float f32_unorder(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// Boxes and the sort keys (world << 32 | ordered lower x bound, rounded down).
template <typename R>
__global__ void __launch_bounds__(256) k_bp_bounds(BroadArgs<R> a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nb) return;
  const int64_t S = a.stride;
  const double cx = double(a.ctransform[9 * S + i]);
  const double cy = double(a.ctransform[10 * S + i]);
  const double cz = double(a.ctransform[11 * S + i]);
  const double r = kBroadGrow * a.brad[i];
  const bool wild = !(isfinite(cx) && isfinite(cy) && isfinite(cz) && isfinite(r));
  float hx = INFINITY;  // a body that is not finite spans the whole axis
  uint32_t kx = f32_order(-INFINITY);
  if (!wild) {
    a.lo[i] = cy - r;
    a.lo[a.nb + i] = cz - r;
    a.hi[i] = cy + r;
    a.hi[a.nb + i] = cz + r;
    hx = f32_up(cx + r);
    kx = f32_order(f32_down(cx - r));
  }
  a.hix[i] = hx;
  a.wild[i] = wild ? 1 : 0;
  a.key_in[i] = (uint64_t(uint32_t(a.body_world[i])) << 32) | kx;
  a.val_in[i] = i;
}

// The body-plane pairs of every body: planes first in each row.
__global__ void __launch_bounds__(256) k_bp_planes(BroadPlanes p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.nb) return;
  const int w = p.body_world[i];
  const int b0 = p.body_off[w], np = p.plane_off[w + 1] - p.plane_off[w];
  const int64_t at = p.plane_pair_base[w] + int64_t(i - b0) * np;
  for (int k = 0; k < np; ++k) p.pkey[at + k] = (uint64_t(uint32_t(i)) << 32) | uint32_t(k);
}

// Sweep along x over each world's sorted boxes: thread p scans forward from
// sorted position p while the next lower bound is within its upper bound,
// tests y and z, and appends both ordered pairs of every overlap.  Run once
// to count (block-aggregated, one atomic per block) and write; entries past
// the capacity are counted and dropped.
template <typename R>
__global__ void __launch_bounds__(256) k_bp_sweep(BroadArgs<R> a) {
  __shared__ int s_wtot[4];
  __shared__ long long s_base;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  int mine = 0;
  int i = -1, w = 0, end = 0;
  float hx = 0.0f;
  if (p < a.nb) {
    i = a.val_out[p];
    w = int(a.key_out[p] >> 32);
    end = a.body_off[w + 1];
    hx = a.hix[i];
    for (int q = p + 1; q < end; ++q) {
      if (f32_unorder(uint32_t(a.key_out[q])) > hx) break;
      const int j = a.val_out[q];
      const bool keep = a.wild[i] || a.wild[j] ||
                        (a.lo[j] <= a.hi[i] && a.lo[i] <= a.hi[j] && a.lo[a.nb + j] <= a.hi[a.nb + i] &&
                         a.lo[a.nb + i] <= a.hi[a.nb + j]);
      mine += keep ? 2 : 0;
    }
  }
  // block-wide reservation (all 4 waves take part)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int incl = wave_incl_scan_i32(mine);
  if (lane == 63) s_wtot[wv] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int u = 0; u < 4; ++u) tot += s_wtot[u];
    s_base = tot > 0 ? (long long)atomicAdd((unsigned long long*)a.pcount, (unsigned long long)tot) : 0;
  }
  __syncthreads();
  int64_t at = int64_t(a.plane_pairs) + s_base;
  for (int u = 0; u < wv; ++u) at += s_wtot[u];
  at += incl - mine;
  if (mine == 0) return;
  const int b0 = a.body_off[w];
  const int np = a.plane_off[w + 1] - a.plane_off[w];
  for (int q = p + 1; q < end; ++q) {
    if (f32_unorder(uint32_t(a.key_out[q])) > hx) break;
    const int j = a.val_out[q];
    const bool keep = a.wild[i] || a.wild[j] ||
                      (a.lo[j] <= a.hi[i] && a.lo[i] <= a.hi[j] && a.lo[a.nb + j] <= a.hi[a.nb + i] &&
                       a.lo[a.nb + i] <= a.hi[a.nb + j]);
    if (!keep) continue;
    if (at < a.pcap) a.pkey[at] = (uint64_t(uint32_t(i)) << 32) | uint32_t(np + (j - b0));
    if (at + 1 < a.pcap) a.pkey[at + 1] = (uint64_t(uint32_t(j)) << 32) | uint32_t(np + (i - b0));
    at += 2;
  }
}

// Sorted pair keys -> the explicit list (one, two); a plane is -(k+1).
__global__ void __launch_bounds__(256) k_bp_unpack(BroadList L) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= L.n) return;
  const uint64_t key = L.pkey[t];
  const int i = int(key >> 32);
  const int jk = int(uint32_t(key));
  const int w = L.body_world[i];
  const int np = L.plane_off[w + 1] - L.plane_off[w];
  L.one[t] = i;
  L.two[t] = jk < np ? -(L.plane_off[w] + jk + 1) : L.body_off[w] + (jk - np);
}

// World w's pairs start at the first key of its first body.
__global__ void __launch_bounds__(256) k_bp_world_off(BroadList L) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w > L.nw) return;
  const uint64_t want = uint64_t(uint32_t(L.body_off[w])) << 32;
  int64_t lo = 0, hi = L.n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (L.pkey[mid] < want)
      lo = mid + 1;
    else
      hi = mid;
  }
  L.list_off[w] = int32_t(lo);
}

template <typename R>
hipError_t launch_bp_bounds_sort(const BroadArgs<R>& a, void* tmp, size_t tmp_bytes, hipStream_t st) {
  if (a.nb <= 0) return hipSuccess;
  k_bp_bounds<R><<<(a.nb + 255) / 256, 256, 0, st>>>(a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return radix_sort_u64(a.key_in, a.key_out, a.val_in, a.val_out, a.nb, tmp, tmp_bytes, st);
}

template <typename R>
hipError_t launch_bp_sweep(const BroadArgs<R>& a, const BroadPlanes& p, hipStream_t st) {
  if (a.nb <= 0) return hipSuccess;
  hipError_t e = launch_fill_i32((int32_t*)a.pcount, 2, 0, st);
  if (e != hipSuccess) return e;
  k_bp_planes<<<(p.nb + 255) / 256, 256, 0, st>>>(p);
  k_bp_sweep<R><<<(a.nb + 255) / 256, 256, 0, st>>>(a);
  return hipGetLastError();
}

hipError_t launch_bp_list(const BroadList& L, void* tmp, size_t tmp_bytes, hipStream_t st) {
  hipError_t e = hipSuccess;
  if (L.n > 0) {
    e = radix_sort_u64(L.pkey_in, L.pkey, nullptr, nullptr, L.n, tmp, tmp_bytes, st);
    if (e != hipSuccess) return e;
    k_bp_unpack<<<int((L.n + 255) / 256), 256, 0, st>>>(L);
  }
  k_bp_world_off<<<(L.nw + 1 + 255) / 256, 256, 0, st>>>(L);
  return hipGetLastError();
}

// Temporary storage the two radix sorts need for n bodies and m pair keys.
hipError_t bp_sort_bytes(int32_t n, int64_t m, size_t* bytes) {
  if (m >= (int64_t(1) << 31)) return hipErrorInvalidValue;
  const size_t b1 = radix_sort_bytes(n, true), b2 = radix_sort_bytes(m, false);
  *bytes = b1 > b2 ? b1 : b2;
  return hipSuccess;
}

template hipError_t launch_bp_bounds_sort<double>(const BroadArgs<double>&, void*, size_t, hipStream_t);
template hipError_t launch_bp_bounds_sort<float>(const BroadArgs<float>&, void*, size_t, hipStream_t);
template hipError_t launch_bp_sweep<double>(const BroadArgs<double>&, const BroadPlanes&, hipStream_t);
template hipError_t launch_bp_sweep<float>(const BroadArgs<float>&, const BroadPlanes&, hipStream_t);

}  // namespace cbz
