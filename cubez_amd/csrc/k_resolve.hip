// k_resolve.hip — launch dispatch of the resolver (K3): picks the
// compile-time kind of the kernel for the requested phases and materials.
// The kernels are in cbz_resolve.h, instantiated by k_resolve_<precision>_<kind>.hip.
#include "cbz_resolve.h"

namespace cbz {

// The worlds the multi-wave kernel takes: of the worlds it can take
// (mw_fits), the mw_max costliest, the cost being the iterations of the
// phases this launch runs in the world's last resolve (world_iters), or its
// contact count where the world has none yet.  The whole last cost bucket
// that reaches mw_max is taken, then capped at mw_max in `order` (larger
// worlds first).  Writes mw_list[0 .. mw_count[0]), mw_flag for every
// world, and resets the queue head mw_count[1].  One workgroup; thread t
// looks at kSelectPer consecutive entries of each chunk.
constexpr int kSelectThreads = 1024, kSelectPer = 16, kSelectBins = 257, kSelectShift = 6;
static_assert(kSelectBins <= kSelectThreads, "one thread per bucket in the threshold scan");
__device__ __forceinline__ int select_bin(const int64_t* world_iters, int32_t mode, int32_t w, int C) {
  int64_t cost = 0;
  if (world_iters) {
    if (mode & RM_POS) cost += world_iters[2 * w];
    if (mode & RM_VEL) cost += world_iters[2 * w + 1];
  }
  if (cost <= 0) cost = C;
  const int64_t b = cost >> kSelectShift;
  return int(b < kSelectBins - 1 ? b : kSelectBins - 1);
}

__global__ void __launch_bounds__(kSelectThreads)
    k_mw_select(const int32_t* order, const int32_t* world_cnt, const int32_t* body_off, const int64_t* world_iters,
                int32_t mode, int32_t n_worlds, int32_t mw_min, int32_t mw_max, int32_t* mw_list, int32_t* mw_count,
                uint8_t* mw_flag) {
  __shared__ int32_t s_hist[kSelectBins];
  __shared__ int32_t s_wave[kSelectThreads / 64];
  __shared__ int32_t s_carry, s_thr;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int i = t; i < kSelectBins; i += kSelectThreads) s_hist[i] = 0;
  if (t == 0) s_carry = 0;
  __syncthreads();
  // cost histogram of the worlds the kernel can take
  for (int32_t w = t; w < n_worlds; w += kSelectThreads) {
    const int C = world_cnt[w];
    if (mw_fits(C, body_off[w + 1] - body_off[w], mw_min)) atomicAdd(&s_hist[select_bin(world_iters, mode, w, C)], 1);
  }
  if (t == 0) s_thr = mw_max <= 0 ? kSelectBins : 0;
  __syncthreads();
  // The threshold: the highest bucket whose suffix count (its worlds and
  // every costlier bucket's) reaches mw_max; 0 when none does, kSelectBins
  // (no world) when mw_max <= 0.  Thread r holds bucket kSelectBins-1-r, so
  // an inclusive scan over threads is the suffix count.
  {
    const int r = t;
    const int inc = wave_incl_scan_i32(r < kSelectBins ? s_hist[kSelectBins - 1 - r] : 0);
    if (lane == 63) s_wave[wv] = inc;
    __syncthreads();
    int suffix = inc;
    for (int v = 0; v < wv; ++v) suffix += s_wave[v];
    if (mw_max > 0 && r < kSelectBins && suffix >= mw_max) atomicMax(&s_thr, kSelectBins - 1 - r);
  }
  __syncthreads();
  const int thr = s_thr;
  for (int32_t b0 = 0; b0 < n_worlds; b0 += kSelectThreads * kSelectPer) {
    // this thread's entries: world ids, and a bit mask of the candidates
    int32_t w[kSelectPer];
    uint32_t cand = 0;
#pragma unroll
    for (int k = 0; k < kSelectPer; ++k) {
      const int32_t i = b0 + t * kSelectPer + k;
      w[k] = i < n_worlds ? (order ? order[i] : i) : -1;
    }
#pragma unroll
    for (int k = 0; k < kSelectPer; ++k) {
      if (w[k] >= 0) {
        const int C = world_cnt[w[k]];
        const bool c = mw_fits(C, body_off[w[k] + 1] - body_off[w[k]], mw_min) &&
                       select_bin(world_iters, mode, w[k], C) >= thr;
        cand |= uint32_t(c) << k;
        if (!c) mw_flag[w[k]] = 0;
      }
    }
    const int cnt = __builtin_popcount(cand);
    const int inc = wave_incl_scan_i32(cnt);
    if (lane == 63) s_wave[wv] = inc;
    __syncthreads();
    int pos = s_carry + inc - cnt;
    for (int v = 0; v < wv; ++v) pos += s_wave[v];
#pragma unroll
    for (int k = 0; k < kSelectPer; ++k) {
      if (!((cand >> k) & 1u)) continue;
      const bool take = pos < mw_max;
      if (take) mw_list[pos] = w[k];
      mw_flag[w[k]] = take ? 1 : 0;
      ++pos;
    }
    __syncthreads();
    if (t == kSelectThreads - 1) s_carry = pos;
    __syncthreads();
  }
  if (t == 0) {
    mw_count[0] = s_carry < mw_max ? s_carry : mw_max;
    mw_count[1] = 0;
  }
}

// prepareContacts alone (a position launch without iterations): one
// workgroup per world, a thread per contact, from the bodies' state in HBM
// (the arithmetic of the worlds' in-kernel prepare), with every contact's
// record for the one-wavefront kernel's wide classes.  A world kernel
// spends this launch in a lane's serial walk over its contacts.
template <typename R, int KIND>
__global__ void __launch_bounds__(256) k_prepare(RS<R> a) {
  const int w = blockIdx.x;
  const int C = a.world_cnt[w];
  if (a.overflow[0] != 0 || C <= 0) return;  // the resolver's guard
  // the bodies' positions at prepare time, which the prepared launches read
  const int b0w = a.body_off[w], n = a.body_off[w + 1] - b0w;
  const int64_t S = a.b.stride;
  for (int i = threadIdx.x; i < n; i += 256) st3(a.pos0, S, b0w + i, ld3(a.b.pos, S, b0w + i));
  const int64_t base = a.world_coff[w];
  for (int j = threadIdx.x; j < C; j += 256) prepare_contact<R, KIND, true>(a, base + j);
}

// With the multi-wave kernel: the launch's stream forks to `side`, which
// selects the worlds and runs the multi-wave kernel; the one-wave kernel
// starts on the launch's stream once the selection is done (so that the
// multi-wave workgroups, which each need a whole CU, are placed first), and
// the launch's stream waits for `side` at the end.
template <typename R, int KIND>
hipError_t launch_both(const RS<R>& a, const ResolveLaunch<R>& L, hipStream_t st) {
  hipError_t e;
  const bool mw = a.mw_list != nullptr;
  if (mw) {
    if ((e = hipEventRecord(L.fork, st)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(L.side, L.fork, 0)) != hipSuccess) return e;
    k_mw_select<<<1, kSelectThreads, 0, L.side>>>(a.order, a.world_cnt, a.body_off, a.world_iters, a.mode, a.n_worlds,
                                                   L.mw_min, L.mw_max, L.mw_list, L.mw_count, L.mw_flag);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipEventRecord(L.sel, L.side)) != hipSuccess) return e;
    int grid = L.mw_max < L.n_cu ? L.mw_max : L.n_cu;
    if (grid > a.n_worlds) grid = a.n_worlds;
    launch_mw_kind<R, KIND>(a, grid, L.side);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipEventRecord(L.join, L.side)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(st, L.sel, 0)) != hipSuccess) return e;
  }
  launch_kind<R, KIND>(a, st);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (mw) return hipStreamWaitEvent(st, L.join, 0);
  return hipSuccess;
}

template <typename R>
hipError_t launch_resolve(const ResolveLaunch<R>& L, hipStream_t st) {
  RS<R> a;
  a.b = L.b;
  a.c = L.c;
  a.world_cnt = L.world_cnt;
  a.world_coff = L.world_coff;
  a.overflow = L.overflow;
  a.order = L.order;
  a.body_off = L.body_off;
  a.world_iters = L.world_iters;
  a.world_time = L.world_time;
  const bool mw = L.mw_min > 0 && L.mw_max > 0 && L.mw_list && L.mw_count && L.mw_flag && L.side && L.fork && L.sel &&
                  L.join;
  a.mw_list = mw ? L.mw_list : nullptr;
  a.mw_count = mw ? L.mw_count : nullptr;
  a.mw_flag = mw ? L.mw_flag : nullptr;
  a.n_worlds = L.n_worlds;
  a.dt = L.dt;
  a.iter_mult = L.iter_mult;
  a.iter_fixed = L.iter_fixed;
  a.mode = L.mode;
  a.world_mask = L.world_mask;
  a.pos0 = L.pos0;
  a.labels = L.labels;
  a.delta = L.delta;
  a.vel0 = L.vel0;
  a.cvel0 = L.cvel0;
  a.polish_tol = L.polish_tol;
  a.polish_mult = L.polish_mult;
  if (a.mode & RM_POLISH) {
    // one phase per polish launch; the velocity polish also needs the
    // contact-velocity snapshot
    const bool pos = (a.mode & RM_POS) != 0, vel = (a.mode & RM_VEL) != 0;
    if (pos == vel || !a.world_mask || !a.delta || !a.vel0 || (vel && !a.cvel0)) return hipErrorInvalidValue;
  }
  if ((a.mode & RM_PREPARED) && !(a.mode & RM_POS)) return hipErrorInvalidValue;
  if ((a.mode & RM_LABEL) && !a.labels) return hipErrorInvalidValue;
  // per-contact materials: the plain resolver only
  if (a.c.fric && (!a.c.rest || (a.mode & (RM_POLISH | RM_LABEL)))) return hipErrorInvalidValue;
  if (a.n_worlds <= 0) return hipSuccess;
  if (a.iter_mult == 0 && a.iter_fixed == 0 && (a.mode & ~RM_FRICTIONLESS2) == RM_POS) {
    if (!a.pos0 || !a.c.rec) return hipErrorInvalidValue;  // k_prepare writes both
    if (a.c.fric)
      k_prepare<R, KIND_MAT><<<a.n_worlds, 256, 0, st>>>(a);
    else
      k_prepare<R, KIND_PLAIN><<<a.n_worlds, 256, 0, st>>>(a);
    return hipGetLastError();
  }
  if (a.c.fric) return launch_both<R, KIND_MAT>(a, L, st);
  if (a.mode & RM_POLISH)
    return (a.mode & RM_POS) ? launch_both<R, KIND_PPOLISH>(a, L, st) : launch_both<R, KIND_POLISH>(a, L, st);
  if (a.mode & RM_LABEL) return launch_both<R, KIND_LABEL>(a, L, st);
  return launch_both<R, KIND_PLAIN>(a, L, st);
}

// Worlds (workgroups) of the plain resolver kernel resident per CU, from
// its register and LDS use.
hipError_t resolver_occupancy(bool f64, int* worlds_per_cu) {
  return f64 ? kind_occupancy<double, KIND_PLAIN>(worlds_per_cu) : kind_occupancy<float, KIND_PLAIN>(worlds_per_cu);
}

template hipError_t launch_resolve<double>(const ResolveLaunch<double>&, hipStream_t);
template hipError_t launch_resolve<float>(const ResolveLaunch<float>&, hipStream_t);

}  // namespace cbz
