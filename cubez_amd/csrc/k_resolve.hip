// k_resolve.hip — launch dispatch of the resolver (K3): picks the
// compile-time kind of the kernel for the requested phases and materials.
// The kernels are in cbz_resolve.h, instantiated by k_resolve_<precision>_<kind>.hip.
#include "cbz_resolve.h"

namespace cbz {

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
  if ((a.mode & RM_POLISH) && (!a.world_mask || !a.delta || !a.vel0 || !a.cvel0)) return hipErrorInvalidValue;
  if ((a.mode & RM_LABEL) && !a.labels) return hipErrorInvalidValue;
  // per-contact materials: the plain resolver only
  if (a.c.fric && (!a.c.rest || (a.mode & (RM_POLISH | RM_LABEL)))) return hipErrorInvalidValue;
  if (a.n_worlds <= 0) return hipSuccess;
  if (a.c.fric)
    launch_kind<R, KIND_MAT>(a, st);
  else if (a.mode & RM_POLISH)
    launch_kind<R, KIND_POLISH>(a, st);
  else if (a.mode & RM_LABEL)
    launch_kind<R, KIND_LABEL>(a, st);
  else
    launch_kind<R, KIND_PLAIN>(a, st);
  return hipGetLastError();
}

// Worlds (workgroups) of the plain resolver kernel resident per CU, from
// its register and LDS use.
hipError_t resolver_occupancy(bool f64, int* worlds_per_cu) {
  return f64 ? kind_occupancy<double, KIND_PLAIN>(worlds_per_cu) : kind_occupancy<float, KIND_PLAIN>(worlds_per_cu);
}

template hipError_t launch_resolve<double>(const ResolveLaunch<double>&, hipStream_t);
template hipError_t launch_resolve<float>(const ResolveLaunch<float>&, hipStream_t);

}  // namespace cbz
