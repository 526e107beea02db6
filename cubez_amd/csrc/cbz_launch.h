// cbz_launch.h — host-side launch entry points of the step kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cbz_types.h"

namespace cbz {

// Resolver phases: RM_POS = prepareContacts + adjustPositions, RM_VEL =
// adjustVelocities (which reads the contact velocities prepare stored).
// Run together they are ResolveContacts; run apart, the velocity phase can
// be given to a subset of worlds.  RM_LABEL records, per contact, the sum of
// the velocity and rotation changes applyVelocityChange gave its two bodies.
// With RM_LABEL the position phase also records, per contact, the sum of
// the penetrations it resolved there (every selection's depth).
// RM_POLISH with RM_VEL: worlds whose mask byte is 0 first take the
// per-body velocity changes in `delta`, then run the velocity loop until
// every contact is within velocityEpsilon + polish_tol, at most
// polish_mult * C iterations; a world that does not get there is restored
// (velocities, rotations, motion, awake flags) and re-solved by the plain
// velocity phase, and its mask byte set to 1.  Worlds whose mask byte is 1
// run the plain velocity phase.
// RM_POLISH with RM_POS: the same for the position phase, from the per-body
// position and orientation changes in `delta`, to positionEpsilon +
// polish_tol; the restore covers positions, orientations, the awake flags
// and the motion, transform and world inverse inertia of the bodies asleep
// when the phase starts (snapshot in vel0).  A single launch polishes one
// phase.
// RM_PREPARED (with RM_POS): prepareContacts already ran, in an earlier
// launch with RM_POS and an iteration budget of 0; this launch reads the
// contact internals and prepare-time positions it stored.
// RM_FRICTIONLESS2: the frictionless impulse of a two-body contact includes
// its second body (the reference omits it, contact.go:512); only contacts
// whose friction is 0 (per-contact materials) take that impulse.
enum : int32_t { RM_POS = 1, RM_VEL = 2, RM_LABEL = 4, RM_POLISH = 8, RM_FRICTIONLESS2 = 16, RM_PREPARED = 32 };

template <typename R>
struct NarrowLaunch {
  Bodies<R> b;
  Planes<R> pl;
  PairSpace ps;
  Contacts<R> c;
  const double* brad;
  int32_t* tile_cnt;
  uint64_t* tile_mask;
  int32_t* tile_off;
  int32_t* world_cnt;
  int64_t* world_coff;
  int64_t* total;
  int32_t* overflow;
  uint32_t* hits;     // [cap] hit pairs as tile * 64 + lane (all-pairs)
  int32_t* n_hits;
  uint32_t* cand;     // [cand_cap] candidate pairs, same coding
  int32_t* n_cand;
  int64_t cand_cap;
  int32_t n_worlds;
  Materials<R> mat{};  // per-body and per-plane materials (read when c.fric is set)
};

template <typename R>
struct ResolveLaunch {
  Bodies<R> b;
  Contacts<R> c;
  const int32_t* world_cnt;
  const int64_t* world_coff;
  const int32_t* overflow;
  const int32_t* order;
  const int32_t* body_off;
  int64_t* world_iters;
  int64_t* world_time = nullptr;         // [4][n_worlds] wall-clock ticks (see the resolver's RS)
  // The multi-wave kernel for the mw_max largest worlds of at least mw_min
  // contacts (either 0: off), on stream `side`, forked from and joined back
  // into the launch's stream by the events; mw_list [n_worlds], mw_count [2]
  // and mw_flag [n_worlds] are its work list.
  int32_t mw_min = 0;
  int32_t mw_max = 0;
  int32_t* mw_list = nullptr;
  int32_t* mw_count = nullptr;
  uint8_t* mw_flag = nullptr;
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, sel = nullptr, join = nullptr;
  int32_t n_cu = 256;
  int32_t n_worlds;
  R dt;
  int64_t iter_mult;
  int64_t iter_fixed;
  int32_t mode = 3;                      // RM_POS | RM_VEL: the whole of ResolveContacts
  uint8_t* world_mask = nullptr;         // velocity phase only where set (RM_POLISH: in/out)
  R* pos0 = nullptr;                     // [3][n_bodies] positions at prepare time
  float* labels = nullptr;               // [cap][13] velocity-phase changes, resolved penetration per contact
  const float* delta = nullptr;          // RM_POLISH: [6][n_bodies] velocity, rotation (or position, orientation) changes
  R* vel0 = nullptr;                     // RM_POLISH: snapshot, [7][n_bodies] (RM_VEL) or [22][n_bodies] (RM_POS)
  R* cvel0 = nullptr;                    // RM_POLISH with RM_VEL: [3][cap] contact velocities after prepare
  R polish_tol = R(0);
  int64_t polish_mult = 0;
};

template <typename R>
hipError_t launch_calculate_derived(const Bodies<R>& b, hipStream_t st);
template <typename R>
hipError_t launch_integrate(const Bodies<R>& b, R dt, R bias, hipStream_t st);
template <typename R>
hipError_t launch_narrowphase(const NarrowLaunch<R>& L, hipStream_t st, hipEvent_t* ev = nullptr);
template <typename R>
hipError_t launch_resolve(const ResolveLaunch<R>& L, hipStream_t st);
hipError_t resolver_occupancy(bool f64, int* worlds_per_cu);

hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st);
hipError_t launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t st);
hipError_t launch_prof_accum(uint32_t* awake, const int64_t* total, const int64_t* iters, int n, int64_t* acc,
                             hipStream_t st);

// Device self-test of the FP64 operations whose rounding the parity claim
// depends on: op 0 = a/b, 1 = sqrt(a), 2 = a*b, 3 = a+b, 4 = a-b.
hipError_t launch_selftest(int op, int64_t n, const double* a, const double* b, double* out, hipStream_t st);

}  // namespace cbz
