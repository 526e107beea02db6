// cbz_resolve.h — K3: ResolveContacts (contact.go:208-222), one wavefront
// per world.  Device code of the resolver kernel; each (precision, kind)
// instantiation is compiled in a translation unit of its own
// (k_resolve_<precision>_<kind>.hip), the launch dispatch is k_resolve.hip.
//
// The reference's resolver is sequential: each iteration picks the worst
// contact (largest penetration, then largest desired velocity change; ties
// go to the lowest index because the scan uses a strict '>'), applies a
// correction to its one or two bodies, and updates every contact that
// shares one of those bodies.  A world therefore runs on one wave: the
// argmax is a lane-strided scan plus a butterfly reduction, the correction
// is computed redundantly by all 64 lanes (lane 0 stores it), and the
// update scans the world's contacts lane-parallel, each contact applying
// its matches in the reference's (b, d) order.  Worlds run concurrently.
//
// The selection keys and the body indices of the world's contacts live in
// LDS when the world has at most kKeyCap contacts, otherwise in HBM.
#pragma once
#include <climits>
#include <utility>

#include "cbz_devmath.h"
#include "cbz_launch.h"
#include "cbz_mlp.h"
#include "cbz_types.h"
#include "cbz_wave.h"

namespace cbz {

constexpr int kKeyCap = 1280;

// Relative energy slack of the surrogate's acceptance test (rounding); the
// absolute slack is 1/2 tol^2 per kilogram: the kinetic energy of the
// residual closing speed the test already accepts.
constexpr double kEnergyRel = 1e-5;

// Compile-time variants of the resolver kernel, so that the plain
// resolver's hot loops carry no code (and no register pressure) of the
// others: KIND_PLAIN (ResolveContacts, or one of its phases), KIND_LABEL
// (also records every impulse's body changes and every resolved
// penetration per contact), KIND_POLISH (the surrogate's velocity polish
// with the plain fallback), KIND_MAT (KIND_PLAIN with per-contact friction
// and restitution, and the frictionless impulse for contacts whose
// friction is 0), KIND_PPOLISH (the surrogate's position polish with the
// plain fallback).
enum : int { KIND_PLAIN = 0, KIND_LABEL = 1, KIND_POLISH = 2, KIND_MAT = 3, KIND_PPOLISH = 4 };

// The position polish's separation bound: a polished world is rejected if a
// contact ends separated by more than kSepFactor times the world's deepest
// penetration at the start (+ the tolerance).  The plain phase stays within
// 1.4 times it (every world of 4096 at six points of a 600-step run; 0.1 %
// of worlds pass 1.0 times).
constexpr double kSepFactor = 2.0;
// A prediction that breaks that bound is tried once more at this fraction of
// its moves before the world is rejected: most over-separations come from
// moves that are too large, and the loop finishes what smaller ones leave.
constexpr double kPosRetryScale = 1.0 / 3.0;

template <typename R>
struct RS {
  Bodies<R> b;
  Contacts<R> c;
  const int32_t* world_cnt;
  const int64_t* world_coff;
  const int32_t* overflow;
  const int32_t* order;
  const int32_t* body_off;
  int64_t* world_iters;  // [2 * n_worlds]: position / velocity iterations
  // [4 * n_worlds] (or null): wall-clock ticks when the world's workgroup
  // started and ended, for a launch without (0, 1) and with (2, 3) RM_VEL
  int64_t* world_time;
  // worlds taken by the multi-wave kernel (resolve_mw), chosen by
  // k_mw_select: mw_list[0 .. mw_count[0]), flagged in mw_flag; mw_count[1]
  // is the kernel's work-queue head (null mw_list: none)
  const int32_t* mw_list;
  int32_t* mw_count;
  const uint8_t* mw_flag;
  int32_t n_worlds;
  R dt;
  int64_t iter_mult;
  int64_t iter_fixed;
  int32_t mode;                // RM_* bits
  uint8_t* world_mask;         // velocity phase only where set (null: every world); RM_POLISH: in/out
  R* pos0;                     // [3][n_bodies] body positions at prepare time
  // [cap][kOut] per contact: accumulated velocity-phase changes (0-11) and
  // resolved penetration (12) (null: off)
  float* labels;
  // RM_POLISH: [6][n_bodies] velocity and rotation changes (RM_VEL) or
  // position and orientation changes (RM_POS)
  const float* delta;
  // RM_POLISH snapshot: [7][n_bodies] vel, rot, motion before the changes
  // (RM_VEL); [22][n_bodies] motion, transform, iit_world of the bodies
  // asleep when the phase starts (RM_POS)
  R* vel0;
  R* cvel0;                    // RM_POLISH with RM_VEL: [3][cap] contact velocities as prepare left them
  R polish_tol;
  int64_t polish_mult;
};


template <typename R>
__device__ __forceinline__ void add_label(const RS<R>& a, int64_t g, const V3<R> vc[2], const V3<R> rc[2]) {
  float* L = a.labels + g * kOut;
  const V3<R>* v[4] = {&vc[0], &rc[0], &vc[1], &rc[1]};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    L[3 * k + 0] += float(v[k]->x);
    L[3 * k + 1] += float(v[k]->y);
    L[3 * k + 2] += float(v[k]->z);
  }
}

// The position phase's label: the penetration resolved at the contact.
template <typename R>
__device__ __forceinline__ void add_pos_label(const RS<R>& a, int64_t g, R depth) {
  a.labels[g * kOut + kOutPos] += float(depth);
}

template <typename R>
__device__ __forceinline__ V3<R> ld3(const R* a, int64_t s, int64_t i) {
  return V3<R>{a[i], a[s + i], a[2 * s + i]};
}
template <typename R>
__device__ __forceinline__ void st3(R* a, int64_t s, int64_t i, const V3<R>& v) {
  a[i] = v.x;
  a[s + i] = v.y;
  a[2 * s + i] = v.z;
}
template <typename R>
__device__ __forceinline__ void ld9(const R* a, int64_t s, int64_t i, R* o) {
#pragma unroll
  for (int k = 0; k < 9; ++k) o[k] = a[k * s + i];
}

template <typename R>
__device__ __forceinline__ bool awake(const Bodies<R>& b, int i) {
  return (b.flags[i] & F_AWAKE) != 0;
}

// calculateLocalVelocity, contact.go:159-182
template <typename R>
__device__ __forceinline__ V3<R> local_velocity(const Bodies<R>& b, int body, const V3<R>& rel, const R* c2w,
                                                R dt) {
  const int64_t S = b.stride;
  V3<R> velocity = add(cross(ld3(b.rot, S, body), rel), ld3(b.vel, S, body));
  V3<R> cv = m3tv(c2w, velocity);
  V3<R> accv = m3tv(c2w, mul(ld3(b.last_acc, S, body), dt));
  accv.x = R(0);
  return add(cv, accv);
}

// Contact.Restitution and Contact.Friction of contact g: the reference's
// constants (colliders.go:201-202) unless the kernel is KIND_MAT.
template <typename R, int KIND>
__device__ __forceinline__ R restitution_of(const Contacts<R>& c, int64_t g) {
  if constexpr (KIND == KIND_MAT) return c.rest[g];
  return R(kRestitution);
}
template <typename R, int KIND>
__device__ __forceinline__ R friction_of(const Contacts<R>& c, int64_t g) {
  if constexpr (KIND == KIND_MAT) return c.fric[g];
  return R(kFriction);
}

// calculateDesiredDeltaVelocity, contact.go:87-112, with the contact's
// restitution e
template <typename R>
__device__ __forceinline__ R desired_dv(const Bodies<R>& b, int b0, int b1, const V3<R>& n, R cvx, R dt, R e) {
  const int64_t S = b.stride;
  R vfa = R(0);
  if (awake(b, b0)) vfa += dot(mul(ld3(b.last_acc, S, b0), dt), n);
  if (b1 >= 0 && awake(b, b1)) vfa -= dot(mul(ld3(b.last_acc, S, b1), dt), n);
  R restitution = e;
  if (dabs(cvx) < R(0.25)) restitution = R(0);
  return -cvx - restitution * (cvx - vfa);
}

// calculateContactBasis, contact.go:118-156, split in two: the scale
// s = 1/sqrt(...) (the only square root and division), and the frame built
// from the normal and s.  Rebuilding the frame from a stored s gives the
// same bits as the full computation.  Both run in FP64 whatever the storage
// type R (the FP32 variant keeps the basis in FP64; for R = double nothing
// is converted).
template <typename R>
__device__ __forceinline__ double basis_scale(const V3<R>& nr) {
  const V3<double> n{double(nr.x), double(nr.y), double(nr.z)};
  if (dabs(n.x) > dabs(n.y)) return 1.0 / dsqrt(n.z * n.z + n.x * n.x);
  return 1.0 / dsqrt(n.z * n.z + n.y * n.y);
}
template <typename R>
__device__ __forceinline__ void basis_from_scale(const V3<R>& nr, double s, R* m) {
  const V3<double> n{double(nr.x), double(nr.y), double(nr.z)};
  V3<double> ty, tz;
  if (dabs(n.x) > dabs(n.y)) {
    ty = V3<double>{n.z * s, 0.0, n.x * -s};
    tz = V3<double>{n.y * ty.x, n.z * ty.x - n.x * ty.z, -n.y * ty.x};
  } else {
    ty = V3<double>{0.0, -n.z * s, n.y * s};
    tz = V3<double>{n.y * ty.z - n.z * ty.y, -n.x * ty.z, n.x * ty.y};
  }
  // SetComponents(normal, tangentY, tangentZ), math/matrix.go:52-56
  m[0] = nr.x; m[1] = nr.y; m[2] = nr.z;
  m[3] = R(ty.x); m[4] = R(ty.y); m[5] = R(ty.z);
  m[6] = R(tz.x); m[7] = R(tz.y); m[8] = R(tz.z);
}
template <typename R>
__device__ __forceinline__ void contact_basis(const V3<R>& n, R* m) {
  basis_from_scale(n, basis_scale(n), m);
}

// A contact's record (Contacts::rec): 16 reals from one 64-byte (FP32) or
// 128-byte (FP64) aligned address, in 16-byte loads.
template <typename R>
__device__ __forceinline__ void load_rec(const R* q, R (&v)[16]) {
  if constexpr (sizeof(R) == 4) {
    const float4* p = reinterpret_cast<const float4*>(q);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float4 x = p[i];
      v[4 * i] = x.x;
      v[4 * i + 1] = x.y;
      v[4 * i + 2] = x.z;
      v[4 * i + 3] = x.w;
    }
  } else {
    const double2* p = reinterpret_cast<const double2*>(q);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const double2 x = p[i];
      v[2 * i] = x.x;
      v[2 * i + 1] = x.y;
    }
  }
}
// contactToWorld and both relative positions
template <typename R>
__device__ __forceinline__ void rec_frame(const R* q, R (&m)[9], V3<R> (&r)[2]) {
  R v[16];
  load_rec(q, v);
#pragma unroll
  for (int k = 0; k < 9; ++k) m[k] = v[k];
  r[0] = V3<R>{v[9], v[10], v[11]};
  r[1] = V3<R>{v[12], v[13], v[14]};
}
// the normal (contactToWorld's first column) and both relative positions
template <typename R>
__device__ __forceinline__ void rec_normal(const R* q, V3<R>& n, V3<R> (&r)[2]) {
  R v[16];
  load_rec(q, v);
  n = V3<R>{v[0], v[1], v[2]};
  r[0] = V3<R>{v[9], v[10], v[11]};
  r[1] = V3<R>{v[12], v[13], v[14]};
}
template <typename R>
__device__ __forceinline__ void store_rec(R* q, const R (&m)[9], const V3<R>& r0, const V3<R>& r1) {
  const R v[16] = {m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, R(0)};
  if constexpr (sizeof(R) == 4) {
    float4* p = reinterpret_cast<float4*>(q);
#pragma unroll
    for (int i = 0; i < 4; ++i) p[i] = float4{v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]};
  } else {
    double2* p = reinterpret_cast<double2*>(q);
#pragma unroll
    for (int i = 0; i < 8; ++i) p[i] = double2{v[2 * i], v[2 * i + 1]};
  }
}

// calculateInternals, contact.go:59-85, for contact g.
// With REC, also the contact's record (Contacts::rec).
template <typename R, int KIND, bool REC = false>
__device__ __forceinline__ void prepare_contact(const RS<R>& a, int64_t g) {
  const Bodies<R>& b = a.b;
  const Contacts<R>& c = a.c;
  const int64_t S = b.stride, CS = c.cap;
  int b0 = c.body0[g], b1 = c.body1[g];
  V3<R> n = ld3(c.normal, CS, g);
  if (b0 < 0) {  // contact.go:61-65 (not produced by the narrowphase)
    n = mul(n, R(-1));
    b0 = b1;
    b1 = -1;
    c.body0[g] = b0;
    c.body1[g] = b1;
    st3(c.normal, CS, g, n);
  }
  R m[9];
  contact_basis(n, m);
  const V3<R> p = ld3(c.point, CS, g);
  const V3<R> r0 = sub(p, ld3(b.pos, S, b0));
  V3<R> cv = local_velocity(b, b0, r0, m, a.dt);
  V3<R> r1{R(0), R(0), R(0)};
  if (b1 >= 0) {
    r1 = sub(p, ld3(b.pos, S, b1));
    cv = sub(cv, local_velocity(b, b1, r1, m, a.dt));
  }
  const R ddv = desired_dv(b, b0, b1, n, cv.x, a.dt, restitution_of<R, KIND>(c, g));
#pragma unroll
  for (int k = 0; k < 9; ++k) c.c2w[k * CS + g] = m[k];
  st3(c.rel, CS, g, r0);
  st3(c.rel + 3 * CS, CS, g, r1);
  st3(c.cvel, CS, g, cv);
  c.ddv[g] = ddv;
  if constexpr (REC) store_rec(c.rec + 16 * g, m, r0, r1);
}

// matchAwakeState, contact.go:185-202.  Returns the awake states after the
// call in a0/a1 (lane 0 stores the change).
template <typename R>
__device__ __forceinline__ void match_awake(const RS<R>& a, int k0, int k1, bool& a0, bool& a1, int lane) {
  a0 = awake(a.b, k0);
  a1 = k1 >= 0 ? awake(a.b, k1) : false;
  if (k1 < 0) return;
  if ((a0 || a1) && !(a0 && a1)) {
    const int t = a0 ? k1 : k0;
    if (lane == 0) {  // SetAwake(true), rigidbody.go:183-186
      a.b.flags[t] = uint8_t(a.b.flags[t] | F_AWAKE);
      a.b.motion[t] = R(kSleepEpsilon * 2.0);
    }
    a0 = a1 = true;
  }
}

// CalculateDerivedData (rigidbody.go:268-272) for a sleeping body moved by
// applyPositionChange (contact.go:380-382).
template <typename R>
__device__ __forceinline__ void rederive(const Bodies<R>& b, int i, const V3<R>& pos, Q4<R>& q, int lane) {
  const int64_t S = b.stride;
  q = q_normalize(q);
  R T[12], IB[9], IW[9];
  set_as_transform(pos, q, T);
  ld9(b.iit_body, S, i, IB);
  transform_inertia(IB, T, IW);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 12; ++k) b.transform[k * S + i] = T[k];
#pragma unroll
    for (int k = 0; k < 9; ++k) b.iit_world[k * S + i] = IW[k];
  }
}

// applyPositionChange, contact.go:286-386
template <typename R>
__device__ __forceinline__ void apply_position_change(const RS<R>& a, const int kb[2], const bool aw[2], const V3<R>& n,
                                      const V3<R> rel[2], R penetration, V3<R> lin[2], V3<R> ang[2], int lane) {
  const Bodies<R>& b = a.b;
  const int64_t S = b.stride;
  const R angularLimit = R(0.2);
  R angI[2] = {R(0), R(0)}, linI[2] = {R(0), R(0)};
  R total = R(0);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    lin[i] = V3<R>{R(0), R(0), R(0)};
    ang[i] = V3<R>{R(0), R(0), R(0)};
    if (kb[i] < 0) continue;
    R iw[9];
    ld9(b.iit_world, S, kb[i], iw);
    V3<R> aiw = cross(rel[i], n);
    aiw = m3v(iw, aiw);
    aiw = cross(aiw, rel[i]);
    angI[i] = dot(aiw, n);
    linI[i] = b.inv_mass[kb[i]];
    total += linI[i] + angI[i];
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int bi = kb[i];
    if (bi < 0) continue;
    R sign = R(1);
    if (i != 0) sign = R(-1);
    R angM = sign * penetration * (angI[i] / total);
    R linM = sign * penetration * (linI[i] / total);
    V3<R> projection = add_scaled(rel[i], n, -dot(rel[i], n));
    R maxMagnitude = angularLimit * mag(projection);
    if (angM < -maxMagnitude) {
      R totalMove = angM + linM;
      angM = -maxMagnitude;
      linM = totalMove - angM;
    } else if (angM > maxMagnitude) {
      R totalMove = angM + linM;
      angM = maxMagnitude;
      linM = totalMove - angM;
    }
    if (angM == R(0)) {
      ang[i] = V3<R>{R(0), R(0), R(0)};
    } else {
      V3<R> tad = cross(rel[i], n);
      R iw[9];
      ld9(b.iit_world, S, bi, iw);
      ang[i] = mul(m3v(iw, tad), angM / angI[i]);
    }
    lin[i] = mul(n, linM);
    V3<R> pos = add_scaled(ld3(b.pos, S, bi), n, linM);
    Q4<R> q{b.quat[bi], b.quat[S + bi], b.quat[2 * S + bi], b.quat[3 * S + bi]};
    q = q_add_scaled_vector(q, ang[i], R(1));
    q = q_normalize(q);
    if (!aw[i]) rederive(b, bi, pos, q, lane);
    if (lane == 0) {
      st3(b.pos, S, bi, pos);
      b.quat[bi] = q.w;
      b.quat[S + bi] = q.x;
      b.quat[2 * S + bi] = q.y;
      b.quat[3 * S + bi] = q.z;
    }
  }
}

// setSkewSymmetric, contact.go:612-616
template <typename R>
__device__ __forceinline__ void skew(const V3<R>& v, R* m) {
  m[0] = R(0); m[3] = -v.z; m[6] = v.y;
  m[1] = v.z;  m[4] = R(0); m[7] = -v.x;
  m[2] = -v.y; m[5] = v.x;  m[8] = R(0);
}

// calculateFrictionlessImpulse, contact.go:498-531, taken by contacts
// whose friction is 0 (contact.go:459-461; only KIND_MAT has such
// contacts): the impulse is (desiredDeltaVelocity / deltaVelocity, 0, 0)
// with this deltaVelocity, which depends only on quantities the velocity
// phase does not change.  ContactNormal is column 0 of contactToWorld.  The
// reference's test at :512 is inverted (Bodies[1] == nil): a two-body
// contact never adds its second body's terms, and a half-space contact
// dereferences nil there (a Go panic; here the absent body adds nothing).
// two_body (RM_FRICTIONLESS2) adds the second body's terms when it exists.
// im: the two bodies' inverse masses (im[1] unused without a second body).
template <typename R>
__device__ __forceinline__ R frictionless_delta_velocity(const R im[2], bool has1, const R* iit0, const R* iit1,
                                                         const V3<R> rel[2], const R* c2w, bool two_body) {
  const V3<R> n{c2w[0], c2w[1], c2w[2]};
  V3<R> dvw = cross(rel[0], n);
  dvw = m3v(iit0, dvw);
  dvw = cross(dvw, rel[0]);
  R deltaVelocity = dot(dvw, n);
  deltaVelocity += im[0];
  if (two_body && has1) {
    dvw = cross(rel[1], n);
    dvw = m3v(iit1, dvw);
    dvw = cross(dvw, rel[1]);
    deltaVelocity += dot(dvw, n);
    deltaVelocity += im[1];
  }
  return deltaVelocity;
}

// calculateFrictionImpulse, contact.go:535-606, in two parts.  The first
// builds the contact-frame matrix deltaVelocity (the velocity change per
// unit impulse, :537-578) and its inverse impulseMatrix (:581); they depend
// only on contactToWorld, the relative contact positions, both bodies'
// world inverse inertia and inverse masses, none of which the velocity
// phase changes, so the register-resident resolver computes them once per
// phase for every contact.  It keeps the inverse (9) and deltaVelocity's
// entries 0, 3, 6 (the only ones the friction cone reads).
template <typename R>
__device__ __forceinline__ void friction_matrices(const R invm[2], bool has1, const R* iit0, const R* iit1,
                                                  const V3<R> rel[2], const R* c2w, R* im, R* dvcol) {
  R inverseMass = invm[0];
  R sk[9], t[9], dvw[9];
  skew(rel[0], sk);
  m3mul(sk, iit0, t);
  m3mul(t, sk, dvw);
#pragma unroll
  for (int k = 0; k < 9; ++k) dvw[k] *= R(-1);
  if (has1) {
    R d2[9];
    skew(rel[1], sk);
    m3mul(sk, iit1, t);
    m3mul(t, sk, d2);
#pragma unroll
    for (int k = 0; k < 9; ++k) d2[k] *= R(-1);
#pragma unroll
    for (int k = 0; k < 9; ++k) dvw[k] += d2[k];
    inverseMass += invm[1];
  }
  R ct[9] = {c2w[0], c2w[3], c2w[6], c2w[1], c2w[4], c2w[7], c2w[2], c2w[5], c2w[8]};
  R dv[9];
  m3mul(ct, dvw, t);
  m3mul(t, c2w, dv);
  dv[0] += inverseMass;
  dv[4] += inverseMass;
  dv[8] += inverseMass;
  m3inv(dv, im);
  dvcol[0] = dv[0];
  dvcol[1] = dv[3];
  dvcol[2] = dv[6];
}

// The second part (:582-606): the impulse that removes the desired normal
// velocity change and the tangential velocity, projected onto the friction
// cone when it leaves it.
template <typename R>
__device__ __forceinline__ V3<R> friction_impulse_from(const R* im, const R* dvcol, R ddv, const V3<R>& cvel,
                                                       R friction) {
  V3<R> velKill{ddv, -cvel.y, -cvel.z};
  V3<R> ic = m3v(im, velKill);
  R planar = dsqrt(ic.y * ic.y + ic.z * ic.z);
  if (planar > ic.x * friction) {
    ic.y /= planar;
    ic.z /= planar;
    ic.x = dvcol[0] + dvcol[1] * friction * ic.y + dvcol[2] * friction * ic.z;
    ic.x = ddv / ic.x;
    ic.y *= friction * ic.x;
    ic.z *= friction * ic.x;
  }
  return ic;
}

// The per-contact data of the velocity phase that the phase does not
// change: impulseMatrix (9) and deltaVelocity's entries 0, 3, 6 (3) of
// calculateFrictionImpulse, or, for a contact of friction 0 (KIND_MAT),
// calculateFrictionlessImpulse's deltaVelocity in entry 0.
constexpr int kImpPlanes = 12;

// From the bodies' world inverse inertia and inverse masses as given.
template <typename R, int KIND>
__device__ __forceinline__ void impulse_data_from(int32_t mode, const R* iit0, const R* iit1, const R invm[2], bool has1,
                                                  const V3<R> rel[2], const R* c2w, R friction, R* out) {
  if (KIND == KIND_MAT && friction == R(0)) {
    out[0] = frictionless_delta_velocity(invm, has1, iit0, iit1, rel, c2w, (mode & RM_FRICTIONLESS2) != 0);
#pragma unroll
    for (int k = 1; k < kImpPlanes; ++k) out[k] = R(0);
  } else {
    friction_matrices(invm, has1, iit0, iit1, rel, c2w, out, out + 9);
  }
}

template <typename R, int KIND>
__device__ __forceinline__ void impulse_data(const RS<R>& a, const int kb[2], const V3<R> rel[2], const R* c2w,
                                             R friction, R* out) {
  const Bodies<R>& b = a.b;
  const int64_t S = b.stride;
  R iit0[9], iit1[9];
  ld9(b.iit_world, S, kb[0], iit0);
#pragma unroll
  for (int k = 0; k < 9; ++k) iit1[k] = R(0);
  if (kb[1] >= 0) ld9(b.iit_world, S, kb[1], iit1);
  const R invm[2] = {b.inv_mass[kb[0]], kb[1] >= 0 ? b.inv_mass[kb[1]] : R(0)};
  impulse_data_from<R, KIND>(a.mode, iit0, iit1, invm, kb[1] >= 0, rel, c2w, friction, out);
}

// applyVelocityChange, contact.go:448-494, for a contact of the given
// friction, from its impulse data (impulse_data): the impulse in contact
// coordinates (the frictionless branch of :459-461 exists in KIND_MAT only),
// then the velocity and rotation changes of both bodies (:463-494).
template <typename R, int KIND>
__device__ __forceinline__ void apply_velocity_change(const RS<R>& a, const int kb[2], const V3<R> rel[2],
                                                      const R* c2w, R ddv, const V3<R>& cvel, const R* imp,
                                                      V3<R> vc[2], V3<R> rc[2], int lane, R friction) {
  const Bodies<R>& b = a.b;
  const int64_t S = b.stride;
  V3<R> ic;
  if (KIND == KIND_MAT && friction == R(0))
    ic = V3<R>{ddv / imp[0], R(0), R(0)};
  else
    ic = friction_impulse_from(imp, imp + 9, ddv, cvel, friction);
  const V3<R> impulse = m3v(c2w, ic);
  const V3<R> zero{R(0), R(0), R(0)};
  R iw[9];
  ld9(b.iit_world, S, kb[0], iw);
  V3<R> torque = cross(rel[0], impulse);
  rc[0] = m3v(iw, torque);
  vc[0] = add_scaled(zero, impulse, b.inv_mass[kb[0]]);
  rc[1] = zero;
  vc[1] = zero;
  const V3<R> v0 = add(ld3(b.vel, S, kb[0]), vc[0]);
  const V3<R> w0 = add(ld3(b.rot, S, kb[0]), rc[0]);
  if (lane == 0) {
    st3(b.vel, S, kb[0], v0);
    st3(b.rot, S, kb[0], w0);
  }
  if (kb[1] >= 0) {
    ld9(b.iit_world, S, kb[1], iw);
    torque = cross(impulse, rel[1]);
    rc[1] = m3v(iw, torque);
    vc[1] = add_scaled(zero, impulse, -b.inv_mass[kb[1]]);
    const V3<R> v1 = add(ld3(b.vel, S, kb[1]), vc[1]);
    const V3<R> w1 = add(ld3(b.rot, S, kb[1]), rc[1]);
    if (lane == 0) {
      st3(b.vel, S, kb[1], v1);
      st3(b.rot, S, kb[1], w1);
    }
  }
}

template <typename R, bool LDS, int KIND>
__device__ __forceinline__ void resolve_world(const RS<R>& a, int w, int C, int64_t base, R* key, int32_t* kb0, int32_t* kb1,
                              int lane) {
  const Bodies<R>& b = a.b;
  const Contacts<R>& c = a.c;
  const int64_t CS = c.cap;
  const int64_t maxIter = a.iter_fixed > 0 ? a.iter_fixed : a.iter_mult * int64_t(C);

  const bool do_pos = (a.mode & RM_POS) != 0;
  const bool do_prep = do_pos && !(a.mode & RM_PREPARED);
  const bool do_vel = (a.mode & RM_VEL) != 0 && (KIND == KIND_POLISH || !a.world_mask || a.world_mask[w]);
  // worlds this path takes are never polished: the plain phase
  if ((KIND == KIND_POLISH || KIND == KIND_PPOLISH) && lane == 0) a.world_mask[w] = 1;
  // prepareContacts, contact.go:225-229
  for (int j = lane; j < C; j += 64) {
    if (do_prep) prepare_contact<R, KIND>(a, base + j);
    if (LDS) {
      key[j] = c.pen[base + j];
      kb0[j] = c.body0[base + j];
      kb1[j] = c.body1[base + j];
    }
  }
  __syncthreads();

  // adjustPositions, contact.go:233-283
  int64_t used = 0;
  while (do_pos && used < maxIter) {
    R best = R(kPositionEpsilon);
    int bi = INT_MAX;
    for (int j = lane; j < C; j += 64) {
      const R p = key[j];
      if (p > best) {
        best = p;
        bi = j;
      }
    }
    wave_argmax_dpp(best, bi);
    if (bi == INT_MAX) break;
    const int64_t g = base + bi;
    const int kb[2] = {kb0[bi], kb1[bi]};
    const V3<R> n = ld3(c.normal, CS, g);
    const V3<R> rel[2] = {ld3(c.rel, CS, g), ld3(c.rel + 3 * CS, CS, g)};
    bool aw[2];
    match_awake(a, kb[0], kb[1], aw[0], aw[1], lane);
    V3<R> lin[2], ang[2];
    apply_position_change(a, kb, aw, n, rel, best, lin, ang, lane);
    if constexpr (KIND == KIND_LABEL)
      if (lane == 0) add_pos_label(a, g, best);
    for (int j = lane; j < C; j += 64) {
      const int jb[2] = {kb0[j], kb1[j]};
      bool any = false;
#pragma unroll
      for (int bb = 0; bb < 2; ++bb)
#pragma unroll
        for (int d = 0; d < 2; ++d) any |= (jb[bb] >= 0 && jb[bb] == kb[d]);
      if (!any) continue;
      R p = key[j];
      const V3<R> nj = ld3(c.normal, CS, base + j);
#pragma unroll
      for (int bb = 0; bb < 2; ++bb) {
        if (jb[bb] < 0) continue;
        const V3<R> rj = ld3(c.rel + 3 * bb * CS, CS, base + j);
        // a side is at most one of the two moved bodies: its change is
        // selected and applied once
        if (jb[bb] == kb[0] || jb[bb] == kb[1]) {
          const bool d0 = jb[bb] == kb[0];
          V3<R> dp = add(cross(sel3(d0, ang[0], ang[1]), rj), sel3(d0, lin[0], lin[1]));
          R sign = R(1);
          if (bb == 0) sign = R(-1);
          p += dot(dp, nj) * sign;
        }
      }
      key[j] = p;
    }
    __syncthreads();
    ++used;
  }
  const int64_t pos_iters = used;
  if (LDS) {
    for (int j = lane; j < C; j += 64) {
      if (do_pos) c.pen[base + j] = key[j];
      key[j] = c.ddv[base + j];
    }
    __syncthreads();
  }
  R* dkey = LDS ? key : c.ddv + base;

  // adjustVelocities, contact.go:390-445
  used = 0;
  while (do_vel && used < maxIter) {
    R best = R(kVelocityEpsilon);
    int bi = INT_MAX;
    for (int j = lane; j < C; j += 64) {
      const R p = dkey[j];
      if (p > best) {
        best = p;
        bi = j;
      }
    }
    wave_argmax_dpp(best, bi);
    if (bi == INT_MAX) break;
    const int64_t g = base + bi;
    const int kb[2] = {kb0[bi], kb1[bi]};
    const V3<R> rel[2] = {ld3(c.rel, CS, g), ld3(c.rel + 3 * CS, CS, g)};
    R m[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = c.c2w[k * CS + g];
    const V3<R> cv = ld3(c.cvel, CS, g);
    bool aw[2];
    match_awake(a, kb[0], kb[1], aw[0], aw[1], lane);
    V3<R> vc[2], rc[2];
    const R fr = friction_of<R, KIND>(c, g);
    R imp[kImpPlanes];
    impulse_data<R, KIND>(a, kb, rel, m, fr, imp);
    apply_velocity_change<R, KIND>(a, kb, rel, m, best, cv, imp, vc, rc, lane, fr);
    if constexpr (KIND == KIND_LABEL)
      if (lane == 0) add_label(a, g, vc, rc);
    for (int j = lane; j < C; j += 64) {
      const int jb[2] = {kb0[j], kb1[j]};
      bool any = false;
#pragma unroll
      for (int bb = 0; bb < 2; ++bb)
#pragma unroll
        for (int d = 0; d < 2; ++d) any |= (jb[bb] >= 0 && jb[bb] == kb[d]);
      if (!any) continue;
      const int64_t gj = base + j;
      R mj[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) mj[k] = c.c2w[k * CS + gj];
      V3<R> cvj = ld3(c.cvel, CS, gj);
#pragma unroll
      for (int bb = 0; bb < 2; ++bb) {
        if (jb[bb] < 0) continue;
        const V3<R> rj = ld3(c.rel + 3 * bb * CS, CS, gj);
        // a side is at most one of the two moved bodies: its change is
        // selected and applied once
        if (jb[bb] == kb[0] || jb[bb] == kb[1]) {
          const bool d0 = jb[bb] == kb[0];
          V3<R> dv = add(cross(sel3(d0, rc[0], rc[1]), rj), sel3(d0, vc[0], vc[1]));
          R sign = R(1);
          if (bb == 1) sign = R(-1);
          cvj = add(cvj, mul(m3tv(mj, dv), sign));
        }
      }
      st3(c.cvel, CS, gj, cvj);
      const V3<R> nj = ld3(c.normal, CS, gj);
      dkey[j] = desired_dv(b, jb[0], jb[1], nj, cvj.x, a.dt, restitution_of<R, KIND>(c, gj));
    }
    __syncthreads();
    ++used;
  }
  if (LDS && do_vel) {
    for (int j = lane; j < C; j += 64) c.ddv[base + j] = dkey[j];
  }
  if (lane == 0 && a.world_iters) {
    if (do_pos) a.world_iters[2 * w] = pos_iters;
    if (do_vel) a.world_iters[2 * w + 1] = used;
  }
}

constexpr int kRegBodies = 256;   // bodies per world for the register-resident kernel
constexpr int kRegMaxSlots = 32;  // contacts per lane (64 * 32 = 2048 per world)

__device__ __forceinline__ bool fits_registers(int C, int n) { return C <= 64 * kRegMaxSlots && n <= kRegBodies; }

// HBM-resident path (worlds too large for the register kernel): selection
// keys and body ids in LDS when the world has at most kKeyCap contacts.
template <typename R, int KIND>
__device__ __forceinline__ void resolve_hbm(const RS<R>& a, int w, int C, char* lds, int lane) {
  if (a.world_iters && lane == 0) {
    if (a.mode & RM_POS) a.world_iters[2 * w] = 0;
    if (a.mode & RM_VEL) a.world_iters[2 * w + 1] = 0;
  }
  // ResolveContacts' guard, contact.go:210-212
  if (a.overflow[0] != 0 || C <= 0 || a.dt <= R(0)) return;
  const int64_t base = a.world_coff[w];
  if (C <= kKeyCap) {
    R* s_key = (R*)lds;
    int32_t* s_b0 = (int32_t*)(lds + kKeyCap * sizeof(R));
    int32_t* s_b1 = s_b0 + kKeyCap;
    resolve_world<R, true, KIND>(a, w, C, base, s_key, s_b0, s_b1, lane);
  } else {
    resolve_world<R, false, KIND>(a, w, C, base, a.c.pen + base, a.c.body0 + base, a.c.body1 + base, lane);
  }
}

// ---------------------------------------------------------------------------
// Register-resident resolver: worlds with at most 64*SLOTS contacts and
// kRegBodies bodies.  Lane l holds contacts j = 64*s + l (s < SLOTS) in
// VGPRs: the selection key, the body ids (world-local, packed 16+16 bits)
// and, in the velocity phase, the contact velocity; the narrow classes
// (SLOTS <= 8) also hold the normal and the point, and keep the contact
// basis scale in LDS (contactToWorld is rebuilt from the normal and that
// scale bit for bit), while the wide classes read the normal,
// contactToWorld and relative positions as prepare stored them in HBM.
//
// LDS holds per body the position at prepare time (relativeContactPosition
// = point - that position, recomputed bit for bit where the reference
// stores it), lastFrameAccelleration * dt, the flag byte, and the state the
// phase changes: position and orientation in the position phase, velocity
// and rotation in the velocity phase, written back to HBM when the phase
// ends.  So every value an iteration writes and a later one reads stays in
// the wave's registers or LDS, whose operations a wave performs in program
// order, and the serial loop needs no barrier; the selected contact's
// register data reaches every lane by readlane.  HBM supplies the
// read-only data: inverse masses, world inverse inertia (in the position
// phase a sleeping body's new tensor goes to HBM, followed by a barrier),
// the velocity phase's impulse data (written when the phase starts).
// ---------------------------------------------------------------------------
constexpr uint32_t kNoBody = 0xFFFFu;

// Kinetic energy of a body, 1/2 (m |v|^2 + w . I w) with I the inverse of
// the world inverse inertia tensor (a zero matrix for an infinite tensor,
// math/matrix.go:133-137), in FP64; 0 for an infinite mass.  Also returns m.
template <typename R>
__device__ __forceinline__ double kinetic_energy(const V3<R>& v, const V3<R>& w, R inv_mass, const R* iit,
                                                 double& mass) {
  mass = 0.0;
  if (!(inv_mass > R(0))) return 0.0;
  mass = 1.0 / double(inv_mass);
  double m[9], I[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) m[k] = double(iit[k]);
  m3inv(m, I);
  const V3<double> wd{double(w.x), double(w.y), double(w.z)};
  const V3<double> Iw = m3v(I, wd);
  const double vv = double(v.x) * double(v.x) + double(v.y) * double(v.y) + double(v.z) * double(v.z);
  return 0.5 * (mass * vv + dot(wd, Iw));
}

template <typename R>
__device__ __forceinline__ V3<R> lds3(const R* a, int i) {
  return V3<R>{a[i], a[kRegBodies + i], a[2 * kRegBodies + i]};
}
template <typename R>
__device__ __forceinline__ void lds3_store(R* a, int i, const V3<R>& v) {
  a[i] = v.x;
  a[kRegBodies + i] = v.y;
  a[2 * kRegBodies + i] = v.z;
}

// calculateDesiredDeltaVelocity (contact.go:87-112) with the awake flags and
// lastFrameAccelleration*dt from LDS, for a contact of restitution e.
template <typename R>
__device__ __forceinline__ R desired_dv_lds(const uint8_t* fl, const R* lad, int l0, int l1, const V3<R>& n,
                                            R cvx, R e) {
  R vfa = R(0);
  if (fl[l0] & F_AWAKE) vfa += dot(lds3(lad, l0), n);
  if (l1 >= 0 && (fl[l1] & F_AWAKE)) vfa -= dot(lds3(lad, l1), n);
  R restitution = e;
  if (dabs(cvx) < R(0.25)) restitution = R(0);
  return -cvx - restitution * (cvx - vfa);
}

// calculateLocalVelocity (contact.go:159-182) with lastFrameAccelleration*dt
// from LDS.
template <typename R>
__device__ __forceinline__ V3<R> local_velocity_lds(const Bodies<R>& b, int body, const V3<R>& lad,
                                                    const V3<R>& rel, const R* c2w) {
  const int64_t S = b.stride;
  V3<R> velocity = add(cross(ld3(b.rot, S, body), rel), ld3(b.vel, S, body));
  V3<R> cv = m3tv(c2w, velocity);
  V3<R> accv = m3tv(c2w, lad);
  accv.x = R(0);
  return add(cv, accv);
}

__device__ __forceinline__ int id_lo(uint32_t id) { return int(id & 0xFFFFu); }
__device__ __forceinline__ int id_hi(uint32_t id) { return (id >> 16) == kNoBody ? -1 : int(id >> 16); }

// Broadcast of lane l's value (l uniform) by readlane.
__device__ __forceinline__ uint32_t lane_value(uint32_t v, int l) {
  return uint32_t(__builtin_amdgcn_readlane(int(v), l));
}
__device__ __forceinline__ float lane_value(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_value(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
template <typename R>
__device__ __forceinline__ V3<R> lane_value(const V3<R>& v, int l) {
  return V3<R>{lane_value(v.x, l), lane_value(v.y, l), lane_value(v.z, l)};
}
// Marks the slots' packed body ids as redefined, so that values derived from
// them (decoded ids, comparison masks: one 64-bit scalar mask per slot) are
// recomputed in each serial iteration instead of being hoisted out of the
// loop, which for 16-32 slots exhausts the register files.
template <int N>
__device__ __forceinline__ void opaque_ids(uint32_t (&ids)[N]) {
#pragma unroll
  for (int k = 0; k < N; ++k) asm volatile("" : "+v"(ids[k]));
}

// The lane index as a value redefined at the point of the call: the
// per-slot addresses computed from it in a serial iteration (hundreds of
// 64-bit values for the wide classes) are then recomputed there instead of
// being hoisted out of the loop.
__device__ __forceinline__ int opaque_lane(int lane) {
  asm volatile("" : "+v"(lane));
  return lane;
}

// Element s (uniform) of a register array, as a chain of selects with
// compile-time indexes (an indexed loop would keep the array in scratch).
template <typename T, int N, int... K>
__device__ __forceinline__ T pick_impl(const T (&a)[N], int s, std::integer_sequence<int, K...>) {
  T v = a[0];
  ((v = (K == s) ? a[K] : v), ...);
  return v;
}
template <typename R, int N, int... K>
__device__ __forceinline__ V3<R> pick_impl(const V3<R> (&a)[N], int s, std::integer_sequence<int, K...>) {
  // component by component: a select of the structs would go through memory
  R x = a[0].x, y = a[0].y, z = a[0].z;
  ((x = (K == s) ? a[K].x : x, y = (K == s) ? a[K].y : y, z = (K == s) ? a[K].z : z), ...);
  return V3<R>{x, y, z};
}
template <typename T, int N>
__device__ __forceinline__ T pick(const T (&a)[N], int s) {
  return pick_impl(a, s, std::make_integer_sequence<int, N>{});
}
// a if c else b, component by component
template <typename R>
__device__ __forceinline__ V3<R> sel3(bool c, const V3<R>& a, const V3<R>& b) {
  return V3<R>{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z};
}


template <typename R, int SLOTS, int KIND>
struct RegLds {
  // the wide classes keep only key, ids and contact velocity in VGPRs; the
  // rest of a contact comes from HBM where an iteration needs it
  static constexpr bool WIDE = SLOTS >= 12;
  static constexpr int CAP = 64 * SLOTS;
  static constexpr size_t pos0 = 0;                                   // R[3][kRegBodies]
  static constexpr size_t lad = pos0 + 3 * kRegBodies * sizeof(R);    // R[3][kRegBodies]
  static constexpr size_t st = lad + 3 * kRegBodies * sizeof(R);      // R[7][kRegBodies] pos+quat / vel+rot
  static constexpr size_t bs = st + 7 * kRegBodies * sizeof(R);       // double[CAP] basis scale (narrow)
  static constexpr size_t flags = bs + (WIDE ? 0 : CAP * sizeof(double));  // uint8[kRegBodies]
  static constexpr size_t bytes = flags + kRegBodies;
};

// Both bodies' world inverse inertia and inverse mass (zeros for an absent
// second body).
template <typename R>
__device__ __forceinline__ void body_inertia(const Bodies<R>& b, const int kb[2], R (&iw)[2][9], R im[2]) {
  const int64_t S = b.stride;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int k = 0; k < 9; ++k) iw[i][k] = R(0);
    im[i] = R(0);
  }
  ld9(b.iit_world, S, kb[0], iw[0]);
  im[0] = b.inv_mass[kb[0]];
  if (kb[1] >= 0) {
    ld9(b.iit_world, S, kb[1], iw[1]);
    im[1] = b.inv_mass[kb[1]];
  }
}

// applyPositionChange (contact.go:286-386) with the bodies' position and
// orientation in LDS (s_st: pos planes 0-2, quat planes 3-6, world-local
// index kl).  Returns whether a sleeping body's derived data were rewritten
// (in HBM).
// The bodies' world inverse inertia iw and inverse masses im are loaded by
// the caller (with the iteration's other loads, before any store).
template <typename R>
__device__ __forceinline__ bool apply_position_change_lds(const RS<R>& a, const int kb[2], const int kl[2],
                                                          const bool aw[2], const V3<R>& n, const V3<R> rel[2],
                                                          R penetration, const R (&iw)[2][9], const R im[2],
                                                          V3<R> lin[2], V3<R> ang[2], R* s_st, int lane) {
  const Bodies<R>& b = a.b;
  constexpr int KB = kRegBodies;
  const R angularLimit = R(0.2);
  R angI[2] = {R(0), R(0)}, linI[2] = {R(0), R(0)};
  R total = R(0);
  bool rederived = false;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    lin[i] = V3<R>{R(0), R(0), R(0)};
    ang[i] = V3<R>{R(0), R(0), R(0)};
    if (kb[i] < 0) continue;
    V3<R> aiw = cross(rel[i], n);
    aiw = m3v(iw[i], aiw);
    aiw = cross(aiw, rel[i]);
    angI[i] = dot(aiw, n);
    linI[i] = im[i];
    total += linI[i] + angI[i];
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int bi = kb[i];
    if (bi < 0) continue;
    R sign = R(1);
    if (i != 0) sign = R(-1);
    R angM = sign * penetration * (angI[i] / total);
    R linM = sign * penetration * (linI[i] / total);
    V3<R> projection = add_scaled(rel[i], n, -dot(rel[i], n));
    R maxMagnitude = angularLimit * mag(projection);
    if (angM < -maxMagnitude) {
      R totalMove = angM + linM;
      angM = -maxMagnitude;
      linM = totalMove - angM;
    } else if (angM > maxMagnitude) {
      R totalMove = angM + linM;
      angM = maxMagnitude;
      linM = totalMove - angM;
    }
    if (angM == R(0)) {
      ang[i] = V3<R>{R(0), R(0), R(0)};
    } else {
      V3<R> tad = cross(rel[i], n);
      ang[i] = mul(m3v(iw[i], tad), angM / angI[i]);
    }
    lin[i] = mul(n, linM);
    const int l = kl[i];
    V3<R> pos = add_scaled(lds3(s_st, l), n, linM);
    Q4<R> q{s_st[3 * KB + l], s_st[4 * KB + l], s_st[5 * KB + l], s_st[6 * KB + l]};
    q = q_add_scaled_vector(q, ang[i], R(1));
    q = q_normalize(q);
    if (!aw[i]) {
      rederive(b, bi, pos, q, lane);
      rederived = true;
    }
    if (lane == 0) {
      lds3_store(s_st, l, pos);
      s_st[3 * KB + l] = q.w;
      s_st[4 * KB + l] = q.x;
      s_st[5 * KB + l] = q.y;
      s_st[6 * KB + l] = q.z;
    }
  }
  return rederived;
}

// applyVelocityChange (contact.go:448-494) from the contact's impulse data,
// with the bodies' velocity and rotation in LDS (s_st: vel planes 0-2, rot
// planes 3-5, world-local index kl).
template <typename R, int KIND>
__device__ __forceinline__ void apply_velocity_change_lds(const int kb[2], const int kl[2], const V3<R> rel[2],
                                                          const R* c2w, R ddv, const V3<R>& cvel, const R* imp,
                                                          const R (&iw)[2][9], const R im[2], V3<R> vc[2],
                                                          V3<R> rc[2], R* s_st, int lane, R friction) {
  R* s_rot = s_st + 3 * kRegBodies;
  V3<R> ic;
  if (KIND == KIND_MAT && friction == R(0))
    ic = V3<R>{ddv / imp[0], R(0), R(0)};
  else
    ic = friction_impulse_from(imp, imp + 9, ddv, cvel, friction);
  const V3<R> impulse = m3v(c2w, ic);
  const V3<R> zero{R(0), R(0), R(0)};
  V3<R> torque = cross(rel[0], impulse);
  rc[0] = m3v(iw[0], torque);
  vc[0] = add_scaled(zero, impulse, im[0]);
  rc[1] = zero;
  vc[1] = zero;
  const V3<R> v0 = add(lds3(s_st, kl[0]), vc[0]);
  const V3<R> w0 = add(lds3(s_rot, kl[0]), rc[0]);
  V3<R> v1 = zero, w1 = zero;
  if (kb[1] >= 0) {
    torque = cross(impulse, rel[1]);
    rc[1] = m3v(iw[1], torque);
    vc[1] = add_scaled(zero, impulse, -im[1]);
    v1 = add(lds3(s_st, kl[1]), vc[1]);
    w1 = add(lds3(s_rot, kl[1]), rc[1]);
  }
  if (lane == 0) {
    lds3_store(s_st, kl[0], v0);
    lds3_store(s_rot, kl[0], w0);
    if (kb[1] >= 0) {
      lds3_store(s_st, kl[1], v1);
      lds3_store(s_rot, kl[1], w1);
    }
  }
}

template <typename R, int SLOTS, int KIND>
__device__ __forceinline__ void resolve_reg(const RS<R>& a, int w, int C, int b0w, int n, char* lds, int lane) {
  using L = RegLds<R, SLOTS, KIND>;
  constexpr bool WIDE = L::WIDE;
  constexpr int KB = kRegBodies;
  R* s_pos0 = (R*)(lds + L::pos0);
  R* s_lad = (R*)(lds + L::lad);
  R* s_st = (R*)(lds + L::st);
  double* s_bs = (double*)(lds + L::bs);  // narrow classes only
  uint8_t* s_flags = (uint8_t*)(lds + L::flags);
  const bool do_pos = (a.mode & RM_POS) != 0;
  const bool do_prep = do_pos && !(a.mode & RM_PREPARED);
  const bool do_vel = (a.mode & RM_VEL) != 0 && (KIND == KIND_POLISH || !a.world_mask || a.world_mask[w]);
  if (a.world_iters && lane == 0) {
    if (do_pos) a.world_iters[2 * w] = 0;
    if (a.mode & RM_VEL) a.world_iters[2 * w + 1] = 0;
  }
  // ResolveContacts' guard, contact.go:210-212
  if (a.overflow[0] != 0 || C <= 0 || a.dt <= R(0)) return;
  if (!do_pos && !do_vel) return;

  const Bodies<R>& b = a.b;
  const Contacts<R>& c = a.c;
  const int64_t S = b.stride, CS = c.cap;
  const int64_t base = a.world_coff[w];
  const int64_t maxIter = a.iter_fixed > 0 ? a.iter_fixed : a.iter_mult * int64_t(C);
  const V3<R> zero{R(0), R(0), R(0)};

  for (int i = lane; i < n; i += 64) {
    const int g = b0w + i;
    // body positions at prepare time: the current ones when this launch
    // prepares, else the snapshot the preparing launch stored
    const V3<R> p = ld3(do_prep ? b.pos : a.pos0, S, g);
    if (do_prep) st3(a.pos0, S, g, p);
    lds3_store(s_pos0, i, p);
    lds3_store(s_lad, i, mul(ld3(b.last_acc, S, g), a.dt));
    s_flags[i] = b.flags[g];
  }
  __syncthreads();

  R key[SLOTS];
  uint32_t ids[SLOTS];
  V3<R> nrm[WIDE ? 1 : SLOTS];
  V3<R> ptr[WIDE ? 1 : SLOTS];

  // prepareContacts / calculateInternals, contact.go:225-229, :59-85, over
  // this lane's contacts (a rolled loop: the body is large), into the
  // contact arrays
  if (do_prep) {
    for (int j = lane; j < C; j += 64) {
      const int64_t g = base + j;
      int c0 = c.body0[g], c1 = c.body1[g];
      V3<R> nn = ld3(c.normal, CS, g);
      if (c0 < 0) {  // contact.go:61-65
        nn = mul(nn, R(-1));
        c0 = c1;
        c1 = -1;
        c.body0[g] = c0;
        c.body1[g] = c1;
        st3(c.normal, CS, g, nn);
      }
      const int l0 = c0 - b0w, l1 = c1 >= 0 ? c1 - b0w : -1;
      R m[9];
      const double sc = basis_scale(nn);
      basis_from_scale(nn, sc, m);
      const V3<R> p = ld3(c.point, CS, g);
      const V3<R> r0 = sub(p, lds3(s_pos0, l0));
      V3<R> cv = local_velocity_lds(b, c0, lds3(s_lad, l0), r0, m);
      V3<R> r1 = zero;
      if (c1 >= 0) {
        r1 = sub(p, lds3(s_pos0, l1));
        cv = sub(cv, local_velocity_lds(b, c1, lds3(s_lad, l1), r1, m));
      }
      const R ddv = desired_dv_lds(s_flags, s_lad, l0, l1, nn, cv.x, restitution_of<R, KIND>(c, g));
#pragma unroll
      for (int k = 0; k < 9; ++k) c.c2w[k * CS + g] = m[k];
      st3(c.rel, CS, g, r0);
      st3(c.rel + 3 * CS, CS, g, r1);
      st3(c.cvel, CS, g, cv);
      c.ddv[g] = ddv;
      if constexpr (WIDE) store_rec(c.rec + 16 * g, m, r0, r1);
      if constexpr (!WIDE) s_bs[j] = sc;
    }
  } else if constexpr (!WIDE) {  // prepared by an earlier launch: the basis scales of prepare's normals
    for (int j = lane; j < C; j += 64) s_bs[j] = basis_scale(ld3(c.normal, CS, base + j));
  }
  // this lane's contacts into registers (the lane wrote what it reads here)
#pragma unroll
  for (int s = 0; s < SLOTS; ++s) {
    const int j = 64 * s + lane;
    key[s] = R(0);
    ids[s] = 0xFFFFFFFFu;
    if constexpr (!WIDE) nrm[s] = ptr[s] = zero;
    if (j < C) {
      const int64_t g = base + j;
      const int c0 = c.body0[g], c1 = c.body1[g];
      ids[s] = uint32_t(c0 - b0w) | (uint32_t(c1 >= 0 ? c1 - b0w : int(kNoBody)) << 16);
      if (do_pos) key[s] = c.pen[g];
      if constexpr (!WIDE) {
        nrm[s] = ld3(c.normal, CS, g);
        ptr[s] = ld3(c.point, CS, g);
      }
    }
  }
  // the position phase's body state into LDS; the wide classes read what
  // prepare stored in HBM below
  if (do_pos) {
    for (int i = lane; i < n; i += 64) {
      const int g = b0w + i;
      lds3_store(s_st, i, ld3(b.pos, S, g));
#pragma unroll
      for (int k = 0; k < 4; ++k) s_st[(3 + k) * KB + i] = b.quat[k * S + g];
    }
  }
  __syncthreads();

  // The surrogate's position polish (KIND_PPOLISH, worlds whose mask byte
  // is 0): snapshot what the loop may write to HBM for the bodies
  // asleep now (the motion matchAwakeState sets, the derived data a
  // sleeping body's move recomputes), add the predicted changes to the
  // awake bodies' positions and orientations, update every contact's
  // penetration for them by the loop's own formula, then run the loop to
  // positionEpsilon + polish_tol within polish_mult * C iterations.
  bool ppolish = false;
  R sep_floor = R(0);  // the polished world's least penetration the acceptance test allows
  // the world's least penetration (every lane gets it)
  auto key_min = [&]() {
    R kmin = R(0);
#pragma unroll
    for (int s = 0; s < SLOTS; ++s)
      if (64 * s + lane < C) kmin = key[s] < kmin ? key[s] : kmin;
    return dpp::allreduce(kmin, R(0), [](R x, R y) { return y < x ? y : x; });
  };
  // Undo a rejected position polish: positions, orientations and awake
  // flags from HBM (untouched until the phase ends), the sleeping bodies'
  // motion and derived data from the snapshot, the penetrations from the
  // contact array; then the plain phase runs.
  auto restore_positions = [&]() {
    // lane 0's HBM stores of the loop (motion, a sleeping body's derived
    // data) complete before the other lanes restore over them
    __syncthreads();
    for (int i = lane; i < n; i += 64) {
      const int g = b0w + i;
      lds3_store(s_st, i, ld3(b.pos, S, g));
#pragma unroll
      for (int k = 0; k < 4; ++k) s_st[(3 + k) * KB + i] = b.quat[k * S + g];
      const uint8_t f = b.flags[g];
      s_flags[i] = f;
      if (!(f & F_AWAKE)) {
        b.motion[g] = a.vel0[g];
#pragma unroll
        for (int k = 0; k < 12; ++k) b.transform[k * S + g] = a.vel0[(1 + k) * S + g];
#pragma unroll
        for (int k = 0; k < 9; ++k) b.iit_world[k * S + g] = a.vel0[(13 + k) * S + g];
      }
    }
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
      const int j = 64 * s + lane;
      key[s] = j < C ? c.pen[base + j] : R(0);
    }
    __syncthreads();
    if (lane == 0) a.world_mask[w] = 1;
  };
  if constexpr (KIND == KIND_PPOLISH) {
    ppolish = do_pos && a.world_mask[w] == 0;
    if (ppolish) {
      R pmax = R(0);
#pragma unroll
      for (int s = 0; s < SLOTS; ++s) pmax = key[s] > pmax ? key[s] : pmax;
      pmax = dpp::allreduce(pmax, R(0), [](R x, R y) { return y > x ? y : x; });
      sep_floor = -(R(kSepFactor) * pmax + a.polish_tol);
      const float* dl = a.delta;
      for (int i = lane; i < n; i += 64) {
        const int g = b0w + i;
        if (!(s_flags[i] & F_AWAKE)) {
          a.vel0[g] = b.motion[g];
#pragma unroll
          for (int k = 0; k < 12; ++k) a.vel0[(1 + k) * S + g] = b.transform[k * S + g];
#pragma unroll
          for (int k = 0; k < 9; ++k) a.vel0[(13 + k) * S + g] = b.iit_world[k * S + g];
        }
      }
      // the predicted moves times `scale`, from the phase's starting state
      // (positions and orientations in HBM, penetrations in the contact
      // array): the awake bodies' state, then every contact's penetration
      // by the loop's own formula
      auto take_moves = [&](R scale) {
        for (int i = lane; i < n; i += 64) {
          const int g = b0w + i;
          if (!(s_flags[i] & F_AWAKE)) continue;
          const V3<R> dx = mul(V3<R>{R(dl[g]), R(dl[S + g]), R(dl[2 * S + g])}, scale);
          const V3<R> dq = mul(V3<R>{R(dl[3 * S + g]), R(dl[4 * S + g]), R(dl[5 * S + g])}, scale);
          if (dx.x == R(0) && dx.y == R(0) && dx.z == R(0) && dq.x == R(0) && dq.y == R(0) && dq.z == R(0)) continue;
          lds3_store(s_st, i, add(ld3(b.pos, S, g), dx));
          Q4<R> q{b.quat[g], b.quat[S + g], b.quat[2 * S + g], b.quat[3 * S + g]};
          q = q_normalize(q_add_scaled_vector(q, dq, R(1)));
          s_st[3 * KB + i] = q.w;
          s_st[4 * KB + i] = q.x;
          s_st[5 * KB + i] = q.y;
          s_st[6 * KB + i] = q.z;
        }
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) {
          asm volatile("" ::: "memory");
          if (64 * s + lane >= C) continue;
          const int64_t gj = base + 64 * s + lane;
          const int jb[2] = {id_lo(ids[s]), id_hi(ids[s])};
          V3<R> nj;
          if constexpr (WIDE)
            nj = ld3(c.normal, CS, gj);
          else
            nj = nrm[s];
          R pen = c.pen[gj];
#pragma unroll
          for (int bb = 0; bb < 2; ++bb) {
            if (jb[bb] < 0) continue;
            const int g = b0w + jb[bb];
            const V3<R> dx = mul(V3<R>{R(dl[g]), R(dl[S + g]), R(dl[2 * S + g])}, scale);
            const V3<R> dq = mul(V3<R>{R(dl[3 * S + g]), R(dl[4 * S + g]), R(dl[5 * S + g])}, scale);
            if (dx.x == R(0) && dx.y == R(0) && dx.z == R(0) && dq.x == R(0) && dq.y == R(0) && dq.z == R(0)) continue;
            V3<R> rj;
            if constexpr (WIDE)
              rj = ld3(c.rel + 3 * bb * CS, CS, gj);
            else
              rj = sub(ptr[s], lds3(s_pos0, jb[bb]));
            const V3<R> dp = add(cross(dq, rj), dx);
            R sign = R(1);
            if (bb == 0) sign = R(-1);
            pen += dot(dp, nj) * sign;
          }
          key[s] = pen;
        }
        __syncthreads();
      };
      // a prediction that separates a contact beyond the bound is retried
      // at a fraction of its moves, then rejected, before the loop spends
      // iterations on it
      take_moves(R(1));
      if (!(key_min() >= sep_floor)) {
        take_moves(R(kPosRetryScale));
        if (!(key_min() >= sep_floor)) {
          restore_positions();
          ppolish = false;
        }
      }
    }
  }

  // adjustPositions, contact.go:233-283
  R peps = ppolish ? R(kPositionEpsilon) + a.polish_tol : R(kPositionEpsilon);
  int64_t pbudget = ppolish ? a.polish_mult * int64_t(C) : maxIter;
  int64_t used = 0;
  // A slot past the world's contacts keeps key 0 (never above the epsilon)
  // and ids 0xFFFFFFFF (never a body of the world), so the loops below need
  // no bounds test.
  while (do_pos) {
    opaque_ids(ids);
    R best = peps;
    int bi = INT_MAX;
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
      if (key[s] > best) {
        best = key[s];
        bi = 64 * s + lane;
      }
    }
    wave_argmax_dpp(best, bi);
    bool reject = false;  // the position polish did not give an acceptable result
    if (bi == INT_MAX) {
      if constexpr (KIND == KIND_PPOLISH) {
        // Every contact is within positionEpsilon + tol.  The plain phase
        // moves bodies only to resolve penetration; a polished world in
        // which a contact ends separated by more than kSepFactor times the
        // world's deepest penetration at the start (+ tol) is rejected: that
        // is the mark of a prediction that pushed bodies apart.
        if (ppolish) reject = !(key_min() >= sep_floor);
      }
      if (!reject) break;
    } else if (used >= pbudget) {
      if (KIND != KIND_PPOLISH || !ppolish) break;
      reject = true;
    }
    if constexpr (KIND == KIND_PPOLISH) {
      if (reject) {
        restore_positions();
        ppolish = false;
        peps = R(kPositionEpsilon);
        pbudget = maxIter;
        used = 0;
        continue;
      }
    }
    if (bi == INT_MAX || used >= pbudget) break;
    const int ss = bi >> 6, sl = bi & 63;
    const uint32_t kid = lane_value(pick(ids, ss), sl);
    const int kl[2] = {id_lo(kid), id_hi(kid)};
    const int kb[2] = {b0w + kl[0], kl[1] >= 0 ? b0w + kl[1] : -1};
    V3<R> n, rel[2];
    if constexpr (WIDE) {
      rec_normal(c.rec + 16 * (base + bi), n, rel);
    } else {
      n = lane_value(pick(nrm, ss), sl);
      const V3<R> p = lane_value(pick(ptr, ss), sl);
      rel[0] = sub(p, lds3(s_pos0, kl[0]));
      rel[1] = kl[1] >= 0 ? sub(p, lds3(s_pos0, kl[1])) : zero;
    }
    // both bodies' inertia, loaded before this iteration's first store
    R iw[2][9], im[2];
    body_inertia(b, kb, iw, im);
    // matchAwakeState, contact.go:185-202
    bool aw[2] = {(s_flags[kl[0]] & F_AWAKE) != 0, kl[1] >= 0 && (s_flags[kl[1]] & F_AWAKE) != 0};
    if (kl[1] >= 0 && (aw[0] || aw[1]) && !(aw[0] && aw[1])) {
      const int t = aw[0] ? kl[1] : kl[0];
      if (lane == 0) {
        s_flags[t] = uint8_t(s_flags[t] | F_AWAKE);
        b.motion[b0w + t] = R(kSleepEpsilon * 2.0);
      }
      aw[0] = aw[1] = true;
    }
    V3<R> lin[2], ang[2];
    const bool rederived = apply_position_change_lds(a, kb, kl, aw, n, rel, best, iw, im, lin, ang, s_st, lane);
    if constexpr (KIND == KIND_LABEL)
      if (lane == 0) add_pos_label(a, base + bi, best);
    const int ln = opaque_lane(lane);
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
      const int jb[2] = {id_lo(ids[s]), id_hi(ids[s])};
      bool any = false;
#pragma unroll
      for (int bb = 0; bb < 2; ++bb)
#pragma unroll
        for (int d = 0; d < 2; ++d) any |= (jb[bb] >= 0 && jb[bb] == kl[d]);
      if (!any) continue;
      const int64_t gj = base + 64 * s + ln;
      R pen = key[s];
      V3<R> nj, rw[2];
      if constexpr (WIDE)
        rec_normal(c.rec + 16 * gj, nj, rw);
      else
        nj = nrm[s];
#pragma unroll
      for (int bb = 0; bb < 2; ++bb) {
        if (jb[bb] < 0) continue;
        V3<R> rj;
        if constexpr (WIDE)
          rj = rw[bb];
        else
          rj = sub(ptr[s], lds3(s_pos0, jb[bb]));
        // a side is at most one of the two moved bodies: its change is
        // selected and applied once
        if (jb[bb] == kl[0] || jb[bb] == kl[1]) {
          const bool d0 = jb[bb] == kl[0];
          V3<R> dp = add(cross(sel3(d0, ang[0], ang[1]), rj), sel3(d0, lin[0], lin[1]));
          R sign = R(1);
          if (bb == 0) sign = R(-1);
          pen += dot(dp, nj) * sign;
        }
      }
      key[s] = pen;
    }
    // a sleeping body's new world inverse inertia went to HBM; the next
    // iteration may read it
    if (rederived) __syncthreads();
    ++used;
  }
  const int64_t pos_iters = used;
  if (do_pos) {
    for (int i = lane; i < n; i += 64) {
      const int g = b0w + i;
      st3(b.pos, S, g, lds3(s_st, i));
#pragma unroll
      for (int k = 0; k < 4; ++k) b.quat[k * S + g] = s_st[(3 + k) * KB + i];
    }
  }

  // adjustVelocities, contact.go:390-445.  The bodies' velocity and
  // rotation live in LDS for the whole phase; the contact velocities in
  // VGPRs where they fit, else in the contact array, where each slot is
  // read and written only by the lane that owns it (the selected one reaches
  // the other lanes by readlane).
  constexpr bool CV_REG = SLOTS * sizeof(R) <= 64;
  R* s_rot = s_st + 3 * KB;
  V3<R> cvr[CV_REG ? SLOTS : 1];
  // slot s of this lane, contact g = base + 64 s + lane
  auto cv_get = [&](int s, int64_t g) -> V3<R> {
    if constexpr (CV_REG)
      return cvr[s];
    else
      return ld3(c.cvel, CS, g);
  };
  auto cv_set = [&](int s, int64_t g, const V3<R>& v) {
    if constexpr (CV_REG)
      cvr[s] = v;
    else
      st3(c.cvel, CS, g, v);
  };
#pragma unroll
  for (int s = 0; s < SLOTS; ++s) {
    const int j = 64 * s + lane;
    if constexpr (CV_REG) cvr[s] = zero;
    if (do_pos && j < C) c.pen[base + j] = key[s];
    if (do_vel && j < C) {
      key[s] = c.ddv[base + j];  // selection keys as prepare stored them
      if constexpr (CV_REG) cvr[s] = ld3(c.cvel, CS, base + j);
    }
  }
  // every contact's impulse data, fixed for the whole phase, from the
  // contactToWorld and relative positions prepare stored (this lane's own
  // contacts)
  if (do_vel) {
    for (int j = lane; j < C; j += 64) {
      const int64_t g = base + j;
      const int jb[2] = {c.body0[g], c.body1[g]};
      R mj[9], imp[kImpPlanes];
      V3<R> rj[2];
      if constexpr (WIDE) {
        rec_frame(c.rec + 16 * g, mj, rj);
      } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) mj[k] = c.c2w[k * CS + g];
        rj[0] = ld3(c.rel, CS, g);
        rj[1] = ld3(c.rel + 3 * CS, CS, g);
      }
      impulse_data<R, KIND>(a, jb, rj, mj, friction_of<R, KIND>(c, g), imp);
#pragma unroll
      for (int k = 0; k < kImpPlanes; ++k) c.imp[k * CS + g] = imp[k];
    }
  }
  // the impulse data every lane reads below, and LDS reuse
  __syncthreads();
  used = 0;
  if (do_vel) {
    for (int i = lane; i < n; i += 64) {
      const int g = b0w + i;
      lds3_store(s_st, i, ld3(b.vel, S, g));
      lds3_store(s_rot, i, ld3(b.rot, S, g));
    }
    bool polishing = KIND == KIND_POLISH && a.world_mask[w] == 0;
    double ke_before = 0.0, world_mass = 0.0;
    if constexpr (KIND == KIND_POLISH) {
      if (polishing) {
        // take the predicted changes: snapshot the motion (matchAwakeState
        // writes it in HBM; velocity and rotation stay in HBM untouched
        // until the phase ends), then add the changes to the bodies
        const float* dl = a.delta;
        if constexpr (!CV_REG)
          for (int j = lane; j < C; j += 64) st3(a.cvel0, CS, base + j, ld3(c.cvel, CS, base + j));
        double ke = 0.0, mass = 0.0;
        for (int i = lane; i < n; i += 64) {
          const int g = b0w + i;
          const V3<R> v = lds3(s_st, i), r = lds3(s_rot, i);
          R iw[9];
          ld9(b.iit_world, S, g, iw);
          double mi;
          ke += kinetic_energy(v, r, b.inv_mass[g], iw, mi);
          mass += mi;
          a.vel0[6 * S + g] = b.motion[g];
          const V3<R> dv{R(dl[g]), R(dl[S + g]), R(dl[2 * S + g])};
          const V3<R> dw{R(dl[3 * S + g]), R(dl[4 * S + g]), R(dl[5 * S + g])};
          if (dv.x != R(0) || dv.y != R(0) || dv.z != R(0) || dw.x != R(0) || dw.y != R(0) || dw.z != R(0)) {
            lds3_store(s_st, i, add(v, dv));
            lds3_store(s_rot, i, add(r, dw));
          }
        }
        // contact velocities after those changes: the update of
        // contact.go:424-441 applied to each body's summed change (a
        // compiler fence per slot keeps the slots' loads from being
        // interleaved, which would need registers for all of them)
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) {
          asm volatile("" ::: "memory");
          if (64 * s + lane >= C) continue;
          const int64_t gj = base + 64 * s + lane;
          const int jb[2] = {id_lo(ids[s]), id_hi(ids[s])};
          R mj[9];
          V3<R> rw[2];
          if constexpr (WIDE) {
            rec_frame(c.rec + 16 * gj, mj, rw);
          } else {
            basis_from_scale(nrm[s], s_bs[64 * s + lane], mj);
          }
          V3<R> cvj = cv_get(s, gj);
          bool moved = false;
#pragma unroll
          for (int bb = 0; bb < 2; ++bb) {
            if (jb[bb] < 0) continue;
            const int g = b0w + jb[bb];
            const V3<R> dv{R(dl[g]), R(dl[S + g]), R(dl[2 * S + g])};
            const V3<R> dw{R(dl[3 * S + g]), R(dl[4 * S + g]), R(dl[5 * S + g])};
            if (dv.x == R(0) && dv.y == R(0) && dv.z == R(0) && dw.x == R(0) && dw.y == R(0) && dw.z == R(0)) continue;
            moved = true;
            V3<R> rj;
            if constexpr (WIDE)
              rj = rw[bb];
            else
              rj = sub(ptr[s], lds3(s_pos0, jb[bb]));
            R sign = R(1);
            if (bb == 1) sign = R(-1);
            cvj = add(cvj, mul(m3tv(mj, add(cross(dw, rj), dv)), sign));
          }
          // a contact neither of whose bodies changed keeps the key prepare
          // gave it, as in the plain loop
          if (moved) {
            cv_set(s, gj, cvj);
            key[s] = desired_dv_lds(s_flags, s_lad, jb[0], jb[1], V3<R>{mj[0], mj[1], mj[2]}, cvj.x,
                                    restitution_of<R, KIND>(c, gj));
          }
        }
        ke_before = wave_sum_f64(ke);
        world_mass = wave_sum_f64(mass);
      }
    }
    // The reference's loop.  A polished world first runs it to
    // velocityEpsilon + polish_tol within polish_mult * C iterations; if that
    // does not get there, everything the prediction and the loop changed is
    // undone and the same loop runs as the plain phase.
    R eps = polishing ? R(kVelocityEpsilon) + a.polish_tol : R(kVelocityEpsilon);
    int64_t budget = polishing ? a.polish_mult * int64_t(C) : maxIter;
    int64_t it = 0;
    while (true) {
      opaque_ids(ids);
      R best = eps;
      int bi = INT_MAX;
#pragma unroll
      for (int s = 0; s < SLOTS; ++s) {
        if (key[s] > best) {
          best = key[s];
          bi = 64 * s + lane;
        }
      }
      wave_argmax_dpp(best, bi);
      bool reject = false;  // the polish did not give an acceptable result
      if (bi == INT_MAX) {
        if constexpr (KIND == KIND_POLISH) {
          // Every contact is within velocityEpsilon + tol of closing.  The
          // plain phase applies only compressive impulses inside the
          // friction cone with restitution below 1, so it never adds kinetic
          // energy; a polished world that ends with more than it started
          // with, beyond a slack of 1/2 tol^2 per kilogram (which also covers
          // the acceleration term of the restitution target,
          // contact.go:87-112), is rejected.  This is what stops a wrong
          // prediction from pushing bodies apart.
          if (polishing) {
            double ke = 0.0, mass = 0.0;
            for (int i = lane; i < n; i += 64) {
              const int g = b0w + i;
              R iw[9];
              ld9(b.iit_world, S, g, iw);
              ke += kinetic_energy(lds3(s_st, i), lds3(s_rot, i), b.inv_mass[g], iw, mass);
            }
            const double ke_after = wave_sum_f64(ke);
            const double tol = double(a.polish_tol);
            reject = !(ke_after <= ke_before * (1.0 + kEnergyRel) + 0.5 * tol * tol * world_mass);
          }
        }
        if (!reject) break;
      } else if (it >= budget) {
        if (KIND != KIND_POLISH || !polishing) break;
        reject = true;
      }
      if (reject) {
        // lane 0's motion stores of the loop complete before other lanes
        // restore the snapshot over them
        __syncthreads();
        for (int i = lane; i < n; i += 64) {
          const int g = b0w + i;
          lds3_store(s_st, i, ld3(b.vel, S, g));
          lds3_store(s_rot, i, ld3(b.rot, S, g));
          b.motion[g] = a.vel0[6 * S + g];
          s_flags[i] = b.flags[g];
        }
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) {
          const int j = 64 * s + lane;
          if (j < C) {
            key[s] = c.ddv[base + j];
            cv_set(s, base + j, ld3(CV_REG ? c.cvel : a.cvel0, CS, base + j));
          }
        }
        __syncthreads();
        if (lane == 0) a.world_mask[w] = 1;
        polishing = false;
        eps = R(kVelocityEpsilon);
        budget = maxIter;
        it = 0;
        continue;
      }
      const int ss = bi >> 6, sl = bi & 63;
      const uint32_t kid = lane_value(pick(ids, ss), sl);
      V3<R> cvk;
      if constexpr (CV_REG) {
        cvk = lane_value(pick(cvr, ss), sl);
      } else {  // the owner lane's own slot
        const int jo = 64 * ss + lane;
        cvk = lane_value(jo < C ? ld3(c.cvel, CS, base + jo) : zero, sl);
      }
      const int kl[2] = {id_lo(kid), id_hi(kid)};
      const int kb[2] = {b0w + kl[0], kl[1] >= 0 ? b0w + kl[1] : -1};
      R m[9], imp[kImpPlanes];
      V3<R> rel[2];
      if constexpr (WIDE) {
        rec_frame(c.rec + 16 * (base + bi), m, rel);
      } else {
        const V3<R> n = lane_value(pick(nrm, ss), sl);
        const V3<R> p = lane_value(pick(ptr, ss), sl);
        basis_from_scale(n, s_bs[bi], m);
        rel[0] = sub(p, lds3(s_pos0, kl[0]));
        rel[1] = kl[1] >= 0 ? sub(p, lds3(s_pos0, kl[1])) : zero;
      }
#pragma unroll
      for (int k = 0; k < kImpPlanes; ++k) imp[k] = c.imp[k * CS + base + bi];
      // both bodies' inertia, loaded with the impulse data before this
      // iteration's first store
      R iw[2][9], im[2];
      body_inertia(b, kb, iw, im);
      // matchAwakeState, contact.go:185-202
      const bool aw0 = (s_flags[kl[0]] & F_AWAKE) != 0;
      const bool aw1 = kl[1] >= 0 && (s_flags[kl[1]] & F_AWAKE) != 0;
      if (kl[1] >= 0 && (aw0 || aw1) && !(aw0 && aw1)) {
        const int t = aw0 ? kl[1] : kl[0];
        if (lane == 0) {
          s_flags[t] = uint8_t(s_flags[t] | F_AWAKE);
          b.motion[b0w + t] = R(kSleepEpsilon * 2.0);
        }
      }
      V3<R> vc[2], rc[2];
      apply_velocity_change_lds<R, KIND>(kb, kl, rel, m, best, cvk, imp, iw, im, vc, rc, s_st, lane,
                                         friction_of<R, KIND>(c, base + bi));
      if constexpr (KIND == KIND_LABEL)
        if (lane == 0) add_label(a, base + bi, vc, rc);
      const int ln = opaque_lane(lane);
#pragma unroll
      for (int s = 0; s < SLOTS; ++s) {
        const int jb[2] = {id_lo(ids[s]), id_hi(ids[s])};
        bool any = false;
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
          for (int d = 0; d < 2; ++d) any |= (jb[bb] >= 0 && jb[bb] == kl[d]);
        if (!any) continue;
        const int64_t gj = base + 64 * s + ln;
        R mj[9];
        V3<R> rw[2];
        if constexpr (WIDE) {
          rec_frame(c.rec + 16 * gj, mj, rw);
        } else {
          basis_from_scale(nrm[s], s_bs[64 * s + lane], mj);
        }
        V3<R> cvj = cv_get(s, gj);
#pragma unroll
        for (int bb = 0; bb < 2; ++bb) {
          if (jb[bb] < 0) continue;
          V3<R> rj;
          if constexpr (WIDE)
            rj = rw[bb];
          else
            rj = sub(ptr[s], lds3(s_pos0, jb[bb]));
          // a side is at most one of the two moved bodies: its change is
          // selected and applied once
          if (jb[bb] == kl[0] || jb[bb] == kl[1]) {
            const bool d0 = jb[bb] == kl[0];
            V3<R> dv = add(cross(sel3(d0, rc[0], rc[1]), rj), sel3(d0, vc[0], vc[1]));
            R sign = R(1);
            if (bb == 1) sign = R(-1);
            cvj = add(cvj, mul(m3tv(mj, dv), sign));
          }
        }
        cv_set(s, gj, cvj);
        key[s] = desired_dv_lds(s_flags, s_lad, jb[0], jb[1], V3<R>{mj[0], mj[1], mj[2]}, cvj.x,
                                restitution_of<R, KIND>(c, gj));
      }
      ++it;
      ++used;
    }
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
      const int j = 64 * s + lane;
      if (j < C) {
        c.ddv[base + j] = key[s];
        if constexpr (CV_REG) st3(c.cvel, CS, base + j, cvr[s]);
      }
    }
    for (int i = lane; i < n; i += 64) {
      const int g = b0w + i;
      st3(b.vel, S, g, lds3(s_st, i));
      st3(b.rot, S, g, lds3(s_rot, i));
    }
  }
  for (int i = lane; i < n; i += 64) b.flags[b0w + i] = s_flags[i];
  if (lane == 0 && a.world_iters) {
    if (do_pos) a.world_iters[2 * w] = pos_iters;
    if (do_vel) a.world_iters[2 * w + 1] = used;
  }
}

// ----------------------------------------------------------------------
// The largest worlds: one workgroup of kMwWaves wavefronts per world.
//
// In a world of many hundred contacts, an iteration of the one-wave kernel
// scans up to 32 contacts per lane and reads the selected contact's impulse
// data, the bodies' inertia and the contact velocities from HBM: a chain
// of dependent memory round trips.  Here thread t (of kMwThreads, one wave
// per SIMD) holds contacts j = t + kMwThreads * s (s < SL, SL = 4 or 8):
// their selection keys, body ids, basis scales, contact velocities and
// impulse data in registers; their normals and points are in LDS by
// contact, beside the bodies' state, world inverse inertia, inverse mass
// and flags.  An iteration:
//   1. each wave takes its best contact (larger key; equal keys -> lower
//      index) by DPP, and the lane that holds it offers its register data
//      in LDS;                                              barrier A
//   2. wave 0 alone takes the best of the offers, computes the correction
//      and stores the bodies' new state (lane 0; wave 0 is the only reader
//      of that state), then publishes the two bodies' changes; barrier C
//   3. every thread updates its own contacts that share one of the bodies.
// The arithmetic, its order and the selection are those of resolve_reg, so
// both give the same bits.
constexpr int kMwWaves = 4;
constexpr int kMwThreads = 64 * kMwWaves;
constexpr int kMwMaxSlots = 8;
constexpr int kMwMaxContacts = kMwThreads * kMwMaxSlots;  // 2048
// per-wave offer: contact velocity 3, impulse data 12, friction 1
constexpr int kMwPay = 16;
// wave 0's result: state, packed body ids, woken body
enum : int { MW_APPLY = 0, MW_DONE = 1, MW_BUDGET = 2 };

// A world the multi-wave kernel can take (the selection then keeps the
// largest of these).
__device__ __forceinline__ bool mw_fits(int C, int n, int mw_min) {
  return mw_min > 0 && C >= mw_min && C <= kMwMaxContacts && n <= kRegBodies;
}

constexpr size_t up8(size_t x) { return (x + 7) & ~size_t(7); }

template <typename R>
struct MwLds {
  static constexpr size_t pos0 = 0;                                   // R[3][kRegBodies]
  static constexpr size_t lad = pos0 + 3 * kRegBodies * sizeof(R);    // R[3][kRegBodies]
  static constexpr size_t st = lad + 3 * kRegBodies * sizeof(R);      // R[7][kRegBodies]
  static constexpr size_t iit = st + 7 * kRegBodies * sizeof(R);      // R[9][kRegBodies]
  static constexpr size_t im = iit + 9 * kRegBodies * sizeof(R);      // R[kRegBodies]
  static constexpr size_t cn = im + kRegBodies * sizeof(R);           // R[3][kMwMaxContacts] normals
  static constexpr size_t cp = cn + 3 * kMwMaxContacts * sizeof(R);   // R[3][kMwMaxContacts] points
  static constexpr size_t pbs = up8(cp + 3 * kMwMaxContacts * sizeof(R));  // double[kMwWaves] offered basis scale
  static constexpr size_t sum = pbs + kMwWaves * sizeof(double);      // double[kMwWaves]
  static constexpr size_t pay = sum + kMwWaves * sizeof(double);      // R[kMwWaves][kMwPay]
  static constexpr size_t res = up8(pay + kMwWaves * kMwPay * sizeof(R));  // R[12] wave 0's result
  static constexpr size_t pkey = up8(res + 12 * sizeof(R));           // R[kMwWaves]
  static constexpr size_t pidx = up8(pkey + kMwWaves * sizeof(R));    // int32[kMwWaves]
  static constexpr size_t pkid = pidx + kMwWaves * sizeof(int32_t);   // uint32[kMwWaves]
  static constexpr size_t resi = pkid + kMwWaves * sizeof(uint32_t);  // int32[4] state, ids, woken body
  static constexpr size_t flags = resi + 4 * sizeof(int32_t);         // uint8[kRegBodies]
  static constexpr size_t bytes = flags + kRegBodies;
};

// The largest (MAX) or least value over the workgroup; every thread gets it.
template <bool MAX>
__device__ __forceinline__ double mw_extreme(double v, double* s_sum, int t) {
  v = dpp::allreduce(v, MAX ? -__builtin_inf() : __builtin_inf(),
                     [](double x, double y) { return MAX ? (y > x ? y : x) : (y < x ? y : x); });
  if ((t & 63) == 0) s_sum[t >> 6] = v;
  __syncthreads();
  double r = s_sum[0];
#pragma unroll
  for (int k = 1; k < kMwWaves; ++k) r = MAX ? (s_sum[k] > r ? s_sum[k] : r) : (s_sum[k] < r ? s_sum[k] : r);
  __syncthreads();
  return r;
}

// Sum over the workgroup in a fixed order (each wave's butterfly, then the
// waves in index order); every thread gets the result.
__device__ __forceinline__ double mw_sum(double v, double* s_sum, int t) {
  v = wave_sum_f64(v);
  if ((t & 63) == 0) s_sum[t >> 6] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int k = 0; k < kMwWaves; ++k) r += s_sum[k];
  __syncthreads();
  return r;
}

// The best of the waves' offers (larger key; equal keys -> lower index);
// gw = -1 when no wave offered a contact.
template <typename R>
__device__ __forceinline__ int mw_best(const R* s_pkey, const int32_t* s_pidx, R& key, int& idx) {
  int gw = -1;
#pragma unroll
  for (int v = 0; v < kMwWaves; ++v) {
    const int i = s_pidx[v];
    if (i == INT_MAX) continue;
    const R k = s_pkey[v];
    if (k > key || (k == key && i < idx)) {
      key = k;
      idx = i;
      gw = v;
    }
  }
  return gw;
}

// applyPositionChange (contact.go:286-386) as apply_position_change_lds,
// from the LDS tables, storing nothing: the bodies' new position and
// orientation come back in np/nq.  Bit i of the result is set when body i
// sleeps (its derived data are rewritten from np/nq, whose orientation has
// then been normalised a second time, as CalculateDerivedData does).
template <typename R>
__device__ __forceinline__ int position_change_mw(const int kl[2], const bool aw[2], const V3<R>& n,
                                                  const V3<R> rel[2], R penetration, V3<R> lin[2], V3<R> ang[2],
                                                  V3<R> np[2], Q4<R> nq[2], const R* s_st, const R* s_iit,
                                                  const R* s_im) {
  constexpr int KB = kRegBodies;
  const R angularLimit = R(0.2);
  R angI[2] = {R(0), R(0)}, linI[2] = {R(0), R(0)};
  R total = R(0);
  int sleeping = 0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    lin[i] = V3<R>{R(0), R(0), R(0)};
    ang[i] = V3<R>{R(0), R(0), R(0)};
    np[i] = V3<R>{R(0), R(0), R(0)};
    nq[i] = Q4<R>{R(1), R(0), R(0), R(0)};
    if (kl[i] < 0) continue;
    R iw[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) iw[k] = s_iit[k * KB + kl[i]];
    V3<R> aiw = cross(rel[i], n);
    aiw = m3v(iw, aiw);
    aiw = cross(aiw, rel[i]);
    angI[i] = dot(aiw, n);
    linI[i] = s_im[kl[i]];
    total += linI[i] + angI[i];
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int l = kl[i];
    if (l < 0) continue;
    R sign = R(1);
    if (i != 0) sign = R(-1);
    R angM = sign * penetration * (angI[i] / total);
    R linM = sign * penetration * (linI[i] / total);
    V3<R> projection = add_scaled(rel[i], n, -dot(rel[i], n));
    R maxMagnitude = angularLimit * mag(projection);
    if (angM < -maxMagnitude) {
      R totalMove = angM + linM;
      angM = -maxMagnitude;
      linM = totalMove - angM;
    } else if (angM > maxMagnitude) {
      R totalMove = angM + linM;
      angM = maxMagnitude;
      linM = totalMove - angM;
    }
    if (angM == R(0)) {
      ang[i] = V3<R>{R(0), R(0), R(0)};
    } else {
      V3<R> tad = cross(rel[i], n);
      R iw[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) iw[k] = s_iit[k * KB + l];
      ang[i] = mul(m3v(iw, tad), angM / angI[i]);
    }
    lin[i] = mul(n, linM);
    np[i] = add_scaled(lds3(s_st, l), n, linM);
    Q4<R> q{s_st[3 * KB + l], s_st[4 * KB + l], s_st[5 * KB + l], s_st[6 * KB + l]};
    q = q_add_scaled_vector(q, ang[i], R(1));
    q = q_normalize(q);
    if (!aw[i]) {
      q = q_normalize(q);
      sleeping |= 1 << i;
    }
    nq[i] = q;
  }
  return sleeping;
}

template <typename R, int KIND, int SL>
__device__ __forceinline__ void resolve_mw(const RS<R>& a, int w, int C, int b0w, int n, char* lds, int t) {
  using L = MwLds<R>;
  constexpr int KB = kRegBodies, T = kMwThreads, KC = kMwMaxContacts;
  constexpr bool MAT = KIND == KIND_MAT;
  R* s_pos0 = (R*)(lds + L::pos0);
  R* s_lad = (R*)(lds + L::lad);
  R* s_st = (R*)(lds + L::st);
  R* s_rot = s_st + 3 * KB;
  R* s_iit = (R*)(lds + L::iit);
  R* s_im = (R*)(lds + L::im);
  R* s_cn = (R*)(lds + L::cn);
  R* s_cp = (R*)(lds + L::cp);
  double* s_pbs = (double*)(lds + L::pbs);
  double* s_sum = (double*)(lds + L::sum);
  R* s_pay = (R*)(lds + L::pay);
  R* s_res = (R*)(lds + L::res);
  R* s_pkey = (R*)(lds + L::pkey);
  int32_t* s_pidx = (int32_t*)(lds + L::pidx);
  uint32_t* s_pkid = (uint32_t*)(lds + L::pkid);
  int32_t* s_resi = (int32_t*)(lds + L::resi);
  uint8_t* s_flags = (uint8_t*)(lds + L::flags);
  auto cnrm = [&](int j) { return V3<R>{s_cn[j], s_cn[KC + j], s_cn[2 * KC + j]}; };
  auto cpt = [&](int j) { return V3<R>{s_cp[j], s_cp[KC + j], s_cp[2 * KC + j]}; };
  // wave 0's result: four vectors (the two bodies' linear and angular
  // changes)
  auto res_store = [&](int k, const V3<R>& v) {
    s_res[3 * k] = v.x;
    s_res[3 * k + 1] = v.y;
    s_res[3 * k + 2] = v.z;
  };
  auto res_load = [&](int k) { return V3<R>{s_res[3 * k], s_res[3 * k + 1], s_res[3 * k + 2]}; };
  const int lane = t & 63, wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const bool do_pos = (a.mode & RM_POS) != 0;
  const bool do_prep = do_pos && !(a.mode & RM_PREPARED);
  const bool do_vel = (a.mode & RM_VEL) != 0 && (KIND == KIND_POLISH || !a.world_mask || a.world_mask[w]);
  if (a.world_iters && t == 0) {
    if (do_pos) a.world_iters[2 * w] = 0;
    if (a.mode & RM_VEL) a.world_iters[2 * w + 1] = 0;
  }
  // ResolveContacts' guard, contact.go:210-212
  if (a.overflow[0] != 0 || C <= 0 || a.dt <= R(0)) return;
  if (!do_pos && !do_vel) return;

  const Bodies<R>& b = a.b;
  const Contacts<R>& c = a.c;
  const int64_t S = b.stride, CS = c.cap;
  const int64_t base = a.world_coff[w];
  const int64_t maxIter = a.iter_fixed > 0 ? a.iter_fixed : a.iter_mult * int64_t(C);
  const V3<R> zero{R(0), R(0), R(0)};

  for (int i = t; i < n; i += T) {
    const int g = b0w + i;
    const V3<R> p = ld3(do_prep ? b.pos : a.pos0, S, g);
    if (do_prep) st3(a.pos0, S, g, p);
    lds3_store(s_pos0, i, p);
    lds3_store(s_lad, i, mul(ld3(b.last_acc, S, g), a.dt));
    s_flags[i] = b.flags[g];
#pragma unroll
    for (int k = 0; k < 9; ++k) s_iit[k * KB + i] = b.iit_world[k * S + g];
    s_im[i] = b.inv_mass[g];
  }
  __syncthreads();

  // prepareContacts / calculateInternals, contact.go:225-229, :59-85, as in
  // resolve_reg
  if (do_prep) {
    for (int j = t; j < C; j += T) {
      const int64_t g = base + j;
      int c0 = c.body0[g], c1 = c.body1[g];
      V3<R> nn = ld3(c.normal, CS, g);
      if (c0 < 0) {  // contact.go:61-65
        nn = mul(nn, R(-1));
        c0 = c1;
        c1 = -1;
        c.body0[g] = c0;
        c.body1[g] = c1;
        st3(c.normal, CS, g, nn);
      }
      const int l0 = c0 - b0w, l1 = c1 >= 0 ? c1 - b0w : -1;
      R m[9];
      basis_from_scale(nn, basis_scale(nn), m);
      const V3<R> p = ld3(c.point, CS, g);
      const V3<R> r0 = sub(p, lds3(s_pos0, l0));
      V3<R> cv = local_velocity_lds(b, c0, lds3(s_lad, l0), r0, m);
      V3<R> r1 = zero;
      if (c1 >= 0) {
        r1 = sub(p, lds3(s_pos0, l1));
        cv = sub(cv, local_velocity_lds(b, c1, lds3(s_lad, l1), r1, m));
      }
      const R ddv = desired_dv_lds(s_flags, s_lad, l0, l1, nn, cv.x, restitution_of<R, KIND>(c, g));
#pragma unroll
      for (int k = 0; k < 9; ++k) c.c2w[k * CS + g] = m[k];
      st3(c.rel, CS, g, r0);
      st3(c.rel + 3 * CS, CS, g, r1);
      st3(c.cvel, CS, g, cv);
      c.ddv[g] = ddv;
    }
  }
  // this thread's contacts into registers and LDS (it wrote what it reads
  // here)
  R key[SL];
  uint32_t ids[SL];
  double bsc[SL];
  R fr[MAT ? SL : 1], rs[MAT ? SL : 1];
#pragma unroll
  for (int s = 0; s < SL; ++s) {
    const int j = T * s + t;
    key[s] = R(0);
    ids[s] = 0xFFFFFFFFu;
    bsc[s] = 0.0;
    if constexpr (MAT) fr[s] = rs[s] = R(0);
    if (j < C) {
      const int64_t g = base + j;
      const int c0 = c.body0[g], c1 = c.body1[g];
      ids[s] = uint32_t(c0 - b0w) | (uint32_t(c1 >= 0 ? c1 - b0w : int(kNoBody)) << 16);
      if (do_pos) key[s] = c.pen[g];
      const V3<R> nn = ld3(c.normal, CS, g), p = ld3(c.point, CS, g);
      s_cn[j] = nn.x; s_cn[KC + j] = nn.y; s_cn[2 * KC + j] = nn.z;
      s_cp[j] = p.x; s_cp[KC + j] = p.y; s_cp[2 * KC + j] = p.z;
      bsc[s] = basis_scale(nn);
      if constexpr (MAT) {
        fr[s] = c.fric[g];
        rs[s] = c.rest[g];
      }
    }
  }
  if (do_pos) {
    for (int i = t; i < n; i += T) {
      const int g = b0w + i;
      lds3_store(s_st, i, ld3(b.pos, S, g));
#pragma unroll
      for (int k = 0; k < 4; ++k) s_st[(3 + k) * KB + i] = b.quat[k * S + g];
    }
  }
  __syncthreads();

  // The surrogate's position polish, as in resolve_reg: snapshot the
  // sleeping bodies' motion and derived data, take the predicted moves, update
  // the penetrations, reject a world they over-separate, then the loop to
  // positionEpsilon + polish_tol within polish_mult * C iterations.
  bool ppolish = KIND == KIND_PPOLISH && do_pos && a.world_mask[w] == 0;
  R sep_floor = R(0);
  auto key_min = [&]() {
    R kmin = R(0);
#pragma unroll
    for (int s = 0; s < SL; ++s)
      if (T * s + t < C) kmin = key[s] < kmin ? key[s] : kmin;
    return R(mw_extreme<false>(double(kmin), s_sum, t));
  };
  auto restore_positions = [&]() {
    // wave 0's HBM stores of the loop (motion, a sleeping body's derived
    // data) complete before the other threads restore over them
    __syncthreads();
    for (int i = t; i < n; i += T) {
      const int g = b0w + i;
      lds3_store(s_st, i, ld3(b.pos, S, g));
#pragma unroll
      for (int k = 0; k < 4; ++k) s_st[(3 + k) * KB + i] = b.quat[k * S + g];
      const uint8_t f = b.flags[g];
      s_flags[i] = f;
      if (!(f & F_AWAKE)) {
        b.motion[g] = a.vel0[g];
#pragma unroll
        for (int k = 0; k < 12; ++k) b.transform[k * S + g] = a.vel0[(1 + k) * S + g];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          const R v = a.vel0[(13 + k) * S + g];
          b.iit_world[k * S + g] = v;
          s_iit[k * KB + i] = v;
        }
      }
    }
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      const int j = T * s + t;
      key[s] = j < C ? c.pen[base + j] : R(0);
    }
    __syncthreads();
    if (t == 0) a.world_mask[w] = 1;
  };
  if constexpr (KIND == KIND_PPOLISH) {
    if (ppolish) {
      R pmax = R(0);
#pragma unroll
      for (int s = 0; s < SL; ++s) pmax = key[s] > pmax ? key[s] : pmax;
      pmax = R(mw_extreme<true>(double(pmax), s_sum, t));
      sep_floor = -(R(kSepFactor) * pmax + a.polish_tol);
      const float* dl = a.delta;
      for (int i = t; i < n; i += T) {
        const int g = b0w + i;
        if (!(s_flags[i] & F_AWAKE)) {
          a.vel0[g] = b.motion[g];
#pragma unroll
          for (int k = 0; k < 12; ++k) a.vel0[(1 + k) * S + g] = b.transform[k * S + g];
#pragma unroll
          for (int k = 0; k < 9; ++k) a.vel0[(13 + k) * S + g] = b.iit_world[k * S + g];
        }
      }
      // the predicted moves times `scale` from the phase's starting state,
      // as in resolve_reg
      auto take_moves = [&](R scale) {
        for (int i = t; i < n; i += T) {
          const int g = b0w + i;
          if (!(s_flags[i] & F_AWAKE)) continue;
          const V3<R> dx = mul(V3<R>{R(dl[g]), R(dl[S + g]), R(dl[2 * S + g])}, scale);
          const V3<R> dq = mul(V3<R>{R(dl[3 * S + g]), R(dl[4 * S + g]), R(dl[5 * S + g])}, scale);
          if (dx.x == R(0) && dx.y == R(0) && dx.z == R(0) && dq.x == R(0) && dq.y == R(0) && dq.z == R(0)) continue;
          lds3_store(s_st, i, add(ld3(b.pos, S, g), dx));
          Q4<R> q{b.quat[g], b.quat[S + g], b.quat[2 * S + g], b.quat[3 * S + g]};
          q = q_normalize(q_add_scaled_vector(q, dq, R(1)));
          s_st[3 * KB + i] = q.w;
          s_st[4 * KB + i] = q.x;
          s_st[5 * KB + i] = q.y;
          s_st[6 * KB + i] = q.z;
        }
#pragma unroll
        for (int s = 0; s < SL; ++s) {
          const int j = T * s + t;
          if (j >= C) continue;
          const int jb[2] = {id_lo(ids[s]), id_hi(ids[s])};
          const V3<R> nj = cnrm(j), ps = cpt(j);
          R pen = c.pen[base + j];
#pragma unroll
          for (int bb = 0; bb < 2; ++bb) {
            if (jb[bb] < 0) continue;
            const int g = b0w + jb[bb];
            const V3<R> dx = mul(V3<R>{R(dl[g]), R(dl[S + g]), R(dl[2 * S + g])}, scale);
            const V3<R> dq = mul(V3<R>{R(dl[3 * S + g]), R(dl[4 * S + g]), R(dl[5 * S + g])}, scale);
            if (dx.x == R(0) && dx.y == R(0) && dx.z == R(0) && dq.x == R(0) && dq.y == R(0) && dq.z == R(0)) continue;
            const V3<R> rj = sub(ps, lds3(s_pos0, jb[bb]));
            const V3<R> dp = add(cross(dq, rj), dx);
            R sign = R(1);
            if (bb == 0) sign = R(-1);
            pen += dot(dp, nj) * sign;
          }
          key[s] = pen;
        }
        __syncthreads();
      };
      take_moves(R(1));
      if (!(key_min() >= sep_floor)) {
        take_moves(R(kPosRetryScale));
        if (!(key_min() >= sep_floor)) {
          restore_positions();
          ppolish = false;
        }
      }
    }
  }

  // adjustPositions, contact.go:233-283.  A slot past the world's contacts
  // keeps key 0 and ids 0xFFFFFFFF, so the loops need no bounds test.
  R peps = ppolish ? R(kPositionEpsilon) + a.polish_tol : R(kPositionEpsilon);
  int64_t pbudget = ppolish ? a.polish_mult * int64_t(C) : maxIter;
  int64_t used = 0;
  while (do_pos) {
    opaque_ids(ids);
    R best = peps;
    int bi = INT_MAX;
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      if (key[s] > best) {
        best = key[s];
        bi = T * s + t;
      }
    }
    wave_argmax_dpp(best, bi);
    if (bi == INT_MAX) {
      if (lane == 0) s_pidx[wv] = INT_MAX;
    } else if ((bi & (T - 1)) == t) {
      const int ss = unsigned(bi) / T;
      s_pkey[wv] = best;
      s_pidx[wv] = bi;
#pragma unroll
      for (int s = 0; s < SL; ++s)
        if (s == ss) s_pkid[wv] = ids[s];
    }
    __syncthreads();  // A
    if (wv == 0) {
      R gk = peps;
      int gi = INT_MAX;
      const int gw = mw_best(s_pkey, s_pidx, gk, gi);
      if (gw < 0) {
        if (lane == 0) s_resi[0] = MW_DONE;
      } else if (used >= pbudget) {
        if (lane == 0) s_resi[0] = MW_BUDGET;
      } else {
        const uint32_t kid = s_pkid[gw];
        const int kl[2] = {id_lo(kid), id_hi(kid)};
        const V3<R> nn = cnrm(gi), p = cpt(gi);
        V3<R> rel[2];
        rel[0] = sub(p, lds3(s_pos0, kl[0]));
        rel[1] = kl[1] >= 0 ? sub(p, lds3(s_pos0, kl[1])) : zero;
        // matchAwakeState, contact.go:185-202
        bool aw[2] = {(s_flags[kl[0]] & F_AWAKE) != 0, kl[1] >= 0 && (s_flags[kl[1]] & F_AWAKE) != 0};
        int woke = -1;
        if (kl[1] >= 0 && (aw[0] || aw[1]) && !(aw[0] && aw[1])) {
          woke = aw[0] ? kl[1] : kl[0];
          aw[0] = aw[1] = true;
        }
        V3<R> lin[2], ang[2], np[2];
        Q4<R> nq[2];
        const int sleeping = position_change_mw(kl, aw, nn, rel, gk, lin, ang, np, nq, s_st, s_iit, s_im);
        if (lane == 0) {
          if (woke >= 0) {  // SetAwake(true), rigidbody.go:183-186
            s_flags[woke] = uint8_t(s_flags[woke] | F_AWAKE);
            b.motion[b0w + woke] = R(kSleepEpsilon * 2.0);
          }
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            const int l = kl[i];
            if (l < 0) continue;
            lds3_store(s_st, l, np[i]);
            s_st[3 * KB + l] = nq[i].w;
            s_st[4 * KB + l] = nq[i].x;
            s_st[5 * KB + l] = nq[i].y;
            s_st[6 * KB + l] = nq[i].z;
            if (sleeping & (1 << i)) {  // CalculateDerivedData, rigidbody.go:268-272
              const int g = b0w + l;
              R Tm[12], IB[9], IW[9];
              set_as_transform(np[i], nq[i], Tm);
              ld9(b.iit_body, S, g, IB);
              transform_inertia(IB, Tm, IW);
#pragma unroll
              for (int k = 0; k < 12; ++k) b.transform[k * S + g] = Tm[k];
#pragma unroll
              for (int k = 0; k < 9; ++k) {
                b.iit_world[k * S + g] = IW[k];
                s_iit[k * KB + l] = IW[k];
              }
            }
          }
          res_store(0, lin[0]);
          res_store(1, lin[1]);
          res_store(2, ang[0]);
          res_store(3, ang[1]);
          if constexpr (KIND == KIND_LABEL) add_pos_label(a, base + gi, gk);
          s_resi[1] = int32_t(kid);
          s_resi[0] = MW_APPLY;
        }
      }
    }
    __syncthreads();  // C
    const int state = s_resi[0];
    if (state != MW_APPLY) {
      bool reject = false;
      if constexpr (KIND == KIND_PPOLISH) {
        // every contact within positionEpsilon + tol: the separation test;
        // out of budget: rejected
        if (ppolish) reject = state == MW_BUDGET || !(key_min() >= sep_floor);
      }
      if (!reject) break;
      restore_positions();
      ppolish = false;
      peps = R(kPositionEpsilon);
      pbudget = maxIter;
      used = 0;
      continue;
    }
    {
      const uint32_t kid = uint32_t(s_resi[1]);
      const int kl[2] = {id_lo(kid), id_hi(kid)};
      const V3<R> lin[2] = {res_load(0), res_load(1)};
      const V3<R> ang[2] = {res_load(2), res_load(3)};
#pragma unroll
      for (int s = 0; s < SL; ++s) {
        const int jb[2] = {id_lo(ids[s]), id_hi(ids[s])};
        bool any = false;
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
          for (int d = 0; d < 2; ++d) any |= (jb[bb] >= 0 && jb[bb] == kl[d]);
        if (!any) continue;
        R pen = key[s];
        const V3<R> ns = cnrm(T * s + t), ps = cpt(T * s + t);
#pragma unroll
        for (int bb = 0; bb < 2; ++bb) {
          if (jb[bb] < 0) continue;
          const V3<R> rj = sub(ps, lds3(s_pos0, jb[bb]));
          // a side is at most one of the two moved bodies: its change is
          // selected and applied once
          if (jb[bb] == kl[0] || jb[bb] == kl[1]) {
            const bool d0 = jb[bb] == kl[0];
            V3<R> dp = add(cross(sel3(d0, ang[0], ang[1]), rj), sel3(d0, lin[0], lin[1]));
            R sign = R(1);
            if (bb == 0) sign = R(-1);
            pen += dot(dp, ns) * sign;
          }
        }
        key[s] = pen;
      }
    }
    ++used;
  }
  const int64_t pos_iters = used;
  __syncthreads();
  if (do_pos) {
    for (int i = t; i < n; i += T) {
      const int g = b0w + i;
      st3(b.pos, S, g, lds3(s_st, i));
#pragma unroll
      for (int k = 0; k < 4; ++k) b.quat[k * S + g] = s_st[(3 + k) * KB + i];
    }
  }

  // adjustVelocities, contact.go:390-445
  V3<R> cv[SL];
  R imp[SL][kImpPlanes];
#pragma unroll
  for (int s = 0; s < SL; ++s) {
    const int j = T * s + t;
    cv[s] = zero;
#pragma unroll
    for (int k = 0; k < kImpPlanes; ++k) imp[s][k] = R(0);
    if (do_pos && j < C) c.pen[base + j] = key[s];
    if (do_vel && j < C) {
      key[s] = c.ddv[base + j];  // selection keys as prepare stored them
      cv[s] = ld3(c.cvel, CS, base + j);
      // the impulse data, fixed for the whole phase (as impulse_data from
      // the contact arrays prepare stored: same values)
      const int jb[2] = {id_lo(ids[s]), id_hi(ids[s])};
      R mj[9], i0[9], i1[9];
      const V3<R> ps = cpt(j);
      basis_from_scale(cnrm(j), bsc[s], mj);
      const V3<R> rj[2] = {sub(ps, lds3(s_pos0, jb[0])), jb[1] >= 0 ? sub(ps, lds3(s_pos0, jb[1])) : zero};
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        i0[k] = s_iit[k * KB + jb[0]];
        i1[k] = jb[1] >= 0 ? s_iit[k * KB + jb[1]] : R(0);
      }
      const R invm[2] = {s_im[jb[0]], jb[1] >= 0 ? s_im[jb[1]] : R(0)};
      R fj = R(kFriction);
      if constexpr (MAT) fj = fr[s];
      impulse_data_from<R, KIND>(a.mode, i0, i1, invm, jb[1] >= 0, rj, mj, fj, imp[s]);
    }
  }
  // the position write-back read the LDS state that the velocity phase
  // reuses
  __syncthreads();
  used = 0;
  if (do_vel) {
    for (int i = t; i < n; i += T) {
      const int g = b0w + i;
      lds3_store(s_st, i, ld3(b.vel, S, g));
      lds3_store(s_rot, i, ld3(b.rot, S, g));
    }
    bool polishing = KIND == KIND_POLISH && a.world_mask[w] == 0;
    double ke_before = 0.0, world_mass = 0.0;
    if constexpr (KIND == KIND_POLISH) {
      if (polishing) {
        // take the predicted changes (as resolve_reg)
        const float* dl = a.delta;
        double ke = 0.0, mass = 0.0;
        for (int i = t; i < n; i += T) {
          const int g = b0w + i;
          const V3<R> v = lds3(s_st, i), r = lds3(s_rot, i);
          R iw[9];
#pragma unroll
          for (int k = 0; k < 9; ++k) iw[k] = s_iit[k * KB + i];
          double mi;
          ke += kinetic_energy(v, r, s_im[i], iw, mi);
          mass += mi;
          a.vel0[6 * S + g] = b.motion[g];
          const V3<R> dv{R(dl[g]), R(dl[S + g]), R(dl[2 * S + g])};
          const V3<R> dw{R(dl[3 * S + g]), R(dl[4 * S + g]), R(dl[5 * S + g])};
          if (dv.x != R(0) || dv.y != R(0) || dv.z != R(0) || dw.x != R(0) || dw.y != R(0) || dw.z != R(0)) {
            lds3_store(s_st, i, add(v, dv));
            lds3_store(s_rot, i, add(r, dw));
          }
        }
#pragma unroll
        for (int s = 0; s < SL; ++s) {
          if (T * s + t >= C) continue;
          const int jb[2] = {id_lo(ids[s]), id_hi(ids[s])};
          R mj[9];
          basis_from_scale(cnrm(T * s + t), bsc[s], mj);
          const V3<R> ps = cpt(T * s + t);
          V3<R> cvj = cv[s];
          bool moved = false;
#pragma unroll
          for (int bb = 0; bb < 2; ++bb) {
            if (jb[bb] < 0) continue;
            const int g = b0w + jb[bb];
            const V3<R> dv{R(dl[g]), R(dl[S + g]), R(dl[2 * S + g])};
            const V3<R> dw{R(dl[3 * S + g]), R(dl[4 * S + g]), R(dl[5 * S + g])};
            if (dv.x == R(0) && dv.y == R(0) && dv.z == R(0) && dw.x == R(0) && dw.y == R(0) && dw.z == R(0)) continue;
            moved = true;
            const V3<R> rj = sub(ps, lds3(s_pos0, jb[bb]));
            R sign = R(1);
            if (bb == 1) sign = R(-1);
            cvj = add(cvj, mul(m3tv(mj, add(cross(dw, rj), dv)), sign));
          }
          if (moved) {
            cv[s] = cvj;
            key[s] = desired_dv_lds(s_flags, s_lad, jb[0], jb[1], V3<R>{mj[0], mj[1], mj[2]}, cvj.x,
                                    R(kRestitution));
          }
        }
        ke_before = mw_sum(ke, s_sum, t);
        world_mass = mw_sum(mass, s_sum, t);
      }
    }
    __syncthreads();
    R eps = polishing ? R(kVelocityEpsilon) + a.polish_tol : R(kVelocityEpsilon);
    int64_t budget = polishing ? a.polish_mult * int64_t(C) : maxIter;
    int64_t it = 0;
    while (true) {
      opaque_ids(ids);
      R best = eps;
      int bi = INT_MAX;
#pragma unroll
      for (int s = 0; s < SL; ++s) {
        if (key[s] > best) {
          best = key[s];
          bi = T * s + t;
        }
      }
      wave_argmax_dpp(best, bi);
      if (bi == INT_MAX) {
        if (lane == 0) s_pidx[wv] = INT_MAX;
      } else if ((bi & (T - 1)) == t) {
        const int ss = unsigned(bi) / T;
        R* pay = s_pay + wv * kMwPay;
        s_pkey[wv] = best;
        s_pidx[wv] = bi;
#pragma unroll
        for (int s = 0; s < SL; ++s) {
          if (s != ss) continue;
          s_pkid[wv] = ids[s];
          s_pbs[wv] = bsc[s];
          pay[0] = cv[s].x; pay[1] = cv[s].y; pay[2] = cv[s].z;
#pragma unroll
          for (int k = 0; k < kImpPlanes; ++k) pay[3 + k] = imp[s][k];
          if constexpr (MAT) pay[15] = fr[s];
        }
      }
      __syncthreads();  // A
      if (wv == 0) {
        R gk = eps;
        int gi = INT_MAX;
        const int gw = mw_best(s_pkey, s_pidx, gk, gi);
        if (gw < 0 || it >= budget) {
          if (lane == 0) s_resi[0] = gw < 0 ? MW_DONE : MW_BUDGET;
        } else {
          const uint32_t kid = s_pkid[gw];
          const R* pay = s_pay + gw * kMwPay;
          const int kl[2] = {id_lo(kid), id_hi(kid)};
          const V3<R> nn = cnrm(gi), p = cpt(gi);
          const V3<R> cvk{pay[0], pay[1], pay[2]};
          R m[9], ik[kImpPlanes];
          basis_from_scale(nn, s_pbs[gw], m);
#pragma unroll
          for (int k = 0; k < kImpPlanes; ++k) ik[k] = pay[3 + k];
          R fk = R(kFriction);
          if constexpr (MAT) fk = pay[15];
          V3<R> rel[2];
          rel[0] = sub(p, lds3(s_pos0, kl[0]));
          rel[1] = kl[1] >= 0 ? sub(p, lds3(s_pos0, kl[1])) : zero;
          // matchAwakeState, contact.go:185-202
          const bool aw0 = (s_flags[kl[0]] & F_AWAKE) != 0;
          const bool aw1 = kl[1] >= 0 && (s_flags[kl[1]] & F_AWAKE) != 0;
          int woke = -1;
          if (kl[1] >= 0 && (aw0 || aw1) && !(aw0 && aw1)) woke = aw0 ? kl[1] : kl[0];
          // applyVelocityChange, contact.go:448-494 (apply_velocity_change_lds)
          V3<R> ic;
          if (MAT && fk == R(0))
            ic = V3<R>{gk / ik[0], R(0), R(0)};
          else
            ic = friction_impulse_from(ik, ik + 9, gk, cvk, fk);
          const V3<R> impulse = m3v(m, ic);
          V3<R> vc[2], rc[2];
          R iw[9];
#pragma unroll
          for (int k = 0; k < 9; ++k) iw[k] = s_iit[k * KB + kl[0]];
          V3<R> torque = cross(rel[0], impulse);
          rc[0] = m3v(iw, torque);
          vc[0] = add_scaled(zero, impulse, s_im[kl[0]]);
          rc[1] = zero;
          vc[1] = zero;
          if (kl[1] >= 0) {
#pragma unroll
            for (int k = 0; k < 9; ++k) iw[k] = s_iit[k * KB + kl[1]];
            torque = cross(impulse, rel[1]);
            rc[1] = m3v(iw, torque);
            vc[1] = add_scaled(zero, impulse, -s_im[kl[1]]);
          }
          if (lane == 0) {
            if (woke >= 0) {  // SetAwake(true), rigidbody.go:183-186
              s_flags[woke] = uint8_t(s_flags[woke] | F_AWAKE);
              b.motion[b0w + woke] = R(kSleepEpsilon * 2.0);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i) {
              if (kl[i] < 0) continue;
              lds3_store(s_st, kl[i], add(lds3(s_st, kl[i]), vc[i]));
              lds3_store(s_rot, kl[i], add(lds3(s_rot, kl[i]), rc[i]));
            }
            if constexpr (KIND == KIND_LABEL) add_label(a, base + gi, vc, rc);
            res_store(0, vc[0]);
            res_store(1, vc[1]);
            res_store(2, rc[0]);
            res_store(3, rc[1]);
            s_resi[1] = int32_t(kid);
            s_resi[0] = MW_APPLY;
          }
        }
      }
      __syncthreads();  // C
      const int state = s_resi[0];
      if (state != MW_APPLY) {
        bool reject = false;  // the polish did not give an acceptable result
        if (state == MW_DONE) {
          if constexpr (KIND == KIND_POLISH) {
            // the acceptance test of resolve_reg
            if (polishing) {
              double ke = 0.0, mass = 0.0;
              for (int i = t; i < n; i += T) {
                R iw[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) iw[k] = s_iit[k * KB + i];
                ke += kinetic_energy(lds3(s_st, i), lds3(s_rot, i), s_im[i], iw, mass);
              }
              const double ke_after = mw_sum(ke, s_sum, t);
              const double tol = double(a.polish_tol);
              reject = !(ke_after <= ke_before * (1.0 + kEnergyRel) + 0.5 * tol * tol * world_mass);
            }
          }
          if (!reject) break;
        } else {  // the iteration budget is used
          if (KIND != KIND_POLISH || !polishing) break;
          reject = true;
        }
        // undo the prediction and the loop; run the plain phase
        __syncthreads();
        for (int i = t; i < n; i += T) {
          const int g = b0w + i;
          lds3_store(s_st, i, ld3(b.vel, S, g));
          lds3_store(s_rot, i, ld3(b.rot, S, g));
          b.motion[g] = a.vel0[6 * S + g];
          s_flags[i] = b.flags[g];
        }
#pragma unroll
        for (int s = 0; s < SL; ++s) {
          const int j = T * s + t;
          if (j < C) {
            key[s] = c.ddv[base + j];
            cv[s] = ld3(c.cvel, CS, base + j);
          }
        }
        __syncthreads();
        if (t == 0) a.world_mask[w] = 1;
        polishing = false;
        eps = R(kVelocityEpsilon);
        budget = maxIter;
        it = 0;
        continue;
      }
      {
        const uint32_t kid = uint32_t(s_resi[1]);
        const int kl[2] = {id_lo(kid), id_hi(kid)};
        const V3<R> vc[2] = {res_load(0), res_load(1)};
        const V3<R> rc[2] = {res_load(2), res_load(3)};
#pragma unroll
        for (int s = 0; s < SL; ++s) {
          const int jb[2] = {id_lo(ids[s]), id_hi(ids[s])};
          bool any = false;
#pragma unroll
          for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int d = 0; d < 2; ++d) any |= (jb[bb] >= 0 && jb[bb] == kl[d]);
          if (!any) continue;
          R mj[9];
          basis_from_scale(cnrm(T * s + t), bsc[s], mj);
          const V3<R> ps = cpt(T * s + t);
          V3<R> cvj = cv[s];
#pragma unroll
          for (int bb = 0; bb < 2; ++bb) {
            if (jb[bb] < 0) continue;
            const V3<R> rj = sub(ps, lds3(s_pos0, jb[bb]));
            // a side is at most one of the two moved bodies: its change is
            // selected and applied once
            if (jb[bb] == kl[0] || jb[bb] == kl[1]) {
              const bool d0 = jb[bb] == kl[0];
              V3<R> dv = add(cross(sel3(d0, rc[0], rc[1]), rj), sel3(d0, vc[0], vc[1]));
              R sign = R(1);
              if (bb == 1) sign = R(-1);
              cvj = add(cvj, mul(m3tv(mj, dv), sign));
            }
          }
          cv[s] = cvj;
          R e = R(kRestitution);
          if constexpr (MAT) e = rs[s];
          key[s] = desired_dv_lds(s_flags, s_lad, jb[0], jb[1], V3<R>{mj[0], mj[1], mj[2]}, cvj.x, e);
        }
      }
      ++it;
      ++used;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      const int j = T * s + t;
      if (j < C) {
        c.ddv[base + j] = key[s];
        st3(c.cvel, CS, base + j, cv[s]);
      }
    }
    for (int i = t; i < n; i += T) {
      const int g = b0w + i;
      st3(b.vel, S, g, lds3(s_st, i));
      st3(b.rot, S, g, lds3(s_rot, i));
    }
  }
  __syncthreads();
  for (int i = t; i < n; i += T) b.flags[b0w + i] = s_flags[i];
  if (t == 0 && a.world_iters) {
    if (do_pos) a.world_iters[2 * w] = pos_iters;
    if (do_vel) a.world_iters[2 * w + 1] = used;
  }
}

template <typename R, int KIND>
__device__ __forceinline__ void resolve_mw_world(const RS<R>& a, int w, int C, int b0w, int n, char* lds, int t) {
  if (C <= kMwThreads * 4)
    resolve_mw<R, KIND, 4>(a, w, C, b0w, n, lds, t);
  else
    resolve_mw<R, KIND, kMwMaxSlots>(a, w, C, b0w, n, lds, t);
}

constexpr size_t cmax(size_t x, size_t y) { return x > y ? x : y; }
template <typename R, int KIND>
constexpr size_t resolve_lds_bytes() {
  return cmax(cmax(RegLds<R, 32, KIND>::bytes, RegLds<R, 8, KIND>::bytes),
              kKeyCap * (sizeof(R) + 2 * sizeof(int32_t)));
}

// The polish kind carries more state through its loop: its widest
// register class is 24 slots (1536 contacts).
template <int KIND>
constexpr int max_slots() {
  return KIND == KIND_POLISH ? 24 : kRegMaxSlots;
}

template <typename R, int KIND>
__device__ __forceinline__ void resolve_world_kind(const RS<R>& a, int w, int C, int b0w, int n, char* lds, int lane) {
  if (!fits_registers(C, n) || C > 64 * max_slots<KIND>()) {
    resolve_hbm<R, KIND>(a, w, C, lds, lane);
  } else if (C <= 64) {
    resolve_reg<R, 1, KIND>(a, w, C, b0w, n, lds, lane);
  } else if (C <= 128) {
    resolve_reg<R, 2, KIND>(a, w, C, b0w, n, lds, lane);
  } else if (C <= 256) {
    resolve_reg<R, 4, KIND>(a, w, C, b0w, n, lds, lane);
  } else if (C <= 512) {
    resolve_reg<R, 8, KIND>(a, w, C, b0w, n, lds, lane);
  } else if (C <= 768) {
    resolve_reg<R, 12, KIND>(a, w, C, b0w, n, lds, lane);
  } else if (C <= 1024) {
    resolve_reg<R, 16, KIND>(a, w, C, b0w, n, lds, lane);
  } else if (C <= 1536) {
    resolve_reg<R, 24, KIND>(a, w, C, b0w, n, lds, lane);
  } else if constexpr (max_slots<KIND>() >= 32) {
    resolve_reg<R, 32, KIND>(a, w, C, b0w, n, lds, lane);
  }
}

// One launch for every world: each workgroup takes the world order[blockIdx]
// (larger worlds first) and runs the variant for its contact count.
template <typename R, int KIND>
__device__ __forceinline__ void resolve_block(const RS<R>& a) {
  __shared__ __align__(16) char lds[resolve_lds_bytes<R, KIND>()];
  const int lane = threadIdx.x;
  const int w = a.order ? a.order[blockIdx.x] : int(blockIdx.x);
  if (w >= a.n_worlds) return;
  const int C = a.world_cnt[w];
  const int b0w = a.body_off[w], n = a.body_off[w + 1] - b0w;
  if (a.mw_list && a.mw_flag[w]) return;  // the multi-wave kernel's
  int64_t* wt = a.world_time ? a.world_time + 4 * int64_t(w) + ((a.mode & RM_VEL) ? 2 : 0) : nullptr;
  if (wt && lane == 0) wt[0] = int64_t(wall_clock64());
  resolve_world_kind<R, KIND>(a, w, C, b0w, n, lds, lane);
  if (wt && lane == 0) wt[1] = int64_t(wall_clock64());
}

// The multi-wave kernel, persistent: each workgroup takes the next listed
// world (largest first) from the queue until the list is done.
template <typename R, int KIND>
__device__ __forceinline__ void resolve_mw_block(const RS<R>& a) {
  __shared__ __align__(16) char lds[MwLds<R>::bytes];
  __shared__ int32_t s_next;
  const int t = threadIdx.x;
  const int nbig = a.mw_count[0];
  while (true) {
    if (t == 0) s_next = atomicAdd(&a.mw_count[1], 1);
    __syncthreads();
    const int k = s_next;
    __syncthreads();
    if (k >= nbig) break;
    const int w = a.mw_list[k];
    const int C = a.world_cnt[w];
    const int b0w = a.body_off[w], n = a.body_off[w + 1] - b0w;
    int64_t* wt = a.world_time ? a.world_time + 4 * int64_t(w) + ((a.mode & RM_VEL) ? 2 : 0) : nullptr;
    if (wt && t == 0) wt[0] = int64_t(wall_clock64());
    resolve_mw_world<R, KIND>(a, w, C, b0w, n, lds, t);
    __syncthreads();
    if (wt && t == 0) wt[1] = int64_t(wall_clock64());
  }
}

// The FP64 kernel keeps a full wave's registers (its working set spills
// heavily below that); its LDS (about 21 KB) lets 4 worlds share a CU.  The
// FP32 kernel is held to two waves per SIMD: its few spills cost less than
// the second wave gains.
template <int KIND>
__global__ void __launch_bounds__(64) k_resolve_f64(RS<double> a) {
  resolve_block<double, KIND>(a);
}
template <int KIND>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) k_resolve_f32(RS<float> a) {
  resolve_block<float, KIND>(a);
}

template <int KIND>
__global__ void __launch_bounds__(kMwThreads) k_resolve_mw_f64(RS<double> a) {
  resolve_mw_block<double, KIND>(a);
}
template <int KIND>
__global__ void __launch_bounds__(kMwThreads) k_resolve_mw_f32(RS<float> a) {
  resolve_mw_block<float, KIND>(a);
}

template <typename R, int KIND>
void launch_kind(const RS<R>& a, hipStream_t st) {
  if constexpr (sizeof(R) == 8)
    k_resolve_f64<KIND><<<a.n_worlds, 64, 0, st>>>(a);
  else
    k_resolve_f32<KIND><<<a.n_worlds, 64, 0, st>>>(a);
}

// The multi-wave kernel over mw_list: at most one workgroup per CU (each
// takes a whole CU's registers).
template <typename R, int KIND>
void launch_mw_kind(const RS<R>& a, int grid, hipStream_t st) {
  if constexpr (sizeof(R) == 8)
    k_resolve_mw_f64<KIND><<<grid, kMwThreads, 0, st>>>(a);
  else
    k_resolve_mw_f32<KIND><<<grid, kMwThreads, 0, st>>>(a);
}

// Worlds (workgroups) of one kind's kernel resident per CU, from its
// register and LDS use.
template <typename R, int KIND>
hipError_t kind_occupancy(int* worlds_per_cu) {
  if constexpr (sizeof(R) == 8)
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(worlds_per_cu, k_resolve_f64<KIND>, 64, 0);
  else
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(worlds_per_cu, k_resolve_f32<KIND>, 64, 0);
}

// Every instantiation lives in one translation unit of its own
// (k_resolve_<precision>_<kind>.hip); other units only declare them.
#define CBZ_RESOLVE_KIND(EXT, R, KIND)                                          \
  EXT template void launch_kind<R, KIND>(const RS<R>&, hipStream_t);           \
  EXT template void launch_mw_kind<R, KIND>(const RS<R>&, int, hipStream_t);   \
  EXT template hipError_t kind_occupancy<R, KIND>(int*);
#ifndef CBZ_RESOLVE_INSTANTIATE
CBZ_RESOLVE_KIND(extern, double, KIND_PLAIN)
CBZ_RESOLVE_KIND(extern, double, KIND_LABEL)
CBZ_RESOLVE_KIND(extern, double, KIND_POLISH)
CBZ_RESOLVE_KIND(extern, double, KIND_MAT)
CBZ_RESOLVE_KIND(extern, double, KIND_PPOLISH)
CBZ_RESOLVE_KIND(extern, float, KIND_PLAIN)
CBZ_RESOLVE_KIND(extern, float, KIND_LABEL)
CBZ_RESOLVE_KIND(extern, float, KIND_POLISH)
CBZ_RESOLVE_KIND(extern, float, KIND_MAT)
CBZ_RESOLVE_KIND(extern, float, KIND_PPOLISH)
#endif

}  // namespace cbz
