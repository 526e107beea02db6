// k_resolve.hip — K3: ResolveContacts (contact.go:208-222), one wavefront
// per world.
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
#include <climits>

#include "cbz_devmath.h"
#include "cbz_launch.h"
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
// (also records every impulse's body changes per contact), KIND_POLISH (the
// surrogate's polish with the plain fallback), KIND_MAT (KIND_PLAIN with
// per-contact friction and restitution, and the frictionless impulse for
// contacts whose friction is 0).
enum : int { KIND_PLAIN = 0, KIND_LABEL = 1, KIND_POLISH = 2, KIND_MAT = 3 };

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
  int32_t n_worlds;
  R dt;
  int64_t iter_mult;
  int64_t iter_fixed;
  int32_t mode;                // RM_* bits
  uint8_t* world_mask;         // velocity phase only where set (null: every world); RM_POLISH: in/out
  R* pos0;                     // [3][n_bodies] body positions at prepare time
  float* labels;               // [cap][12] accumulated velocity-phase changes (null: off)
  const float* delta;          // RM_POLISH: [6][n_bodies] velocity and rotation changes
  R* vel0;                     // RM_POLISH: [7][n_bodies] vel, rot, motion before them
  R* cvel0;                    // RM_POLISH: [3][cap] contact velocities as prepare left them
  R polish_tol;
  int64_t polish_mult;
};


template <typename R>
__device__ __forceinline__ void add_label(const RS<R>& a, int64_t g, const V3<R> vc[2], const V3<R> rc[2]) {
  float* L = a.labels + g * 12;
  const V3<R>* v[4] = {&vc[0], &rc[0], &vc[1], &rc[1]};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    L[3 * k + 0] += float(v[k]->x);
    L[3 * k + 1] += float(v[k]->y);
    L[3 * k + 2] += float(v[k]->z);
  }
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

// calculateInternals, contact.go:59-85, for contact g.
template <typename R, int KIND>
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
// contacts).  ContactNormal is column 0 of contactToWorld.  The reference's
// test at :512 is inverted (Bodies[1] == nil): a two-body contact never adds
// its second body's terms, and a half-space contact dereferences nil there
// (a Go panic; here the absent body adds nothing).  two_body
// (RM_FRICTIONLESS2) adds the second body's terms when it exists.
template <typename R>
__device__ __forceinline__ V3<R> frictionless_impulse(const Bodies<R>& b, const int kb[2], const R* iit0,
                                                      const R* iit1, const V3<R> rel[2], const R* c2w, R ddv,
                                                      bool two_body) {
  const V3<R> n{c2w[0], c2w[1], c2w[2]};
  V3<R> dvw = cross(rel[0], n);
  dvw = m3v(iit0, dvw);
  dvw = cross(dvw, rel[0]);
  R deltaVelocity = dot(dvw, n);
  deltaVelocity += b.inv_mass[kb[0]];
  if (two_body && kb[1] >= 0) {
    dvw = cross(rel[1], n);
    dvw = m3v(iit1, dvw);
    dvw = cross(dvw, rel[1]);
    deltaVelocity += dot(dvw, n);
    deltaVelocity += b.inv_mass[kb[1]];
  }
  return V3<R>{ddv / deltaVelocity, R(0), R(0)};
}

// calculateFrictionImpulse, contact.go:535-606
template <typename R>
__device__ __forceinline__ V3<R> friction_impulse(const Bodies<R>& b, const int kb[2], const R* iit0, const R* iit1,
                                                  const V3<R> rel[2], const R* c2w, R ddv, const V3<R>& cvel,
                                                  R friction) {
  R inverseMass = b.inv_mass[kb[0]];
  R sk[9], t[9], dvw[9];
  skew(rel[0], sk);
  m3mul(sk, iit0, t);
  m3mul(t, sk, dvw);
#pragma unroll
  for (int k = 0; k < 9; ++k) dvw[k] *= R(-1);
  if (kb[1] >= 0) {
    R d2[9];
    skew(rel[1], sk);
    m3mul(sk, iit1, t);
    m3mul(t, sk, d2);
#pragma unroll
    for (int k = 0; k < 9; ++k) d2[k] *= R(-1);
#pragma unroll
    for (int k = 0; k < 9; ++k) dvw[k] += d2[k];
    inverseMass += b.inv_mass[kb[1]];
  }
  R ct[9] = {c2w[0], c2w[3], c2w[6], c2w[1], c2w[4], c2w[7], c2w[2], c2w[5], c2w[8]};
  R dv[9];
  m3mul(ct, dvw, t);
  m3mul(t, c2w, dv);
  dv[0] += inverseMass;
  dv[4] += inverseMass;
  dv[8] += inverseMass;
  R im[9];
  m3inv(dv, im);
  V3<R> velKill{ddv, -cvel.y, -cvel.z};
  V3<R> ic = m3v(im, velKill);
  R planar = dsqrt(ic.y * ic.y + ic.z * ic.z);
  if (planar > ic.x * friction) {
    ic.y /= planar;
    ic.z /= planar;
    ic.x = dv[0] + dv[3] * friction * ic.y + dv[6] * friction * ic.z;
    ic.x = ddv / ic.x;
    ic.y *= friction * ic.x;
    ic.z *= friction * ic.x;
  }
  return ic;
}

// applyVelocityChange, contact.go:448-494, for a contact of the given
// friction (the frictionless branch of :459-461 exists in KIND_MAT only)
template <typename R, int KIND>
__device__ __forceinline__ void apply_velocity_change(const RS<R>& a, const int kb[2], const V3<R> rel[2], const R* c2w, R ddv,
                                      const V3<R>& cvel, V3<R> vc[2], V3<R> rc[2], int lane, R friction) {
  const Bodies<R>& b = a.b;
  const int64_t S = b.stride;
  R iit0[9], iit1[9];
  ld9(b.iit_world, S, kb[0], iit0);
#pragma unroll
  for (int k = 0; k < 9; ++k) iit1[k] = R(0);
  if (kb[1] >= 0) ld9(b.iit_world, S, kb[1], iit1);
  V3<R> ic;
  if (KIND == KIND_MAT && friction == R(0))
    ic = frictionless_impulse(b, kb, iit0, iit1, rel, c2w, ddv, (a.mode & RM_FRICTIONLESS2) != 0);
  else
    ic = friction_impulse(b, kb, iit0, iit1, rel, c2w, ddv, cvel, friction);
  const V3<R> impulse = m3v(c2w, ic);
  const V3<R> zero{R(0), R(0), R(0)};
  V3<R> torque = cross(rel[0], impulse);
  rc[0] = m3v(iit0, torque);
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
    torque = cross(impulse, rel[1]);
    rc[1] = m3v(iit1, torque);
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
  const bool do_vel = (a.mode & RM_VEL) != 0 && (KIND == KIND_POLISH || !a.world_mask || a.world_mask[w]);
  // worlds this path takes are never polished: plain velocity phase
  if (KIND == KIND_POLISH && lane == 0) a.world_mask[w] = 1;
  // prepareContacts, contact.go:225-229
  for (int j = lane; j < C; j += 64) {
    if (do_pos) prepare_contact<R, KIND>(a, base + j);
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
#pragma unroll
        for (int d = 0; d < 2; ++d) {
          if (jb[bb] == kb[d]) {
            V3<R> dp = add(cross(ang[d], rj), lin[d]);
            R sign = R(1);
            if (bb == 0) sign = R(-1);
            p += dot(dp, nj) * sign;
          }
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
    apply_velocity_change<R, KIND>(a, kb, rel, m, best, cv, vc, rc, lane, friction_of<R, KIND>(c, g));
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
#pragma unroll
        for (int d = 0; d < 2; ++d) {
          if (jb[bb] == kb[d]) {
            V3<R> dv = add(cross(rc[d], rj), vc[d]);
            R sign = R(1);
            if (bb == 1) sign = R(-1);
            cvj = add(cvj, mul(m3tv(mj, dv), sign));
          }
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
constexpr int kRegMaxSlots = 16;  // contacts per lane (64 * 16 = 1024 per world)

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
// VGPRs -- selection key, body ids (world-local, packed 16+16 bits), normal,
// point and (velocity phase) contact velocity -- so the argmax and the
// shared-body update read no memory.  LDS holds per body the position at
// prepare time (relativeContactPosition = point - that position, recomputed
// bit for bit where the reference stores it), lastFrameAccelleration * dt
// and the flag byte (matchAwakeState writes it).  contactToWorld is rebuilt
// from the normal where needed (calculateContactBasis reads nothing else).
// The two bodies a correction moves are read from and written to HBM, one
// round trip per iteration.
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

// Restitution of world-local contact j: from LDS in KIND_MAT, else the
// reference's constant.
template <typename R, int KIND>
__device__ __forceinline__ R lds_restitution(const R* s_rest, int j) {
  if constexpr (KIND == KIND_MAT) return s_rest[j];
  return R(kRestitution);
}

__device__ __forceinline__ int id_lo(uint32_t id) { return int(id & 0xFFFFu); }
__device__ __forceinline__ int id_hi(uint32_t id) { return (id >> 16) == kNoBody ? -1 : int(id >> 16); }

template <typename R, int SLOTS, int KIND>
struct RegLds {
  // the widest variants read contact points from HBM (c.point) where they
  // need them: in VGPRs they would spill, and in LDS they would cut the
  // worlds resident per CU
  static constexpr bool PT_HBM = SLOTS >= 12;
  static constexpr int CAP = 64 * SLOTS;
  static constexpr bool MAT = KIND == KIND_MAT;
  static constexpr size_t pos0 = 0;                                   // R[3][kRegBodies]
  static constexpr size_t lad = pos0 + 3 * kRegBodies * sizeof(R);    // R[3][kRegBodies]
  static constexpr size_t bs = lad + 3 * kRegBodies * sizeof(R);      // double[CAP] basis scale
  static constexpr size_t bc = bs + CAP * sizeof(double);             // R[9] broadcast
  static constexpr size_t flags = bc + 16 * sizeof(R);                // uint8[kRegBodies]
  static constexpr size_t bcid = flags + kRegBodies;                  // uint32
  static constexpr size_t rest = bcid + 16;                           // R[CAP] restitution (MAT)
  static constexpr size_t bytes = rest + (MAT ? CAP * sizeof(R) : 0);
};

template <typename R, int SLOTS, int KIND>
__device__ __forceinline__ void resolve_reg(const RS<R>& a, int w, int C, int b0w, int n, char* lds, int lane) {
  using L = RegLds<R, SLOTS, KIND>;
  constexpr bool PT_HBM = L::PT_HBM;
  constexpr int CAP = L::CAP;
  R* s_pos0 = (R*)(lds + L::pos0);
  R* s_lad = (R*)(lds + L::lad);
  double* s_bs = (double*)(lds + L::bs);
  R* s_bc = (R*)(lds + L::bc);
  uint8_t* s_flags = (uint8_t*)(lds + L::flags);
  uint32_t& s_bcid = *(uint32_t*)(lds + L::bcid);
  R* s_rest = (R*)(lds + L::rest);  // KIND_MAT: restitution of contact j (world-local)
  const bool do_pos = (a.mode & RM_POS) != 0;
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

  for (int i = lane; i < n; i += 64) {
    const int g = b0w + i;
    // body positions at prepare time: the current ones when this launch
    // prepares, else the snapshot the preparing launch stored
    const V3<R> p = ld3(do_pos ? b.pos : a.pos0, S, g);
    if (do_pos) st3(a.pos0, S, g, p);
    const V3<R> la = mul(ld3(b.last_acc, S, g), a.dt);
    s_pos0[i] = p.x;
    s_pos0[kRegBodies + i] = p.y;
    s_pos0[2 * kRegBodies + i] = p.z;
    s_lad[i] = la.x;
    s_lad[kRegBodies + i] = la.y;
    s_lad[2 * kRegBodies + i] = la.z;
    s_flags[i] = b.flags[g];
  }
  __syncthreads();

  R key[SLOTS];
  uint32_t ids[SLOTS];
  V3<R> nrm[SLOTS];
  V3<R> ptr[PT_HBM ? 1 : SLOTS];
  auto get_pt = [&](int s) -> V3<R> {
    if constexpr (PT_HBM) {
      return ld3(c.point, CS, base + 64 * s + lane);
    } else {
      return ptr[s];
    }
  };

  // prepareContacts / calculateInternals, contact.go:225-229, :59-85
#pragma unroll
  for (int s = 0; s < SLOTS; ++s) {
    const int j = 64 * s + lane;
    key[s] = R(0);
    ids[s] = 0xFFFFFFFFu;
    nrm[s] = V3<R>{R(0), R(0), R(0)};
    if constexpr (!PT_HBM) ptr[s] = nrm[s];
    if (j < C && !do_pos) {  // velocity phase only: reload what prepare stored
      const int64_t g = base + j;
      const int c0 = c.body0[g], c1 = c.body1[g];
      const V3<R> nn = ld3(c.normal, CS, g);
      const V3<R> p = ld3(c.point, CS, g);
      s_bs[j] = basis_scale(nn);
      if constexpr (KIND == KIND_MAT) s_rest[j] = c.rest[g];
      ids[s] = uint32_t(c0 - b0w) | (uint32_t(c1 >= 0 ? c1 - b0w : int(kNoBody)) << 16);
      nrm[s] = nn;
      if constexpr (!PT_HBM) ptr[s] = p;
    } else if (j < C) {
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
      s_bs[j] = sc;
      const V3<R> p = ld3(c.point, CS, g);
      const V3<R> r0 = sub(p, lds3(s_pos0, l0));
      V3<R> cv = local_velocity_lds(b, c0, lds3(s_lad, l0), r0, m);
      V3<R> r1{R(0), R(0), R(0)};
      if (c1 >= 0) {
        r1 = sub(p, lds3(s_pos0, l1));
        cv = sub(cv, local_velocity_lds(b, c1, lds3(s_lad, l1), r1, m));
      }
      if constexpr (KIND == KIND_MAT) s_rest[j] = c.rest[g];
      const R ddv = desired_dv_lds(s_flags, s_lad, l0, l1, nn, cv.x, lds_restitution<R, KIND>(s_rest, j));
#pragma unroll
      for (int k = 0; k < 9; ++k) c.c2w[k * CS + g] = m[k];
      st3(c.rel, CS, g, r0);
      st3(c.rel + 3 * CS, CS, g, r1);
      st3(c.cvel, CS, g, cv);
      c.ddv[g] = ddv;
      key[s] = c.pen[g];
      ids[s] = uint32_t(l0) | (uint32_t(l1 >= 0 ? l1 : int(kNoBody)) << 16);
      nrm[s] = nn;
      if constexpr (!PT_HBM) ptr[s] = p;
    }
  }
  __syncthreads();

  // adjustPositions, contact.go:233-283
  int64_t used = 0;
  while (do_pos && used < maxIter) {
    R best = R(kPositionEpsilon);
    int bi = INT_MAX;
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
      if (64 * s + lane < C && key[s] > best) {
        best = key[s];
        bi = 64 * s + lane;
      }
    }
    wave_argmax_dpp(best, bi);
    if (bi == INT_MAX) break;
    if (lane == (bi & 63)) {
#pragma unroll
      for (int s = 0; s < SLOTS; ++s) {
        if (s == (bi >> 6)) {
          s_bc[0] = nrm[s].x; s_bc[1] = nrm[s].y; s_bc[2] = nrm[s].z;
          if constexpr (!PT_HBM) {  // the wide classes read the point from HBM below
            const V3<R> q = get_pt(s);
            s_bc[3] = q.x; s_bc[4] = q.y; s_bc[5] = q.z;
          }
          s_bcid = ids[s];
        }
      }
    }
    __syncthreads();
    const V3<R> n{s_bc[0], s_bc[1], s_bc[2]};
    const V3<R> p = PT_HBM ? ld3(c.point, CS, base + bi) : V3<R>{s_bc[3], s_bc[4], s_bc[5]};
    const uint32_t kid = s_bcid;
    const int kl[2] = {id_lo(kid), id_hi(kid)};
    const int kb[2] = {b0w + kl[0], kl[1] >= 0 ? b0w + kl[1] : -1};
    V3<R> rel[2] = {sub(p, lds3(s_pos0, kl[0])), V3<R>{R(0), R(0), R(0)}};
    if (kl[1] >= 0) rel[1] = sub(p, lds3(s_pos0, kl[1]));
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
    apply_position_change(a, kb, aw, n, rel, best, lin, ang, lane);
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
      if (64 * s + lane >= C) continue;
      const int jb[2] = {id_lo(ids[s]), id_hi(ids[s])};
      bool any = false;
#pragma unroll
      for (int bb = 0; bb < 2; ++bb)
#pragma unroll
        for (int d = 0; d < 2; ++d) any |= (jb[bb] >= 0 && jb[bb] == kl[d]);
      if (!any) continue;
      R pen = key[s];
#pragma unroll
      for (int bb = 0; bb < 2; ++bb) {
        if (jb[bb] < 0) continue;
        const V3<R> rj = sub(get_pt(s), lds3(s_pos0, jb[bb]));
#pragma unroll
        for (int d = 0; d < 2; ++d) {
          if (jb[bb] == kl[d]) {
            V3<R> dp = add(cross(ang[d], rj), lin[d]);
            R sign = R(1);
            if (bb == 0) sign = R(-1);
            pen += dot(dp, nrm[s]) * sign;
          }
        }
      }
      key[s] = pen;
    }
    __syncthreads();
    ++used;
  }
  const int64_t pos_iters = used;

  // adjustVelocities, contact.go:390-445.  The contact velocities stay in
  // HBM (c.cvel, as prepare stored them): an iteration touches only the
  // resolved contact's and the one or two slots that share its bodies, and
  // keeping them out of VGPRs keeps the wide classes from spilling.
#pragma unroll
  for (int s = 0; s < SLOTS; ++s) {
    const int j = 64 * s + lane;
    if (do_pos && j < C) c.pen[base + j] = key[s];
    if (do_vel && j < C) key[s] = c.ddv[base + j];  // selection keys as prepare stored them
  }
  used = 0;
  if (do_vel) {
    bool polishing = KIND == KIND_POLISH && a.world_mask[w] == 0;
    double ke_before = 0.0, world_mass = 0.0;
    if constexpr (KIND == KIND_POLISH) {
      if (polishing) {
      // take the predicted changes: snapshot, then add them to the bodies
      const float* dl = a.delta;
      for (int j = lane; j < C; j += 64) st3(a.cvel0, CS, base + j, ld3(c.cvel, CS, base + j));
      double ke = 0.0, mass = 0.0;
      for (int i = lane; i < n; i += 64) {
        const int g = b0w + i;
        const V3<R> v = ld3(b.vel, S, g), r = ld3(b.rot, S, g);
        R iw[9];
        ld9(b.iit_world, S, g, iw);
        double mi;
        ke += kinetic_energy(v, r, b.inv_mass[g], iw, mi);
        mass += mi;
        st3(a.vel0, S, g, v);
        st3(a.vel0 + 3 * S, S, g, r);
        a.vel0[6 * S + g] = b.motion[g];
        const V3<R> dv{R(dl[g]), R(dl[S + g]), R(dl[2 * S + g])};
        const V3<R> dw{R(dl[3 * S + g]), R(dl[4 * S + g]), R(dl[5 * S + g])};
        if (dv.x != R(0) || dv.y != R(0) || dv.z != R(0) || dw.x != R(0) || dw.y != R(0) || dw.z != R(0)) {
          st3(b.vel, S, g, add(v, dv));
          st3(b.rot, S, g, add(r, dw));
        }
      }
      // contact velocities after those changes: the update of
      // contact.go:424-441 applied to each body's summed change
#pragma unroll
      for (int s = 0; s < SLOTS; ++s) {
        if (64 * s + lane >= C) continue;
        const int jb[2] = {id_lo(ids[s]), id_hi(ids[s])};
        R mj[9];
        basis_from_scale(nrm[s], s_bs[64 * s + lane], mj);
        V3<R> cvj = ld3(c.cvel, CS, base + 64 * s + lane);
        bool moved = false;
#pragma unroll
        for (int bb = 0; bb < 2; ++bb) {
          if (jb[bb] < 0) continue;
          const int g = b0w + jb[bb];
          const V3<R> dv{R(dl[g]), R(dl[S + g]), R(dl[2 * S + g])};
          const V3<R> dw{R(dl[3 * S + g]), R(dl[4 * S + g]), R(dl[5 * S + g])};
          if (dv.x == R(0) && dv.y == R(0) && dv.z == R(0) && dw.x == R(0) && dw.y == R(0) && dw.z == R(0)) continue;
          moved = true;
          const V3<R> rj = sub(get_pt(s), lds3(s_pos0, jb[bb]));
          R sign = R(1);
          if (bb == 1) sign = R(-1);
          cvj = add(cvj, mul(m3tv(mj, add(cross(dw, rj), dv)), sign));
        }
        // a contact neither of whose bodies changed keeps the key prepare
        // gave it, as in the plain loop
        if (moved) {
          st3(c.cvel, CS, base + 64 * s + lane, cvj);
          key[s] = desired_dv_lds(s_flags, s_lad, jb[0], jb[1], nrm[s], cvj.x, lds_restitution<R, KIND>(s_rest, 64 * s + lane));
        }
      }
      __syncthreads();
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
      R best = eps;
      int bi = INT_MAX;
#pragma unroll
      for (int s = 0; s < SLOTS; ++s) {
        if (64 * s + lane < C && key[s] > best) {
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
              ke += kinetic_energy(ld3(b.vel, S, g), ld3(b.rot, S, g), b.inv_mass[g], iw, mass);
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
        for (int i = lane; i < n; i += 64) {
          const int g = b0w + i;
          st3(b.vel, S, g, ld3(a.vel0, S, g));
          st3(b.rot, S, g, ld3(a.vel0 + 3 * S, S, g));
          b.motion[g] = a.vel0[6 * S + g];
          s_flags[i] = b.flags[g];
        }
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) {
          const int j = 64 * s + lane;
          if (j < C) {
            key[s] = c.ddv[base + j];
            st3(c.cvel, CS, base + j, ld3(a.cvel0, CS, base + j));
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
      if (lane == (bi & 63)) {
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) {
          if (s == (bi >> 6)) {
            s_bc[0] = nrm[s].x; s_bc[1] = nrm[s].y; s_bc[2] = nrm[s].z;
            if constexpr (!PT_HBM) {  // the wide classes read the point from HBM below
              const V3<R> q = get_pt(s);
              s_bc[3] = q.x; s_bc[4] = q.y; s_bc[5] = q.z;
            }
            s_bcid = ids[s];
          }
        }
      }
      __syncthreads();
      const V3<R> n{s_bc[0], s_bc[1], s_bc[2]};
      const V3<R> p = PT_HBM ? ld3(c.point, CS, base + bi) : V3<R>{s_bc[3], s_bc[4], s_bc[5]};
      const V3<R> cvk = ld3(c.cvel, CS, base + bi);
      const uint32_t kid = s_bcid;
      const int kl[2] = {id_lo(kid), id_hi(kid)};
      const int kb[2] = {b0w + kl[0], kl[1] >= 0 ? b0w + kl[1] : -1};
      V3<R> rel[2] = {sub(p, lds3(s_pos0, kl[0])), V3<R>{R(0), R(0), R(0)}};
      if (kl[1] >= 0) rel[1] = sub(p, lds3(s_pos0, kl[1]));
      bool aw0 = (s_flags[kl[0]] & F_AWAKE) != 0;
      bool aw1 = kl[1] >= 0 && (s_flags[kl[1]] & F_AWAKE) != 0;
      if (kl[1] >= 0 && (aw0 || aw1) && !(aw0 && aw1)) {
        const int t = aw0 ? kl[1] : kl[0];
        if (lane == 0) {
          s_flags[t] = uint8_t(s_flags[t] | F_AWAKE);
          b.motion[b0w + t] = R(kSleepEpsilon * 2.0);
        }
      }
      __syncthreads();
      R m[9];
      basis_from_scale(n, s_bs[bi], m);
      V3<R> vc[2], rc[2];
      apply_velocity_change<R, KIND>(a, kb, rel, m, best, cvk, vc, rc, lane, friction_of<R, KIND>(c, base + bi));
      if constexpr (KIND == KIND_LABEL)
        if (lane == 0) add_label(a, base + bi, vc, rc);
#pragma unroll
      for (int s = 0; s < SLOTS; ++s) {
        if (64 * s + lane >= C) continue;
        const int jb[2] = {id_lo(ids[s]), id_hi(ids[s])};
        bool any = false;
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
          for (int d = 0; d < 2; ++d) any |= (jb[bb] >= 0 && jb[bb] == kl[d]);
        if (!any) continue;
        R mj[9];
        basis_from_scale(nrm[s], s_bs[64 * s + lane], mj);
        V3<R> cvj = ld3(c.cvel, CS, base + 64 * s + lane);
#pragma unroll
        for (int bb = 0; bb < 2; ++bb) {
          if (jb[bb] < 0) continue;
          const V3<R> rj = sub(get_pt(s), lds3(s_pos0, jb[bb]));
#pragma unroll
          for (int d = 0; d < 2; ++d) {
            if (jb[bb] == kl[d]) {
              V3<R> dv = add(cross(rc[d], rj), vc[d]);
              R sign = R(1);
              if (bb == 1) sign = R(-1);
              cvj = add(cvj, mul(m3tv(mj, dv), sign));
            }
          }
        }
        st3(c.cvel, CS, base + 64 * s + lane, cvj);
        key[s] = desired_dv_lds(s_flags, s_lad, jb[0], jb[1], nrm[s], cvj.x, lds_restitution<R, KIND>(s_rest, 64 * s + lane));
      }
      __syncthreads();
      ++it;
      ++used;
    }
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
      const int j = 64 * s + lane;
      if (j < C) c.ddv[base + j] = key[s];
    }
  }
  for (int i = lane; i < n; i += 64) b.flags[b0w + i] = s_flags[i];
  if (lane == 0 && a.world_iters) {
    if (do_pos) a.world_iters[2 * w] = pos_iters;
    if (do_vel) a.world_iters[2 * w + 1] = used;
  }
}

constexpr size_t cmax(size_t x, size_t y) { return x > y ? x : y; }
template <typename R, int KIND>
constexpr size_t resolve_lds_bytes() {
  return cmax(cmax(RegLds<R, 16, KIND>::bytes, RegLds<R, 12, KIND>::bytes),
              cmax(RegLds<R, 8, KIND>::bytes, kKeyCap * (sizeof(R) + 2 * sizeof(int32_t))));
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
  if (!fits_registers(C, n)) {
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
  } else {
    resolve_reg<R, 16, KIND>(a, w, C, b0w, n, lds, lane);
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

template <typename R, int KIND>
void launch_kind(const RS<R>& a, hipStream_t st) {
  if constexpr (sizeof(R) == 8)
    k_resolve_f64<KIND><<<a.n_worlds, 64, 0, st>>>(a);
  else
    k_resolve_f32<KIND><<<a.n_worlds, 64, 0, st>>>(a);
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
  return f64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(worlds_per_cu, k_resolve_f64<KIND_PLAIN>, 64, 0)
             : hipOccupancyMaxActiveBlocksPerMultiprocessor(worlds_per_cu, k_resolve_f32<KIND_PLAIN>, 64, 0);
}

template hipError_t launch_resolve<double>(const ResolveLaunch<double>&, hipStream_t);
template hipError_t launch_resolve<float>(const ResolveLaunch<float>&, hipStream_t);

}  // namespace cbz
