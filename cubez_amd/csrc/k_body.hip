// k_body.hip — K1: Integrate + CalculateDerivedData, one thread per body.
//
// HBM-bound: an awake body reads 29 values (pos 3, quat 4, vel 3, rot 3,
// acc 3, inv_mass 1, iit_body 9, motion 1, pow 2) plus its flag byte and
// writes 50 (pos, quat, vel, rot, last_acc, transform 12, iit_world 9,
// ctransform 12, motion) plus the flag byte.  A sleeping body reads its
// transform and writes the collider transform.
#include "cbz_devmath.h"
#include "cbz_launch.h"
#include "cbz_types.h"

namespace cbz {

template <typename R>
__device__ __forceinline__ V3<R> ld3(const R* a, int64_t s, int i) {
  return V3<R>{a[i], a[s + i], a[2 * s + i]};
}
template <typename R>
__device__ __forceinline__ void st3(R* a, int64_t s, int i, const V3<R>& v) {
  a[i] = v.x;
  a[s + i] = v.y;
  a[2 * s + i] = v.z;
}
template <typename R, int K>
__device__ __forceinline__ void ldk(const R* a, int64_t s, int i, R* o) {
#pragma unroll
  for (int c = 0; c < K; ++c) o[c] = a[c * s + i];
}
template <typename R, int K>
__device__ __forceinline__ void stk(R* a, int64_t s, int i, const R* o) {
#pragma unroll
  for (int c = 0; c < K; ++c) a[c * s + i] = o[c];
}

// Collider CalculateDerivedData, colliders.go:173-176 / :302-304:
// collider transform = body transform * Offset.
template <typename R>
__device__ __forceinline__ void collider_derived(const Bodies<R>& b, int i, uint8_t f, const R* T) {
  R C[12];
  if (f & F_OFFSET_ID) {
    m34mul_identity(T, C);
  } else {
    R O[12];
    ldk<R, 12>(b.offset, b.stride, i, O);
    m34mul(T, O, C);
  }
  stk<R, 12>(b.ctransform, b.stride, i, C);
}

// RigidBody.CalculateDerivedData (rigidbody.go:268-272) followed by the
// collider's, for every body: the set-up pass of examples/cubedrop.go:171-172.
template <typename R>
__global__ void __launch_bounds__(256) k_calculate_derived(Bodies<R> b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  const int64_t S = b.stride;
  const uint8_t f = b.flags[i];
  Q4<R> q{b.quat[i], b.quat[S + i], b.quat[2 * S + i], b.quat[3 * S + i]};
  V3<R> p = ld3(b.pos, S, i);
  q = q_normalize(q);
  R T[12], IB[9], IW[9];
  set_as_transform(p, q, T);
  ldk<R, 9>(b.iit_body, S, i, IB);
  transform_inertia(IB, T, IW);
  b.quat[i] = q.w;
  b.quat[S + i] = q.x;
  b.quat[2 * S + i] = q.y;
  b.quat[3 * S + i] = q.z;
  stk<R, 12>(b.transform, S, i, T);
  stk<R, 9>(b.iit_world, S, i, IW);
  collider_derived(b, i, f, T);
}

// RigidBody.Integrate, rigidbody.go:213-259, then the collider's
// CalculateDerivedData as in examples/cubedrop.go:31-33.  The three
// math.Pow values arrive precomputed (pow_lin, pow_ang per body; bias per
// step).  forceAccum and torqueAccum are always zero in the reference (no
// API adds to them), so lastFrameAccelleration = Acceleration + 0*invMass
// and the angular acceleration is iitWorld * 0.
template <typename R>
__global__ void __launch_bounds__(256) k_integrate(Bodies<R> b, R dt, R bias) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  uint8_t f = i < b.n ? b.flags[i] : uint8_t(F_STATIC);
  if (b.n_awake) {  // kAwakeSlots counters, so the waves' atomics do not queue on one address
    const uint64_t m = __ballot(!(f & F_STATIC) && (f & F_AWAKE));
    const int wv = int(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(b.n_awake + (wv & (kAwakeSlots - 1)), uint32_t(__popcll(m)));
  }
  if (i >= b.n) return;
  const int64_t S = b.stride;
  if (f & F_STATIC) return;
  R T[12];
  if (f & F_AWAKE) {
    const R im = b.inv_mass[i];
    V3<R> acc = ld3(b.acc, S, i);
    // lastFrameAccelleration.AddScaled(&forceAccum, inverseMass)  (:219-220)
    V3<R> la{acc.x + R(0) * im, acc.y + R(0) * im, acc.z + R(0) * im};
    V3<R> vel = ld3(b.vel, S, i);
    V3<R> rot = ld3(b.rot, S, i);
    // angularAcceleration = iitWorld * torqueAccum (:223).  Each component
    // is a signed zero for a finite tensor (which changes only a zero
    // Rotation component's sign) and NaN for a tensor with an infinite or
    // NaN entry, so the product is always evaluated.
    {
      R IWo[9];
      ldk<R, 9>(b.iit_world, S, i, IWo);
      V3<R> aa = m3v(IWo, V3<R>{R(0), R(0), R(0)});
      rot = add_scaled(rot, aa, dt);
    }
    vel = add_scaled(vel, la, dt);                 // :227
    vel = mul(vel, b.pow_lin[i]);                  // :233
    rot = mul(rot, b.pow_ang[i]);                  // :234
    V3<R> pos = add_scaled(ld3(b.pos, S, i), vel, dt);  // :238
    Q4<R> q{b.quat[i], b.quat[S + i], b.quat[2 * S + i], b.quat[3 * S + i]};
    q = q_add_scaled_vector(q, rot, dt);           // :241
    // CalculateDerivedData (:244 -> :268-272)
    q = q_normalize(q);
    set_as_transform(pos, q, T);
    R IB[9], IW[9];
    ldk<R, 9>(b.iit_body, S, i, IB);
    transform_inertia(IB, T, IW);
    R motion = b.motion[i];
    if (f & F_CAN_SLEEP) {                         // :248-258
      R cur = dot(vel, vel) + dot(rot, rot);
      motion = bias * motion + (R(1) - bias) * cur;
      if (motion < R(kSleepEpsilon)) {
        f = uint8_t(f & ~F_AWAKE);                 // SetAwake(false), :182-192
        vel = V3<R>{R(0), R(0), R(0)};
        rot = V3<R>{R(0), R(0), R(0)};
      } else if (motion > R(10 * kSleepEpsilon)) {
        motion = R(10 * kSleepEpsilon);
      }
      b.motion[i] = motion;
      b.flags[i] = f;
    }
    st3(b.pos, S, i, pos);
    b.quat[i] = q.w;
    b.quat[S + i] = q.x;
    b.quat[2 * S + i] = q.y;
    b.quat[3 * S + i] = q.z;
    st3(b.vel, S, i, vel);
    st3(b.rot, S, i, rot);
    st3(b.last_acc, S, i, la);
    stk<R, 12>(b.transform, S, i, T);
    stk<R, 9>(b.iit_world, S, i, IW);
  } else {
    ldk<R, 12>(b.transform, S, i, T);
  }
  collider_derived(b, i, f, T);
}

template <typename R>
hipError_t launch_calculate_derived(const Bodies<R>& b, hipStream_t st) {
  if (b.n <= 0) return hipSuccess;
  k_calculate_derived<R><<<(b.n + 255) / 256, 256, 0, st>>>(b);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_integrate(const Bodies<R>& b, R dt, R bias, hipStream_t st) {
  if (b.n <= 0) return hipSuccess;
  k_integrate<R><<<(b.n + 255) / 256, 256, 0, st>>>(b, dt, bias);
  return hipGetLastError();
}

template hipError_t launch_calculate_derived<double>(const Bodies<double>&, hipStream_t);
template hipError_t launch_calculate_derived<float>(const Bodies<float>&, hipStream_t);
template hipError_t launch_integrate<double>(const Bodies<double>&, double, double, hipStream_t);
template hipError_t launch_integrate<float>(const Bodies<float>&, float, float, hipStream_t);

}  // namespace cbz
