// cbz_devmath.h — device math for the cubez step, templated on the
// arithmetic type R (double for the exact path, float for the FP32 variant).
//
// Every expression keeps the Go reference's evaluation order: Go evaluates
// a*b + c*d + e*f as ((a*b) + (c*d)) + (e*f) and never fuses a multiply-add
// on amd64; all sources are built with -ffp-contract=off so the device code
// does the same.  Citations are to the reference's math package.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cbz {

template <typename R> struct V3 { R x, y, z; };
template <typename R> struct Q4 { R w, x, y, z; };
template <typename R> struct M3 { R m[9]; };    // column-major
template <typename R> struct M34 { R m[12]; };  // column-major, [9..11] = translation

template <typename R> struct Lim;
template <> struct Lim<double> {
  static constexpr double eps = 1e-7;                       // math/math.go:27
  static constexpr double min_normal = 1.1754943508222875e-38;  // math/math.go:30
  static constexpr double max_value = 1.7976931348623157e308;   // math/math.go:32
};
template <> struct Lim<float> {
  static constexpr float eps = 1e-7f;
  static constexpr float min_normal = 1.17549435e-38f;
  static constexpr float max_value = 3.40282347e38f;
};

__device__ __forceinline__ double dabs(double a) { return __builtin_fabs(a); }
__device__ __forceinline__ float dabs(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ double dsqrt(double a) { return __builtin_sqrt(a); }
__device__ __forceinline__ float dsqrt(float a) { return __builtin_sqrtf(a); }

// RealEqual, math/math.go:64-78
template <typename R> __device__ __forceinline__ bool real_equal(R a, R b) {
  if (a == b) return true;
  R diff = dabs(a - b);
  if (a * b == R(0) || diff < Lim<R>::min_normal) return diff < Lim<R>::eps * Lim<R>::eps;
  return diff / (dabs(a) + dabs(b)) < Lim<R>::eps;
}

template <typename R> __device__ __forceinline__ V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
// Cross, math/vector.go:33-39
template <typename R> __device__ __forceinline__ V3<R> cross(const V3<R>& a, const V3<R>& b) {
  return V3<R>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Dot, math/vector.go:42-44
template <typename R> __device__ __forceinline__ R dot(const V3<R>& a, const V3<R>& b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
// SquareMagnitude / Magnitude, math/vector.go:47-54
template <typename R> __device__ __forceinline__ R sqmag(const V3<R>& a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <typename R> __device__ __forceinline__ R mag(const V3<R>& a) { return dsqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// Add / Sub / MulWith / AddScaled, math/vector.go
template <typename R> __device__ __forceinline__ V3<R> add(const V3<R>& a, const V3<R>& b) { return V3<R>{a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> __device__ __forceinline__ V3<R> sub(const V3<R>& a, const V3<R>& b) { return V3<R>{a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> __device__ __forceinline__ V3<R> mul(const V3<R>& a, R s) { return V3<R>{a.x * s, a.y * s, a.z * s}; }
template <typename R> __device__ __forceinline__ V3<R> add_scaled(const V3<R>& a, const V3<R>& b, R s) {
  return V3<R>{a.x + b.x * s, a.y + b.y * s, a.z + b.z * s};
}
// Normalize, math/vector.go:64-72
template <typename R> __device__ __forceinline__ V3<R> normalize(V3<R> a) {
  R m = mag(a);
  if (!real_equal(m, R(0))) {
    R l = R(1) / m;
    a.x *= l; a.y *= l; a.z *= l;
  }
  return a;
}

// Quat.Mul, math/quaternion.go:46-53 (q = a * b)
template <typename R> __device__ __forceinline__ Q4<R> qmul(const Q4<R>& a, const Q4<R>& b) {
  Q4<R> r;
  r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
  r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
  r.y = a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z;
  r.z = a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x;
  return r;
}
// Quat.AddScaledVector, math/quaternion.go:20-33
template <typename R> __device__ __forceinline__ Q4<R> q_add_scaled_vector(const Q4<R>& q, const V3<R>& v, R s) {
  Q4<R> t{R(0), v.x * s, v.y * s, v.z * s};
  t = qmul(t, q);
  return Q4<R>{q.w + t.w * R(0.5), q.x + t.x * R(0.5), q.y + t.y * R(0.5), q.z + t.z * R(0.5)};
}
// Quat.Normalize, math/quaternion.go:78-99
template <typename R> __device__ __forceinline__ Q4<R> q_normalize(Q4<R> q) {
  R length = dsqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  if (real_equal(R(1), length)) return q;
  if (length == R(0)) return Q4<R>{R(1), R(0), R(0), R(0)};
  if (length == R(__builtin_inf())) length = Lim<R>::max_value;
  R inv = R(1) / length;
  return Q4<R>{q.w * inv, q.x * inv, q.y * inv, q.z * inv};
}

// Matrix3.MulVector3, math/matrix.go:80-86
template <typename R> __device__ __forceinline__ V3<R> m3v(const R* m, const V3<R>& v) {
  return V3<R>{m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
               m[2] * v.x + m[5] * v.y + m[8] * v.z};
}
// Matrix3.TransformTranspose, math/matrix.go:157-163
template <typename R> __device__ __forceinline__ V3<R> m3tv(const R* m, const V3<R>& v) {
  return V3<R>{v.x * m[0] + v.y * m[1] + v.z * m[2], v.x * m[3] + v.y * m[4] + v.z * m[5],
               v.x * m[6] + v.y * m[7] + v.z * m[8]};
}
// Matrix3.MulMatrix3, math/matrix.go:89-101  (o = a * b)
template <typename R> __device__ __forceinline__ void m3mul(const R* a, const R* b, R* o) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    o[3 * c + 0] = a[0] * b[3 * c + 0] + a[3] * b[3 * c + 1] + a[6] * b[3 * c + 2];
    o[3 * c + 1] = a[1] * b[3 * c + 0] + a[4] * b[3 * c + 1] + a[7] * b[3 * c + 2];
    o[3 * c + 2] = a[2] * b[3 * c + 0] + a[5] * b[3 * c + 1] + a[8] * b[3 * c + 2];
  }
}
// Matrix3.Determinant, math/matrix.go:127-129
template <typename R> __device__ __forceinline__ R m3det(const R* m) {
  return m[0] * m[4] * m[8] + m[3] * m[7] * m[2] + m[6] * m[1] * m[5] - m[6] * m[4] * m[2] - m[3] * m[1] * m[8] -
         m[0] * m[7] * m[5];
}
// Matrix3.Invert, math/matrix.go:133-153
template <typename R> __device__ __forceinline__ void m3inv(const R* m, R* o) {
  R det = m3det(m);
  if (real_equal(det, R(0))) {
#pragma unroll
    for (int k = 0; k < 9; ++k) o[k] = R(0);
    return;
  }
  o[0] = m[4] * m[8] - m[5] * m[7];
  o[1] = m[2] * m[7] - m[1] * m[8];
  o[2] = m[1] * m[5] - m[2] * m[4];
  o[3] = m[5] * m[6] - m[3] * m[8];
  o[4] = m[0] * m[8] - m[2] * m[6];
  o[5] = m[2] * m[3] - m[0] * m[5];
  o[6] = m[3] * m[7] - m[4] * m[6];
  o[7] = m[1] * m[6] - m[0] * m[7];
  o[8] = m[0] * m[4] - m[1] * m[3];
  R s = R(1) / det;
#pragma unroll
  for (int k = 0; k < 9; ++k) o[k] *= s;
}
// Matrix3x4.SetAsTransform, math/matrix.go:167-185
template <typename R> __device__ __forceinline__ void set_as_transform(const V3<R>& p, const Q4<R>& q, R* m) {
  const R w = q.w, x = q.x, y = q.y, z = q.z;
  m[0] = R(1) - R(2) * y * y - R(2) * z * z;
  m[1] = R(2) * x * y + R(2) * w * z;
  m[2] = R(2) * x * z - R(2) * w * y;
  m[3] = R(2) * x * y - R(2) * w * z;
  m[4] = R(1) - R(2) * x * x - R(2) * z * z;
  m[5] = R(2) * y * z + R(2) * w * x;
  m[6] = R(2) * x * z + R(2) * w * y;
  m[7] = R(2) * y * z - R(2) * w * x;
  m[8] = R(1) - R(2) * x * x - R(2) * y * y;
  m[9] = p.x;
  m[10] = p.y;
  m[11] = p.z;
}
// Matrix3x4.MulVector3, math/matrix.go:188-194
template <typename R> __device__ __forceinline__ V3<R> m34v(const R* m, const V3<R>& v) {
  return V3<R>{v.x * m[0] + v.y * m[3] + v.z * m[6] + m[9], v.x * m[1] + v.y * m[4] + v.z * m[7] + m[10],
               v.x * m[2] + v.y * m[5] + v.z * m[8] + m[11]};
}
// Matrix3x4.MulMatrix3x4, math/matrix.go:198-217 (o = m * b)
template <typename R> __device__ __forceinline__ void m34mul(const R* m, const R* b, R* o) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    o[3 * c + 0] = m[0] * b[3 * c + 0] + m[3] * b[3 * c + 1] + m[6] * b[3 * c + 2];
    o[3 * c + 1] = m[1] * b[3 * c + 0] + m[4] * b[3 * c + 1] + m[7] * b[3 * c + 2];
    o[3 * c + 2] = m[2] * b[3 * c + 0] + m[5] * b[3 * c + 1] + m[8] * b[3 * c + 2];
  }
  o[9] = m[0] * b[9] + m[3] * b[10] + m[6] * b[11] + m[9];
  o[10] = m[1] * b[9] + m[4] * b[10] + m[7] * b[11] + m[10];
  o[11] = m[2] * b[9] + m[5] * b[10] + m[8] * b[11] + m[11];
}
// Matrix3x4.MulMatrix3x4 with the identity as right operand: the same
// products and sums as m34mul against the constants 1 and 0 (kept so that
// signed zeros and NaNs behave as in the reference).
template <typename R> __device__ __forceinline__ void m34mul_identity(const R* m, R* o) {
  const R one = R(1), zero = R(0);
  const R b[12] = {one, zero, zero, zero, one, zero, zero, zero, one, zero, zero, zero};
  m34mul(m, b, o);
}
// Matrix3x4.TransformInverse, math/matrix.go:222-233
template <typename R> __device__ __forceinline__ V3<R> m34_tinv(const R* m, const V3<R>& v) {
  V3<R> t{v.x - m[9], v.y - m[10], v.z - m[11]};
  return V3<R>{t.x * m[0] + t.y * m[1] + t.z * m[2], t.x * m[3] + t.y * m[4] + t.z * m[5],
               t.x * m[6] + t.y * m[7] + t.z * m[8]};
}
// Matrix3x4.GetAxis, math/matrix.go:235-247
template <typename R> __device__ __forceinline__ V3<R> axis(const R* m, int col) {
  int i = col > 3 ? 3 : col;
  return V3<R>{m[3 * i + 0], m[3 * i + 1], m[3 * i + 2]};
}

// transformInertiaTensor, rigidbody.go:275-299
template <typename R> __device__ __forceinline__ void transform_inertia(const R* ib, const R* r, R* iw) {
  R t4 = r[0] * ib[0] + r[3] * ib[1] + r[6] * ib[2];
  R t9 = r[0] * ib[3] + r[3] * ib[4] + r[6] * ib[5];
  R t14 = r[0] * ib[6] + r[3] * ib[7] + r[6] * ib[8];
  R t28 = r[1] * ib[0] + r[4] * ib[1] + r[7] * ib[2];
  R t33 = r[1] * ib[3] + r[4] * ib[4] + r[7] * ib[5];
  R t38 = r[1] * ib[6] + r[4] * ib[7] + r[7] * ib[8];
  R t52 = r[2] * ib[0] + r[5] * ib[1] + r[8] * ib[2];
  R t57 = r[2] * ib[3] + r[5] * ib[4] + r[8] * ib[5];
  R t62 = r[2] * ib[6] + r[5] * ib[7] + r[8] * ib[8];
  iw[0] = t4 * r[0] + t9 * r[3] + t14 * r[6];
  iw[3] = t4 * r[1] + t9 * r[4] + t14 * r[7];
  iw[6] = t4 * r[2] + t9 * r[5] + t14 * r[8];
  iw[1] = t28 * r[0] + t33 * r[3] + t38 * r[6];
  iw[4] = t28 * r[1] + t33 * r[4] + t38 * r[7];
  iw[7] = t28 * r[2] + t33 * r[5] + t38 * r[8];
  iw[2] = t52 * r[0] + t57 * r[3] + t62 * r[6];
  iw[5] = t52 * r[1] + t57 * r[4] + t62 * r[7];
  iw[8] = t52 * r[2] + t57 * r[5] + t62 * r[8];
}

}  // namespace cbz
