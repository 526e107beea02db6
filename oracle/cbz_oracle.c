/*
 * cbz_oracle.c — CPU FP64 restatement of the cubez simulation step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product never does.
 *
 * Each function restates one Go function of the reference (cited as
 * file:line) with the same statements, branches and floating-point
 * evaluation order.  Go evaluates a+b+c as (a+b)+c and a*b*c as (a*b)*c,
 * and on amd64 (GOAMD64=v1) never fuses a multiply-add; this file is built
 * with -O2 -ffp-contract=off -fno-fast-math to match.
 *
 * Objects mirror the Go ones: RigidBody, colliders that point at a body, and
 * contacts that point at bodies, so the pointer-identity tests of the
 * resolver (contact.go:264, :424) are restated literally.
 *
 * math.Pow (rigidbody.go:233-234, 250) is not evaluated here: the three
 * values are inputs (pow_lin/pow_ang per body, sleep_bias per step), so this
 * oracle and the product consume identical bits for them.
 */
#include "cbz_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double Real;
typedef struct { Real v[3]; } Vector3;
typedef struct { Real v[4]; } Quat; /* w,x,y,z  (math/math.go:58-59) */
typedef struct { Real m[9]; } Matrix3; /* column-major (math/math.go:41-42) */
typedef struct { Real m[12]; } Matrix3x4; /* column-major (math/math.go:44-47) */

/* math/math.go:27-33 */
static const Real Epsilon = 1e-7;
static const Real MinNormal = 1.1754943508222875e-38;
#define MaxValue DBL_MAX

/* ===================================================================== */
/* L0 math                                                                */
/* ===================================================================== */

/* RealEqual, math/math.go:64-78 */
static int RealEqual(Real a, Real b) {
    if (a == b) return 1;
    Real diff = fabs(a - b);
    if (a * b == 0 || diff < MinNormal) return diff < Epsilon * Epsilon;
    return diff / (fabs(a) + fabs(b)) < Epsilon;
}

/* math/vector.go:7-30, 57-61, 75-86 */
static void v3_add(Vector3 *v, const Vector3 *v2) {
    v->v[0] += v2->v[0];
    v->v[1] += v2->v[1];
    v->v[2] += v2->v[2];
}
static void v3_add_scaled(Vector3 *v, const Vector3 *v2, Real scale) {
    v->v[0] += v2->v[0] * scale;
    v->v[1] += v2->v[1] * scale;
    v->v[2] += v2->v[2] * scale;
}
static void v3_clear(Vector3 *v) { v->v[0] = v->v[1] = v->v[2] = 0.0; }
static void v3_component_product(Vector3 *v, const Vector3 *v2) {
    v->v[0] *= v2->v[0];
    v->v[1] *= v2->v[1];
    v->v[2] *= v2->v[2];
}
/* Cross, math/vector.go:33-39 */
static Vector3 v3_cross(const Vector3 *v, const Vector3 *v2) {
    Vector3 r;
    r.v[0] = v->v[1] * v2->v[2] - v->v[2] * v2->v[1];
    r.v[1] = v->v[2] * v2->v[0] - v->v[0] * v2->v[2];
    r.v[2] = v->v[0] * v2->v[1] - v->v[1] * v2->v[0];
    return r;
}
/* Dot / Magnitude / SquareMagnitude, math/vector.go:42-54 */
static Real v3_dot(const Vector3 *v, const Vector3 *v2) {
    return v->v[0] * v2->v[0] + v->v[1] * v2->v[1] + v->v[2] * v2->v[2];
}
static Real v3_magnitude(const Vector3 *v) {
    return sqrt(v->v[0] * v->v[0] + v->v[1] * v->v[1] + v->v[2] * v->v[2]);
}
static Real v3_square_magnitude(const Vector3 *v) {
    return v->v[0] * v->v[0] + v->v[1] * v->v[1] + v->v[2] * v->v[2];
}
static void v3_mul_with(Vector3 *v, Real r) {
    v->v[0] *= r;
    v->v[1] *= r;
    v->v[2] *= r;
}
/* Normalize, math/vector.go:64-72 */
static void v3_normalize(Vector3 *v) {
    Real m = v3_magnitude(v);
    if (!RealEqual(m, 0.0)) {
        Real l = 1.0 / m;
        v->v[0] *= l;
        v->v[1] *= l;
        v->v[2] *= l;
    }
}
static void v3_sub(Vector3 *v, const Vector3 *v2) {
    v->v[0] -= v2->v[0];
    v->v[1] -= v2->v[1];
    v->v[2] -= v2->v[2];
}
static Vector3 v3(Real x, Real y, Real z) {
    Vector3 r = {{x, y, z}};
    return r;
}

/* Quat.SetIdentity, math/quaternion.go:36-38 */
static void q_set_identity(Quat *q) {
    q->v[0] = 1.0;
    q->v[1] = q->v[2] = q->v[3] = 0.0;
}
/* Quat.Len, math/quaternion.go:41-43 */
static Real q_len(const Quat *q) {
    return sqrt(q->v[0] * q->v[0] + q->v[1] * q->v[1] + q->v[2] * q->v[2] + q->v[3] * q->v[3]);
}
/* Quat.Mul, math/quaternion.go:46-53 */
static void q_mul(Quat *q, const Quat *q2) {
    const Real *a = q->v, *b = q2->v;
    Real w = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    Real x = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    Real y = a[0] * b[2] + a[2] * b[0] + a[3] * b[1] - a[1] * b[3];
    Real z = a[0] * b[3] + a[3] * b[0] + a[1] * b[2] - a[2] * b[1];
    q->v[0] = w;
    q->v[1] = x;
    q->v[2] = y;
    q->v[3] = z;
}
/* Quat.AddScaledVector, math/quaternion.go:20-33 */
static void q_add_scaled_vector(Quat *q, const Vector3 *v, Real scale) {
    Quat temp;
    temp.v[0] = 0.0;
    temp.v[1] = v->v[0] * scale;
    temp.v[2] = v->v[1] * scale;
    temp.v[3] = v->v[2] * scale;
    q_mul(&temp, q);
    q->v[0] += temp.v[0] * 0.5;
    q->v[1] += temp.v[1] * 0.5;
    q->v[2] += temp.v[2] * 0.5;
    q->v[3] += temp.v[3] * 0.5;
}
/* Quat.Normalize, math/quaternion.go:78-99 */
static void q_normalize(Quat *q) {
    Real length = q_len(q);
    if (RealEqual(1.0, length)) return;
    if (length == 0.0) {
        q_set_identity(q);
        return;
    }
    if (length == INFINITY) length = MaxValue;
    Real invLength = 1.0 / length;
    q->v[0] *= invLength;
    q->v[1] *= invLength;
    q->v[2] *= invLength;
    q->v[3] *= invLength;
}
/* Quat.Rotate, math/quaternion.go:56-70 (used only by known-answer tests) */
static Vector3 q_rotate(const Quat *q, const Vector3 *v) {
    Vector3 qvec = v3(q->v[1], q->v[2], q->v[3]);
    Vector3 cross = v3_cross(&qvec, v);
    Vector3 result = *v;
    v3_mul_with(&qvec, 2.0);
    Vector3 c2 = v3_cross(&qvec, &cross);
    v3_add(&result, &c2);
    v3_mul_with(&cross, 2.0 * q->v[0]);
    v3_add(&result, &cross);
    return result;
}
/* QuatFromAxis, math/quaternion.go:9-16 (setup / known-answer tests only) */
static Quat q_from_axis(Real angle, Real x, Real y, Real z) {
    Real s = sin(angle / 2.0);
    Real c = cos(angle / 2.0);
    Quat r = {{c, x * s, y * s, z * s}};
    q_normalize(&r);
    return r;
}

/* Matrix3.SetIdentity, math/matrix.go:15-19 */
static void m3_set_identity(Matrix3 *m) {
    memset(m, 0, sizeof *m);
    m->m[0] = m->m[4] = m->m[8] = 1.0;
}
/* Matrix3x4.SetIdentity, math/matrix.go:22-26 */
static void m34_set_identity(Matrix3x4 *m) {
    memset(m, 0, sizeof *m);
    m->m[0] = m->m[4] = m->m[8] = 1.0;
}
/* Matrix3.Add, math/matrix.go:37-47 */
static void m3_add(Matrix3 *m, const Matrix3 *m2) {
    for (int i = 0; i < 9; i++) m->m[i] += m2->m[i];
}
/* Matrix3.SetComponents, math/matrix.go:52-56 */
static void m3_set_components(Matrix3 *m, const Vector3 *v1, const Vector3 *v2, const Vector3 *v3c) {
    m->m[0] = v1->v[0]; m->m[3] = v2->v[0]; m->m[6] = v3c->v[0];
    m->m[1] = v1->v[1]; m->m[4] = v2->v[1]; m->m[7] = v3c->v[1];
    m->m[2] = v1->v[2]; m->m[5] = v2->v[2]; m->m[8] = v3c->v[2];
}
/* Matrix3.SetInertiaTensorCoeffs, math/matrix.go:59-63 */
static void m3_set_inertia_coeffs(Matrix3 *m, Real ix, Real iy, Real iz, Real ixy, Real ixz, Real iyz) {
    m->m[0] = ix;   m->m[3] = -ixy; m->m[6] = -ixz;
    m->m[1] = -ixy; m->m[4] = iy;   m->m[7] = -iyz;
    m->m[2] = -ixz; m->m[5] = -iyz; m->m[8] = iz;
}
/* Matrix3.SetBlockInertiaTensor, math/matrix.go:68-77 */
static void m3_set_block_inertia(Matrix3 *m, const Vector3 *halfSize, Real mass) {
    Vector3 squares = *halfSize;
    v3_component_product(&squares, halfSize);
    m3_set_inertia_coeffs(m, 0.3 * mass * (squares.v[1] + squares.v[2]),
                          0.3 * mass * (squares.v[0] + squares.v[2]),
                          0.3 * mass * (squares.v[0] + squares.v[1]), 0.0, 0.0, 0.0);
}
/* Matrix3.MulVector3, math/matrix.go:80-86 */
static Vector3 m3_mul_vec(const Matrix3 *m, const Vector3 *v) {
    const Real *a = m->m;
    Vector3 r;
    r.v[0] = a[0] * v->v[0] + a[3] * v->v[1] + a[6] * v->v[2];
    r.v[1] = a[1] * v->v[0] + a[4] * v->v[1] + a[7] * v->v[2];
    r.v[2] = a[2] * v->v[0] + a[5] * v->v[1] + a[8] * v->v[2];
    return r;
}
/* Matrix3.MulMatrix3, math/matrix.go:89-101 */
static Matrix3 m3_mul(const Matrix3 *m1, const Matrix3 *m2) {
    const Real *a = m1->m, *b = m2->m;
    Matrix3 r;
    r.m[0] = a[0] * b[0] + a[3] * b[1] + a[6] * b[2];
    r.m[1] = a[1] * b[0] + a[4] * b[1] + a[7] * b[2];
    r.m[2] = a[2] * b[0] + a[5] * b[1] + a[8] * b[2];
    r.m[3] = a[0] * b[3] + a[3] * b[4] + a[6] * b[5];
    r.m[4] = a[1] * b[3] + a[4] * b[4] + a[7] * b[5];
    r.m[5] = a[2] * b[3] + a[5] * b[4] + a[8] * b[5];
    r.m[6] = a[0] * b[6] + a[3] * b[7] + a[6] * b[8];
    r.m[7] = a[1] * b[6] + a[4] * b[7] + a[7] * b[8];
    r.m[8] = a[2] * b[6] + a[5] * b[7] + a[8] * b[8];
    return r;
}
/* Matrix3.MulWith, math/matrix.go:104-114 */
static void m3_mul_with(Matrix3 *m, Real s) {
    for (int i = 0; i < 9; i++) m->m[i] *= s;
}
/* Matrix3.Transpose, math/matrix.go:117-123 */
static Matrix3 m3_transpose(const Matrix3 *m) {
    const Real *a = m->m;
    Matrix3 r = {{a[0], a[3], a[6], a[1], a[4], a[7], a[2], a[5], a[8]}};
    return r;
}
/* Matrix3.Determinant, math/matrix.go:127-129 */
static Real m3_det(const Matrix3 *m) {
    const Real *a = m->m;
    return a[0] * a[4] * a[8] + a[3] * a[7] * a[2] + a[6] * a[1] * a[5] - a[6] * a[4] * a[2] -
           a[3] * a[1] * a[8] - a[0] * a[7] * a[5];
}
/* Matrix3.Invert, math/matrix.go:133-153 */
static Matrix3 m3_invert(const Matrix3 *m) {
    Matrix3 r;
    Real det = m3_det(m);
    if (RealEqual(det, 0.0)) {
        memset(&r, 0, sizeof r);
        return r;
    }
    const Real *a = m->m;
    r.m[0] = a[4] * a[8] - a[5] * a[7];
    r.m[1] = a[2] * a[7] - a[1] * a[8];
    r.m[2] = a[1] * a[5] - a[2] * a[4];
    r.m[3] = a[5] * a[6] - a[3] * a[8];
    r.m[4] = a[0] * a[8] - a[2] * a[6];
    r.m[5] = a[2] * a[3] - a[0] * a[5];
    r.m[6] = a[3] * a[7] - a[4] * a[6];
    r.m[7] = a[1] * a[6] - a[0] * a[7];
    r.m[8] = a[0] * a[4] - a[1] * a[3];
    m3_mul_with(&r, 1 / det);
    return r;
}
/* Matrix3.TransformTranspose, math/matrix.go:157-163 */
static Vector3 m3_transform_transpose(const Matrix3 *m, const Vector3 *v) {
    const Real *a = m->m;
    Vector3 r;
    r.v[0] = v->v[0] * a[0] + v->v[1] * a[1] + v->v[2] * a[2];
    r.v[1] = v->v[0] * a[3] + v->v[1] * a[4] + v->v[2] * a[5];
    r.v[2] = v->v[0] * a[6] + v->v[1] * a[7] + v->v[2] * a[8];
    return r;
}
/* Matrix3x4.SetAsTransform, math/matrix.go:167-185 */
static void m34_set_as_transform(Matrix3x4 *m, const Vector3 *pos, const Quat *rot) {
    Real w = rot->v[0], x = rot->v[1], y = rot->v[2], z = rot->v[3];
    m->m[0] = 1 - 2 * y * y - 2 * z * z;
    m->m[1] = 2 * x * y + 2 * w * z;
    m->m[2] = 2 * x * z - 2 * w * y;
    m->m[3] = 2 * x * y - 2 * w * z;
    m->m[4] = 1 - 2 * x * x - 2 * z * z;
    m->m[5] = 2 * y * z + 2 * w * x;
    m->m[6] = 2 * x * z + 2 * w * y;
    m->m[7] = 2 * y * z - 2 * w * x;
    m->m[8] = 1 - 2 * x * x - 2 * y * y;
    m->m[9] = pos->v[0];
    m->m[10] = pos->v[1];
    m->m[11] = pos->v[2];
}
/* Matrix3x4.MulVector3, math/matrix.go:188-194 */
static Vector3 m34_mul_vec(const Matrix3x4 *m, const Vector3 *v) {
    const Real *a = m->m;
    Vector3 r;
    r.v[0] = v->v[0] * a[0] + v->v[1] * a[3] + v->v[2] * a[6] + a[9];
    r.v[1] = v->v[0] * a[1] + v->v[1] * a[4] + v->v[2] * a[7] + a[10];
    r.v[2] = v->v[0] * a[2] + v->v[1] * a[5] + v->v[2] * a[8] + a[11];
    return r;
}
/* Matrix3x4.MulMatrix3x4, math/matrix.go:198-217 */
static Matrix3x4 m34_mul(const Matrix3x4 *mm, const Matrix3x4 *oo) {
    const Real *m = mm->m, *o = oo->m;
    Matrix3x4 r;
    r.m[0] = m[0] * o[0] + m[3] * o[1] + m[6] * o[2];
    r.m[1] = m[1] * o[0] + m[4] * o[1] + m[7] * o[2];
    r.m[2] = m[2] * o[0] + m[5] * o[1] + m[8] * o[2];
    r.m[3] = m[0] * o[3] + m[3] * o[4] + m[6] * o[5];
    r.m[4] = m[1] * o[3] + m[4] * o[4] + m[7] * o[5];
    r.m[5] = m[2] * o[3] + m[5] * o[4] + m[8] * o[5];
    r.m[6] = m[0] * o[6] + m[3] * o[7] + m[6] * o[8];
    r.m[7] = m[1] * o[6] + m[4] * o[7] + m[7] * o[8];
    r.m[8] = m[2] * o[6] + m[5] * o[7] + m[8] * o[8];
    r.m[9] = m[0] * o[9] + m[3] * o[10] + m[6] * o[11] + m[9];
    r.m[10] = m[1] * o[9] + m[4] * o[10] + m[7] * o[11] + m[10];
    r.m[11] = m[2] * o[9] + m[5] * o[10] + m[8] * o[11] + m[11];
    return r;
}
/* Matrix3x4.TransformInverse, math/matrix.go:222-233 */
static Vector3 m34_transform_inverse(const Matrix3x4 *m, const Vector3 *v) {
    const Real *a = m->m;
    Vector3 tmp = *v;
    tmp.v[0] -= a[9];
    tmp.v[1] -= a[10];
    tmp.v[2] -= a[11];
    Vector3 r;
    r.v[0] = tmp.v[0] * a[0] + tmp.v[1] * a[1] + tmp.v[2] * a[2];
    r.v[1] = tmp.v[0] * a[3] + tmp.v[1] * a[4] + tmp.v[2] * a[5];
    r.v[2] = tmp.v[0] * a[6] + tmp.v[1] * a[7] + tmp.v[2] * a[8];
    return r;
}
/* Matrix3x4.GetAxis, math/matrix.go:235-247 */
static Vector3 m34_get_axis(const Matrix3x4 *m, int col) {
    int i = col > 3 ? 3 : col;
    return v3(m->m[i * 3 + 0], m->m[i * 3 + 1], m->m[i * 3 + 2]);
}

/* ===================================================================== */
/* L1 rigid body                                                          */
/* ===================================================================== */

/* RigidBody, rigidbody.go:23-101 (forceAccum/torqueAccum are always zero:
 * no API adds to them, rigidbody.go:206-209) */
typedef struct RigidBody {
    Real LinearDamping, AngularDamping;
    Vector3 Position;
    Quat Orientation;
    Vector3 Velocity, Acceleration, Rotation;
    Matrix3 InverseInertiaTensor;
    int IsAwake, CanSleep;
    Matrix3 inverseInertiaTensorWorld;
    Real inverseMass;
    Matrix3x4 transform;
    Vector3 forceAccum, torqueAccum, lastFrameAccelleration;
    Real motion;
    /* oracle bookkeeping */
    Real powLin, powAng;
    int isStatic;
    int index;
} RigidBody;

/* rigidbody.go:11-15 */
static const Real sleepEpsilon = 0.3;

/* SetAwake, rigidbody.go:182-192 */
static void rb_set_awake(RigidBody *b, int awake) {
    if (awake) {
        b->IsAwake = 1;
        b->motion = 0.6; /* sleepEpsilon * 2.0 as a Go constant expression */
    } else {
        b->IsAwake = 0;
        v3_clear(&b->Velocity);
        v3_clear(&b->Rotation);
    }
}

/* transformInertiaTensor, rigidbody.go:275-299 */
static void transform_inertia_tensor(Matrix3 *iitWorld, const Matrix3 *iitBody, const Matrix3x4 *rotmat) {
    const Real *r = rotmat->m, *ib = iitBody->m;
    Real *iw = iitWorld->m;
    Real t4 = r[0] * ib[0] + r[3] * ib[1] + r[6] * ib[2];
    Real t9 = r[0] * ib[3] + r[3] * ib[4] + r[6] * ib[5];
    Real t14 = r[0] * ib[6] + r[3] * ib[7] + r[6] * ib[8];
    Real t28 = r[1] * ib[0] + r[4] * ib[1] + r[7] * ib[2];
    Real t33 = r[1] * ib[3] + r[4] * ib[4] + r[7] * ib[5];
    Real t38 = r[1] * ib[6] + r[4] * ib[7] + r[7] * ib[8];
    Real t52 = r[2] * ib[0] + r[5] * ib[1] + r[8] * ib[2];
    Real t57 = r[2] * ib[3] + r[5] * ib[4] + r[8] * ib[5];
    Real t62 = r[2] * ib[6] + r[5] * ib[7] + r[8] * ib[8];
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

/* CalculateDerivedData, rigidbody.go:268-272 */
static void rb_calculate_derived(RigidBody *b) {
    q_normalize(&b->Orientation);
    m34_set_as_transform(&b->transform, &b->Position, &b->Orientation);
    transform_inertia_tensor(&b->inverseInertiaTensorWorld, &b->InverseInertiaTensor, &b->transform);
}

/* ClearAccumulators, rigidbody.go:206-209 */
static void rb_clear_accumulators(RigidBody *b) {
    v3_clear(&b->forceAccum);
    v3_clear(&b->torqueAccum);
}

/* Integrate, rigidbody.go:213-259.  powLin/powAng/bias are the values of the
 * three math.Pow calls (lines 233, 234, 250), supplied by the caller. */
static void rb_integrate(RigidBody *b, Real duration, Real bias) {
    if (b->IsAwake == 0) return;
    b->lastFrameAccelleration = b->Acceleration;
    v3_add_scaled(&b->lastFrameAccelleration, &b->forceAccum, b->inverseMass);
    Vector3 angularAcceleration = m3_mul_vec(&b->inverseInertiaTensorWorld, &b->torqueAccum);
    v3_add_scaled(&b->Velocity, &b->lastFrameAccelleration, duration);
    v3_add_scaled(&b->Rotation, &angularAcceleration, duration);
    v3_mul_with(&b->Velocity, b->powLin);
    v3_mul_with(&b->Rotation, b->powAng);
    v3_add_scaled(&b->Position, &b->Velocity, duration);
    q_add_scaled_vector(&b->Orientation, &b->Rotation, duration);
    rb_calculate_derived(b);
    rb_clear_accumulators(b);
    if (b->CanSleep) {
        Real currentMotion = v3_dot(&b->Velocity, &b->Velocity) + v3_dot(&b->Rotation, &b->Rotation);
        b->motion = bias * b->motion + (1.0 - bias) * currentMotion;
        if (b->motion < sleepEpsilon) {
            rb_set_awake(b, 0);
        } else if (b->motion > 3.0) { /* 10*sleepEpsilon, an exact Go constant */
            b->motion = 3.0;
        }
    }
}

/* ===================================================================== */
/* L2 colliders                                                           */
/* ===================================================================== */

/* CollisionCube / CollisionSphere, colliders.go:39-71 */
typedef struct Collider {
    int type;
    RigidBody *Body;
    Matrix3x4 Offset;
    Matrix3x4 transform;
    Vector3 HalfSize;
    Real Radius;
} Collider;

/* CollisionPlane, colliders.go:29-35 */
typedef struct CollisionPlane {
    Vector3 Normal;
    Real Offset;
} CollisionPlane;

/* Contact, contact.go:17-51 */
typedef struct Contact {
    RigidBody *Bodies[2];
    Real Friction, Restitution;
    Vector3 ContactPoint, ContactNormal;
    Real Penetration;
    Matrix3 contactToWorld;
    Vector3 relativeContactPosition[2];
    Vector3 contactVelocity;
    Real desiredDeltaVelocity;
} Contact;

typedef struct ContactList {
    Contact *c;
    int64_t n, cap;
    int overflow;
} ContactList;

static Contact *new_contact(ContactList *l) {
    if (l->n >= l->cap) {
        l->overflow = 1;
        l->n++;
        return NULL;
    }
    Contact *c = &l->c[l->n++];
    memset(c, 0, sizeof *c);
    return c;
}

/* Collider.CalculateDerivedData, colliders.go:173-176 (sphere), :302-304 (cube) */
static void collider_calculate_derived(Collider *c) {
    c->transform = m34_mul(&c->Body->transform, &c->Offset);
}

/* transformToAxis, colliders.go:762-770 */
static Real transform_to_axis(const Collider *cube, const Vector3 *axis) {
    Vector3 x = m34_get_axis(&cube->transform, 0);
    Vector3 y = m34_get_axis(&cube->transform, 1);
    Vector3 z = m34_get_axis(&cube->transform, 2);
    return cube->HalfSize.v[0] * fabs(v3_dot(axis, &x)) + cube->HalfSize.v[1] * fabs(v3_dot(axis, &y)) +
           cube->HalfSize.v[2] * fabs(v3_dot(axis, &z));
}

/* intersectCubeAndHalfSpace, colliders.go:750-760 */
static int intersect_cube_half_space(const Collider *cube, const CollisionPlane *plane) {
    Real projectedRadius = transform_to_axis(cube, &plane->Normal);
    Vector3 axis = m34_get_axis(&cube->transform, 3);
    Real cubeDistance = v3_dot(&plane->Normal, &axis) - projectedRadius;
    return cubeDistance <= plane->Offset;
}

/* CollisionSphere.CheckAgainstHalfSpace, colliders.go:180-207 */
static int sphere_vs_half_space(const Collider *s, const CollisionPlane *plane, ContactList *out) {
    Vector3 positionAxis = m34_get_axis(&s->transform, 3);
    Real distance = v3_dot(&plane->Normal, &positionAxis) - s->Radius;
    if ((distance <= plane->Offset) == 0) return 0;
    Contact *c = new_contact(out);
    if (!c) return 1;
    c->ContactPoint = plane->Normal;
    v3_mul_with(&c->ContactPoint, distance + s->Radius * -1.0);
    v3_add(&c->ContactPoint, &positionAxis);
    c->ContactNormal = plane->Normal;
    c->Penetration = -distance;
    c->Bodies[0] = s->Body;
    c->Bodies[1] = NULL;
    c->Friction = 0.9;
    c->Restitution = 0.1;
    return 1;
}

/* CollisionSphere.CheckAgainstSphere, colliders.go:216-254 */
static int sphere_vs_sphere(const Collider *s, const Collider *second, ContactList *out) {
    Vector3 positionOne = m34_get_axis(&s->transform, 3);
    Vector3 positionTwo = m34_get_axis(&second->transform, 3);
    Vector3 midline = positionOne;
    v3_sub(&midline, &positionTwo);
    Real size = v3_magnitude(&midline);
    if (size <= 0.0 || size >= s->Radius + second->Radius) return 0;
    Contact *c = new_contact(out);
    if (!c) return 1;
    c->ContactPoint = midline;
    v3_mul_with(&c->ContactPoint, 0.5);
    v3_add(&c->ContactPoint, &positionOne);
    c->ContactNormal = midline;
    v3_mul_with(&c->ContactNormal, 1.0 / size);
    c->Penetration = s->Radius + second->Radius - size;
    c->Bodies[0] = s->Body;
    c->Bodies[1] = second->Body;
    c->Friction = 0.9;
    c->Restitution = 0.1;
    return 1;
}

/* CollisionCube.CheckAgainstHalfSpace, colliders.go:308-366 */
static int cube_vs_half_space(const Collider *cube, const CollisionPlane *plane, ContactList *out) {
    if (!intersect_cube_half_space(cube, plane)) return 0;
    static const Real mults[8][3] = {{1, 1, 1},  {-1, 1, 1},  {1, -1, 1},  {-1, -1, 1},
                                     {1, 1, -1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};
    int contactDetected = 0;
    for (int k = 0; k < 8; k++) {
        Vector3 v = v3(mults[k][0], mults[k][1], mults[k][2]);
        v3_component_product(&v, &cube->HalfSize);
        Vector3 vertexPos = m34_mul_vec(&cube->transform, &v);
        Real vertexDistance = v3_dot(&vertexPos, &plane->Normal);
        if (vertexDistance <= plane->Offset) {
            Contact *c = new_contact(out);
            contactDetected = 1;
            if (!c) continue;
            c->ContactPoint = plane->Normal;
            v3_mul_with(&c->ContactPoint, vertexDistance - plane->Offset);
            v3_add(&c->ContactPoint, &vertexPos);
            c->ContactNormal = plane->Normal;
            c->Penetration = plane->Offset - vertexDistance;
            c->Bodies[0] = cube->Body;
            c->Bodies[1] = NULL;
            c->Friction = 0.9;
            c->Restitution = 0.1;
        }
    }
    return contactDetected;
}

/* CollisionCube.CheckAgainstSphere, colliders.go:369-441 */
static int cube_vs_sphere(const Collider *cube, const Collider *sphere, ContactList *out) {
    Vector3 position = m34_get_axis(&sphere->transform, 3);
    Vector3 relCenter = m34_transform_inverse(&cube->transform, &position);
    if (fabs(relCenter.v[0]) - sphere->Radius > cube->HalfSize.v[0] ||
        fabs(relCenter.v[1]) - sphere->Radius > cube->HalfSize.v[1] ||
        fabs(relCenter.v[2]) - sphere->Radius > cube->HalfSize.v[2])
        return 0;
    Vector3 closestPoint;
    for (int i = 0; i < 3; i++) {
        Real dist = relCenter.v[i];
        if (dist > cube->HalfSize.v[i]) {
            dist = cube->HalfSize.v[i];
        } else if (dist < -cube->HalfSize.v[i]) {
            dist = -cube->HalfSize.v[i];
        }
        closestPoint.v[i] = dist;
    }
    Vector3 distCheck = closestPoint;
    v3_sub(&distCheck, &relCenter);
    Real dist = v3_square_magnitude(&distCheck);
    if (dist > sphere->Radius * sphere->Radius) return 0;
    Vector3 closestPointWorld = m34_mul_vec(&cube->transform, &closestPoint);
    Contact *c = new_contact(out);
    if (!c) return 1;
    c->ContactPoint = closestPointWorld;
    c->ContactNormal = closestPointWorld;
    v3_sub(&c->ContactNormal, &position);
    if (RealEqual(v3_magnitude(&c->ContactNormal), 0.0)) {
        c->ContactNormal = sphere->Body->Velocity;
    }
    v3_normalize(&c->ContactNormal);
    c->Penetration = sphere->Radius;
    if (!RealEqual(dist, 0.0)) {
        c->Penetration -= sqrt(dist);
    } else {
        c->Penetration = 0.0;
    }
    c->Bodies[0] = cube->Body;
    c->Bodies[1] = sphere->Body;
    c->Friction = 0.9;
    c->Restitution = 0.1;
    return 1;
}

/* penetrationOnAxis, colliders.go:445-456 */
static Real penetration_on_axis(const Collider *one, const Collider *two, const Vector3 *axis,
                                const Vector3 *toCenter) {
    Real oneProject = transform_to_axis(one, axis);
    Real twoProject = transform_to_axis(two, axis);
    Real distance = fabs(v3_dot(toCenter, axis));
    return oneProject + twoProject - distance;
}

/* tryAxis, colliders.go:458-477.  Returns the continue flag; updates the
 * running smallest penetration and its case index in place. */
static int try_axis(const Collider *one, const Collider *two, Vector3 axis, const Vector3 *toCenter, int index,
                    Real *smallestPenetration, int *smallestCase, cbzo_counters *cnt) {
    if (v3_square_magnitude(&axis) < Epsilon) {
        if (cnt) cnt->sat_skipped++;
        return 1;
    }
    if (cnt) cnt->sat_axes++;
    v3_normalize(&axis);
    Real penetration = penetration_on_axis(one, two, &axis, toCenter);
    if (penetration < 0) return 0;
    if (penetration < *smallestPenetration) {
        *smallestPenetration = penetration;
        *smallestCase = index;
    }
    return 1;
}

/* fillPointFaceBoxBox, colliders.go:481-517 */
static void fill_point_face_box_box(const Collider *one, const Collider *two, const Vector3 *toCenter, int best,
                                    Real pen, ContactList *out) {
    Vector3 normal = m34_get_axis(&one->transform, best);
    if (v3_dot(&normal, toCenter) > 0) v3_mul_with(&normal, -1.0);
    Vector3 v = two->HalfSize;
    Vector3 a0 = m34_get_axis(&two->transform, 0);
    if (v3_dot(&a0, &normal) < 0) v.v[0] = -v.v[0];
    Vector3 a1 = m34_get_axis(&two->transform, 1);
    if (v3_dot(&a1, &normal) < 0) v.v[1] = -v.v[1];
    Vector3 a2 = m34_get_axis(&two->transform, 2);
    if (v3_dot(&a2, &normal) < 0) v.v[2] = -v.v[2];
    Contact *c = new_contact(out);
    if (!c) return;
    c->ContactNormal = normal;
    c->Penetration = pen;
    c->ContactPoint = m34_mul_vec(&two->transform, &v);
    c->Bodies[0] = one->Body;
    c->Bodies[1] = two->Body;
    c->Friction = 0.9;
    c->Restitution = 0.1;
}

/* contactPoint, colliders.go:519-573 */
static Vector3 contact_point(const Vector3 *pOne, const Vector3 *dOne, Real oneSize, const Vector3 *pTwo,
                             const Vector3 *dTwo, Real twoSize, int useOne) {
    Real smOne = v3_square_magnitude(dOne);
    Real smTwo = v3_square_magnitude(dTwo);
    Real dpOneTwo = v3_dot(dTwo, dOne);
    Vector3 toSt = *pOne;
    v3_sub(&toSt, pTwo);
    Real dpStaOne = v3_dot(dOne, &toSt);
    Real dpStaTwo = v3_dot(dTwo, &toSt);
    Real denom = smOne * smTwo - dpOneTwo * dpOneTwo;
    if (fabs(denom) < Epsilon) return useOne ? *pOne : *pTwo;
    Real mua = (dpOneTwo * dpStaTwo - smTwo * dpStaOne) / denom;
    Real mub = (smOne * dpStaTwo - dpOneTwo * dpStaOne) / denom;
    if (mua > oneSize || mua < -oneSize || mub > twoSize || mub < -twoSize) return useOne ? *pOne : *pTwo;
    Vector3 cOne = *dOne;
    v3_mul_with(&cOne, mua);
    v3_add(&cOne, pOne);
    Vector3 cTwo = *dTwo;
    v3_mul_with(&cTwo, mub);
    v3_add(&cTwo, pTwo);
    v3_mul_with(&cOne, 0.5);
    v3_mul_with(&cTwo, 0.5);
    v3_add(&cOne, &cTwo);
    return cOne;
}

/* CollisionCube.CheckAgainstCube, colliders.go:576-710 */
static int cube_vs_cube(const Collider *cube, const Collider *second, ContactList *out, cbzo_counters *cnt) {
    Vector3 toCenter = m34_get_axis(&second->transform, 3);
    Vector3 oneAxis3 = m34_get_axis(&cube->transform, 3);
    v3_sub(&toCenter, &oneAxis3);
    Real pen = MaxValue;
    int best = 0xffffff;
    for (int i = 0; i <= 2; i++)
        if (!try_axis(cube, second, m34_get_axis(&cube->transform, i), &toCenter, i, &pen, &best, cnt)) return 0;
    for (int i = 0; i <= 2; i++)
        if (!try_axis(cube, second, m34_get_axis(&second->transform, i), &toCenter, i + 3, &pen, &best, cnt))
            return 0;
    int bestSingleAxis = best;
    for (int i = 0; i <= 2; i++) {
        for (int k = 0; k <= 2; k++) {
            Vector3 a1 = m34_get_axis(&cube->transform, i);
            Vector3 a2 = m34_get_axis(&second->transform, k);
            Vector3 cross = v3_cross(&a1, &a2);
            if (!try_axis(cube, second, cross, &toCenter, (i * 3) + 6 + k, &pen, &best, cnt)) return 0;
        }
    }
    if (best < 3) {
        fill_point_face_box_box(cube, second, &toCenter, best, pen, out);
        return 1;
    } else if (best < 6) {
        Vector3 newCenter = toCenter;
        v3_mul_with(&newCenter, -1.0);
        fill_point_face_box_box(second, cube, &newCenter, best - 3, pen, out);
        return 1;
    }
    best -= 6;
    int oneAxisIndex = best / 3;
    int twoAxisIndex = best % 3;
    Vector3 oneAxis = m34_get_axis(&cube->transform, oneAxisIndex);
    Vector3 twoAxis = m34_get_axis(&second->transform, twoAxisIndex);
    Vector3 axis = v3_cross(&oneAxis, &twoAxis);
    v3_normalize(&axis);
    if (v3_dot(&axis, &toCenter) > 0) v3_mul_with(&axis, -1.0);
    Vector3 ptOnOneEdge = cube->HalfSize;
    Vector3 ptOnTwoEdge = second->HalfSize;
    for (int i = 0; i < 3; i++) {
        if (i == oneAxisIndex) {
            ptOnOneEdge.v[i] = 0;
        } else {
            Vector3 a = m34_get_axis(&cube->transform, i);
            if (v3_dot(&a, &axis) > 0) ptOnOneEdge.v[i] = -ptOnOneEdge.v[i];
        }
        if (i == twoAxisIndex) {
            ptOnTwoEdge.v[i] = 0;
        } else {
            Vector3 a = m34_get_axis(&second->transform, i);
            if (v3_dot(&a, &axis) < 0) ptOnTwoEdge.v[i] = -ptOnTwoEdge.v[i];
        }
    }
    ptOnOneEdge = m34_mul_vec(&cube->transform, &ptOnOneEdge);
    ptOnTwoEdge = m34_mul_vec(&second->transform, &ptOnTwoEdge);
    int useOne = bestSingleAxis > 2;
    /* An index past 2 arises only when no axis was ever accepted (every
     * penetration NaN); Go then indexes HalfSize out of range and panics. */
    Real oneSize = 0.0, twoSize = 0.0;
    if (oneAxisIndex < 3) oneSize = cube->HalfSize.v[oneAxisIndex];
    else if (cnt) cnt->panics++;
    if (twoAxisIndex < 3) twoSize = second->HalfSize.v[twoAxisIndex];
    Vector3 contactVertex = contact_point(&ptOnOneEdge, &oneAxis, oneSize, &ptOnTwoEdge, &twoAxis, twoSize, useOne);
    Contact *c = new_contact(out);
    if (!c) return 1;
    c->ContactNormal = axis;
    c->Penetration = pen;
    c->ContactPoint = contactVertex;
    c->Bodies[0] = cube->Body;
    c->Bodies[1] = second->Body;
    c->Friction = 0.9;
    c->Restitution = 0.1;
    return 1;
}

/* CheckForCollisions, colliders.go:720-747: dispatch on the type of the
 * second collider; the sphere/plane delegations of colliders.go:116-125 and
 * :210-213 are folded in. */
static int check_for_collisions(const Collider *one, const Collider *two, ContactList *out, cbzo_counters *cnt) {
    if (cnt) cnt->pairs++;
    if (two->type == CBZO_SPHERE) {
        if (one->type == CBZO_SPHERE) return sphere_vs_sphere(one, two, out);
        return cube_vs_sphere(one, two, out); /* cube.CheckAgainstSphere */
    }
    if (one->type == CBZO_SPHERE) return cube_vs_sphere(two, one, out); /* sphere.CheckAgainstCube */
    return cube_vs_cube(one, two, out, cnt);
}

static int check_half_space(const Collider *one, const CollisionPlane *p, ContactList *out, cbzo_counters *cnt) {
    if (cnt) cnt->plane_checks++;
    if (one->type == CBZO_SPHERE) return sphere_vs_half_space(one, p, out);
    return cube_vs_half_space(one, p, out);
}

/* ===================================================================== */
/* L3 contact resolution                                                  */
/* ===================================================================== */

static const Real velocityEpsilon = 0.01; /* contact.go:11-14 */
static const Real positionEpsilon = 0.01;

/* calculateContactBasis, contact.go:118-156 */
static void contact_basis(Contact *c) {
    Vector3 ty, tz;
    const Real *n = c->ContactNormal.v;
    Real ax = fabs(n[0]);
    Real ay = fabs(n[1]);
    if (ax > ay) {
        Real s = (Real)1.0 / sqrt(n[2] * n[2] + n[0] * n[0]);
        ty.v[0] = n[2] * s;
        ty.v[1] = 0;
        ty.v[2] = n[0] * -s;
        tz.v[0] = n[1] * ty.v[0];
        tz.v[1] = n[2] * ty.v[0] - n[0] * ty.v[2];
        tz.v[2] = -n[1] * ty.v[0];
    } else {
        Real s = (Real)1.0 / sqrt(n[2] * n[2] + n[1] * n[1]);
        ty.v[0] = 0;
        ty.v[1] = -n[2] * s;
        ty.v[2] = n[1] * s;
        tz.v[0] = n[1] * ty.v[2] - n[2] * ty.v[1];
        tz.v[1] = -n[0] * ty.v[2];
        tz.v[2] = n[0] * ty.v[1];
    }
    m3_set_components(&c->contactToWorld, &c->ContactNormal, &ty, &tz);
}

/* calculateLocalVelocity, contact.go:159-182 */
static Vector3 local_velocity(const Contact *c, int bodyIndex, Real duration) {
    const RigidBody *body = c->Bodies[bodyIndex];
    Vector3 velocity = v3_cross(&body->Rotation, &c->relativeContactPosition[bodyIndex]);
    v3_add(&velocity, &body->Velocity);
    Vector3 contactVelocity = m3_transform_transpose(&c->contactToWorld, &velocity);
    Vector3 accVelocity = body->lastFrameAccelleration;
    v3_mul_with(&accVelocity, duration);
    accVelocity = m3_transform_transpose(&c->contactToWorld, &accVelocity);
    accVelocity.v[0] = 0.0;
    v3_add(&contactVelocity, &accVelocity);
    return contactVelocity;
}

/* calculateDesiredDeltaVelocity, contact.go:87-112 */
static void desired_delta_velocity(Contact *c, Real duration) {
    const Real velocityLimit = 0.25;
    Real velocityFromAcc = 0;
    Vector3 tempVelocity;
    if (c->Bodies[0]->IsAwake) {
        tempVelocity = c->Bodies[0]->lastFrameAccelleration;
        v3_mul_with(&tempVelocity, duration);
        velocityFromAcc += v3_dot(&tempVelocity, &c->ContactNormal);
    }
    if (c->Bodies[1] != NULL && c->Bodies[1]->IsAwake) {
        tempVelocity = c->Bodies[1]->lastFrameAccelleration;
        v3_mul_with(&tempVelocity, duration);
        velocityFromAcc -= v3_dot(&tempVelocity, &c->ContactNormal);
    }
    Real restitution = c->Restitution;
    if (fabs(c->contactVelocity.v[0]) < velocityLimit) restitution = 0.0;
    c->desiredDeltaVelocity = -c->contactVelocity.v[0] - restitution * (c->contactVelocity.v[0] - velocityFromAcc);
}

/* calculateInternals, contact.go:59-85 */
static void calculate_internals(Contact *c, Real duration) {
    if (c->Bodies[0] == NULL) {
        v3_mul_with(&c->ContactNormal, -1.0);
        c->Bodies[0] = c->Bodies[1];
        c->Bodies[1] = NULL;
    }
    contact_basis(c);
    c->relativeContactPosition[0] = c->ContactPoint;
    v3_sub(&c->relativeContactPosition[0], &c->Bodies[0]->Position);
    c->contactVelocity = local_velocity(c, 0, duration);
    if (c->Bodies[1] != NULL) {
        c->relativeContactPosition[1] = c->ContactPoint;
        v3_sub(&c->relativeContactPosition[1], &c->Bodies[1]->Position);
        Vector3 cv1 = local_velocity(c, 1, duration);
        v3_sub(&c->contactVelocity, &cv1);
    }
    desired_delta_velocity(c, duration);
}

/* matchAwakeState, contact.go:185-202 */
static void match_awake_state(Contact *c) {
    if (c->Bodies[1] == NULL) return;
    int b0Awake = c->Bodies[0]->IsAwake;
    int b1Awake = c->Bodies[1]->IsAwake;
    if ((b0Awake || b1Awake) && !(b0Awake && b1Awake)) {
        if (b0Awake)
            rb_set_awake(c->Bodies[1], 1);
        else
            rb_set_awake(c->Bodies[0], 1);
    }
}

/* applyPositionChange, contact.go:286-386 */
static void apply_position_change(Contact *c, Real penetration, Vector3 linearChange[2], Vector3 angularChange[2]) {
    const Real angularLimit = 0.2;
    Real angularInertia[2] = {0, 0}, linearInertia[2] = {0, 0}, angularMove[2] = {0, 0}, linearMove[2] = {0, 0};
    Real totalInertia = 0;
    memset(linearChange, 0, 2 * sizeof(Vector3));
    memset(angularChange, 0, 2 * sizeof(Vector3));
    for (int i = 0; i < 2; i++) {
        RigidBody *body = c->Bodies[i];
        if (body == NULL) continue;
        Matrix3 iit = body->inverseInertiaTensorWorld;
        Vector3 aiw = v3_cross(&c->relativeContactPosition[i], &c->ContactNormal);
        aiw = m3_mul_vec(&iit, &aiw);
        aiw = v3_cross(&aiw, &c->relativeContactPosition[i]);
        angularInertia[i] = v3_dot(&aiw, &c->ContactNormal);
        linearInertia[i] = body->inverseMass;
        totalInertia += linearInertia[i] + angularInertia[i];
    }
    for (int i = 0; i < 2; i++) {
        RigidBody *body = c->Bodies[i];
        if (body == NULL) continue;
        Real sign = 1.0;
        if (i != 0) sign = -1.0;
        angularMove[i] = sign * penetration * (angularInertia[i] / totalInertia);
        linearMove[i] = sign * penetration * (linearInertia[i] / totalInertia);
        Vector3 projection = c->relativeContactPosition[i];
        v3_add_scaled(&projection, &c->ContactNormal, -v3_dot(&c->relativeContactPosition[i], &c->ContactNormal));
        Real maxMagnitude = angularLimit * v3_magnitude(&projection);
        if (angularMove[i] < -maxMagnitude) {
            Real totalMove = angularMove[i] + linearMove[i];
            angularMove[i] = -maxMagnitude;
            linearMove[i] = totalMove - angularMove[i];
        } else if (angularMove[i] > maxMagnitude) {
            Real totalMove = angularMove[i] + linearMove[i];
            angularMove[i] = maxMagnitude;
            linearMove[i] = totalMove - angularMove[i];
        }
        if (angularMove[i] == 0.0) {
            v3_clear(&angularChange[i]);
        } else {
            Vector3 targetAngularDirection = v3_cross(&c->relativeContactPosition[i], &c->ContactNormal);
            Matrix3 iit = body->inverseInertiaTensorWorld;
            angularChange[i] = m3_mul_vec(&iit, &targetAngularDirection);
            v3_mul_with(&angularChange[i], angularMove[i] / angularInertia[i]);
        }
        linearChange[i] = c->ContactNormal;
        v3_mul_with(&linearChange[i], linearMove[i]);
        v3_add_scaled(&body->Position, &c->ContactNormal, linearMove[i]);
        q_add_scaled_vector(&body->Orientation, &angularChange[i], 1.0);
        q_normalize(&body->Orientation);
        if (body->IsAwake == 0) rb_calculate_derived(body);
    }
}

/* adjustPositions, contact.go:233-283 */
static void adjust_positions(int64_t maxIterations, Contact *contacts, int64_t n, cbzo_counters *cnt) {
    int64_t iterationsUsed = 0;
    while (iterationsUsed < maxIterations) {
        Real max = positionEpsilon;
        int64_t index = n;
        for (int64_t i = 0; i < n; i++) {
            if (contacts[i].Penetration > max) {
                max = contacts[i].Penetration;
                index = i;
            }
        }
        if (index == n) break;
        Contact *contact = &contacts[index];
        match_awake_state(contact);
        Vector3 linearChange[2], angularChange[2];
        apply_position_change(contact, max, linearChange, angularChange);
        for (int64_t i = 0; i < n; i++) {
            Contact *c = &contacts[i];
            for (int b = 0; b < 2; b++) {
                if (c->Bodies[b] == NULL) continue;
                for (int d = 0; d < 2; d++) {
                    if (c->Bodies[b] == contact->Bodies[d]) {
                        Vector3 deltaPosition = v3_cross(&angularChange[d], &c->relativeContactPosition[b]);
                        v3_add(&deltaPosition, &linearChange[d]);
                        Real sign = 1.0;
                        if (b == 0) sign = -1.0;
                        c->Penetration += v3_dot(&deltaPosition, &c->ContactNormal) * sign;
                    }
                }
            }
        }
        iterationsUsed++;
    }
    if (cnt) cnt->pos_iters += iterationsUsed;
}

/* calculateFrictionlessImpulse, contact.go:498-531.  Unreachable in the
 * reference while every contact carries Friction 0.9 (colliders.go:201 etc.);
 * reached here by contacts whose friction a caller set to 0.  The
 * reference's test at :512 is inverted (Bodies[1] == nil): a two-body
 * contact never adds its second body's terms, and a half-space contact would
 * dereference nil (the oracle counts that as a panic and treats the missing
 * body as zero).  With CBZO_FRICTIONLESS_TWO_BODY the second body's terms
 * are added when it exists (the evident intent of :512-524); this option has
 * no counterpart in the reference. */
static Vector3 frictionless_impulse(Contact *c, const Matrix3 iit[2], int32_t flags, cbzo_counters *cnt) {
    Vector3 dvw = v3_cross(&c->relativeContactPosition[0], &c->ContactNormal);
    dvw = m3_mul_vec(&iit[0], &dvw);
    dvw = v3_cross(&dvw, &c->relativeContactPosition[0]);
    Real deltaVelocity = v3_dot(&dvw, &c->ContactNormal);
    deltaVelocity += c->Bodies[0]->inverseMass;
    if ((flags & CBZO_FRICTIONLESS_TWO_BODY) && c->Bodies[1] != NULL) {
        dvw = v3_cross(&c->relativeContactPosition[1], &c->ContactNormal);
        dvw = m3_mul_vec(&iit[1], &dvw);
        dvw = v3_cross(&dvw, &c->relativeContactPosition[1]);
        deltaVelocity += v3_dot(&dvw, &c->ContactNormal);
        deltaVelocity += c->Bodies[1]->inverseMass;
    } else if (c->Bodies[1] == NULL) {
        if (cnt) cnt->panics++;
        dvw = v3_cross(&c->relativeContactPosition[1], &c->ContactNormal);
        dvw = m3_mul_vec(&iit[1], &dvw);
        dvw = v3_cross(&dvw, &c->relativeContactPosition[1]);
        deltaVelocity += v3_dot(&dvw, &c->ContactNormal);
    }
    Vector3 r = {{c->desiredDeltaVelocity / deltaVelocity, 0, 0}};
    return r;
}

/* setSkewSymmetric, contact.go:612-616 */
static void set_skew_symmetric(Matrix3 *m, const Vector3 *v) {
    m->m[0] = 0.0;      m->m[3] = -v->v[2]; m->m[6] = v->v[1];
    m->m[1] = v->v[2];  m->m[4] = 0.0;      m->m[7] = -v->v[0];
    m->m[2] = -v->v[1]; m->m[5] = v->v[0];  m->m[8] = 0.0;
}

/* calculateFrictionImpulse, contact.go:535-606 */
static Vector3 friction_impulse(Contact *c, const Matrix3 iit[2]) {
    Real inverseMass = c->Bodies[0]->inverseMass;
    Matrix3 impulseToTorque;
    set_skew_symmetric(&impulseToTorque, &c->relativeContactPosition[0]);
    Matrix3 deltaVelWorld = m3_mul(&impulseToTorque, &iit[0]);
    deltaVelWorld = m3_mul(&deltaVelWorld, &impulseToTorque);
    m3_mul_with(&deltaVelWorld, -1.0);
    if (c->Bodies[1] != NULL) {
        set_skew_symmetric(&impulseToTorque, &c->relativeContactPosition[1]);
        Matrix3 dvw2 = m3_mul(&impulseToTorque, &iit[1]);
        dvw2 = m3_mul(&dvw2, &impulseToTorque);
        m3_mul_with(&dvw2, -1.0);
        m3_add(&deltaVelWorld, &dvw2);
        inverseMass += c->Bodies[1]->inverseMass;
    }
    Matrix3 deltaVelocity = m3_transpose(&c->contactToWorld);
    deltaVelocity = m3_mul(&deltaVelocity, &deltaVelWorld);
    deltaVelocity = m3_mul(&deltaVelocity, &c->contactToWorld);
    deltaVelocity.m[0] += inverseMass;
    deltaVelocity.m[4] += inverseMass;
    deltaVelocity.m[8] += inverseMass;
    Matrix3 impulseMatrix = m3_invert(&deltaVelocity);
    Vector3 velKill = v3(c->desiredDeltaVelocity, -c->contactVelocity.v[1], -c->contactVelocity.v[2]);
    Vector3 ic = m3_mul_vec(&impulseMatrix, &velKill);
    Real planarImpulse = sqrt(ic.v[1] * ic.v[1] + ic.v[2] * ic.v[2]);
    if (planarImpulse > ic.v[0] * c->Friction) {
        ic.v[1] /= planarImpulse;
        ic.v[2] /= planarImpulse;
        ic.v[0] = deltaVelocity.m[0] + deltaVelocity.m[3] * c->Friction * ic.v[1] +
                  deltaVelocity.m[6] * c->Friction * ic.v[2];
        ic.v[0] = c->desiredDeltaVelocity / ic.v[0];
        ic.v[1] *= c->Friction * ic.v[0];
        ic.v[2] *= c->Friction * ic.v[0];
    }
    return ic;
}

/* applyVelocityChange, contact.go:448-494 */
static void apply_velocity_change(Contact *c, Vector3 velocityChange[2], Vector3 rotationChange[2],
                                  int32_t flags, cbzo_counters *cnt) {
    Matrix3 iit[2];
    memset(iit, 0, sizeof iit);
    memset(velocityChange, 0, 2 * sizeof(Vector3));
    memset(rotationChange, 0, 2 * sizeof(Vector3));
    iit[0] = c->Bodies[0]->inverseInertiaTensorWorld;
    if (c->Bodies[1] != NULL) iit[1] = c->Bodies[1]->inverseInertiaTensorWorld;
    Vector3 impulseContact;
    if (c->Friction == 0.0)
        impulseContact = frictionless_impulse(c, iit, flags, cnt);
    else
        impulseContact = friction_impulse(c, iit);
    Vector3 impulse = m3_mul_vec(&c->contactToWorld, &impulseContact);
    Vector3 impulsiveTorque = v3_cross(&c->relativeContactPosition[0], &impulse);
    rotationChange[0] = m3_mul_vec(&iit[0], &impulsiveTorque);
    v3_clear(&velocityChange[0]);
    v3_add_scaled(&velocityChange[0], &impulse, c->Bodies[0]->inverseMass);
    v3_add(&c->Bodies[0]->Velocity, &velocityChange[0]);
    v3_add(&c->Bodies[0]->Rotation, &rotationChange[0]);
    if (c->Bodies[1] != NULL) {
        impulsiveTorque = v3_cross(&impulse, &c->relativeContactPosition[1]);
        rotationChange[1] = m3_mul_vec(&iit[1], &impulsiveTorque);
        v3_clear(&velocityChange[1]);
        v3_add_scaled(&velocityChange[1], &impulse, -c->Bodies[1]->inverseMass);
        v3_add(&c->Bodies[1]->Velocity, &velocityChange[1]);
        v3_add(&c->Bodies[1]->Rotation, &rotationChange[1]);
    }
}

/* adjustVelocities, contact.go:390-445 */
static void adjust_velocities(int64_t maxIterations, Contact *contacts, int64_t n, Real duration,
                              int32_t flags, cbzo_counters *cnt) {
    int64_t iterationsUsed = 0;
    while (iterationsUsed < maxIterations) {
        Real max = velocityEpsilon;
        int64_t index = n;
        for (int64_t i = 0; i < n; i++) {
            if (contacts[i].desiredDeltaVelocity > max) {
                max = contacts[i].desiredDeltaVelocity;
                index = i;
            }
        }
        if (index == n) break;
        Contact *contact = &contacts[index];
        match_awake_state(contact);
        Vector3 velocityChange[2], rotationChange[2];
        apply_velocity_change(contact, velocityChange, rotationChange, flags, cnt);
        for (int64_t i = 0; i < n; i++) {
            Contact *c2 = &contacts[i];
            for (int b = 0; b < 2; b++) {
                if (c2->Bodies[b] == NULL) continue;
                for (int d = 0; d < 2; d++) {
                    if (c2->Bodies[b] == contact->Bodies[d]) {
                        Vector3 deltaVel = v3_cross(&rotationChange[d], &c2->relativeContactPosition[b]);
                        v3_add(&deltaVel, &velocityChange[d]);
                        Real sign = 1.0;
                        if (b == 1) sign = -1.0;
                        Vector3 tmp = m3_transform_transpose(&c2->contactToWorld, &deltaVel);
                        v3_mul_with(&tmp, sign);
                        v3_add(&c2->contactVelocity, &tmp);
                        desired_delta_velocity(c2, duration);
                    }
                }
            }
        }
        iterationsUsed++;
    }
    if (cnt) cnt->vel_iters += iterationsUsed;
}

/* ResolveContacts, contact.go:208-222 */
static void resolve_contacts(int64_t maxIterations, Contact *contacts, int64_t n, Real duration,
                             int32_t flags, cbzo_counters *cnt) {
    if (duration <= 0.0 || contacts == NULL || n == 0) return;
    for (int64_t i = 0; i < n; i++) calculate_internals(&contacts[i], duration); /* prepareContacts :225 */
    adjust_positions(maxIterations, contacts, n, cnt);
    adjust_velocities(maxIterations, contacts, n, duration, flags, cnt);
}

/* ===================================================================== */
/* World <-> SoA scene                                                    */
/* ===================================================================== */

typedef struct World {
    int n, np;
    RigidBody *bodies;
    Collider *colliders;
    CollisionPlane *planes;
} World;

static void world_load(World *w, const cbzo_scene *s) {
    w->n = s->n;
    w->np = s->np;
    w->bodies = (RigidBody *)calloc((size_t)(s->n > 0 ? s->n : 1), sizeof(RigidBody));
    w->colliders = (Collider *)calloc((size_t)(s->n > 0 ? s->n : 1), sizeof(Collider));
    w->planes = (CollisionPlane *)calloc((size_t)(s->np > 0 ? s->np : 1), sizeof(CollisionPlane));
    for (int i = 0; i < s->n; i++) {
        RigidBody *b = &w->bodies[i];
        memcpy(b->Position.v, s->pos + 3 * i, 3 * sizeof(Real));
        memcpy(b->Orientation.v, s->quat + 4 * i, 4 * sizeof(Real));
        memcpy(b->Velocity.v, s->vel + 3 * i, 3 * sizeof(Real));
        memcpy(b->Rotation.v, s->rot + 3 * i, 3 * sizeof(Real));
        memcpy(b->Acceleration.v, s->acc + 3 * i, 3 * sizeof(Real));
        memcpy(b->InverseInertiaTensor.m, s->iit_body + 9 * i, 9 * sizeof(Real));
        memcpy(b->inverseInertiaTensorWorld.m, s->iit_world + 9 * i, 9 * sizeof(Real));
        memcpy(b->transform.m, s->transform + 12 * i, 12 * sizeof(Real));
        memcpy(b->lastFrameAccelleration.v, s->last_acc + 3 * i, 3 * sizeof(Real));
        b->inverseMass = s->inv_mass[i];
        b->motion = s->motion[i];
        b->powLin = s->pow_lin ? s->pow_lin[i] : 1.0;
        b->powAng = s->pow_ang ? s->pow_ang[i] : 1.0;
        b->IsAwake = (s->flags[i] & CBZO_AWAKE) != 0;
        b->CanSleep = (s->flags[i] & CBZO_CAN_SLEEP) != 0;
        b->isStatic = (s->flags[i] & CBZO_STATIC) != 0;
        b->index = i;
        Collider *c = &w->colliders[i];
        c->type = s->shape[i];
        c->Body = b;
        memcpy(c->Offset.m, s->offset + 12 * i, 12 * sizeof(Real));
        memcpy(c->transform.m, s->ctransform + 12 * i, 12 * sizeof(Real));
        memcpy(c->HalfSize.v, s->half + 3 * i, 3 * sizeof(Real));
        c->Radius = s->half[3 * i];
    }
    for (int k = 0; k < s->np; k++) {
        memcpy(w->planes[k].Normal.v, s->plane_normal + 3 * k, 3 * sizeof(Real));
        w->planes[k].Offset = s->plane_offset[k];
    }
}

static void world_store(const World *w, cbzo_scene *s) {
    for (int i = 0; i < s->n; i++) {
        const RigidBody *b = &w->bodies[i];
        memcpy(s->pos + 3 * i, b->Position.v, 3 * sizeof(Real));
        memcpy(s->quat + 4 * i, b->Orientation.v, 4 * sizeof(Real));
        memcpy(s->vel + 3 * i, b->Velocity.v, 3 * sizeof(Real));
        memcpy(s->rot + 3 * i, b->Rotation.v, 3 * sizeof(Real));
        memcpy(s->iit_world + 9 * i, b->inverseInertiaTensorWorld.m, 9 * sizeof(Real));
        memcpy(s->transform + 12 * i, b->transform.m, 12 * sizeof(Real));
        memcpy(s->ctransform + 12 * i, w->colliders[i].transform.m, 12 * sizeof(Real));
        memcpy(s->last_acc + 3 * i, b->lastFrameAccelleration.v, 3 * sizeof(Real));
        s->motion[i] = b->motion;
        uint8_t f = s->flags[i] & ~(uint8_t)CBZO_AWAKE;
        if (b->IsAwake) f |= CBZO_AWAKE;
        s->flags[i] = f;
    }
}

static void world_free(World *w) {
    free(w->bodies);
    free(w->colliders);
    free(w->planes);
}

static void contact_to_record(const World *w, const Contact *c, cbzo_contact *r) {
    r->body0 = c->Bodies[0] ? c->Bodies[0]->index : -1;
    r->body1 = c->Bodies[1] ? c->Bodies[1]->index : -1;
    (void)w;
    r->friction = c->Friction;
    r->restitution = c->Restitution;
    memcpy(r->point, c->ContactPoint.v, sizeof r->point);
    memcpy(r->normal, c->ContactNormal.v, sizeof r->normal);
    r->pen = c->Penetration;
    memcpy(r->c2w, c->contactToWorld.m, sizeof r->c2w);
    memcpy(r->rel[0], c->relativeContactPosition[0].v, 3 * sizeof(Real));
    memcpy(r->rel[1], c->relativeContactPosition[1].v, 3 * sizeof(Real));
    memcpy(r->cvel, c->contactVelocity.v, sizeof r->cvel);
    r->ddv = c->desiredDeltaVelocity;
}

static void record_to_contact(World *w, const cbzo_contact *r, Contact *c) {
    c->Bodies[0] = r->body0 >= 0 ? &w->bodies[r->body0] : NULL;
    c->Bodies[1] = r->body1 >= 0 ? &w->bodies[r->body1] : NULL;
    c->Friction = r->friction;
    c->Restitution = r->restitution;
    memcpy(c->ContactPoint.v, r->point, sizeof r->point);
    memcpy(c->ContactNormal.v, r->normal, sizeof r->normal);
    c->Penetration = r->pen;
    memcpy(c->contactToWorld.m, r->c2w, sizeof r->c2w);
    memcpy(c->relativeContactPosition[0].v, r->rel[0], 3 * sizeof(Real));
    memcpy(c->relativeContactPosition[1].v, r->rel[1], 3 * sizeof(Real));
    memcpy(c->contactVelocity.v, r->cvel, sizeof r->cvel);
    c->desiredDeltaVelocity = r->ddv;
}

/* updateObjects, examples/cubedrop.go:29-39: Integrate then the collider's
 * CalculateDerivedData for every (non-static) body. */
static void world_integrate(World *w, Real dt, Real bias, cbzo_counters *cnt) {
    for (int i = 0; i < w->n; i++) {
        RigidBody *b = &w->bodies[i];
        if (b->isStatic) continue;
        if (cnt && b->IsAwake) cnt->integrated++;
        rb_integrate(b, dt, bias);
        collider_calculate_derived(&w->colliders[i]);
    }
}

/* generateContacts, examples/cubedrop.go:42-67 generalised to any planes and
 * both shapes: for each body i, every plane in order, then every j != i. */
static void world_narrowphase(World *w, ContactList *l, cbzo_counters *cnt) {
    for (int i = 0; i < w->n; i++) {
        for (int k = 0; k < w->np; k++) check_half_space(&w->colliders[i], &w->planes[k], l, cnt);
        for (int j = 0; j < w->n; j++) {
            if (j == i) continue;
            check_for_collisions(&w->colliders[i], &w->colliders[j], l, cnt);
        }
    }
}

/* ===================================================================== */
/* exported entry points                                                  */
/* ===================================================================== */

void cbzo_calculate_derived(cbzo_scene *s) {
    World w;
    world_load(&w, s);
    for (int i = 0; i < w.n; i++) {
        rb_calculate_derived(&w.bodies[i]);
        collider_calculate_derived(&w.colliders[i]);
    }
    world_store(&w, s);
    world_free(&w);
}

void cbzo_integrate(cbzo_scene *s, double dt, double sleep_bias, cbzo_counters *cnt) {
    World w;
    world_load(&w, s);
    world_integrate(&w, dt, sleep_bias, cnt);
    world_store(&w, s);
    world_free(&w);
}

static int64_t emit_records(World *w, ContactList *l, cbzo_contact *out, int64_t cap, cbzo_counters *cnt) {
    int64_t m = l->n < cap ? l->n : cap;
    for (int64_t k = 0; k < m && k < l->cap; k++) contact_to_record(w, &l->c[k], &out[k]);
    if (cnt) cnt->contacts += l->n;
    return l->n;
}

int64_t cbzo_narrowphase(cbzo_scene *s, cbzo_contact *out, int64_t cap, cbzo_counters *cnt) {
    World w;
    world_load(&w, s);
    ContactList l = {(Contact *)calloc((size_t)(cap > 0 ? cap : 1), sizeof(Contact)), 0, cap, 0};
    world_narrowphase(&w, &l, cnt);
    int64_t n = emit_records(&w, &l, out, cap, cnt);
    free(l.c);
    world_free(&w);
    return n;
}

int64_t cbzo_narrowphase_pairs(cbzo_scene *s, int64_t npairs, const int32_t *one, const int32_t *two,
                               cbzo_contact *out, int64_t cap, cbzo_counters *cnt) {
    World w;
    world_load(&w, s);
    ContactList l = {(Contact *)calloc((size_t)(cap > 0 ? cap : 1), sizeof(Contact)), 0, cap, 0};
    for (int64_t p = 0; p < npairs; p++) {
        int32_t a = one[p], b = two[p];
        if (a < 0 && b < 0) continue; /* plane vs plane: colliders.go:111-113 */
        if (b < 0) {
            check_half_space(&w.colliders[a], &w.planes[-b - 1], &l, cnt);
        } else if (a < 0) {
            check_half_space(&w.colliders[b], &w.planes[-a - 1], &l, cnt); /* colliders.go:116-125 */
        } else {
            check_for_collisions(&w.colliders[a], &w.colliders[b], &l, cnt);
        }
    }
    int64_t n = emit_records(&w, &l, out, cap, cnt);
    free(l.c);
    world_free(&w);
    return n;
}

void cbzo_resolve(cbzo_scene *s, cbzo_contact *c, int64_t nc, int64_t max_iter, double dt, cbzo_counters *cnt) {
    cbzo_resolve_ex(s, c, nc, max_iter, dt, 0, cnt);
}

void cbzo_resolve_ex(cbzo_scene *s, cbzo_contact *c, int64_t nc, int64_t max_iter, double dt, int32_t flags,
                     cbzo_counters *cnt) {
    World w;
    world_load(&w, s);
    Contact *cs = (Contact *)calloc((size_t)(nc > 0 ? nc : 1), sizeof(Contact));
    for (int64_t k = 0; k < nc; k++) record_to_contact(&w, &c[k], &cs[k]);
    resolve_contacts(max_iter, cs, nc, dt, flags, cnt);
    for (int64_t k = 0; k < nc; k++) contact_to_record(&w, &cs[k], &c[k]);
    world_store(&w, s);
    free(cs);
    world_free(&w);
}

int cbzo_step(cbzo_scene *s, double dt, double sleep_bias, int32_t iter_mult, int32_t n_steps, int64_t *counts,
              cbzo_counters *cnt) {
    World w;
    world_load(&w, s);
    int64_t cap = 64 + 16 * (int64_t)w.n * (w.np + 1);
    ContactList l = {(Contact *)calloc((size_t)cap, sizeof(Contact)), 0, cap, 0};
    int rc = 0;
    for (int step = 0; step < n_steps; step++) {
        world_integrate(&w, dt, sleep_bias, cnt);
        l.n = 0;
        l.overflow = 0;
        world_narrowphase(&w, &l, cnt);
        if (l.overflow) {
            /* grow and redo the narrowphase; it does not mutate state */
            cap = l.n * 2;
            free(l.c);
            l.c = (Contact *)calloc((size_t)cap, sizeof(Contact));
            l.cap = cap;
            l.n = 0;
            l.overflow = 0;
            world_narrowphase(&w, &l, cnt);
            if (l.overflow) {
                rc = -1;
                break;
            }
        }
        if (counts) counts[step] = l.n;
        if (cnt) cnt->contacts += l.n;
        if (l.n > 0) resolve_contacts((int64_t)iter_mult * l.n, l.c, l.n, dt, 0, cnt);
    }
    world_store(&w, s);
    free(l.c);
    world_free(&w);
    return rc;
}

/* --- math entry points ------------------------------------------------- */
int cbzo_real_equal(double a, double b) { return RealEqual(a, b); }
void cbzo_vec_cross(const double *a, const double *b, double *out) {
    Vector3 x, y;
    memcpy(x.v, a, 24);
    memcpy(y.v, b, 24);
    Vector3 r = v3_cross(&x, &y);
    memcpy(out, r.v, 24);
}
double cbzo_vec_dot(const double *a, const double *b) {
    Vector3 x, y;
    memcpy(x.v, a, 24);
    memcpy(y.v, b, 24);
    return v3_dot(&x, &y);
}
double cbzo_vec_magnitude(const double *a) {
    Vector3 x;
    memcpy(x.v, a, 24);
    return v3_magnitude(&x);
}
void cbzo_vec_normalize(double *a) {
    Vector3 x;
    memcpy(x.v, a, 24);
    v3_normalize(&x);
    memcpy(a, x.v, 24);
}
void cbzo_quat_mul(double *q, const double *q2) {
    Quat a, b;
    memcpy(a.v, q, 32);
    memcpy(b.v, q2, 32);
    q_mul(&a, &b);
    memcpy(q, a.v, 32);
}
double cbzo_quat_len(const double *q) {
    Quat a;
    memcpy(a.v, q, 32);
    return q_len(&a);
}
void cbzo_quat_normalize(double *q) {
    Quat a;
    memcpy(a.v, q, 32);
    q_normalize(&a);
    memcpy(q, a.v, 32);
}
void cbzo_quat_add_scaled_vector(double *q, const double *v, double scale) {
    Quat a;
    Vector3 x;
    memcpy(a.v, q, 32);
    memcpy(x.v, v, 24);
    q_add_scaled_vector(&a, &x, scale);
    memcpy(q, a.v, 32);
}
void cbzo_quat_rotate(const double *q, const double *v, double *out) {
    Quat a;
    Vector3 x;
    memcpy(a.v, q, 32);
    memcpy(x.v, v, 24);
    Vector3 r = q_rotate(&a, &x);
    memcpy(out, r.v, 24);
}
void cbzo_quat_from_axis(double angle, double x, double y, double z, double *out) {
    Quat r = q_from_axis(angle, x, y, z);
    memcpy(out, r.v, 32);
}
void cbzo_mat3_mul(const double *a, const double *b, double *out) {
    Matrix3 x, y;
    memcpy(x.m, a, 72);
    memcpy(y.m, b, 72);
    Matrix3 r = m3_mul(&x, &y);
    memcpy(out, r.m, 72);
}
void cbzo_mat3_mul_vec(const double *a, const double *v, double *out) {
    Matrix3 x;
    Vector3 y;
    memcpy(x.m, a, 72);
    memcpy(y.v, v, 24);
    Vector3 r = m3_mul_vec(&x, &y);
    memcpy(out, r.v, 24);
}
void cbzo_mat3_invert(const double *a, double *out) {
    Matrix3 x;
    memcpy(x.m, a, 72);
    Matrix3 r = m3_invert(&x);
    memcpy(out, r.m, 72);
}
double cbzo_mat3_det(const double *a) {
    Matrix3 x;
    memcpy(x.m, a, 72);
    return m3_det(&x);
}
void cbzo_mat3_block_inertia(const double *half, double mass, double *out) {
    Vector3 h;
    Matrix3 m;
    memcpy(h.v, half, 24);
    m3_set_block_inertia(&m, &h, mass);
    memcpy(out, m.m, 72);
}
void cbzo_mat34_set_transform(const double *pos, const double *q, double *out) {
    Vector3 p;
    Quat r;
    Matrix3x4 m;
    memcpy(p.v, pos, 24);
    memcpy(r.v, q, 32);
    m34_set_as_transform(&m, &p, &r);
    memcpy(out, m.m, 96);
}
void cbzo_mat34_mul(const double *a, const double *b, double *out) {
    Matrix3x4 x, y;
    memcpy(x.m, a, 96);
    memcpy(y.m, b, 96);
    Matrix3x4 r = m34_mul(&x, &y);
    memcpy(out, r.m, 96);
}
void cbzo_mat34_mul_vec(const double *a, const double *v, double *out) {
    Matrix3x4 x;
    Vector3 y;
    memcpy(x.m, a, 96);
    memcpy(y.v, v, 24);
    Vector3 r = m34_mul_vec(&x, &y);
    memcpy(out, r.v, 24);
}
void cbzo_mat34_transform_inverse(const double *a, const double *v, double *out) {
    Matrix3x4 x;
    Vector3 y;
    memcpy(x.m, a, 96);
    memcpy(y.v, v, 24);
    Vector3 r = m34_transform_inverse(&x, &y);
    memcpy(out, r.v, 24);
}
void cbzo_transform_inertia(const double *iit_body, const double *rot, double *out) {
    Matrix3 ib, iw;
    Matrix3x4 r;
    memcpy(ib.m, iit_body, 72);
    memcpy(r.m, rot, 96);
    transform_inertia_tensor(&iw, &ib, &r);
    memcpy(out, iw.m, 72);
}
void cbzo_contact_basis(const double *normal, double *c2w) {
    Contact c;
    memset(&c, 0, sizeof c);
    memcpy(c.ContactNormal.v, normal, 24);
    contact_basis(&c);
    memcpy(c2w, c.contactToWorld.m, 72);
}
void cbzo_skew(const double *v, double *out) {
    Vector3 x;
    Matrix3 m;
    memcpy(x.v, v, 24);
    set_skew_symmetric(&m, &x);
    memcpy(out, m.m, 72);
}
void cbzo_contact_point(const double *p1, const double *d1, double s1, const double *p2, const double *d2,
                        double s2, int use_one, double *out) {
    Vector3 a, b, c, d;
    memcpy(a.v, p1, 24);
    memcpy(b.v, d1, 24);
    memcpy(c.v, p2, 24);
    memcpy(d.v, d2, 24);
    Vector3 r = contact_point(&a, &b, s1, &c, &d, s2, use_one);
    memcpy(out, r.v, 24);
}

/* unused-function guards for helpers only the known-answer tests reach */
void cbzo__keep(void) {
    Matrix3 m;
    m3_set_identity(&m);
    Matrix3x4 t;
    m34_set_identity(&t);
    (void)m;
    (void)t;
}
