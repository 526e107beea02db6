"""Pins of the oracle's L0 math against the reference's own known-answer
tests (math/vector_test.go, math/quaternion_test.go, math/matrix_test.go),
plus exactness where the answer is exactly representable and numpy
cross-checks for the matrix products and inverse."""
import ctypes
import math

import numpy as np
import pytest

from oracle import cbz_oracle

MAXV = 1.7976931348623157e308


@pytest.fixture(scope="module")
def L():
    return cbz_oracle.load()


def arr(*v):
    return np.array(v, dtype=np.float64)


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def real_equal(L, a, b):
    return bool(L.cbzo_real_equal(float(a), float(b)))


def test_real_equal_semantics(L):
    # math/math.go:64-78: exact equality (incl. inf), absolute 1e-14 near 0, relative 1e-7 otherwise
    assert real_equal(L, math.inf, math.inf)
    assert real_equal(L, 1.0, 1.0 + 5e-8)
    assert not real_equal(L, 1.0, 1.0 + 3e-7)
    assert real_equal(L, 0.0, 9e-15)
    assert not real_equal(L, 0.0, 2e-14)
    assert real_equal(L, 1e-13, 1e-13)
    assert not real_equal(L, math.nan, math.nan)


def test_vector_cross_dot_magnitude(L):
    # math/vector_test.go:74-132
    out = np.zeros(3)
    L.cbzo_vec_cross(p(arr(1, 2, 3)), p(arr(10, 11, 12)), p(out))
    assert list(out) == [-9.0, 18.0, -9.0]  # exactly representable
    assert L.cbzo_vec_dot(p(arr(-1, -5, -7)), p(arr(10, 20, 30))) == -320.0
    m = L.cbzo_vec_magnitude(p(arr(2, -5, 4)))
    assert real_equal(L, m, 6.708203932499369)
    assert m == math.sqrt(45.0)


def test_vector_normalize(L):
    # math/vector_test.go:144-150; and the RealEqual(|v|, 0) no-op (math/vector.go:64-72)
    v = arr(1.0, 2.5, 3.5)
    L.cbzo_vec_normalize(p(v))
    assert real_equal(L, L.cbzo_vec_magnitude(p(v)), 1.0)
    tiny = arr(1e-16, 0, 0)
    L.cbzo_vec_normalize(p(tiny))
    assert tiny[0] == 1e-16


def test_quat_mul(L):
    # math/quaternion_test.go:75-84
    q = arr(1.0, 0.5, -3.0, 4.0)
    L.cbzo_quat_mul(p(q), p(arr(6.0, 2.0, 1.0, -9.0)))
    assert list(q) == [44.0, 28.0, -4.5, 21.5]
    i1 = arr(1, 0, 0, 0)
    L.cbzo_quat_mul(p(i1), p(arr(1, 0, 0, 0)))
    assert list(i1) == [1, 0, 0, 0]


def test_quat_len(L):
    # math/quaternion_test.go:23-47
    assert real_equal(L, L.cbzo_quat_len(p(arr(0, 1, 0, 0))), 1.0)
    assert real_equal(L, L.cbzo_quat_len(p(arr(0, 1e-13, 0, 0))), 1e-13)
    assert real_equal(L, L.cbzo_quat_len(p(arr(0, MAXV, 1, 0))), math.inf)
    assert real_equal(L, L.cbzo_quat_len(p(arr(4, 1, 2, 3))), math.sqrt(30.0))


def test_quat_normalize(L):
    # math/quaternion_test.go:49-73
    cases = [((0, 0, 0, 0), (1, 0, 0, 0)), ((0, 1, 0, 0), (0, 1, 0, 0)),
             ((0, 1e-13, 0, 0), (0, 1, 0, 0)), ((0, MAXV, 1, 0), (0, 1, 0, 0))]
    for q0, want in cases:
        q = arr(*q0)
        L.cbzo_quat_normalize(p(q))
        for a, b in zip(q, want):
            assert real_equal(L, a, b), (q0, q)
    # MaxValue case: 1/MaxValue is subnormal, so the y component is 1/MaxValue, not 0
    q = arr(0, MAXV, 1, 0)
    L.cbzo_quat_normalize(p(q))
    assert q[2] == 1.0 / MAXV and q[2] > 0


ROTATE_CASES = [
    # (angle deg or None for the first case, axis, v, expected)  math/quaternion_test.go:86-181
    (0, (0, 1, 0), (1, 0, 0), (1, 0, 0)),
    (90, (0, 1, 0), (1, 0, 0), (0, 0, -1)),
    (180, (0, 1, 0), (1, 0, 0), (-1, 0, 0)),
    (270, (0, 1, 0), (0, 0, 1), (-1, 0, 0)),
    (0, (0, 0, 1), (0, 1, 0), (0, 1, 0)),
    (90, (1, 0, 0), (0, 1, 0), (0, 0, 1)),
    (180, (1, 0, 0), (0, 1, 0), (0, -1, 0)),
    (270, (1, 0, 0), (0, 1, 0), (0, 0, -1)),
    (0, (0, 0, 1), (0, 1, 0), (0, 1, 0)),
    (90, (0, 0, 1), (0, 1, 0), (-1, 0, 0)),
    (180, (0, 0, 1), (0, 1, 0), (0, -1, 0)),
    (270, (0, 0, 1), (0, 1, 0), (1, 0, 0)),
]


def test_quat_rotate(L):
    q = arr(1.0, 0.0, 1.0, 0.0)
    L.cbzo_quat_normalize(p(q))
    out = np.zeros(3)
    L.cbzo_quat_rotate(p(q), p(arr(1, 0, 0)), p(out))
    assert all(real_equal(L, a, b) for a, b in zip(out, (0, 0, -1)))
    for deg, ax, v, want in ROTATE_CASES:
        q = np.zeros(4)
        L.cbzo_quat_from_axis(deg * math.pi / 180.0, *map(float, ax), p(q))
        L.cbzo_quat_rotate(p(q), p(arr(*v)), p(out))
        assert all(real_equal(L, a, b) for a, b in zip(out, want)), (deg, ax, out)


def test_mat3_identity_and_inverse(L):
    # math/matrix_test.go:10-53
    I = np.eye(3).reshape(-1, order="F").copy()
    out = np.zeros(9)
    L.cbzo_mat3_mul(p(I), p(I), p(out))
    assert np.array_equal(out, I)
    m = arr(0.6, 0.2, 0.3, 0.2, 0.7, 0.5, 0.3, 0.5, 0.7)
    inv = np.zeros(9)
    L.cbzo_mat3_invert(p(m), p(inv))
    L.cbzo_mat3_mul(p(m), p(inv), p(out))
    for a, b in zip(out, I):
        assert real_equal(L, a, b)
    ref = np.linalg.inv(m.reshape(3, 3, order="F")).reshape(-1, order="F")
    assert np.allclose(inv, ref, rtol=1e-13, atol=1e-13)
    T = np.eye(4)[:3].T.reshape(-1).copy()  # 3x4 identity, column-major
    o12 = np.zeros(12)
    L.cbzo_mat34_mul(p(T), p(T), p(o12))
    assert np.array_equal(o12, T)


def test_mat3_singular_inverse_is_zero(L):
    # math/matrix.go:135-137
    out = np.ones(9)
    L.cbzo_mat3_invert(p(arr(1, 2, 3, 2, 4, 6, 1, 1, 1)), p(out))
    assert np.all(out == 0)


def test_mat3_products_vs_numpy(L):
    rng = np.random.default_rng(0)
    for _ in range(200):
        a = rng.standard_normal(9)
        b = rng.standard_normal(9)
        v = rng.standard_normal(3)
        out = np.zeros(9)
        L.cbzo_mat3_mul(p(a), p(b), p(out))
        A, B = a.reshape(3, 3, order="F"), b.reshape(3, 3, order="F")
        assert np.allclose(out.reshape(3, 3, order="F"), A @ B, rtol=1e-14, atol=1e-14)
        ov = np.zeros(3)
        L.cbzo_mat3_mul_vec(p(a), p(v), p(ov))
        assert np.allclose(ov, A @ v, rtol=1e-14, atol=1e-14)
        assert math.isclose(L.cbzo_mat3_det(p(a)), np.linalg.det(A), rel_tol=1e-11, abs_tol=1e-12)


def test_block_inertia_closed_form(L):
    # math/matrix.go:68-77 uses 0.3*m*(h^2 + h^2) (not 1/3)
    out = np.zeros(9)
    L.cbzo_mat3_block_inertia(p(arr(1.0, 2.0, 3.0)), 8.0, p(out))
    M = out.reshape(3, 3, order="F")
    assert math.isclose(M[0, 0], 0.3 * 8 * 13, rel_tol=1e-15)
    assert math.isclose(M[1, 1], 0.3 * 8 * 10, rel_tol=1e-15)
    assert math.isclose(M[2, 2], 0.3 * 8 * 5, rel_tol=1e-15)
    assert np.all(out[[1, 2, 3, 5, 6, 7]] == 0)


def test_transform_matches_rotate_and_is_orthonormal(L):
    # SetAsTransform (math/matrix.go:167-185) against Quat.Rotate (pinned above)
    rng = np.random.default_rng(2)
    for _ in range(100):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        pos = rng.standard_normal(3)
        T = np.zeros(12)
        L.cbzo_mat34_set_transform(p(pos), p(q), p(T))
        R = T[:9].reshape(3, 3, order="F")
        assert np.allclose(R.T @ R, np.eye(3), atol=1e-14)
        assert math.isclose(np.linalg.det(R), 1.0, rel_tol=1e-13)
        assert np.array_equal(T[9:], pos)
        v = rng.standard_normal(3)
        rv = np.zeros(3)
        L.cbzo_quat_rotate(p(q), p(v), p(rv))
        assert np.allclose(R @ v, rv, atol=1e-13)
        mv = np.zeros(3)
        L.cbzo_mat34_mul_vec(p(T), p(v), p(mv))
        back = np.zeros(3)
        L.cbzo_mat34_transform_inverse(p(T), p(mv), p(back))
        assert np.allclose(back, v, atol=1e-13)
        # world inverse inertia = R I^-1 R^T (rigidbody.go:275-299)
        ib = rng.standard_normal(9)
        iw = np.zeros(9)
        L.cbzo_transform_inertia(p(ib), p(T), p(iw))
        IB = ib.reshape(3, 3, order="F")
        assert np.allclose(iw.reshape(3, 3, order="F"), R @ IB @ R.T, atol=1e-13)


def test_contact_basis(L):
    # calculateContactBasis (contact.go:118-156).  First column = normal,
    # second = a unit tangent orthogonal to it.  The y-dominant branch is an
    # orthonormal right-handed frame.  The x-dominant branch sets
    # tangentZ.x = n.y * tangentY.x (contact.go:136) where n x tangentY has
    # n.y * tangentY.z, so its third column equals n x tangentY except in x:
    # the reference's frame is not orthonormal there, and the oracle keeps it.
    rng = np.random.default_rng(3)
    seen = [0, 0]
    for k in range(400):
        n = rng.standard_normal(3)
        if k < 6:
            n = np.eye(3)[k % 3] * (1 if k < 3 else -1)
        n /= np.linalg.norm(n)
        m = np.zeros(9)
        L.cbzo_contact_basis(p(n), p(m))
        M = m.reshape(3, 3, order="F")
        t, z = M[:, 1], M[:, 2]
        assert np.array_equal(M[:, 0], n)
        assert math.isclose(np.linalg.norm(t), 1.0, rel_tol=1e-14)
        assert abs(np.dot(t, n)) < 1e-15
        c = np.cross(n, t)
        if abs(n[0]) > abs(n[1]):
            seen[0] += 1
            assert t[1] == 0
            assert np.allclose(z[1:], c[1:], atol=1e-15)
            assert math.isclose(z[0] - c[0], n[1] * (t[0] - t[2]), abs_tol=1e-15)
        else:
            seen[1] += 1
            assert t[0] == 0
            assert np.allclose(M.T @ M, np.eye(3), atol=1e-14)
            assert np.allclose(z, c, atol=1e-15)
    assert min(seen) > 50


def test_skew_is_cross(L):
    rng = np.random.default_rng(4)
    for _ in range(50):
        v, w = rng.standard_normal(3), rng.standard_normal(3)
        m = np.zeros(9)
        L.cbzo_skew(p(v), p(m))
        assert np.allclose(m.reshape(3, 3, order="F") @ w, np.cross(v, w), atol=1e-14)


def test_add_scaled_vector_is_first_order_rotation(L):
    # AddScaledVector (math/quaternion.go:20-33): q += 0.5 (0, w dt) q, i.e. a
    # world-frame angular velocity multiplied on the left
    rng = np.random.default_rng(5)
    for _ in range(50):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        w = rng.standard_normal(3)
        dt = 1e-3
        q1 = q.copy()
        L.cbzo_quat_add_scaled_vector(p(q1), p(w), dt)
        wq = np.array([0.0, *(w * dt)])
        ref = np.array([  # Hamilton product (0, w dt) * q
            wq[0] * q[0] - wq[1] * q[1] - wq[2] * q[2] - wq[3] * q[3],
            wq[0] * q[1] + wq[1] * q[0] + wq[2] * q[3] - wq[3] * q[2],
            wq[0] * q[2] - wq[1] * q[3] + wq[2] * q[0] + wq[3] * q[1],
            wq[0] * q[3] + wq[1] * q[2] - wq[2] * q[1] + wq[3] * q[0]])
        assert np.allclose(q1, q + 0.5 * ref, atol=1e-15)


def test_contact_point_closest_points(L):
    # contactPoint (colliders.go:519-573): midpoint of the closest points of two
    # crossing edges; parallel or out-of-range edges fall back to one endpoint
    out = np.zeros(3)
    L.cbzo_contact_point(p(arr(0, 0, 0)), p(arr(1, 0, 0)), 2.0, p(arr(0, 1, 1)), p(arr(0, 0, 1)), 2.0, 0, p(out))
    assert np.allclose(out, [0, 0.5, 0])
    L.cbzo_contact_point(p(arr(0, 0, 0)), p(arr(1, 0, 0)), 2.0, p(arr(0, 1, 0)), p(arr(1, 0, 0)), 2.0, 1, p(out))
    assert np.array_equal(out, [0, 0, 0])  # parallel -> useOne -> pOne
    L.cbzo_contact_point(p(arr(0, 0, 0)), p(arr(1, 0, 0)), 0.5, p(arr(3, 1, 1)), p(arr(0, 0, 1)), 2.0, 0, p(out))
    assert np.array_equal(out, [3, 1, 1])  # mua out of range -> pTwo
