"""Seeded scene generators for the benchmark configurations.

A Scene is a set of concatenated worlds in the host layout of the C ABI
(body-major numpy arrays).  Random numbers come from SplitMix64 (counter
form: output k of stream s is mix(s + (k+1)*0x9E3779B97F4A7C15)), so a seed
fixes every bit of the scene on any machine that shares the C library's
log/cos/sqrt; golden fixtures store the generated arrays themselves.

Body set-up follows the reference's own set-up calls, in its operation
order:
  SetMass               rigidbody.go:124-127           inv_mass = 1/m
  SetBlockInertiaTensor math/matrix.go:68-77           0.3*m*(h_j^2 + h_k^2)
  SetInertiaTensorCoeffs math/matrix.go:59-63 (sphere, examples/ballistic.go:217-219)
  SetInertiaTensor      rigidbody.go:176-178 -> Matrix3.Invert math/matrix.go:133-153
"""
import math
from dataclasses import dataclass, field

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Counter-based SplitMix64 stream."""

    def __init__(self, seed, stream=0):
        with np.errstate(over="ignore"):
            self.state = (np.uint64(seed) * np.uint64(0x2545F4914F6CDD1D)
                          + np.uint64(stream) * GAMMA + np.uint64(0x632BE59BD9B4E019))
        self.k = 0

    def u64(self, n):
        with np.errstate(over="ignore"):
            idx = np.arange(self.k + 1, self.k + 1 + n, dtype=np.uint64)
            z = self.state + idx * GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        self.k += n
        return z

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n, sigma=1.0):
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        return sigma * r * np.cos(2.0 * math.pi * u2)

    def randint(self, n, lo, hi):
        """Integers uniform in [lo, hi]."""
        span = np.uint64(hi - lo + 1)
        return (self.u64(n) % span).astype(np.int64) + lo


def go_invert3(m):
    """Matrix3.Invert (math/matrix.go:133-153) on [N,9] column-major arrays,
    elementwise in the reference's operation order."""
    m = np.asarray(m, dtype=np.float64)
    a = [m[:, k] for k in range(9)]
    det = (a[0] * a[4] * a[8] + a[3] * a[7] * a[2] + a[6] * a[1] * a[5]
           - a[6] * a[4] * a[2] - a[3] * a[1] * a[8] - a[0] * a[7] * a[5])
    out = np.stack([
        a[4] * a[8] - a[5] * a[7],
        a[2] * a[7] - a[1] * a[8],
        a[1] * a[5] - a[2] * a[4],
        a[5] * a[6] - a[3] * a[8],
        a[0] * a[8] - a[2] * a[6],
        a[2] * a[3] - a[0] * a[5],
        a[3] * a[7] - a[4] * a[6],
        a[1] * a[6] - a[0] * a[7],
        a[0] * a[4] - a[1] * a[3],
    ], axis=1)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        s = 1 / det
        out = out * s[:, None]  # singular rows are zeroed below
    out[go_real_equal(det, 0.0)] = 0.0
    return out


def go_real_equal(a, b):
    """RealEqual (math/math.go:64-78), elementwise."""
    a = np.asarray(a, dtype=np.float64)
    b = np.broadcast_to(np.asarray(b, dtype=np.float64), a.shape)
    with np.errstate(invalid="ignore", over="ignore"):
        diff = np.abs(a - b)
        small = (a * b == 0) | (diff < 1.1754943508222875e-38)
        rel = diff / (np.abs(a) + np.abs(b))
        eps = 1e-7
        return (a == b) | np.where(small, diff < eps * eps, rel < eps)


def block_inertia(half, mass):
    """SetBlockInertiaTensor (math/matrix.go:68-77) via SetInertiaTensorCoeffs
    (:59-63): [N,9] column-major; the off-diagonals are -0.0 as in Go."""
    half = np.asarray(half, dtype=np.float64)
    mass = np.asarray(mass, dtype=np.float64)
    sq = half * half
    ix = 0.3 * mass * (sq[:, 1] + sq[:, 2])
    iy = 0.3 * mass * (sq[:, 0] + sq[:, 2])
    iz = 0.3 * mass * (sq[:, 0] + sq[:, 1])
    return inertia_coeffs(ix, iy, iz)


def inertia_coeffs(ix, iy, iz):
    z = np.zeros_like(ix)
    nz = -z
    return np.stack([ix, nz, nz, nz, iy, nz, nz, nz, iz], axis=1)


@dataclass
class Scene:
    """Concatenated worlds in the C ABI's host layout."""
    body_off: np.ndarray
    plane_off: np.ndarray
    plane_normal: np.ndarray
    plane_offset: np.ndarray
    fields: dict = field(default_factory=dict)
    name: str = ""
    dt: float = 1.0 / 60.0
    steps: int = 600
    iter_mult: int = 4
    meta: dict = field(default_factory=dict)

    @property
    def n_worlds(self):
        return len(self.body_off) - 1

    @property
    def n_bodies(self):
        return int(self.body_off[-1])

    def world_slice(self, w):
        return slice(int(self.body_off[w]), int(self.body_off[w + 1]))

    def subset(self, worlds):
        """A new Scene holding only the given worlds (in the given order)."""
        worlds = list(worlds)
        bo, po = [0], [0]
        bidx, pidx = [], []
        for w in worlds:
            b0, b1 = int(self.body_off[w]), int(self.body_off[w + 1])
            p0, p1 = int(self.plane_off[w]), int(self.plane_off[w + 1])
            bidx.extend(range(b0, b1))
            pidx.extend(range(p0, p1))
            bo.append(bo[-1] + b1 - b0)
            po.append(po[-1] + p1 - p0)
        bidx = np.asarray(bidx, dtype=np.int64)
        pidx = np.asarray(pidx, dtype=np.int64)
        return Scene(np.asarray(bo, np.int32), np.asarray(po, np.int32),
                     self.plane_normal[pidx].copy(), self.plane_offset[pidx].copy(),
                     {k: v[bidx].copy() for k, v in self.fields.items()},
                     name=self.name, dt=self.dt, steps=self.steps, iter_mult=self.iter_mult)


def concat(parts):
    """One Scene holding the worlds of several Scenes, in order (the first
    one's dt, steps and iteration multiplier)."""
    parts = list(parts)
    bo, po = [0], [0]
    for sc in parts:
        bo.extend(bo[-1] + np.asarray(sc.body_off[1:], np.int64))
        po.extend(po[-1] + np.asarray(sc.plane_off[1:], np.int64))
    f0 = parts[0]
    return Scene(np.asarray(bo, np.int32), np.asarray(po, np.int32),
                 np.concatenate([sc.plane_normal for sc in parts]), np.concatenate([sc.plane_offset for sc in parts]),
                 {k: np.concatenate([sc.fields[k] for sc in parts]) for k in f0.fields},
                 name="+".join(sc.name for sc in parts[:3]) + ("..." if len(parts) > 3 else ""),
                 dt=f0.dt, steps=f0.steps, iter_mult=f0.iter_mult)


def _new_bodies(n):
    """NewRigidBody (rigidbody.go:104-114) defaults for n bodies."""
    f = {
        "shape": np.zeros(n, np.int32),
        "half": np.zeros((n, 3)),
        "pos": np.zeros((n, 3)),
        "quat": np.tile(np.array([1.0, 0.0, 0.0, 0.0]), (n, 1)),
        "vel": np.zeros((n, 3)),
        "rot": np.zeros((n, 3)),
        "acc": np.tile(np.array([0.0, -9.78, 0.0]), (n, 1)),
        "inv_mass": np.zeros(n),
        "iit_body": np.zeros((n, 9)),
        "offset": np.tile(np.array([1.0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]), (n, 1)),
        "motion": np.full(n, 0.6),
        "lin_damp": np.full(n, 0.95),
        "ang_damp": np.full(n, 0.95),
        "flags": np.full(n, 1 | 2, np.uint8),
    }
    return f


def _random_quat(rng, n):
    g = rng.normal(4 * n).reshape(n, 4)
    return g / np.sqrt((g * g).sum(axis=1))[:, None]


def _finish(fields, shape, half, mass, pos, quat, vel, rot, gravity=-9.8):
    n = len(mass)
    fields["shape"][:] = shape
    fields["half"][:] = half
    fields["pos"][:] = pos
    fields["quat"][:] = quat
    fields["vel"][:] = vel
    fields["rot"][:] = rot
    fields["acc"][:] = np.array([0.0, gravity, 0.0])
    fields["inv_mass"][:] = 1.0 / mass
    cube = shape == 0
    iner = np.zeros((n, 9))
    if cube.any():
        iner[cube] = block_inertia(half[cube], mass[cube])
    sph = ~cube
    if sph.any():
        r = half[sph, 0]
        coeff = 0.4 * mass[sph] * r * r  # examples/ballistic.go:218
        iner[sph] = inertia_coeffs(coeff, coeff, coeff)
    fields["iit_body"][:] = go_invert3(iner)


def _ground(n_worlds):
    return (np.arange(n_worlds + 1, dtype=np.int32),
            np.tile(np.array([0.0, 1.0, 0.0]), (n_worlds, 1)), np.zeros(n_worlds))


def cubedrop(seed=0, n=64):
    """Plain method, 1 world: ground y=0 plus n cubes, half-extent 1, mass 8,
    x,z ~ U[-20,20], y ~ U[5,50], random unit quaternions, v = w = 0,
    g = (0,-9.8,0); dt = 1/60, 600 steps, maxIterations = 4 x contacts."""
    rng = SplitMix64(seed, stream=0)
    f = _new_bodies(n)
    pos = np.stack([rng.uniform(n, -20, 20), rng.uniform(n, 5, 50), rng.uniform(n, -20, 20)], axis=1)
    quat = _random_quat(rng, n)
    _finish(f, np.zeros(n, np.int32), np.ones((n, 3)), np.full(n, 8.0), pos, quat,
            np.zeros((n, 3)), np.zeros((n, 3)))
    po, pn, pof = _ground(1)
    return Scene(np.array([0, n], np.int32), po, pn, pof, f, name=f"cubedrop-{n}-s{seed}",
                 dt=1.0 / 60.0, steps=600, iter_mult=4)


def ballistic(seed=0, n_spheres=512, n_cubes=512):
    """Ballistic mix, 1 world: plane plus spheres (r ~ U[0.5,2]) and cubes
    (half-extents ~ U[0.5,2] per axis), density 1, uniform in a 100x40x100 box
    above the plane, v ~ N(0,10^2), w ~ N(0,2^2) per axis; 300 steps at 1/60."""
    rng = SplitMix64(seed, stream=1)
    n = n_spheres + n_cubes
    f = _new_bodies(n)
    shape = np.concatenate([np.ones(n_spheres, np.int32), np.zeros(n_cubes, np.int32)])
    r = rng.uniform(n_spheres, 0.5, 2.0)
    hc = rng.uniform(3 * n_cubes, 0.5, 2.0).reshape(n_cubes, 3)
    half = np.zeros((n, 3))
    half[:n_spheres, 0] = r
    half[n_spheres:] = hc
    mass = np.empty(n)
    mass[:n_spheres] = 4.0 / 3.0 * math.pi * r ** 3
    mass[n_spheres:] = 8.0 * hc[:, 0] * hc[:, 1] * hc[:, 2]
    ext = np.concatenate([r, hc.max(axis=1)])
    pos = np.stack([rng.uniform(n, -50, 50), ext + rng.uniform(n, 0, 40), rng.uniform(n, -50, 50)], axis=1)
    quat = _random_quat(rng, n)
    vel = rng.normal(3 * n, 10.0).reshape(n, 3)
    rot = rng.normal(3 * n, 2.0).reshape(n, 3)
    _finish(f, shape, half, mass, pos, quat, vel, rot)
    po, pn, pof = _ground(1)
    return Scene(np.array([0, n], np.int32), po, pn, pof, f, name=f"ballistic-{n}-s{seed}",
                 dt=1.0 / 60.0, steps=300, iter_mult=4)


def ballistic_example(n_bullets=12):
    """The scene of examples/ballistic.go, 1 world: the cube (half-extent 1,
    mass 8, at (0,5,0), default acceleration, :162-171), the backboard
    (half-extents (0.5,2,0.25) at (0,2,-10), infinite mass, never integrated:
    :180-185, and updateObjects :24-41 integrates only the cube and the
    bullets) and n_bullets bullets as fire() makes them (:205-227: spheres of
    radius 0.2, mass 1.5, velocity (0,0,-40), acceleration (0,-2.5,0)).  The
    example fires a bullet per key press from (0,1.5,20); here all bullets
    exist from step 0, spread across the backboard and the cube (x in
    {-0.9,-0.45,0,0.45,0.9}) and staggered in z (20 + 2.5 k), as if fired
    one after another.  Step it with the example's pair order
    (ballistic_example_pairs) and ResolveContacts(8 * contacts)."""
    n = 2 + n_bullets
    f = _new_bodies(n)  # NewRigidBody defaults: acceleration (0,-9.78,0)
    shape = np.array([0, 0] + [1] * n_bullets, np.int32)
    half = np.zeros((n, 3))
    half[0] = 1.0
    half[1] = (0.5, 2.0, 0.25)
    half[2:, 0] = 0.2
    pos = np.zeros((n, 3))
    pos[0] = (0.0, 5.0, 0.0)
    pos[1] = (0.0, 2.0, -10.0)
    k = np.arange(n_bullets)
    pos[2:, 0] = ((k % 5) - 2) * 0.45
    pos[2:, 1] = 1.5
    pos[2:, 2] = 20.0 + 2.5 * k
    f["shape"][:] = shape
    f["half"][:] = half
    f["pos"][:] = pos
    f["vel"][2:] = (0.0, 0.0, -40.0)
    f["acc"][2:] = (0.0, -2.5, 0.0)
    iner = np.zeros((n, 9))
    iner[0] = block_inertia(half[:1], np.array([8.0]))[0]
    coeff = 0.4 * 1.5 * 0.2 * 0.2
    iner[2:] = inertia_coeffs(np.full(n_bullets, coeff), np.full(n_bullets, coeff), np.full(n_bullets, coeff))
    f["inv_mass"][0] = 1.0 / 8.0
    f["inv_mass"][1] = 0.0  # SetInfiniteMass, rigidbody.go:130-134
    f["inv_mass"][2:] = 1.0 / 1.5
    f["iit_body"][:] = go_invert3(iner)  # the backboard's tensor is never set: zero
    f["iit_body"][1] = 0.0
    f["flags"][1] |= 4  # the caller never integrates it
    po, pn, pof = _ground(1)
    return Scene(np.array([0, n], np.int32), po, pn, pof, f, name=f"ballistic-example-{n_bullets}",
                 dt=1.0 / 60.0, steps=300, iter_mult=8)


def ballistic_example_pairs(n_bullets):
    """generateContacts of examples/ballistic.go:43-83 as an ordered pair list
    for cbz_narrowphase_pairs: (cube, ground), (cube, backboard), then per
    bullet b: (b, ground), (cube, b), (backboard, b), and (b2, b) for every
    other bullet b2.  Bodies: 0 cube, 1 backboard, 2.. bullets; -1 is the
    ground plane."""
    one, two = [0, 0], [-1, 1]
    bullets = list(range(2, 2 + n_bullets))
    for b in bullets:
        one += [b, 0, 1]
        two += [-1, b, b]
        for b2 in bullets:
            if b2 != b:
                one.append(b2)
                two.append(b)
    return np.array(one, np.int32), np.array(two, np.int32)


def batched(seed=0, n_worlds=4096, n_min=8, n_max=256):
    """Batched worlds: each a plane and n ~ U{n_min..n_max} cubes, half-extents
    ~ U[0.5,1.5]^3, mass ~ U[1,10], y ~ U[2,30], random quaternions, angular
    velocity N(0,1) per axis; x,z ~ U[-L,L] with L = 2 sqrt(n) (the
    configuration leaves x,z open; this keeps the density comparable across
    n).  600 steps at 1/60."""
    rng = SplitMix64(seed, stream=2)
    ns = rng.randint(n_worlds, n_min, n_max)
    body_off = np.zeros(n_worlds + 1, np.int32)
    body_off[1:] = np.cumsum(ns)
    n = int(body_off[-1])
    f = _new_bodies(n)
    L = np.repeat(2.0 * np.sqrt(ns.astype(np.float64)), ns)
    half = rng.uniform(3 * n, 0.5, 1.5).reshape(n, 3)
    mass = rng.uniform(n, 1.0, 10.0)
    pos = np.stack([L * rng.uniform(n, -1, 1), rng.uniform(n, 2, 30), L * rng.uniform(n, -1, 1)], axis=1)
    quat = _random_quat(rng, n)
    rot = rng.normal(3 * n, 1.0).reshape(n, 3)
    _finish(f, np.zeros(n, np.int32), half, mass, pos, quat, np.zeros((n, 3)), rot)
    po, pn, pof = _ground(n_worlds)
    return Scene(body_off, po, pn, pof, f, name=f"batched-{n_worlds}-s{seed}",
                 dt=1.0 / 60.0, steps=600, iter_mult=4)


def towers(seed=0, n_worlds=16384, pitch=4.0):
    """Stacked resting contacts: each world a plane with 4-16 towers of 4-12
    cubes (half-extent 1, mass 5) at y = 1 + 2k, x,z jitter N(0,0.05^2), yaw
    N(0,0.1^2); towers on a square grid of the given pitch (the configuration
    leaves the spacing open).  1200 steps at 1/120."""
    rng = SplitMix64(seed, stream=4)
    n_towers = rng.randint(n_worlds, 4, 16)
    heights = rng.randint(int(n_towers.sum()), 4, 12)
    ns = np.zeros(n_worlds, np.int64)
    t0 = 0
    for w in range(n_worlds):
        ns[w] = heights[t0:t0 + n_towers[w]].sum()
        t0 += n_towers[w]
    body_off = np.zeros(n_worlds + 1, np.int32)
    body_off[1:] = np.cumsum(ns)
    n = int(body_off[-1])
    f = _new_bodies(n)
    pos = np.zeros((n, 3))
    yaw = np.zeros(n)
    b = 0
    t0 = 0
    jit = rng.normal(2 * n, 0.05).reshape(n, 2)
    yj = rng.normal(n, 0.1)
    for w in range(n_worlds):
        nt = int(n_towers[w])
        side = int(math.ceil(math.sqrt(nt)))
        for t in range(nt):
            gx = (t % side - (side - 1) / 2.0) * pitch
            gz = (t // side - (side - 1) / 2.0) * pitch
            for k in range(int(heights[t0 + t])):
                pos[b] = (gx + jit[b, 0], 1.0 + 2.0 * k, gz + jit[b, 1])
                yaw[b] = yj[b]
                b += 1
        t0 += nt
    quat = np.stack([np.cos(yaw / 2), np.zeros(n), np.sin(yaw / 2), np.zeros(n)], axis=1)
    _finish(f, np.zeros(n, np.int32), np.ones((n, 3)), np.full(n, 5.0), pos, quat,
            np.zeros((n, 3)), np.zeros((n, 3)))
    po, pn, pof = _ground(n_worlds)
    tops = np.cumsum(heights[:int(n_towers.sum())]) - 1  # the top cube of every tower
    return Scene(body_off, po, pn, pof, f, name=f"towers-{n_worlds}-s{seed}",
                 dt=1.0 / 120.0, steps=1200, iter_mult=4, meta={"tower_top": tops})
