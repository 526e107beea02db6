// k_narrow.hip — K2: the narrowphase over ordered pairs, with contacts
// compacted in the reference's emission order.
//
// Emission order (examples/cubedrop.go:47-64): for each body i of a world,
// first its half-space checks (planes in order, up to 8 cube vertices each
// in the `mults` order of colliders.go:319-327), then CheckForCollisions(i, j)
// for every j != i in ascending order.  Row r of a world has np + n columns
// (planes, then bodies; column j == r is never a pair), cut into 64-column
// tiles; tiles are ordered row by row, so scanning their contact totals in
// tile order gives every contact its slot in the reference's order.  An
// explicit pair list (any caller loop) uses its own order instead, 64
// consecutive entries per tile.
//
// Launches (all-pairs):
//   cull        — one wave per (column tile, 16 rows): clears the tiles and
//                 lists the candidate pairs.  Cube-cube and cube-sphere
//                 pairs whose bounding spheres are far apart are rejected:
//                 the bound guarantees that the reference's own test reports
//                 no contact (see far_apart below), so the result is
//                 unchanged.
//   eval        — one thread per candidate: the reference's test; per tile
//                 a hit bitmask and contact total, and the hit list.
//   scan_tiles  — one wave per world: exclusive scan of tile totals.
//   scan_worlds — one block: exclusive scan of world totals.
//   emit_hits   — one thread per hit pair: recompute it and write its
//                 contacts at world offset + tile offset + in-tile prefix.
// (The explicit-list form counts and emits one wave per tile.)
// All geometry is evaluated in FP64 whatever the storage type R.
#include "cbz_devmath.h"
#include "cbz_launch.h"
#include "cbz_types.h"
#include "cbz_wave.h"

namespace cbz {

using D3 = V3<double>;

struct Col {
  double T[12];
  double h[3];
  int shape;
};

template <typename R>
__device__ __forceinline__ void load_col(const Bodies<R>& b, int i, Col& c) {
  const int64_t S = b.stride;
#pragma unroll
  for (int k = 0; k < 12; ++k) c.T[k] = double(b.ctransform[k * S + i]);
  c.h[0] = double(b.half[i]);
  c.h[1] = double(b.half[S + i]);
  c.h[2] = double(b.half[2 * S + i]);
  c.shape = b.shape[i];
}

template <typename R>
struct NP {
  Bodies<R> b;
  Planes<R> pl;
  PairSpace ps;
  Contacts<R> c;
  const double* brad;  // conservative bounding radius per body
  int32_t* tile_cnt;
  uint64_t* tile_mask;
  int32_t* tile_off;
  int32_t* world_cnt;
  int64_t* world_coff;
  int64_t* total;
  int32_t* overflow;
  uint32_t* hits;
  int32_t* n_hits;
  uint32_t* cand;
  int32_t* n_cand;
  int64_t cand_cap;
  int32_t n_worlds;
  Materials<R> mat;  // read when c.fric is set
};

// Writes one contact.  m1 names the contact's second material: body b1 for
// a body pair, -(k+1) for plane k (Bodies[1] is nil there).  With materials
// on, the contact's friction and restitution are the means of body b0's and
// m1's values (the reference gives every contact 0.9 and 0.1,
// colliders.go:201-202, 248-249, 360-361, 437-438, 511-512, 704-705).
template <typename R>
__device__ __forceinline__ void put_contact(const NP<R>& a, int64_t slot, int b0, int b1, int m1, const D3& p,
                                            const D3& n, double pen) {
  const Contacts<R>& c = a.c;
  if (slot >= c.cap) return;
  const int64_t S = c.cap;
  c.body0[slot] = b0;
  c.body1[slot] = b1;
  c.point[slot] = R(p.x);
  c.point[S + slot] = R(p.y);
  c.point[2 * S + slot] = R(p.z);
  c.normal[slot] = R(n.x);
  c.normal[S + slot] = R(n.y);
  c.normal[2 * S + slot] = R(n.z);
  c.pen[slot] = R(pen);
  if (c.fric) {
    const Materials<R>& m = a.mat;
    const R f1 = m1 >= 0 ? m.body_fric[m1] : m.plane_fric[-m1 - 1];
    const R e1 = m1 >= 0 ? m.body_rest[m1] : m.plane_rest[-m1 - 1];
    c.fric[slot] = (m.body_fric[b0] + f1) * R(0.5);
    c.rest[slot] = (m.body_rest[b0] + e1) * R(0.5);
  }
}

// GetAxis / HalfSize[i] with a run-time index, re-read from the body arrays
// (a dynamic index into a register array would put the whole collider in
// scratch memory).
template <typename R>
__device__ __forceinline__ D3 axis_g(const NP<R>& a, int body, int col) {
  const int c = col > 3 ? 3 : col;
  const int64_t S = a.b.stride;
  return D3{double(a.b.ctransform[(3 * c) * S + body]), double(a.b.ctransform[(3 * c + 1) * S + body]),
            double(a.b.ctransform[(3 * c + 2) * S + body])};
}
template <typename R>
__device__ __forceinline__ double half_g(const NP<R>& a, int body, int k) {
  return double(a.b.half[k * a.b.stride + body]);
}

// transformToAxis, colliders.go:762-770
__device__ __forceinline__ double transform_to_axis(const Col& c, const D3& a) {
  return c.h[0] * dabs(dot(a, axis(c.T, 0))) + c.h[1] * dabs(dot(a, axis(c.T, 1))) +
         c.h[2] * dabs(dot(a, axis(c.T, 2)));
}

// CollisionSphere.CheckAgainstHalfSpace, colliders.go:180-207
template <typename R, bool EMIT>
__device__ __forceinline__ int sphere_half_space(const NP<R>& a, const Col& s, int si, int pk, const D3& pn, double po,
                                                 int64_t slot) {
  D3 positionAxis = axis(s.T, 3);
  double r = s.h[0];
  double distance = dot(pn, positionAxis) - r;
  if ((distance <= po) == false) return 0;
  if (EMIT) {
    D3 p = add(mul(pn, distance + r * -1.0), positionAxis);
    put_contact(a, slot, si, -1, -pk - 1, p, pn, -distance);
  }
  return 1;
}

// CollisionCube.CheckAgainstHalfSpace, colliders.go:308-366, with
// intersectCubeAndHalfSpace (:750-760) as the early-out.
template <typename R, bool EMIT>
__device__ __forceinline__ int cube_half_space(const NP<R>& a, const Col& c, int ci, int pk, const D3& pn, double po,
                                               int64_t slot) {
  double projectedRadius = transform_to_axis(c, pn);
  double cubeDistance = dot(pn, axis(c.T, 3)) - projectedRadius;
  if (!(cubeDistance <= po)) return 0;
  int k = 0;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    // mults[v] = (+-1, +-1, +-1): x flips on odd v, y on v&2, z on v&4
    D3 vert{(v & 1) ? -c.h[0] : c.h[0], (v & 2) ? -c.h[1] : c.h[1], (v & 4) ? -c.h[2] : c.h[2]};
    D3 vertexPos = m34v(c.T, vert);
    double vertexDistance = dot(vertexPos, pn);
    if (vertexDistance <= po) {
      if (EMIT) {
        D3 p = add(mul(pn, vertexDistance - po), vertexPos);
        put_contact(a, slot + k, ci, -1, -pk - 1, p, pn, po - vertexDistance);
      }
      ++k;
    }
  }
  return k;
}

// CollisionSphere.CheckAgainstSphere, colliders.go:216-254
template <typename R, bool EMIT>
__device__ __forceinline__ int sphere_sphere(const NP<R>& a, const Col& s1, int i1, const Col& s2, int i2, int64_t slot) {
  D3 p1 = axis(s1.T, 3), p2 = axis(s2.T, 3);
  D3 midline = sub(p1, p2);
  double size = mag(midline);
  if (size <= 0.0 || size >= s1.h[0] + s2.h[0]) return 0;
  if (EMIT) {
    D3 p = add(mul(midline, 0.5), p1);
    D3 n = mul(midline, 1.0 / size);
    put_contact(a, slot, i1, i2, i2, p, n, s1.h[0] + s2.h[0] - size);
  }
  return 1;
}

// CollisionCube.CheckAgainstSphere, colliders.go:369-441
template <typename R, bool EMIT>
__device__ __forceinline__ int cube_sphere(const NP<R>& a, const Col& c, int ci, const Col& s, int si, int64_t slot) {
  const double r = s.h[0];
  D3 position = axis(s.T, 3);
  D3 rc = m34_tinv(c.T, position);
  if (dabs(rc.x) - r > c.h[0] || dabs(rc.y) - r > c.h[1] || dabs(rc.z) - r > c.h[2]) return 0;
  double cp[3];
  const double rcv[3] = {rc.x, rc.y, rc.z};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double dist = rcv[k];
    if (dist > c.h[k]) {
      dist = c.h[k];
    } else if (dist < -c.h[k]) {
      dist = -c.h[k];
    }
    cp[k] = dist;
  }
  D3 closest{cp[0], cp[1], cp[2]};
  double dist = sqmag(sub(closest, rc));
  if (dist > r * r) return 0;
  if (EMIT) {
    D3 cw = m34v(c.T, closest);
    D3 n = sub(cw, position);
    if (real_equal(mag(n), 0.0)) {
      const int64_t S = a.b.stride;
      n = D3{double(a.b.vel[si]), double(a.b.vel[S + si]), double(a.b.vel[2 * S + si])};
    }
    n = normalize(n);
    double pen = r;
    if (!real_equal(dist, 0.0)) {
      pen -= dsqrt(dist);
    } else {
      pen = 0.0;
    }
    put_contact(a, slot, ci, si, si, cw, n, pen);
  }
  return 1;
}

// tryAxis, colliders.go:458-477 with penetrationOnAxis (:445-456)
__device__ __forceinline__ bool try_axis(const Col& one, const Col& two, D3 ax, const D3& toCenter, int index,
                                         double& smallest, int& best) {
  if (sqmag(ax) < Lim<double>::eps) return true;
  ax = normalize(ax);
  double oneProject = transform_to_axis(one, ax);
  double twoProject = transform_to_axis(two, ax);
  double distance = dabs(dot(toCenter, ax));
  double penetration = oneProject + twoProject - distance;
  if (penetration < 0) return false;
  if (penetration < smallest) {
    smallest = penetration;
    best = index;
  }
  return true;
}

// fillPointFaceBoxBox, colliders.go:481-517 (i1 = one's body index)
template <typename R>
__device__ __forceinline__ void point_face(const NP<R>& a, const Col& one, int i1, const Col& two,
                                           const D3& toCenter, int best, D3& n, D3& p) {
  n = axis_g(a, i1, best);
  if (dot(n, toCenter) > 0) n = mul(n, -1.0);
  D3 v{two.h[0], two.h[1], two.h[2]};
  if (dot(axis(two.T, 0), n) < 0) v.x = -v.x;
  if (dot(axis(two.T, 1), n) < 0) v.y = -v.y;
  if (dot(axis(two.T, 2), n) < 0) v.z = -v.z;
  p = m34v(two.T, v);
}

// contactPoint, colliders.go:519-573
__device__ __forceinline__ D3 contact_point(const D3& pOne, const D3& dOne, double oneSize, const D3& pTwo,
                                            const D3& dTwo, double twoSize, bool useOne) {
  double smOne = sqmag(dOne);
  double smTwo = sqmag(dTwo);
  double dpOneTwo = dot(dTwo, dOne);
  D3 toSt = sub(pOne, pTwo);
  double dpStaOne = dot(dOne, toSt);
  double dpStaTwo = dot(dTwo, toSt);
  double denom = smOne * smTwo - dpOneTwo * dpOneTwo;
  if (dabs(denom) < Lim<double>::eps) return useOne ? pOne : pTwo;
  double mua = (dpOneTwo * dpStaTwo - smTwo * dpStaOne) / denom;
  double mub = (smOne * dpStaTwo - dpOneTwo * dpStaOne) / denom;
  if (mua > oneSize || mua < -oneSize || mub > twoSize || mub < -twoSize) return useOne ? pOne : pTwo;
  D3 cOne = add(mul(dOne, mua), pOne);
  D3 cTwo = add(mul(dTwo, mub), pTwo);
  cOne = mul(cOne, 0.5);
  cTwo = mul(cTwo, 0.5);
  return add(cOne, cTwo);
}

// CollisionCube.CheckAgainstCube, colliders.go:576-710
template <typename R, bool EMIT>
__device__ __forceinline__ int cube_cube(const NP<R>& a, const Col& one, int i1, const Col& two, int i2, int64_t slot) {
  D3 toCenter = sub(axis(two.T, 3), axis(one.T, 3));
  double pen = Lim<double>::max_value;
  int best = 0xffffff;
  for (int i = 0; i <= 2; ++i)
    if (!try_axis(one, two, axis(one.T, i), toCenter, i, pen, best)) return 0;
  for (int i = 0; i <= 2; ++i)
    if (!try_axis(one, two, axis(two.T, i), toCenter, i + 3, pen, best)) return 0;
  const int bestSingleAxis = best;
  for (int i = 0; i <= 2; ++i)
    for (int k = 0; k <= 2; ++k)
      if (!try_axis(one, two, cross(axis(one.T, i), axis(two.T, k)), toCenter, i * 3 + 6 + k, pen, best)) return 0;
  if (!EMIT) return 1;
  D3 n, p;
  if (best < 3) {
    point_face(a, one, i1, two, toCenter, best, n, p);
    put_contact(a, slot, i1, i2, i2, p, n, pen);
  } else if (best < 6) {
    point_face(a, two, i2, one, mul(toCenter, -1.0), best - 3, n, p);
    put_contact(a, slot, i2, i1, i1, p, n, pen);
  } else {
    best -= 6;
    const int oneAxisIndex = best / 3;
    const int twoAxisIndex = best % 3;
    D3 oneAxis = axis_g(a, i1, oneAxisIndex);
    D3 twoAxis = axis_g(a, i2, twoAxisIndex);
    D3 ax = normalize(cross(oneAxis, twoAxis));
    if (dot(ax, toCenter) > 0) ax = mul(ax, -1.0);
    double p1[3] = {one.h[0], one.h[1], one.h[2]};
    double p2[3] = {two.h[0], two.h[1], two.h[2]};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (i == oneAxisIndex) {
        p1[i] = 0;
      } else if (dot(axis(one.T, i), ax) > 0) {
        p1[i] = -p1[i];
      }
      if (i == twoAxisIndex) {
        p2[i] = 0;
      } else if (dot(axis(two.T, i), ax) < 0) {
        p2[i] = -p2[i];
      }
    }
    D3 ptOnOneEdge = m34v(one.T, D3{p1[0], p1[1], p1[2]});
    D3 ptOnTwoEdge = m34v(two.T, D3{p2[0], p2[1], p2[2]});
    const bool useOne = bestSingleAxis > 2;
    // an index past 2 only arises from a NaN state; GetAxis clamps it to 3
    const double oneSize = oneAxisIndex < 3 ? half_g(a, i1, oneAxisIndex) : 0.0;
    const double twoSize = twoAxisIndex < 3 ? half_g(a, i2, twoAxisIndex) : 0.0;
    D3 cv = contact_point(ptOnOneEdge, oneAxis, oneSize, ptOnTwoEdge, twoAxis, twoSize, useOne);
    put_contact(a, slot, i1, i2, i2, cv, ax, pen);
  }
  return 1;
}

// Conservative rejection for body pairs.  For orthonormal box axes a_k,
// max_k |t.a_k| >= |t|/sqrt(3), while the projections of the two shapes on
// a_k are bounded by their bounding radii; so |t|^2 > 3 (rA+rB)^2 makes a
// face axis of A separate (tryAxis returns false) and a sphere centre
// farther than rA + rB from a box centre fails the closest-point test.
// The factors carry a 1e-3 margin for rounding and for the 1e-7 slack the
// quaternion normalisation leaves in the axes.
template <typename R>
__device__ __forceinline__ bool far_apart(const NP<R>& a, int i, int j, int si, int sj) {
  const int64_t S = a.b.stride;
  double dx = double(a.b.ctransform[9 * S + j]) - double(a.b.ctransform[9 * S + i]);
  double dy = double(a.b.ctransform[10 * S + j]) - double(a.b.ctransform[10 * S + i]);
  double dz = double(a.b.ctransform[11 * S + j]) - double(a.b.ctransform[11 * S + i]);
  double d2 = dx * dx + dy * dy + dz * dz;
  double rr = a.brad[i] + a.brad[j];
  double k = (si == SHAPE_CUBE && sj == SHAPE_CUBE) ? 3.003 : 1.001;
  return d2 > k * rr * rr;
}

// CheckForCollisions, colliders.go:720-747, including the plane delegations
// (:116-125) and the sphere-to-cube delegation (:210-213).
template <typename R, bool EMIT>
__device__ __forceinline__ int pair_eval(const NP<R>& a, int one, int two, int64_t slot) {
  if (one < 0 && two < 0) return 0;
  if (one < 0 || two < 0) {
    const int bi = one < 0 ? two : one;
    const int pk = one < 0 ? -one - 1 : -two - 1;
    const int64_t PS = a.pl.stride;
    D3 pn{double(a.pl.normal[pk]), double(a.pl.normal[PS + pk]), double(a.pl.normal[2 * PS + pk])};
    double po = double(a.pl.offset[pk]);
    Col c;
    load_col(a.b, bi, c);
    if (c.shape == SHAPE_SPHERE) return sphere_half_space<R, EMIT>(a, c, bi, pk, pn, po, slot);
    return cube_half_space<R, EMIT>(a, c, bi, pk, pn, po, slot);
  }
  const int s1 = a.b.shape[one], s2 = a.b.shape[two];
  if (!(s1 == SHAPE_SPHERE && s2 == SHAPE_SPHERE) && far_apart(a, one, two, s1, s2)) return 0;
  Col c1, c2;
  load_col(a.b, one, c1);
  load_col(a.b, two, c2);
  if (s2 == SHAPE_SPHERE) {
    if (s1 == SHAPE_SPHERE) return sphere_sphere<R, EMIT>(a, c1, one, c2, two, slot);
    return cube_sphere<R, EMIT>(a, c1, one, c2, two, slot);
  }
  if (s1 == SHAPE_SPHERE) return cube_sphere<R, EMIT>(a, c2, two, c1, one, slot);
  return cube_cube<R, EMIT>(a, c1, one, c2, two, slot);
}

// Entry q of world w's explicit pair list -> (one, two).
__device__ __forceinline__ void pair_of(const PairSpace& ps, int w, int q, int& one, int& two) {
  const int p = ps.list_off[w] + q;
  one = ps.list_one[p];
  two = ps.list_two[p];
}

// ---------------------------------------------------------------------------
// all-pairs space
// ---------------------------------------------------------------------------
// Block-wide append to a global list: every thread of the block calls it
// with its entry count; returns the list index of the thread's first entry.
// One atomic per block (a single shared counter hit by every wave of the
// grid would serialise in the L2).
__device__ __forceinline__ int64_t block_reserve(int mine, int32_t* counter, int* s_wtot, int* s_base) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int incl = wave_incl_scan_i32(mine);
  if (lane == 63) s_wtot[wv] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int u = 0; u < nw; ++u) tot += s_wtot[u];
    *s_base = tot > 0 ? atomicAdd(counter, tot) : 0;
  }
  __syncthreads();
  int64_t base = *s_base;
  for (int u = 0; u < wv; ++u) base += s_wtot[u];
  __syncthreads();  // s_wtot / s_base are reused by the next call
  return base + incl - mine;
}

// tasks per wave of the cull: their candidates share one block reservation
constexpr int kCullTasks = 4;

// Cull: each wave takes kCullTasks consecutive tasks (world w, column tile
// c, rows r0 .. r0+15).  Lane l owns column 64c + l for the whole task: a plane, or body j whose centre,
// bounding radius and shape stay in registers; the 16 row bodies are loaded
// once (lane k holds row r0 + k) and broadcast.  Each tile is cleared here;
// its plane pairs and the body pairs that pass the conservative bound (see
// far_apart) go to the candidate list.
template <typename R>
__global__ void __launch_bounds__(256) k_np_cull(NP<R> a) {
  __shared__ int s_wtot[4];
  __shared__ int s_base;
  const int lane = threadIdx.x & 63;
  const int task0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * kCullTasks;
  // bit 16(q%4) + k of word q/4: row r0(q) + k of the wave's task q is a
  // candidate for this lane
  uint64_t cm[(kCullTasks + 3) / 4];
#pragma unroll
  for (int u = 0; u < (kCullTasks + 3) / 4; ++u) cm[u] = 0;
  int t0[kCullTasks], r0[kCullTasks], nch[kCullTasks], cc[kCullTasks];
#pragma unroll
  for (int q = 0; q < kCullTasks; ++q) {
    t0[q] = 0;
    r0[q] = 0;
    nch[q] = 1;
    cc[q] = 0;
    const int task = task0 + q;
    if (task >= a.ps.n_tasks) continue;  // the wave still joins the block's reservation
    const int w = a.ps.task_world[task];
    const int b0 = a.ps.body_off[w], n = a.ps.body_off[w + 1] - b0;
    const int p0 = a.ps.plane_off[w], np = a.ps.plane_off[w + 1] - p0;
    nch[q] = (np + n + 63) >> 6;
    const int k = task - a.ps.world_task0[w];
    const int c = k % nch[q];
    cc[q] = c;
    r0[q] = (k / nch[q]) * kRowBlock;
    const int r1 = min(n, r0[q] + kRowBlock);
    const int col = 64 * c + lane;
    const bool is_plane = col < np;
    const int j = col - np;
    const bool is_body = !is_plane && j < n;
    const int64_t S = a.b.stride;
    double tj[3] = {0, 0, 0}, rj = 0;
    int sj = 0;
    if (is_body) {
      const int g = b0 + j;
      tj[0] = double(a.b.ctransform[9 * S + g]);
      tj[1] = double(a.b.ctransform[10 * S + g]);
      tj[2] = double(a.b.ctransform[11 * S + g]);
      rj = a.brad[g];
      sj = a.b.shape[g];
    }
    // row bodies r0 .. r1-1: lane k < 16 loads row r0 + k
    double ti[3] = {0, 0, 0}, ri = 0;
    int si = 0;
    if (lane < r1 - r0[q]) {
      const int g = b0 + r0[q] + lane;
      ti[0] = double(a.b.ctransform[9 * S + g]);
      ti[1] = double(a.b.ctransform[10 * S + g]);
      ti[2] = double(a.b.ctransform[11 * S + g]);
      ri = a.brad[g];
      si = a.b.shape[g];
    }
    t0[q] = a.ps.world_tile0[w];
    for (int r = r0[q]; r < r1; ++r) {
      const int src = r - r0[q];
      const double x = __shfl(ti[0], src, 64), y = __shfl(ti[1], src, 64), z = __shfl(ti[2], src, 64);
      const double rr = __shfl(ri, src, 64) + rj;
      const int s_i = __shfl(si, src, 64);
      bool cand = is_plane;
      if (is_body && j != r) {
        const double dx = tj[0] - x, dy = tj[1] - y, dz = tj[2] - z;
        const double d2 = dx * dx + dy * dy + dz * dz;
        const bool spheres = s_i == SHAPE_SPHERE && sj == SHAPE_SPHERE;
        const double kf = (s_i == SHAPE_CUBE && sj == SHAPE_CUBE) ? 3.003 : 1.001;
        cand = spheres || !(d2 > kf * rr * rr);
      }
      if (cand) cm[q >> 2] |= 1ull << (16 * (q & 3) + src);
      if (lane == 0) {
        const int t = t0[q] + r * nch[q] + c;
        a.tile_cnt[t] = 0;
        a.tile_mask[t] = 0ull;
      }
    }
  }
  // one reservation per block for all of its waves' tasks
  int ncand = 0;
#pragma unroll
  for (int u = 0; u < (kCullTasks + 3) / 4; ++u) ncand += __popcll(cm[u]);
  int64_t pos = block_reserve(ncand, a.n_cand, s_wtot, &s_base);
#pragma unroll
  for (int q = 0; q < kCullTasks; ++q) {
    for (uint32_t m = uint32_t(cm[q >> 2] >> (16 * (q & 3))) & 0xFFFFu; m; m &= m - 1u, ++pos) {
      const int r = r0[q] + __ffs(m) - 1;
      if (pos < a.cand_cap) a.cand[pos] = uint32_t(t0[q] + r * nch[q] + cc[q]) * 64u + uint32_t(lane);
    }
  }
}

// Eval: one thread per candidate pair: the reference's test (contact count
// only); hits add to their tile's total and mask (integer atomics: the sums
// do not depend on the order) and join the hit list.  Too many candidates
// for the list raise the overflow flags (the step reports a capacity error).
template <typename R>
__global__ void __launch_bounds__(256) k_np_eval(NP<R> a) {
  __shared__ int s_wtot[4];
  __shared__ int s_base;
  const int64_t nc = a.n_cand[0];
  if (nc > a.cand_cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      a.overflow[0] = 1;
      a.overflow[1] = 1;
    }
    return;  // uniform over the grid
  }
  for (int64_t e0 = int64_t(blockIdx.x) * blockDim.x; e0 < nc; e0 += int64_t(gridDim.x) * blockDim.x) {
    const int64_t e = e0 + threadIdx.x;
    int cnt = 0;
    uint32_t code = 0;
    if (e < nc) {
      code = a.cand[e];
      const int t = int(code >> 6), l = int(code & 63u);
      const int w = a.ps.tile_world[t];
      const int b0 = a.ps.body_off[w], n = a.ps.body_off[w + 1] - b0;
      const int p0 = a.ps.plane_off[w], np = a.ps.plane_off[w + 1] - p0;
      const int nch = (np + n + 63) >> 6;
      const int k = t - a.ps.world_tile0[w];
      const int r = k / nch, c = k - r * nch;
      const int col = 64 * c + l;
      const int i = b0 + r;
      cnt = col < np ? pair_eval<R, false>(a, i, -(p0 + col) - 1, 0) : pair_eval<R, false>(a, i, b0 + (col - np), 0);
      if (cnt > 0) {
        atomicAdd(&a.tile_cnt[t], cnt);
        atomicOr((unsigned long long*)&a.tile_mask[t], 1ull << l);
      }
    }
    const int64_t idx = block_reserve(cnt > 0 ? 1 : 0, a.n_hits, s_wtot, &s_base);
    if (cnt > 0 && idx < a.c.cap) a.hits[idx] = code;
  }
}

// Emit: one thread per hit pair.  Its first contact goes to world offset +
// tile offset + the contacts of the tile's earlier hit pairs: one for a body
// pair (a body pair that hit holds exactly one contact), a recount for a
// plane column (up to 8 cube vertices).
template <typename R>
__global__ void __launch_bounds__(256) k_np_emit_hits(NP<R> a) {
  if (a.overflow[0] != 0) return;
  const int H = min(a.n_hits[0], int(min(a.c.cap, int64_t(INT32_MAX))));
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < H; e += gridDim.x * blockDim.x) {
    const uint32_t code = a.hits[e];
    const int t = int(code >> 6), lane = int(code & 63u);
    const int w = a.ps.tile_world[t];
    const int b0 = a.ps.body_off[w], n = a.ps.body_off[w + 1] - b0;
    const int p0 = a.ps.plane_off[w], np = a.ps.plane_off[w + 1] - p0;
    const int nch = (np + n + 63) >> 6;
    const int k = t - a.ps.world_tile0[w];
    const int r = k / nch, c = k - r * nch;
    const int i = b0 + r;
    const uint64_t below = a.tile_mask[t] & ((1ull << lane) - 1ull);
    // lanes of this tile that are plane columns
    const int np_here = min(max(np - 64 * c, 0), 64);
    const uint64_t plane_lanes = np_here >= 64 ? ~0ull : ((1ull << np_here) - 1ull);
    int64_t prefix = __popcll(below & ~plane_lanes);
    for (uint64_t pb = below & plane_lanes; pb; pb &= pb - 1ull) {
      const int l = __ffsll((unsigned long long)pb) - 1;
      prefix += pair_eval<R, false>(a, i, -(p0 + 64 * c + l) - 1, 0);
    }
    const int64_t slot = a.world_coff[w] + a.tile_off[t] + prefix;
    const int col = 64 * c + lane;
    if (col < np)
      pair_eval<R, true>(a, i, -(p0 + col) - 1, slot);
    else
      pair_eval<R, true>(a, i, b0 + (col - np), slot);
  }
}

template <typename R>
__global__ void __launch_bounds__(256) k_np_count(NP<R> a) {
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= a.ps.n_tiles) return;
  const int w = a.ps.tile_world[t];
  const int q = (t - a.ps.world_tile0[w]) * 64 + lane;
  int cnt = 0;
  if (q < a.ps.world_q[w]) {
    int one, two;
    pair_of(a.ps, w, q, one, two);
    cnt = pair_eval<R, false>(a, one, two, 0);
    if (a.ps.list_cnt) a.ps.list_cnt[a.ps.list_off[w] + q] = cnt;
  }
  const uint64_t m = __ballot(cnt > 0);
  const int tot = wave_sum_i32(cnt);
  if (lane == 0) {
    a.tile_cnt[t] = tot;
    a.tile_mask[t] = m;
  }
}

// above this many tiles in one world the 16-wave tile scan is launched
constexpr int kWideScanTiles = 4096;

template <typename R>
__global__ void __launch_bounds__(64) k_np_scan_tiles(NP<R> a) {
  const int w = blockIdx.x;
  const int lane = threadIdx.x;
  const int t0 = a.ps.world_tile0[w], nt = a.ps.world_tile0[w + 1] - t0;
  int carry = 0;
  for (int base = 0; base < nt; base += 64) {
    const int k = base + lane;
    const int v = k < nt ? a.tile_cnt[t0 + k] : 0;
    const int inc = wave_incl_scan_i32(v);
    if (k < nt) a.tile_off[t0 + k] = carry + inc - v;
    carry += __shfl(inc, 63, 64);
  }
  if (lane == 0) a.world_cnt[w] = carry;
}

// The same scan for worlds of many tiles (thousands of bodies): 16 waves,
// each thread summing kScanPer consecutive tiles, take 8,192 tiles per
// pass, so the serial carry chain is 128x shorter than the one-wave scan's.
// Integer prefix sums, so the offsets equal the one-wave scan's.
constexpr int kScanPer = 8;
template <typename R>
__global__ void __launch_bounds__(1024) k_np_scan_tiles_wide(NP<R> a) {
  __shared__ int32_t wsum[16];
  __shared__ int32_t s_carry;
  const int w = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int t0 = a.ps.world_tile0[w], nt = a.ps.world_tile0[w + 1] - t0;
  if (tid == 0) s_carry = 0;
  __syncthreads();
  for (int base = 0; base < nt; base += 1024 * kScanPer) {
    const int k0 = base + kScanPer * tid;
    int v[kScanPer];
    int sum = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
      v[j] = k0 + j < nt ? a.tile_cnt[t0 + k0 + j] : 0;
      sum += v[j];
    }
    const int inc = wave_incl_scan_i32(sum);
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    int before = 0;
    for (int u = 0; u < wv; ++u) before += wsum[u];
    const int carry = s_carry;
    int run = carry + before + inc - sum;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
      if (k0 + j < nt) a.tile_off[t0 + k0 + j] = run;
      run += v[j];
    }
    __syncthreads();
    if (tid == 1023) s_carry = carry + before + inc;
    __syncthreads();
  }
  if (tid == 0) a.world_cnt[w] = s_carry;
}

template <typename R>
__global__ void __launch_bounds__(1024) k_np_scan_worlds(NP<R> a) {
  __shared__ int64_t wsum[16];
  __shared__ int64_t s_carry;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid == 0) s_carry = 0;
  __syncthreads();
  for (int base = 0; base < a.n_worlds; base += 1024) {
    const int k = base + tid;
    const int64_t v = k < a.n_worlds ? int64_t(a.world_cnt[k]) : 0;
    const int64_t inc = wave_incl_scan_i64(v);
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    int64_t before = 0;
    for (int u = 0; u < wv; ++u) before += wsum[u];
    const int64_t carry = s_carry;
    if (k < a.n_worlds) a.world_coff[k] = carry + before + inc - v;
    __syncthreads();
    if (tid == 1023) s_carry = carry + before + inc;
    __syncthreads();
  }
  if (tid == 0) {
    a.total[0] = s_carry;
    if (s_carry > a.c.cap) {
      a.overflow[0] = 1;  // this narrowphase (read by emit and the resolver)
      a.overflow[1] = 1;  // sticky until the host clears it
    }
  }
}

template <typename R>
__global__ void __launch_bounds__(256) k_np_emit(NP<R> a) {
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= a.ps.n_tiles) return;
  const uint64_t m = a.tile_mask[t];
  if (m == 0) return;
  const int w = a.ps.tile_world[t];
  const int q = (t - a.ps.world_tile0[w]) * 64 + lane;
  const bool hit = (m >> lane) & 1ull;
  int one = 0, two = 0, cnt = 0;
  if (hit) {
    pair_of(a.ps, w, q, one, two);
    // a body pair that hit holds exactly one contact; only a half-space
    // check can hold several (up to 8 cube vertices)
    cnt = (one < 0 || two < 0) ? pair_eval<R, false>(a, one, two, 0) : 1;
  }
  const int inc = wave_incl_scan_i32(cnt);
  const int64_t slot = a.world_coff[w] + a.tile_off[t] + (inc - cnt);
  if (hit) pair_eval<R, true>(a, one, two, slot);
}

template <typename R>
hipError_t launch_narrowphase(const NarrowLaunch<R>& L, hipStream_t st, hipEvent_t* ev) {
  NP<R> a;
  a.b = L.b;
  a.pl = L.pl;
  a.ps = L.ps;
  a.c = L.c;
  a.brad = L.brad;
  a.tile_cnt = L.tile_cnt;
  a.tile_mask = L.tile_mask;
  a.tile_off = L.tile_off;
  a.world_cnt = L.world_cnt;
  a.world_coff = L.world_coff;
  a.total = L.total;
  a.overflow = L.overflow;
  a.hits = L.hits;
  a.n_hits = L.n_hits;
  a.cand = L.cand;
  a.n_cand = L.n_cand;
  a.cand_cap = L.cand_cap;
  a.n_worlds = L.n_worlds;
  a.mat = L.mat;
  if (a.n_worlds <= 0) return hipSuccess;
  if (a.c.fric && (!a.c.rest || !a.mat.body_fric || !a.mat.body_rest || !a.mat.plane_fric || !a.mat.plane_rest))
    return hipErrorInvalidValue;
  const bool list = a.ps.list_one != nullptr;
  if (!list && (!a.hits || !a.n_hits || !a.cand || !a.n_cand)) return hipErrorInvalidValue;
  // ev (optional): recorded before the count, after it, after the scans
  // and after the emit, for per-kernel timing on this stream.
  const int nb = (a.ps.n_tiles + 3) / 4;
  if (ev) (void)hipEventRecord(ev[0], st);
  if (list) {
    if (nb > 0) k_np_count<R><<<nb, 256, 0, st>>>(a);
  } else {
    hipError_t e = launch_fill_i32(a.n_hits, 1, 0, st);
    if (e == hipSuccess) e = launch_fill_i32(a.n_cand, 1, 0, st);
    if (e != hipSuccess) return e;
    const int nbt = (a.ps.n_tasks + 4 * kCullTasks - 1) / (4 * kCullTasks);
    if (nbt > 0) k_np_cull<R><<<nbt, 256, 0, st>>>(a);
    const int64_t want = (a.cand_cap + 255) / 256;
    k_np_eval<R><<<int(want < 8192 ? (want > 0 ? want : 1) : 8192), 256, 0, st>>>(a);
  }
  if (ev) (void)hipEventRecord(ev[1], st);
  if (a.ps.max_world_tiles > kWideScanTiles)
    k_np_scan_tiles_wide<R><<<a.n_worlds, 1024, 0, st>>>(a);
  else
    k_np_scan_tiles<R><<<a.n_worlds, 64, 0, st>>>(a);
  k_np_scan_worlds<R><<<1, 1024, 0, st>>>(a);
  if (ev) (void)hipEventRecord(ev[2], st);
  if (list) {
    if (nb > 0) k_np_emit<R><<<nb, 256, 0, st>>>(a);
  } else {
    const int64_t want = (a.c.cap + 255) / 256;
    const int nbe = int(want < 8192 ? (want > 0 ? want : 1) : 8192);
    k_np_emit_hits<R><<<nbe, 256, 0, st>>>(a);
  }
  if (ev) (void)hipEventRecord(ev[3], st);
  return hipGetLastError();
}

template hipError_t launch_narrowphase<double>(const NarrowLaunch<double>&, hipStream_t, hipEvent_t*);
template hipError_t launch_narrowphase<float>(const NarrowLaunch<float>&, hipStream_t, hipEvent_t*);

}  // namespace cbz
