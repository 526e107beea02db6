// cbz_types.h — device-side data layout of the cubez step.
//
// Body state is structure-of-arrays, component-major: field f with k
// components occupies k planes of `stride` elements, so lane i of a
// thread-per-body kernel reads f[c*stride + i] (fully coalesced).
// All worlds are concatenated; world w owns bodies [body_off[w], body_off[w+1])
// and planes [plane_off[w], plane_off[w+1]).
#pragma once
#include <stdint.h>

namespace cbz {

// Body flag bits (uint8 per body).
enum : uint8_t {
  F_AWAKE = 1,        // RigidBody.IsAwake   (rigidbody.go:62)
  F_CAN_SLEEP = 2,    // RigidBody.CanSleep  (rigidbody.go:67)
  F_STATIC = 4,       // never integrated, collider transform never refreshed
  F_OFFSET_ID = 8,    // collider Offset is bitwise the identity (set by the host)
};

enum : int32_t { SHAPE_CUBE = 0, SHAPE_SPHERE = 1 };

// Constants of the reference.
constexpr double kFriction = 0.9;      // colliders.go:201 (and every other contact site)
constexpr double kRestitution = 0.1;   // colliders.go:202
constexpr double kSleepEpsilon = 0.3;  // rigidbody.go:14
constexpr double kVelocityEpsilon = 0.01;  // contact.go:12
constexpr double kPositionEpsilon = 0.01;  // contact.go:13

template <typename R>
struct Bodies {
  int32_t n;
  int64_t stride;
  const int32_t* shape;
  const R* half;      // [3] cube half-size; sphere radius in plane 0
  R* pos;             // [3]
  R* quat;            // [4] (w,x,y,z)
  R* vel;             // [3]
  R* rot;             // [3] angular velocity ("Rotation")
  const R* acc;       // [3]
  const R* inv_mass;  // [1]
  const R* iit_body;  // [9] body-space inverse inertia, column-major
  R* iit_world;       // [9]
  R* transform;       // [12] body transform
  R* ctransform;      // [12] collider transform (body transform * Offset)
  const R* offset;    // [12] collider Offset
  R* last_acc;        // [3] lastFrameAccelleration
  R* motion;          // [1]
  const R* pow_lin;   // [1] Pow(LinearDamping, dt), host-evaluated
  const R* pow_ang;   // [1] Pow(AngularDamping, dt), host-evaluated
  uint8_t* flags;     // [1]
  uint32_t* n_awake;  // optional counter of bodies integrated (profiling)
};

template <typename R>
struct Planes {
  const R* normal;  // [3][np]
  const R* offset;  // [np]
  int64_t stride;
};

// Contact records, structure-of-arrays with capacity `cap`.
template <typename R>
struct Contacts {
  int64_t cap;
  int32_t* body0;  // global body index
  int32_t* body1;  // global body index or -1
  R* point;        // [3][cap]
  R* normal;       // [3][cap]
  R* pen;          // [cap]
  R* c2w;          // [9][cap] contactToWorld
  R* rel;          // [6][cap] relativeContactPosition[0], [1]
  R* cvel;         // [3][cap] contactVelocity
  R* ddv;          // [cap] desiredDeltaVelocity
  // Contact.Friction / Contact.Restitution (contact.go:21-25) per contact
  // when materials are on; null: every contact carries kFriction and
  // kRestitution, as in the reference.
  R* fric;         // [cap]
  R* rest;         // [cap]
  // resolver scratch: per contact the impulse data of the velocity phase
  // (impulseMatrix 9, deltaVelocity entries 0, 3, 6), written when the
  // phase starts
  R* imp;          // [12][cap]
  // per contact one 16-value record, contactToWorld 9,
  // relativeContactPosition[0] 3, [1] 3, 1 unused: a copy of c2w and rel
  // that the one-wavefront resolver's wide classes write in prepare and read
  // in their loops from one address
  R* rec;          // [cap][16]
};

// Sources of the per-contact materials the narrowphase writes: friction and
// restitution per body and per plane.  A contact takes the mean of its two
// sides' values.
template <typename R>
struct Materials {
  const R* body_fric;   // [n_bodies]
  const R* body_rest;   // [n_bodies]
  const R* plane_fric;  // [n_planes]
  const R* plane_rest;  // [n_planes]
};

// Pair-space description for the narrowphase: the ordered pairs of world w
// are split into 64-wide tiles; tile t belongs to world tile_world[t] and
// covers that world's pair slots [(t - world_tile0[w])*64, +64).
// All-pairs space: row r of world w (body b0+r) has np + n columns (planes,
// then bodies j, j == r skipped), cut into 64-column tiles, so tile
// (r, c) = world_tile0[w] + r * nch + c with nch = ceil((np + n) / 64).
// The count pass runs one wave per task = (column tile c, block of
// kRowBlock rows).  An explicit pair list uses 64 consecutive list entries
// per tile instead.
constexpr int kRowBlock = 16;

// counters of awake bodies the integrate kernel fills when asked (summed by
// the reader)
constexpr int kAwakeSlots = 256;

struct PairSpace {
  int32_t n_tiles;
  int32_t max_world_tiles = 0;  // the largest world's tile count (host side: picks the scan's block size)
  const int32_t* tile_world;
  const int32_t* world_tile0;
  const int32_t* world_q;     // pair slots of each world (explicit list)
  int32_t n_tasks;            // all-pairs count tasks
  const int32_t* task_world;
  const int32_t* world_task0;
  const int32_t* body_off;
  const int32_t* plane_off;
  // explicit pair list (null for all-pairs): pairs of world w are
  // list_one/two[list_off[w] + q]; a negative index -(k+1) names plane k.
  const int32_t* list_off;
  const int32_t* list_one;
  const int32_t* list_two;
  int32_t* list_cnt;  // out (or null): contacts each listed pair produced
};

}  // namespace cbz
