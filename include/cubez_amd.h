/*
 * cubez_amd.h — C ABI of the MI355X implementation of the cubez simulation
 * step (Integrate + CalculateDerivedData, the narrowphase, ResolveContacts).
 *
 * The reference (tbogdala/cubez, pure Go) has no FFI: its boundary is the
 * exported Go API.  Each entry point below names the Go function(s) whose
 * behaviour it provides, as file:line of the reference.  A Go package keeps
 * those names and calls these functions through cgo, one call per phase per
 * step (or one per N steps via cbz_step); see INTEGRATION.md.
 *
 * Conventions
 *  - Every call returns 0 (CBZ_OK) or a negative CBZ_ERR_* code; the message
 *    of the last failure is cbz_last_error(ctx).
 *  - Host buffers are caller-owned and body-major (body i's k components at
 *    [i*k .. i*k+k)); device buffers are library-owned.  Downloads are
 *    synchronous; other calls are ordered on the context's HIP stream.
 *  - A context serves one host thread.
 *  - Bodies of all worlds are concatenated: world w owns bodies
 *    [body_off[w], body_off[w+1]) and planes [plane_off[w], plane_off[w+1]).
 *    Each world is one independent scene: one ResolveContacts call per world.
 */
#ifndef CUBEZ_AMD_H
#define CUBEZ_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CBZ_OK 0
#define CBZ_ERR_ARG (-1)
#define CBZ_ERR_HIP (-2)
#define CBZ_ERR_CAPACITY (-3)
#define CBZ_ERR_STATE (-4)

/* Arithmetic of the stored state.  CBZ_F64 is the exact path (the Go
 * reference's float64 Real, math/math.go:23).  CBZ_F32 stores and resolves in
 * float32 and keeps FP64 for the narrowphase and the contact basis. */
#define CBZ_F64 0
#define CBZ_F32 1

/* Shape codes (CollisionCube colliders.go:39, CollisionSphere :57). */
#define CBZ_CUBE 0
#define CBZ_SPHERE 1

/* Body flag bits (CBZ_FLAGS field, one byte per body). */
#define CBZ_FLAG_AWAKE 1     /* RigidBody.IsAwake,  rigidbody.go:62 */
#define CBZ_FLAG_CAN_SLEEP 2 /* RigidBody.CanSleep, rigidbody.go:67 */
#define CBZ_FLAG_STATIC 4    /* a body the caller never integrates (examples/ballistic.go:27-44 backboard) */

/* Per-body fields: id, components, element type.  Names follow RigidBody
 * (rigidbody.go:23-101) and the colliders (colliders.go:29-71). */
enum cbz_field {
    CBZ_SHAPE = 0,       /* 1 int32  */
    CBZ_HALF = 1,        /* 3 double CollisionCube.HalfSize; CollisionSphere.Radius in [0] */
    CBZ_POS = 2,         /* 3 double Position */
    CBZ_QUAT = 3,        /* 4 double Orientation (w,x,y,z) */
    CBZ_VEL = 4,         /* 3 double Velocity */
    CBZ_ROT = 5,         /* 3 double Rotation (angular velocity) */
    CBZ_ACC = 6,         /* 3 double Acceleration */
    CBZ_INV_MASS = 7,    /* 1 double inverseMass (SetMass rigidbody.go:124) */
    CBZ_IIT_BODY = 8,    /* 9 double InverseInertiaTensor, column-major */
    CBZ_IIT_WORLD = 9,   /* 9 double inverseInertiaTensorWorld */
    CBZ_TRANSFORM = 10,  /* 12 double body transform (Matrix3x4, column-major) */
    CBZ_CTRANSFORM = 11, /* 12 double collider transform */
    CBZ_OFFSET = 12,     /* 12 double collider Offset */
    CBZ_LAST_ACC = 13,   /* 3 double lastFrameAccelleration */
    CBZ_MOTION = 14,     /* 1 double motion */
    CBZ_LIN_DAMP = 15,   /* 1 double LinearDamping */
    CBZ_ANG_DAMP = 16,   /* 1 double AngularDamping */
    CBZ_FLAGS = 17,      /* 1 uint8  CBZ_FLAG_* */
    CBZ_FRICTION = 18,   /* 1 double contact friction of the body's collider (default 0.9; see cbz_set_materials) */
    CBZ_RESTITUTION = 19,/* 1 double contact restitution of the body's collider (default 0.1) */
    CBZ_NUM_FIELDS = 20
};

typedef struct cbz_config {
    int32_t precision;        /* CBZ_F64 or CBZ_F32 */
    int32_t device;           /* HIP device ordinal */
    int64_t contact_capacity; /* contact slots; 0 = 16 per body + 4096 */
} cbz_config;

/* Contact readback (Contact, contact.go:17-51).  Any pointer may be NULL.
 * body0/body1 are global body indices, body1 = -1 for a half-space contact.
 * The resolver internals (c2w .. ddv) are valid after cbz_resolve/cbz_step. */
typedef struct cbz_contact_view {
    int32_t *body0;  /* [n]   Bodies[0] */
    int32_t *body1;  /* [n]   Bodies[1] */
    double *point;   /* [n*3] ContactPoint */
    double *normal;  /* [n*3] ContactNormal */
    double *pen;     /* [n]   Penetration */
    double *c2w;     /* [n*9] contactToWorld */
    double *rel;     /* [n*6] relativeContactPosition[0], [1] */
    double *cvel;    /* [n*3] contactVelocity */
    double *ddv;     /* [n]   desiredDeltaVelocity */
    double *friction;    /* [n] Friction (0.9 on every contact unless materials are on) */
    double *restitution; /* [n] Restitution (0.1 unless materials are on) */
} cbz_contact_view;

typedef struct cbz_context cbz_ctx;

/* --- lifecycle ----------------------------------------------------------- */
int cbz_create(const cbz_config *cfg, cbz_ctx **out);
int cbz_destroy(cbz_ctx *ctx);
const char *cbz_last_error(const cbz_ctx *ctx);
/* components of a field, or -1 for an unknown id */
int cbz_field_width(int32_t field);

/* --- scene --------------------------------------------------------------- */
/* Allocates the state of body_off[n_worlds] bodies and plane_off[n_worlds]
 * planes.  Every body starts as NewRigidBody() leaves it (rigidbody.go:104-114:
 * identity orientation, damping 0.95/0.95, Acceleration (0,-9.78,0),
 * world inverse inertia identity, CanSleep, awake with motion 0.6) inside a
 * cube collider with identity Offset (NewCollisionCube colliders.go:265-274).
 * Planes: NewCollisionPlane(normal, offset), colliders.go:81-86. */
int cbz_set_scene(cbz_ctx *ctx, int32_t n_worlds, const int32_t *body_off, const int32_t *plane_off,
                  const double *plane_normal, const double *plane_offset);
/* Upload / download one field for every body (host layout [n_bodies][width]).
 * Setting a field is assigning the exported Go field (or calling SetMass /
 * SetInertiaTensor / SetAwake with their results). */
int cbz_set_field(cbz_ctx *ctx, int32_t field, const void *host);
int cbz_get_field(cbz_ctx *ctx, int32_t field, void *host);
int32_t cbz_num_worlds(const cbz_ctx *ctx);
int64_t cbz_num_bodies(const cbz_ctx *ctx);

/* --- the step, phase by phase -------------------------------------------- */
/* RigidBody.CalculateDerivedData (rigidbody.go:268-272) followed by the
 * collider's CalculateDerivedData (colliders.go:173-176, :302-304), for every
 * body: the set-up step of examples/cubedrop.go:171-172. */
int cbz_calculate_derived(cbz_ctx *ctx);
/* RigidBody.Integrate(dt) (rigidbody.go:213-259) then the collider's
 * CalculateDerivedData, for every non-static body: updateObjects of
 * examples/cubedrop.go:29-39.  The three math.Pow values (rigidbody.go:233,
 * 234, 250) are evaluated on the host with the C library's pow(). */
int cbz_integrate(cbz_ctx *ctx, double dt);
/* generateContacts of examples/cubedrop.go:42-67 for every world: for each
 * body i, CheckAgainstHalfSpace against each plane, then
 * CheckForCollisions(i, j) (colliders.go:720-747) for every j != i, with the
 * contacts of a world in exactly that order.  *total receives the contact
 * count over all worlds (may be NULL). */
int cbz_narrowphase(cbz_ctx *ctx, int64_t *total);
/* CheckForCollisions(one[k], two[k]) for an explicit ordered pair list (any
 * caller loop, e.g. examples/ballistic.go:47-96).  Indices are global body
 * indices; -(p+1) names plane p.  Pairs must be grouped by world in
 * non-decreasing world order; both operands of a pair in the same world. */
int cbz_narrowphase_pairs(cbz_ctx *ctx, int64_t n_pairs, const int32_t *one, const int32_t *two,
                          int64_t *total);
/* The contacts each pair of the last cbz_narrowphase_pairs call appended,
 * in list order (n_pairs must be that call's): counts[k] > 0 is the bool
 * CheckForCollisions returns for pair k ("at least one contact appended",
 * colliders.go:720-747), and pair k's contacts are the counts[k] that follow
 * the earlier pairs' in cbz_get_contacts order -- what a Go caller needs to
 * answer its per-pair calls from one batched narrowphase. */
int cbz_pair_contact_counts(cbz_ctx *ctx, int64_t n_pairs, int32_t *counts);
/* cbz_narrowphase with a sort-and-sweep broadphase in front, the first item
 * of the reference's roadmap (README.md:98-99): per world, the bodies are
 * sorted by the lower x bound of a conservative box (sqrt(3.003) times the
 * bounding radius, which covers every pair the narrowphase's own far-pair
 * test keeps; a body whose state is not finite overlaps everything), swept
 * along x with y and z tested, and only the overlapping ordered pairs plus
 * every body-plane pair are checked.  The pairs are sorted into the
 * all-pairs order of examples/cubedrop.go:47-64, so the contacts are
 * exactly cbz_narrowphase's, in the same order.  Worth it for worlds of
 * thousands of bodies, where the all-pairs space grows as n^2.
 * *pairs_checked receives the number of pairs the narrowphase evaluated
 * (either output may be NULL). */
int cbz_narrowphase_broad(cbz_ctx *ctx, int64_t *total, int64_t *pairs_checked);
/* contact count of every world after the last narrowphase ([n_worlds]) */
int cbz_contact_counts(cbz_ctx *ctx, int64_t *counts);
/* all contacts of one world, in emission order; world = -1: the contacts of
 * every world, concatenated in world order (the total of cbz_narrowphase) */
int cbz_get_contacts(cbz_ctx *ctx, int32_t world, cbz_contact_view *out);
/* Replace the contact lists (counts[n_worlds], then the records of all
 * worlds concatenated): contacts built by the caller, as when a Go caller
 * appends its own *Contact values before ResolveContacts. */
int cbz_set_contacts(cbz_ctx *ctx, const int64_t *counts, const int32_t *body0, const int32_t *body1,
                     const double *point, const double *normal, const double *pen);
/* --- contact materials ------------------------------------------------------
 * The reference gives every contact Friction 0.9 and Restitution 0.1
 * (colliders.go:201-202, 248-249, 360-361, 437-438, 511-512, 704-705); its
 * README lists per-contact values as a missing feature (README.md:89).
 * With CBZ_MATERIALS on:
 *  - each body (fields CBZ_FRICTION, CBZ_RESTITUTION) and each plane
 *    (cbz_set_plane_materials) has a friction and a restitution, 0.9 and 0.1
 *    by default, and every contact the narrowphase writes carries the means
 *    of its two sides' values (so the defaults reproduce the reference
 *    exactly);
 *  - cbz_set_contact_materials overrides the values of the current contacts,
 *    as a Go caller assigns Contact.Friction / Contact.Restitution
 *    (contact.go:21-25) between CheckForCollisions and ResolveContacts;
 *  - ResolveContacts uses each contact's values, and a contact of friction 0
 *    takes calculateFrictionlessImpulse (contact.go:459-461, :498-531) with
 *    the reference's own behaviour: the test at :512 is inverted, so a
 *    two-body contact's impulse leaves out its second body's terms (a
 *    half-space contact, on which Go would dereference nil, gets the
 *    one-body impulse).  CBZ_MATERIALS_FRICTIONLESS_TWO_BODY adds the second
 *    body's terms.
 * Changing the flags drops the current contacts (run a narrowphase again).
 * The learned resolver (cbz_sur_*) refuses contexts with materials on. */
#define CBZ_MATERIALS 1
#define CBZ_MATERIALS_FRICTIONLESS_TWO_BODY 2
int cbz_set_materials(cbz_ctx *ctx, int32_t flags);
/* friction[n_planes] (>= 0) and restitution[n_planes] of the planes; either
 * may be NULL */
int cbz_set_plane_materials(cbz_ctx *ctx, const double *friction, const double *restitution);
/* friction and restitution of every current contact, all worlds
 * concatenated in emission order ([total] each; either may be NULL);
 * materials must be on */
int cbz_set_contact_materials(cbz_ctx *ctx, const double *friction, const double *restitution);

/* --- state checkpoints -------------------------------------------------------
 * The whole simulation state of a context as one byte string, for resume and
 * for moving a run between machines.  The reference keeps this state in
 * RigidBody's exported and unexported fields (rigidbody.go:23-101) and has no
 * serialisation of its own.  Saved: the scene tables (world and plane
 * offsets, planes with their materials), the material flags and every field
 * of enum cbz_field in cbz_get_field's host layout, which includes what the
 * next step reads besides the pose: motion, IsAwake (CBZ_FLAGS),
 * lastFrameAccelleration and the derived transforms and world inverse
 * inertia.  Real fields are stored as doubles whatever the precision (an
 * FP32 state round-trips exactly).  Loading replaces the context's scene
 * (as cbz_set_scene) with that state, so stepping on gives the same result,
 * bit for bit, as never having stopped.  Contacts are not saved: the next
 * step's narrowphase makes them.  The layout (little-endian): a 64-byte
 * header (magic "CBZSTAT1", version 1, sizes, material flags, field count,
 * saving precision, payload size and FNV-1a hash), the world and plane
 * offsets, the plane data, then one record (id, width, kind, byte count,
 * data) per field; a mismatched magic, version, size or hash is rejected
 * with CBZ_ERR_ARG before the context changes. */
int cbz_checkpoint_size(cbz_ctx *ctx, int64_t *bytes);
int cbz_checkpoint_save(cbz_ctx *ctx, void *buf, int64_t bytes);
int cbz_checkpoint_load(cbz_ctx *ctx, const void *buf, int64_t bytes);

/* ResolveContacts(maxIterations, contacts, dt) (contact.go:208-222) on every
 * world's list, with maxIterations = iter_per_contact * C_w when iter_fixed
 * is 0 (the examples pass 8*len(contacts), examples/cubedrop.go:73), else
 * iter_fixed. */
int cbz_resolve(cbz_ctx *ctx, double dt, int64_t iter_per_contact, int64_t iter_fixed);
/* iterations the last resolve used per world: [2*n_worlds] = positions,
 * velocities (the loop counters of contact.go:235 and :392) */
int cbz_resolver_iterations(cbz_ctx *ctx, int64_t *iters);
/* When each world's resolver workgroup ran in the last resolve(s):
 * [n_worlds][4] microseconds -- start and end of the launch that ran the
 * position phase alone, then start and end of the launch that ran the
 * velocity phase (cbz_resolve runs both in one launch, recorded in the
 * second pair) -- from the GPU's constant-rate clock, relative to the
 * earliest recorded start; -1 where no such launch ran since cbz_set_scene.
 * A world's end minus its start is the latency of its serial chain; no
 * counterpart in the reference (a measurement of this implementation). */
int cbz_resolver_world_times(cbz_ctx *ctx, double *us);
/* Worlds the resolver keeps resident per CU (the plain kernel of the
 * context's precision: one wavefront per world, bounded by its registers
 * and LDS) and the device's CU count; the bench derives the time per
 * serial iteration from them. */
int cbz_resolver_occupancy(cbz_ctx *ctx, int32_t *worlds_per_cu, int32_t *n_cu);
/* The resolver's split between its two kernels.  Of the worlds of at least
 * min_contacts contacts (and at most 2048 contacts and 256 bodies), the
 * max_worlds largest resolve on a persistent kernel of four wavefronts per
 * world, on a second stream beside the one-wavefront kernel that takes the
 * others; both give the same bits.  Either value 0 sends every world to the
 * one-wavefront kernel; the defaults are 257 and 32 (cbz_autotune_resolver
 * picks it by measurement).  A scheduling choice of this implementation, with
 * no counterpart in the reference (ResolveContacts, contact.go:208-222, is
 * the same loop either way). */
int cbz_set_resolver_split(cbz_ctx *ctx, int32_t min_contacts, int32_t max_worlds);
int cbz_get_resolver_split(cbz_ctx *ctx, int32_t *min_contacts, int32_t *max_worlds);
/* The autotuner of that split: from the current state, times n_steps of
 * cbz_step (surrogate = 0) or of cbz_step_sur (surrogate = 1, with tol and
 * polish_per_contact) for each candidate split (cand_min[i], cand_max[i]),
 * restoring the body state before each run and at the end, then keeps the
 * fastest: its index in *best, the times in ms_per_step[n_cand] (may be
 * NULL).  Profiling must be off.  No counterpart in the reference. */
int cbz_autotune_resolver(cbz_ctx *ctx, double dt, int64_t iter_per_contact, int32_t n_steps, int32_t surrogate,
                          double tol, int64_t polish_per_contact, int32_t n_cand, const int32_t *cand_min,
                          const int32_t *cand_max, double *ms_per_step, int32_t *best);

/* --- fused stepping --------------------------------------------------------
 * n_steps of updateCallback (examples/cubedrop.go:69-75): integrate, the
 * all-pairs narrowphase and ResolveContacts(iter_per_contact * C_w) per
 * world, with no host synchronisation between steps.  counts (may be NULL)
 * receives [n_steps][n_worlds] contact counts.  Without counts and with
 * profiling off, the step is captured once as a HIP graph (per dt and
 * iteration multiplier) and replayed n_steps times. */
int cbz_step(cbz_ctx *ctx, double dt, int32_t n_steps, int64_t iter_per_contact, int64_t *counts);
int cbz_synchronize(cbz_ctx *ctx);

/* --- measurement ----------------------------------------------------------
 * When enabled, cbz_step records HIP events between its kernels on the
 * context's stream (and waits for each step); cbz_phase_times returns the
 * accumulated milliseconds [integrate, narrowphase count, narrowphase scans,
 * narrowphase emit, resolve, whole steps] and the counters [steps, bodies
 * integrated (awake at Integrate), contacts, resolver iterations], and
 * resets the accumulators. */
int cbz_set_profiling(cbz_ctx *ctx, int32_t on);
int cbz_phase_times(cbz_ctx *ctx, double *ms6, int64_t *counts4);
/* n back-to-back cbz_integrate(dt) launches timed by two HIP events on the
 * context's stream: the average ms per launch and the bodies integrated
 * (awake at Integrate) over the n launches.  Per-kernel event pairs inside a
 * step add tens of microseconds of their own to a kernel this short; this
 * is how the bench times K1.  Advances the state by n integrations. */
int cbz_integrate_timed(cbz_ctx *ctx, double dt, int32_t n, double *ms_per_launch, int64_t *integrated);
/* Device FP64 arithmetic self-test (the rounding the exact path relies on):
 * op 0: a/b, 1: sqrt(a), 2: a*b, 3: a+b, 4: a-b, elementwise over n. */
int cbz_selftest_f64(cbz_ctx *ctx, int32_t op, int64_t n, const double *a, const double *b, double *out);
/* Self-test of the broadphase's device sort (a stable radix sort of 64-bit
 * keys, values carried along): keys_out = keys sorted, equal keys in input
 * order, vals_out the values in the same permutation (vals and vals_out
 * both NULL: keys only).  Host buffers, n < 2^31.  No counterpart in the
 * reference. */
int cbz_selftest_sort(cbz_ctx *ctx, int64_t n, const uint64_t *keys, const int32_t *vals, uint64_t *keys_out,
                      int32_t *vals_out);

/* --- learned resolver surrogate --------------------------------------------
 * A 40 -> 256 -> 256 -> 256 -> 13 MLP (SiLU, BF16 weights and activations,
 * FP32 accumulation, on the matrix cores) predicts, per contact, the
 * velocity and rotation changes the velocity phase of ResolveContacts
 * (adjustVelocities, contact.go:390-445) gives the contact's two bodies,
 * and the penetration the position phase (adjustPositions,
 * contact.go:233-283) resolves at the contact, whose moves are
 * applyPositionChange's (contact.go:286-386).  No counterpart in the
 * reference: it is the learned part of this implementation.  The exact
 * method stays the judge: per world and phase, the reference's loop runs
 * from the predicted state until every contact is within the phase's
 * epsilon (positionEpsilon, velocityEpsilon: 0.01, contact.go:12-13) plus a
 * tolerance; a world that does not get there is restored and runs the exact
 * phase from its unchanged state.
 *
 * Features (40, per contact, after prepareContacts): normal 3,
 * relativeContactPosition[0] 3, [1] 3, body0 - body1 velocity 3 and
 * rotation 3, inverse masses 2, world inverse inertias 9 + 9, penetration,
 * desiredDeltaVelocity, contactVelocity.x, 1/(contacts of body0),
 * 1/(contacts of body1).
 * Outputs (13): dv0 3, dw0 3, dv1 3, dw1 3, resolved penetration 1.
 * Parameters: per layer W [out x in] row-major then b [out], FP32.
 * Normalisation (106 floats): in_mean 40, in_std 40, out_mean 13,
 * out_std 13. */
int64_t cbz_sur_num_params(void);
int cbz_sur_set_params(cbz_ctx *ctx, const float *params);
int cbz_sur_get_params(cbz_ctx *ctx, float *params);
int cbz_sur_set_norm(cbz_ctx *ctx, const float *norm106);
int cbz_sur_get_norm(cbz_ctx *ctx, float *norm106);
/* Training data: cbz_sur_collect resolves the current contacts exactly (the
 * same result as cbz_resolve) and appends every contact's features and
 * labels (the summed changes the exact velocity phase gave its bodies, the
 * penetration the exact position phase resolved at it) to a device dataset
 * of cbz_sur_reserve rows. */
int cbz_sur_reserve(cbz_ctx *ctx, int64_t rows);
int cbz_sur_collect(cbz_ctx *ctx, double dt, int64_t iter_per_contact, int64_t *rows_added);
int cbz_sur_dataset(cbz_ctx *ctx, int64_t *rows, float *features, float *labels);
int cbz_sur_load_dataset(cbz_ctx *ctx, int64_t rows, const float *features, const float *labels);
/* normalisation from the dataset, He-uniform weights from the seed */
int cbz_sur_init(cbz_ctx *ctx, uint64_t seed);
/* Adam on the mean-squared error of normalised labels, batches sampled from
 * the dataset; losses[steps] (may be NULL) */
int cbz_sur_train(cbz_ctx *ctx, int32_t steps, int32_t batch, float lr, uint64_t seed, float *losses);
/* loss and parameter gradient on given rows (raw features and labels) */
int cbz_sur_grad(cbz_ctx *ctx, int32_t rows, const float *features, const float *labels, float *grad, float *loss);
/* predictions (denormalised) for given rows of raw features, by the fused
 * inference kernel (all layers in one launch), or with
 * CBZ_SUR_PREDICT_LAYERED by the per-layer kernels training uses (the two
 * give the same bits) */
#define CBZ_SUR_PREDICT_LAYERED 1
int cbz_sur_predict(cbz_ctx *ctx, int32_t rows, const float *features, float *out, int32_t flags);
/* cbz_sur_predict (fused inference) on device-resident rows, allocating
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  features: device pointer to
 * [rows][40] raw FP32 features; out: device pointer to [rows][13] FP32
 * predictions (denormalised), both caller-owned.  rows <= 65536
 * (CBZ_ERR_ARG beyond).  The normalised BF16 rows go to a fixed 6 MB
 * device array the library holds per GPU for the process's lifetime:
 * calls that overlap on one GPU (several threads or contexts) race on it.
 * Ordered on the context's stream (the caller's writes of `features` must
 * be complete); returns after `out` is written.  The results are those of
 * cbz_sur_predict, bit for bit. */
int cbz_sur_predict_device(cbz_ctx *ctx, int32_t rows, const float *features, float *out);
/* Host-only consistency check of the surrogate's matrix products (needs no
 * GPU): every GEMM of one training step and of one layered inference pass
 * over `rows` rows is compared with the sizes the library allocates for its
 * operands; returns CBZ_OK, or CBZ_ERR_STATE with the first GEMM whose reads
 * or writes would leave a buffer described in msg[msg_len]. */
int cbz_sur_check_gemms(int32_t rows, char *msg, int32_t msg_len);
/* ResolveContacts with the surrogate.  prepareContacts runs, then the
 * network on the prepared contacts; its per-contact predictions are summed
 * per body.
 * Position phase: per world, the predicted moves are applied and the
 * reference's worst-first position loop runs from there until every
 * contact's penetration is within positionEpsilon + the position tolerance
 * (cbz_sur_set_position), for at most its budget per contact * C
 * iterations; the polished world is accepted only if, besides, no contact
 * ends separated by more than the world's deepest penetration at the start
 * plus that tolerance.  Worlds of more than 1536 contacts, and every world
 * when the position tolerance is negative, take the plain position phase.
 * Velocity phase: likewise from the predicted changes, to velocityEpsilon +
 * tol (tol in m/s) within polish_per_contact * C iterations, accepted only
 * if the world's kinetic energy did not grow (beyond 1/2 tol^2 per kg).
 * A world that a polish does not bring there is restored and re-solved by
 * that plain phase (a fallback).  tol < 0 sends every world to both plain
 * phases; flags CBZ_SUR_ZERO_PREDICTION polishes from zero predictions (the
 * same method without the network).  *n_fallback receives the number of
 * worlds that fell back in either phase; cbz_sur_world_mask gives the
 * velocity phase's fallbacks per world, cbz_sur_position_mask the position
 * phase's (1: the plain phase ran).
 * Replaces ResolveContacts (contact.go:208-222) for callers that accept the
 * tolerances. */
#define CBZ_SUR_ZERO_PREDICTION 1
int cbz_sur_resolve(cbz_ctx *ctx, double dt, int64_t iter_per_contact, double tol, int64_t polish_per_contact,
                    int32_t flags, int64_t *n_fallback);
int cbz_sur_world_mask(cbz_ctx *ctx, uint8_t *mask);
int cbz_sur_position_mask(cbz_ctx *ctx, uint8_t *mask);
/* The position phase's polish: tolerance in m above positionEpsilon
 * (default 0: the plain phase's own stopping rule; negative: the plain
 * position phase everywhere) and budget per contact (default 4, the plain
 * phase's own).  No counterpart in the reference. */
int cbz_sur_set_position(cbz_ctx *ctx, double tol, int64_t polish_per_contact);
int cbz_sur_get_position(cbz_ctx *ctx, double *tol, int64_t *polish_per_contact);
/* With profiling on (cbz_set_profiling), every surrogate resolve records HIP
 * events on the context's stream; this returns the accumulated ms
 * [prepareContacts, features, network forward, per-body sums, position phase
 * (polish and fallbacks), velocity phase (polish and fallbacks)] and counts
 * [resolves, contacts, worlds that fell back in either phase, velocity-phase
 * iterations, worlds that fell back in the position phase, position-phase
 * iterations], and resets them. */
int cbz_sur_phase_times(cbz_ctx *ctx, double *ms6, int64_t *counts6);
/* cbz_step with the surrogate resolver; fallbacks[n_steps] (may be NULL)
 * receives the fallback count of every step */
int cbz_step_sur(cbz_ctx *ctx, double dt, int32_t n_steps, int64_t iter_per_contact, double tol,
                 int64_t polish_per_contact, int32_t flags, int64_t *counts, int64_t *fallbacks);

#ifdef __cplusplus
}
#endif
#endif
