/*
 * cbz_oracle.h — CPU FP64 restatement of the cubez simulation step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product (cubez_amd) never
 * links or calls it.
 *
 * Every function follows one function of the Go reference (file:line cited
 * at each definition in cbz_oracle.c), statement by statement and in the
 * same floating-point evaluation order, compiled with -ffp-contract=off so
 * no multiply-add is fused (Go on amd64 at GOAMD64=v1 does not fuse).
 *
 * Scene interface: structure-of-arrays, one world at a time, the same field
 * layout the C-ABI of the product uses (include/cubez_amd.h).
 */
#ifndef CBZ_ORACLE_H
#define CBZ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Shape codes, same as the product. */
#define CBZO_CUBE 0
#define CBZO_SPHERE 1

/* Body flag bits. */
#define CBZO_AWAKE 1u
#define CBZO_CAN_SLEEP 2u
#define CBZO_STATIC 4u /* never integrated, collider transform never refreshed */

/* One world's state, SoA; n bodies, np planes. Arrays are owned by the caller
 * and updated in place by the step functions. */
typedef struct cbzo_scene {
    int32_t n;
    int32_t np;
    const int32_t *shape;  /* [n]            */
    const double *half;    /* [n*3] cube half-size; sphere radius in [0] */
    double *pos;           /* [n*3] */
    double *quat;          /* [n*4] (w,x,y,z) */
    double *vel;           /* [n*3] */
    double *rot;           /* [n*3] */
    const double *acc;     /* [n*3] */
    const double *inv_mass;/* [n]   */
    const double *iit_body;/* [n*9] column-major */
    double *iit_world;     /* [n*9] */
    double *transform;     /* [n*12] body transform */
    double *ctransform;    /* [n*12] collider transform */
    const double *offset;  /* [n*12] collider offset (identity in every constructor) */
    double *last_acc;      /* [n*3] */
    double *motion;        /* [n]   */
    const double *pow_lin; /* [n]   Pow(LinearDamping, dt)  */
    const double *pow_ang; /* [n]   Pow(AngularDamping, dt) */
    uint8_t *flags;        /* [n]   CBZO_* bits */
    const double *plane_normal; /* [np*3] */
    const double *plane_offset; /* [np]   */
} cbzo_scene;

/* Contact record as produced by the narrowphase, plus the resolver internals
 * after cbzo_resolve. body1 == -1 means "no second body" (half-space). */
typedef struct cbzo_contact {
    int32_t body0, body1;
    double friction, restitution;
    double point[3], normal[3], pen;
    double c2w[9];
    double rel[2][3];
    double cvel[3];
    double ddv;
} cbzo_contact;

/* Counters for the algorithmic work of the hot path. */
typedef struct cbzo_counters {
    int64_t pairs;          /* CheckAgainst* calls between two bodies */
    int64_t plane_checks;   /* body/half-space checks */
    int64_t sat_axes;       /* tryAxis calls that passed the parallel skip */
    int64_t sat_skipped;    /* tryAxis calls skipped as near-parallel */
    int64_t contacts;
    int64_t pos_iters;
    int64_t vel_iters;
    int64_t integrated;
    int64_t panics;         /* paths on which Go would have panicked (nil body) */
} cbzo_counters;

/* --- step pieces (one world) ------------------------------------------ */
void cbzo_calculate_derived(cbzo_scene *s);              /* body + collider derived for all bodies */
void cbzo_integrate(cbzo_scene *s, double dt, double sleep_bias, cbzo_counters *cnt);
/* all-pairs narrowphase in the order of examples/cubedrop.go:47-64 (planes
 * first within each i, then j != i ascending). Returns the contact count;
 * writes at most cap contacts. */
int64_t cbzo_narrowphase(cbzo_scene *s, cbzo_contact *out, int64_t cap, cbzo_counters *cnt);
/* explicit ordered pair list; a negative index -(k+1) names plane k. */
int64_t cbzo_narrowphase_pairs(cbzo_scene *s, int64_t npairs, const int32_t *one,
                               const int32_t *two, cbzo_contact *out, int64_t cap,
                               cbzo_counters *cnt);
void cbzo_resolve(cbzo_scene *s, cbzo_contact *c, int64_t nc, int64_t max_iter, double dt,
                  cbzo_counters *cnt);
/* cbzo_resolve with options: CBZO_FRICTIONLESS_TWO_BODY adds the second
 * body's terms to the frictionless impulse of a two-body contact (the
 * reference omits them, contact.go:512). */
#define CBZO_FRICTIONLESS_TWO_BODY 1
void cbzo_resolve_ex(cbzo_scene *s, cbzo_contact *c, int64_t nc, int64_t max_iter, double dt, int32_t flags,
                     cbzo_counters *cnt);
/* n_steps of: integrate, all-pairs narrowphase, ResolveContacts(mult*C).
 * counts[k] receives the contact count of step k (may be NULL). Returns 0, or
 * -1 if a step exceeded the scratch contact capacity. */
int cbzo_step(cbzo_scene *s, double dt, double sleep_bias, int32_t iter_mult, int32_t n_steps,
              int64_t *counts, cbzo_counters *cnt);

/* --- math entry points for known-answer tests --------------------------- */
int cbzo_real_equal(double a, double b);
void cbzo_vec_cross(const double *a, const double *b, double *out);
double cbzo_vec_dot(const double *a, const double *b);
double cbzo_vec_magnitude(const double *a);
void cbzo_vec_normalize(double *a);
void cbzo_quat_mul(double *q, const double *q2);
double cbzo_quat_len(const double *q);
void cbzo_quat_normalize(double *q);
void cbzo_quat_add_scaled_vector(double *q, const double *v, double scale);
void cbzo_quat_rotate(const double *q, const double *v, double *out);
void cbzo_quat_from_axis(double angle, double x, double y, double z, double *out);
void cbzo_mat3_mul(const double *a, const double *b, double *out);
void cbzo_mat3_mul_vec(const double *a, const double *v, double *out);
void cbzo_mat3_invert(const double *a, double *out);
double cbzo_mat3_det(const double *a);
void cbzo_mat3_block_inertia(const double *half, double mass, double *out);
void cbzo_mat34_set_transform(const double *pos, const double *q, double *out);
void cbzo_mat34_mul(const double *a, const double *b, double *out);
void cbzo_mat34_mul_vec(const double *a, const double *v, double *out);
void cbzo_mat34_transform_inverse(const double *a, const double *v, double *out);
void cbzo_transform_inertia(const double *iit_body, const double *rot, double *out);
void cbzo_contact_basis(const double *normal, double *c2w);
void cbzo_skew(const double *v, double *out);
/* contactPoint (colliders.go:519) */
void cbzo_contact_point(const double *p1, const double *d1, double s1, const double *p2,
                        const double *d2, double s2, int use_one, double *out);

#ifdef __cplusplus
}
#endif
#endif
