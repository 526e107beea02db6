// cbz_api.hip — the C ABI (include/cubez_amd.h): context, device state,
// scene tables and the phase launches.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "cbz_broad.h"
#include "cbz_sort.h"
#include "cbz_launch.h"
#include "cbz_mlp.h"
#include "cbz_types.h"
#include "cubez_amd.h"

using namespace cbz;

namespace {

struct FieldInfo {
  int width;
  int kind;  // 0 = real, 1 = int32, 2 = uint8, 3 = host-only real
};

const FieldInfo kFields[CBZ_NUM_FIELDS] = {
    {1, 1},  {3, 0}, {3, 0}, {4, 0}, {3, 0},  {3, 0}, {3, 0}, {1, 0}, {9, 0},
    {9, 0},  {12, 0}, {12, 0}, {12, 0}, {3, 0}, {1, 0}, {1, 3}, {1, 3}, {1, 2},
    {1, 0},  {1, 0},
};

constexpr uint8_t kUserFlags = CBZ_FLAG_AWAKE | CBZ_FLAG_CAN_SLEEP | CBZ_FLAG_STATIC;
// real planes per contact slot (see cbz_context::c_real)
constexpr int kContactPlanes = 40;
// The resolver's multi-wave kernel takes the 32 largest worlds of more than
// 256 contacts.
constexpr int32_t kDefaultMwMin = 257;
constexpr int32_t kDefaultMwMax = 32;

}  // namespace

// ---------------------------------------------------------------------------
// The allocation-free inference entry point (cbz_sur_predict_device) keeps
// the normalised BF16 rows it feeds the network in this file-scope device
// array: one copy per GPU, at an address each device resolves once (in
// cbz_create) with hipGetSymbolAddress.  Host code uses that address, never
// the array's name.  Every call on a GPU shares its copy.
// ---------------------------------------------------------------------------
constexpr int kNoAllocRows = 65536;
static __device__ __align__(256) __bf16 scratch[kNoAllocRows * kFeatPad];

namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // per device ordinal; null until resolved

// The device's copy of `scratch`, resolved on first use with that device
// current (hipSetDevice done by the caller).
hipError_t scratch_address(int device, void** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}
}  // namespace

struct cbz_context {
  int precision = CBZ_F64;
  int device = 0;
  int64_t cap_cfg = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // the resolver's multi-wave kernel: the mw_max largest worlds of at
  // least mw_min contacts run on stream `side` beside the one-wave kernel
  hipStream_t side = nullptr;
  hipEvent_t mw_fork = nullptr, mw_sel = nullptr, mw_join = nullptr;
  int32_t mw_min = kDefaultMwMin;
  int32_t mw_max = kDefaultMwMax;
  int32_t n_cu = 256;

  // scene
  int32_t nw = 0;
  int64_t nb = 0;
  int32_t np = 0;
  std::vector<int32_t> body_off, plane_off;
  std::vector<double> lin_damp, ang_damp;
  std::vector<int32_t> h_shape;
  std::vector<double> h_half;
  std::vector<uint8_t> h_offset_id;
  bool brad_dirty = true;
  bool pow_dirty = true;
  double pow_dt = std::numeric_limits<double>::quiet_NaN();
  double sleep_bias = 1.0;

  // device state
  void* field[CBZ_NUM_FIELDS] = {};
  void* pow_lin = nullptr;
  void* pow_ang = nullptr;
  double* brad = nullptr;
  void* plane_normal = nullptr;
  void* plane_offset = nullptr;
  // contact materials (cbz_set_materials)
  int32_t mat_flags = 0;
  void* plane_fric = nullptr;  // [np]
  void* plane_rest = nullptr;  // [np]
  int32_t* d_body_off = nullptr;
  int32_t* d_plane_off = nullptr;

  // all-pairs pair space
  int32_t n_tiles = 0, n_tasks = 0, max_world_tiles = 0;
  int32_t* d_tile_world = nullptr;
  int32_t* d_world_tile0 = nullptr;
  int32_t* d_task_world = nullptr;
  int32_t* d_world_task0 = nullptr;
  uint32_t* hits = nullptr;  // [cap] hit pairs of the last narrowphase
  int32_t* n_hits = nullptr;
  uint32_t* cand = nullptr;  // [cand_cap] candidate pairs of the last cull
  int32_t* n_cand = nullptr;
  int64_t cand_cap = 0;
  // explicit-list pair space (grown on demand)
  int32_t l_tiles_cap = 0, l_pairs_cap = 0;
  int32_t* l_tile_world = nullptr;
  int32_t* l_world_tile0 = nullptr;
  int32_t* l_world_q = nullptr;
  int32_t* l_off = nullptr;
  int32_t* l_one = nullptr;
  int32_t* l_two = nullptr;
  int32_t* l_cnt = nullptr;   // contacts per listed pair of the last list narrowphase
  int64_t l_last = -1;        // pairs in that list (-1: the last narrowphase had none)
  // sort-and-sweep broadphase (cbz_narrowphase_broad), allocated on first use
  struct Broad {
    bool ready = false;
    int32_t* body_world = nullptr;     // [nb]
    double* lo = nullptr;              // [2][nb]
    double* hi = nullptr;              // [2][nb]
    float* hix = nullptr;              // [nb]
    uint8_t* wild = nullptr;           // [nb]
    uint64_t* key_in = nullptr;        // [nb]
    uint64_t* key_out = nullptr;
    int32_t* val_in = nullptr;
    int32_t* val_out = nullptr;
    int64_t* plane_base = nullptr;     // [nw]
    int64_t plane_pairs = 0;
    uint64_t* pkey_in = nullptr;       // [pcap]
    uint64_t* pkey = nullptr;
    int64_t pcap = 0;
    int64_t* pcount = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
  } bp;
  // tile scratch, sized for the larger pair space
  int32_t tiles_cap = 0;
  int32_t* tile_cnt = nullptr;
  int32_t* tile_off = nullptr;
  uint64_t* tile_mask = nullptr;

  // contacts
  int64_t cap = 0;
  void* c_body0 = nullptr;
  void* c_body1 = nullptr;
  // point 3, normal 3, pen 1, c2w 9, rel 6, cvel 3, ddv 1, friction 1,
  // restitution 1, then the resolver's impulse data 12 (40 planes)
  void* c_real = nullptr;
  void* c_rec = nullptr;  // [cap][16] reals: Contacts::rec
  int32_t* world_cnt = nullptr;
  int64_t* world_coff = nullptr;
  int64_t* total = nullptr;
  int32_t* overflow = nullptr;  // [0] this narrowphase, [1] sticky since the host cleared it
  int64_t* world_iters = nullptr;
  int64_t* world_time = nullptr;  // [4 * nw] resolver workgroup start/end ticks (position, velocity launch)
  int32_t* mw_list = nullptr;     // [nw] worlds of the multi-wave kernel (resolver launch)
  int32_t* mw_count = nullptr;    // [2] list length, work-queue head
  uint8_t* mw_flag = nullptr;     // [nw] listed
  int32_t* order = nullptr;
  bool have_contacts = false;

  // resolver surrogate (see cbz_sur_*)
  struct Sur {
    bool have_model = false;
    float* params = nullptr;  // FP32 master weights: per layer W [N x K] then b [N]
    float* grad = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    int64_t adam_t = 0;
    __bf16* Wnk[cbz::kLayers] = {};  // [Npad x Kpad] forward operand
    __bf16* Wkn[cbz::kLayers] = {};  // [Kpad x Npad] backward operand
    float* norm = nullptr;  // in_mean 40, in_inv_std 40, out_mean 13, out_std 13, out_inv_std 13
    // dataset of (raw features, labels)
    float* dsX = nullptr;
    float* dsY = nullptr;
    int64_t ds_rows = 0, ds_cap = 0;
    // inference buffers
    int64_t inf_cap = 0, lay_cap = 0;
  int64_t generation = 0;  // bumped whenever the inference buffers are reallocated
    __bf16* X = nullptr;
    __bf16* H[2] = {};
    float* Yp = nullptr;
    int32_t* deg = nullptr;
    uint8_t* mask = nullptr;
    int32_t* nfall = nullptr;
    float* delta = nullptr;  // [6][n_bodies] predicted change per body
    void* vel0 = nullptr;    // [7][n_bodies] velocity, rotation, motion before it
    void* cvel0 = nullptr;   // [3][cap] contact velocities as prepare left them
    // the position phase's polish: per-world mask, predicted change per body
    // [6][n_bodies], snapshot [22][n_bodies] of the sleeping bodies
    uint8_t* pmask = nullptr;
    float* dpos = nullptr;
    void* psnap = nullptr;
    // position polish tolerance (m) above positionEpsilon and budget per
    // contact; a negative tolerance gives every world the plain position phase
    double pos_tol = 0.0;
    int64_t pos_polish = 4;
    bool last_pos = false;  // the last surrogate resolve polished positions (pmask is valid)
    // training buffers
    int batch_cap = 0;
    int64_t* idx = nullptr;
    __bf16* bX = nullptr;
    float* bT = nullptr;
    float* Z[cbz::kLayers] = {};
    __bf16* Hb[cbz::kLayers - 1] = {};
    __bf16* At[cbz::kLayers] = {};  // transposed layer inputs: X^T, H0^T, H1^T, H2^T
    float* Yout = nullptr;
    __bf16* dZ[cbz::kLayers] = {};
    __bf16* dZt[cbz::kLayers] = {};
    float* dA = nullptr;
    float* row_loss = nullptr;
    float* losses = nullptr;
    int losses_cap = 0;
  } sur;

  // captured step graph (cbz_step without counts or profiling)
  hipGraphExec_t gexec = nullptr;
  double g_dt = 0;
  int64_t g_mult = -1;
  // captured surrogate step (cbz_step_sur without counts or profiling),
  // keyed by its arguments and the inference buffers' generation; each
  // replay appends the step's fallback count to fb_hist at fb_ctr
  hipGraphExec_t sgexec = nullptr;
  double sg_dt = 0, sg_tol = 0;
  int64_t sg_mult = -1, sg_polish = -1, sg_gen = -1;
  int32_t sg_flags = -1;
  int32_t* fb_hist = nullptr;
  int32_t* fb_ctr = nullptr;
  int64_t fb_cap = 0;

  // profiling
  bool prof = false;
  // ev[0..7]: cbz_step phases; ev[8..14]: surrogate resolve phases
  hipEvent_t ev[15] = {};
  double acc_ms[6] = {0, 0, 0, 0, 0, 0};
  int64_t acc_cnt[4] = {0, 0, 0, 0};  // steps, bodies integrated, contacts, resolver iterations
  // surrogate resolves: ms [prepare, features, network, per-body sums,
  // position phase, velocity phase], counts [resolves, contacts, worlds
  // given a plain phase, velocity iterations, worlds given the plain
  // position phase, position iterations]
  double sur_ms[6] = {0, 0, 0, 0, 0, 0};
  int64_t sur_cnt[6] = {0, 0, 0, 0, 0, 0};
  uint32_t* d_awake = nullptr;
  bool count_awake = false;
  int64_t* d_prof = nullptr;           // [4] device-side step counters (k_prof_accum)
  std::vector<hipEvent_t> pev;         // per-step event pool of profiled cbz_step calls
  void* pos0 = nullptr;  // body positions at the last prepareContacts
  void* scratch = nullptr;  // this device's copy of the file-scope scratch array
  // Shape check (cbz_sur_check_gemms): when set, the surrogate's GEMMs are
  // checked against these [begin, end) byte ranges instead of launched, and
  // its other kernels are skipped.
  const std::vector<std::pair<uintptr_t, uintptr_t>>* dry_regions = nullptr;
};

namespace {

int fail(cbz_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, x)                                                                            \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess) return fail((ctx), CBZ_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

size_t real_size(const cbz_ctx* c) { return c->precision == CBZ_F64 ? sizeof(double) : sizeof(float); }

template <typename T>
int dalloc(cbz_ctx* c, T** p, size_t count) {
  if (*p) {
    HIPCHK(c, hipFree(*p));
    *p = nullptr;
  }
  if (count == 0) count = 1;
  HIPCHK(c, hipMalloc((void**)p, count * sizeof(T)));
  return CBZ_OK;
}

int dalloc_bytes(cbz_ctx* c, void** p, size_t bytes) {
  if (*p) {
    HIPCHK(c, hipFree(*p));
    *p = nullptr;
  }
  if (bytes == 0) bytes = 8;
  HIPCHK(c, hipMalloc(p, bytes));
  return CBZ_OK;
}

void dfree(void*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}
template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// Host [n][k] doubles -> device [k][n] of the context's real type.
int upload_real(cbz_ctx* c, void* dst, const double* src, int64_t n, int k) {
  if (n == 0) return CBZ_OK;
  if (c->precision == CBZ_F64) {
    std::vector<double> t((size_t)n * k);
    for (int64_t i = 0; i < n; ++i)
      for (int j = 0; j < k; ++j) t[(size_t)j * n + i] = src[(size_t)i * k + j];
    HIPCHK(c, hipMemcpy(dst, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> t((size_t)n * k);
    for (int64_t i = 0; i < n; ++i)
      for (int j = 0; j < k; ++j) t[(size_t)j * n + i] = (float)src[(size_t)i * k + j];
    HIPCHK(c, hipMemcpy(dst, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  return CBZ_OK;
}

// Device [k][stride] (from element `first`, n elements) -> host [n][k] doubles.
int download_real(cbz_ctx* c, const void* srcbase, int64_t stride, int64_t first, int64_t n, int k, double* dst) {
  if (n == 0) return CBZ_OK;
  if (c->precision == CBZ_F64) {
    std::vector<double> t((size_t)n);
    for (int j = 0; j < k; ++j) {
      const double* s = (const double*)srcbase + (size_t)j * stride + first;
      HIPCHK(c, hipMemcpy(t.data(), s, n * sizeof(double), hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < n; ++i) dst[(size_t)i * k + j] = t[i];
    }
  } else {
    std::vector<float> t((size_t)n);
    for (int j = 0; j < k; ++j) {
      const float* s = (const float*)srcbase + (size_t)j * stride + first;
      HIPCHK(c, hipMemcpy(t.data(), s, n * sizeof(float), hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < n; ++i) dst[(size_t)i * k + j] = (double)t[i];
    }
  }
  return CBZ_OK;
}

void drop_graph(cbz_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  c->gexec = nullptr;
  c->g_mult = -1;
  if (c->sgexec) (void)hipGraphExecDestroy(c->sgexec);
  c->sgexec = nullptr;
  c->sg_mult = -1;
}

void free_sur_buffers(cbz_ctx* c) {
  auto& u = c->sur;
  dfree(u.X);
  dfree(u.H[0]);
  dfree(u.H[1]);
  dfree(u.Yp);
  dfree(u.deg);
  dfree(u.mask);
  dfree(u.nfall);
  dfree(u.delta);
  dfree(u.vel0);
  dfree(u.cvel0);
  dfree(u.pmask);
  dfree(u.dpos);
  dfree(u.psnap);
  u.inf_cap = 0;
  u.lay_cap = 0;
  u.generation += 1;
}

void free_scene(cbz_ctx* c) {
  drop_graph(c);
  free_sur_buffers(c);
  dfree(c->fb_hist);
  dfree(c->fb_ctr);
  c->fb_cap = 0;
  for (auto& f : c->field) dfree(f);
  dfree(c->pow_lin);
  dfree(c->pow_ang);
  dfree(c->brad);
  dfree(c->plane_normal);
  dfree(c->plane_offset);
  dfree(c->plane_fric);
  dfree(c->plane_rest);
  dfree(c->d_body_off);
  dfree(c->d_plane_off);
  dfree(c->d_tile_world);
  dfree(c->d_world_tile0);
  dfree(c->d_task_world);
  dfree(c->d_world_task0);
  dfree(c->hits);
  dfree(c->n_hits);
  dfree(c->cand);
  dfree(c->n_cand);
  {
    auto& bp = c->bp;
    dfree(bp.body_world);
    dfree(bp.lo);
    dfree(bp.hi);
    dfree(bp.hix);
    dfree(bp.wild);
    dfree(bp.key_in);
    dfree(bp.key_out);
    dfree(bp.val_in);
    dfree(bp.val_out);
    dfree(bp.plane_base);
    dfree(bp.pkey_in);
    dfree(bp.pkey);
    dfree(bp.pcount);
    dfree(bp.tmp);
    bp = cbz_context::Broad{};
  }
  dfree(c->l_tile_world);
  dfree(c->l_world_tile0);
  dfree(c->l_world_q);
  dfree(c->l_off);
  dfree(c->l_one);
  dfree(c->l_two);
  dfree(c->l_cnt);
  dfree(c->tile_cnt);
  dfree(c->tile_off);
  dfree(c->tile_mask);
  dfree(c->c_body0);
  dfree(c->c_body1);
  dfree(c->c_real);
  dfree(c->c_rec);
  dfree(c->world_cnt);
  dfree(c->world_coff);
  dfree(c->total);
  dfree(c->overflow);
  dfree(c->world_iters);
  dfree(c->world_time);
  dfree(c->mw_list);
  dfree(c->mw_count);
  dfree(c->mw_flag);
  dfree(c->order);
  dfree(c->d_awake);
  dfree(c->d_prof);
  dfree(c->pos0);
  c->l_tiles_cap = c->l_pairs_cap = c->tiles_cap = 0;
  c->have_contacts = false;
}

template <typename R>
Bodies<R> bodies_view(cbz_ctx* c) {
  Bodies<R> b;
  b.n = (int32_t)c->nb;
  b.stride = c->nb;
  b.shape = (const int32_t*)c->field[CBZ_SHAPE];
  b.half = (const R*)c->field[CBZ_HALF];
  b.pos = (R*)c->field[CBZ_POS];
  b.quat = (R*)c->field[CBZ_QUAT];
  b.vel = (R*)c->field[CBZ_VEL];
  b.rot = (R*)c->field[CBZ_ROT];
  b.acc = (const R*)c->field[CBZ_ACC];
  b.inv_mass = (const R*)c->field[CBZ_INV_MASS];
  b.iit_body = (const R*)c->field[CBZ_IIT_BODY];
  b.iit_world = (R*)c->field[CBZ_IIT_WORLD];
  b.transform = (R*)c->field[CBZ_TRANSFORM];
  b.ctransform = (R*)c->field[CBZ_CTRANSFORM];
  b.offset = (const R*)c->field[CBZ_OFFSET];
  b.last_acc = (R*)c->field[CBZ_LAST_ACC];
  b.motion = (R*)c->field[CBZ_MOTION];
  b.pow_lin = (const R*)c->pow_lin;
  b.pow_ang = (const R*)c->pow_ang;
  b.flags = (uint8_t*)c->field[CBZ_FLAGS];
  b.n_awake = c->count_awake ? c->d_awake : nullptr;
  return b;
}

template <typename R>
Contacts<R> contacts_view(cbz_ctx* c) {
  Contacts<R> v;
  const int64_t S = c->cap;
  R* base = (R*)c->c_real;
  v.cap = S;
  v.body0 = (int32_t*)c->c_body0;
  v.body1 = (int32_t*)c->c_body1;
  v.point = base;
  v.normal = base + 3 * S;
  v.pen = base + 6 * S;
  v.c2w = base + 7 * S;
  v.rel = base + 16 * S;
  v.cvel = base + 22 * S;
  v.ddv = base + 25 * S;
  const bool mat = (c->mat_flags & CBZ_MATERIALS) != 0;
  v.fric = mat ? base + 26 * S : nullptr;
  v.rest = mat ? base + 27 * S : nullptr;
  v.imp = base + 28 * S;
  v.rec = (R*)c->c_rec;
  return v;
}

int ensure_tiles(cbz_ctx* c, int32_t n) {
  if (n <= c->tiles_cap) return CBZ_OK;
  int rc;
  if ((rc = dalloc(c, &c->tile_cnt, (size_t)n))) return rc;
  if ((rc = dalloc(c, &c->tile_off, (size_t)n))) return rc;
  if ((rc = dalloc(c, &c->tile_mask, (size_t)n))) return rc;
  c->tiles_cap = n;
  return CBZ_OK;
}

int need_scene(cbz_ctx* c) {
  if (!c) return CBZ_ERR_ARG;
  if (c->nw <= 0) return fail(c, CBZ_ERR_STATE, "no scene: call cbz_set_scene first");
  return CBZ_OK;
}

// Conservative bounding radii for the narrowphase's far-pair rejection.
int refresh_brad(cbz_ctx* c) {
  if (!c->brad_dirty) return CBZ_OK;
  std::vector<double> r((size_t)c->nb);
  for (int64_t i = 0; i < c->nb; ++i) {
    const double* h = &c->h_half[(size_t)i * 3];
    double v = c->h_shape[i] == CBZ_SPHERE ? std::fabs(h[0])
                                           : std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    r[i] = v * (1.0 + 1e-9);
    if (!std::isfinite(r[i])) r[i] = std::numeric_limits<double>::infinity();
  }
  if (c->nb > 0) HIPCHK(c, hipMemcpy(c->brad, r.data(), r.size() * sizeof(double), hipMemcpyHostToDevice));
  c->brad_dirty = false;
  return CBZ_OK;
}

// Pow(LinearDamping, dt), Pow(AngularDamping, dt) per body and Pow(0.5, dt)
// (rigidbody.go:233-234, 250), evaluated with the C library.
int refresh_pow(cbz_ctx* c, double dt) {
  if (!c->pow_dirty && dt == c->pow_dt) return CBZ_OK;
  std::vector<double> pl((size_t)c->nb), pa((size_t)c->nb);
  double last_l = std::numeric_limits<double>::quiet_NaN(), last_a = last_l, vl = 0, va = 0;
  for (int64_t i = 0; i < c->nb; ++i) {
    if (!(c->lin_damp[i] == last_l)) {
      last_l = c->lin_damp[i];
      vl = std::pow(last_l, dt);
    }
    if (!(c->ang_damp[i] == last_a)) {
      last_a = c->ang_damp[i];
      va = std::pow(last_a, dt);
    }
    pl[i] = vl;
    pa[i] = va;
  }
  int rc;
  if ((rc = upload_real(c, c->pow_lin, pl.data(), c->nb, 1))) return rc;
  if ((rc = upload_real(c, c->pow_ang, pa.data(), c->nb, 1))) return rc;
  c->sleep_bias = std::pow(0.5, dt);
  c->pow_dt = dt;
  c->pow_dirty = false;
  return CBZ_OK;
}


template <typename R>
int do_integrate(cbz_ctx* c, double dt) {
  int rc = refresh_pow(c, dt);
  if (rc) return rc;
  HIPCHK(c, launch_integrate<R>(bodies_view<R>(c), (R)dt, (R)c->sleep_bias, c->stream));
  return CBZ_OK;
}

// list_tiles < 0 selects the all-pairs space.
template <typename R>
int do_narrowphase(cbz_ctx* c, int32_t list_tiles, hipEvent_t* ev = nullptr) {
  int rc = refresh_brad(c);
  if (rc) return rc;
  if (list_tiles < 0) c->l_last = -1;  // no pair list: no per-pair counts
  NarrowLaunch<R> L;
  L.b = bodies_view<R>(c);
  L.pl.normal = (const R*)c->plane_normal;
  L.pl.offset = (const R*)c->plane_offset;
  L.pl.stride = c->np;
  L.ps.body_off = c->d_body_off;
  L.ps.plane_off = c->d_plane_off;
  if (list_tiles < 0) {
    L.ps.n_tiles = c->n_tiles;
    L.ps.max_world_tiles = c->max_world_tiles;
    L.ps.tile_world = c->d_tile_world;
    L.ps.world_tile0 = c->d_world_tile0;
    L.ps.world_q = nullptr;
    L.ps.n_tasks = c->n_tasks;
    L.ps.task_world = c->d_task_world;
    L.ps.world_task0 = c->d_world_task0;
    L.ps.list_off = nullptr;
    L.ps.list_one = nullptr;
    L.ps.list_cnt = nullptr;
    L.ps.list_two = nullptr;
  } else {
    L.ps.n_tiles = list_tiles;
    L.ps.max_world_tiles = 0;  // list worlds are short: the one-wave scan
    L.ps.tile_world = c->l_tile_world;
    L.ps.world_tile0 = c->l_world_tile0;
    L.ps.world_q = c->l_world_q;
    L.ps.n_tasks = 0;
    L.ps.task_world = nullptr;
    L.ps.world_task0 = nullptr;
    L.ps.list_off = c->l_off;
    L.ps.list_one = c->l_one;
    L.ps.list_two = c->l_two;
    L.ps.list_cnt = c->l_cnt;
  }
  L.c = contacts_view<R>(c);
  L.brad = c->brad;
  L.tile_cnt = c->tile_cnt;
  L.tile_mask = c->tile_mask;
  L.tile_off = c->tile_off;
  L.world_cnt = c->world_cnt;
  L.world_coff = c->world_coff;
  L.total = c->total;
  L.overflow = c->overflow;
  L.hits = c->hits;
  L.n_hits = c->n_hits;
  L.cand = c->cand;
  L.n_cand = c->n_cand;
  L.cand_cap = c->cand_cap;
  L.n_worlds = c->nw;
  L.mat.body_fric = (const R*)c->field[CBZ_FRICTION];
  L.mat.body_rest = (const R*)c->field[CBZ_RESTITUTION];
  L.mat.plane_fric = (const R*)c->plane_fric;
  L.mat.plane_rest = (const R*)c->plane_rest;
  HIPCHK(c, launch_fill_i32(c->overflow, 1, 0, c->stream));
  HIPCHK(c, launch_narrowphase<R>(L, c->stream, ev));
  c->have_contacts = true;
  return CBZ_OK;
}

template <typename R>
int do_resolve(cbz_ctx* c, double dt, int64_t mult, int64_t fixed, int32_t mode = 3,
               uint8_t* mask = nullptr, float* labels = nullptr, double polish_tol = 0, int64_t polish_mult = 0) {
  ResolveLaunch<R> L;
  L.mode = mode;
  if (c->mat_flags & CBZ_MATERIALS_FRICTIONLESS_TWO_BODY) L.mode |= RM_FRICTIONLESS2;
  L.world_mask = mask;
  L.labels = labels;
  if (mode & RM_POLISH) {
    if (mode & RM_POS) {
      L.delta = c->sur.dpos;
      L.vel0 = (R*)c->sur.psnap;
    } else {
      L.delta = c->sur.delta;
      L.vel0 = (R*)c->sur.vel0;
      L.cvel0 = (R*)c->sur.cvel0;
    }
    L.polish_tol = (R)polish_tol;
    L.polish_mult = polish_mult;
  }
  L.pos0 = (R*)c->pos0;
  L.b = bodies_view<R>(c);
  L.c = contacts_view<R>(c);
  L.world_cnt = c->world_cnt;
  L.world_coff = c->world_coff;
  L.overflow = c->overflow;
  L.order = c->order;
  L.body_off = c->d_body_off;
  L.world_iters = c->world_iters;
  L.world_time = c->world_time;
  L.mw_min = c->mw_min;
  L.mw_max = c->mw_max;
  // A launch without iterations (prepareContacts only) gains nothing from the
  // multi-wave kernel, and leaves the worlds' iteration counts (the next
  // position launch picks its multi-wave worlds by the last ones) alone.
  if (mult == 0 && fixed == 0) {
    L.mw_max = 0;
    L.world_iters = nullptr;
  }
  L.mw_list = c->mw_list;
  L.mw_count = c->mw_count;
  L.mw_flag = c->mw_flag;
  L.side = c->side;
  L.fork = c->mw_fork;
  L.sel = c->mw_sel;
  L.join = c->mw_join;
  L.n_cu = c->n_cu;
  L.n_worlds = c->nw;
  L.dt = (R)dt;
  L.iter_mult = mult;
  L.iter_fixed = fixed;
  HIPCHK(c, launch_resolve<R>(L, c->stream));
  return CBZ_OK;
}

// List buffers for n_pairs explicit pairs.
int ensure_list(cbz_ctx* c, int64_t n_pairs) {
  int rc;
  if (n_pairs > c->l_pairs_cap) {
    if ((rc = dalloc(c, &c->l_one, (size_t)n_pairs))) return rc;
    if ((rc = dalloc(c, &c->l_two, (size_t)n_pairs))) return rc;
    if ((rc = dalloc(c, &c->l_cnt, (size_t)n_pairs))) return rc;
    c->l_pairs_cap = (int32_t)n_pairs;
  }
  if (!c->l_world_tile0) {
    if ((rc = dalloc(c, &c->l_world_tile0, (size_t)c->nw + 1))) return rc;
    if ((rc = dalloc(c, &c->l_world_q, (size_t)c->nw))) return rc;
    if ((rc = dalloc(c, &c->l_off, (size_t)c->nw + 1))) return rc;
  }
  return CBZ_OK;
}

int check_overflow(cbz_ctx* c);

// The explicit-list narrowphase over l_one/l_two (world w's pairs from
// l_off[w], wq[w] of them): tiles of 64 consecutive pairs per world.
int run_list_narrowphase(cbz_ctx* c, const std::vector<int32_t>& wq, int64_t* total) {
  std::vector<int32_t> wt0(c->nw + 1), tw;
  int64_t nt = 0;
  for (int32_t w = 0; w < c->nw; ++w) {
    wt0[w] = (int32_t)nt;
    nt += (wq[w] + 63) / 64;
  }
  wt0[c->nw] = (int32_t)nt;
  if (nt > INT32_MAX / 64) return fail(c, CBZ_ERR_ARG, "narrowphase: pair list too long");
  tw.resize((size_t)nt);
  for (int32_t w = 0; w < c->nw; ++w)
    for (int32_t t = wt0[w]; t < wt0[w + 1]; ++t) tw[t] = w;
  int rc;
  if (nt > c->l_tiles_cap) {
    if ((rc = dalloc(c, &c->l_tile_world, (size_t)nt))) return rc;
    c->l_tiles_cap = (int32_t)nt;
  }
  if ((rc = ensure_tiles(c, (int32_t)std::max<int64_t>(nt, 1)))) return rc;
  if (nt > 0) HIPCHK(c, hipMemcpy(c->l_tile_world, tw.data(), nt * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->l_world_tile0, wt0.data(), wt0.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->l_world_q, wq.data(), wq.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(c, launch_fill_i32(c->overflow, 2, 0, c->stream));
  rc = c->precision == CBZ_F64 ? do_narrowphase<double>(c, (int32_t)nt) : do_narrowphase<float>(c, (int32_t)nt);
  if (rc) return rc;
  if ((rc = check_overflow(c))) return rc;
  if (total) HIPCHK(c, hipMemcpy(total, c->total, sizeof(int64_t), hipMemcpyDeviceToHost));
  return CBZ_OK;
}

// Broadphase buffers sized for the scene (and pair keys for pcap pairs).
int bp_ensure(cbz_ctx* c, int64_t pcap) {
  auto& bp = c->bp;
  int rc;
  const size_t nb = (size_t)std::max<int64_t>(c->nb, 1);
  if (!bp.ready) {
    std::vector<int32_t> bw(nb, 0);
    std::vector<int64_t> base((size_t)c->nw, 0);
    int64_t acc = 0;
    for (int32_t w = 0; w < c->nw; ++w) {
      for (int32_t i = c->body_off[w]; i < c->body_off[w + 1]; ++i) bw[i] = w;
      base[w] = acc;
      acc += (int64_t)(c->body_off[w + 1] - c->body_off[w]) * (c->plane_off[w + 1] - c->plane_off[w]);
    }
    if ((rc = dalloc(c, &bp.body_world, nb))) return rc;
    if ((rc = dalloc(c, &bp.lo, 2 * nb))) return rc;
    if ((rc = dalloc(c, &bp.hi, 2 * nb))) return rc;
    if ((rc = dalloc(c, &bp.hix, nb))) return rc;
    if ((rc = dalloc(c, &bp.wild, nb))) return rc;
    if ((rc = dalloc(c, &bp.key_in, nb))) return rc;
    if ((rc = dalloc(c, &bp.key_out, nb))) return rc;
    if ((rc = dalloc(c, &bp.val_in, nb))) return rc;
    if ((rc = dalloc(c, &bp.val_out, nb))) return rc;
    if ((rc = dalloc(c, &bp.plane_base, (size_t)std::max(c->nw, 1)))) return rc;
    if ((rc = dalloc(c, &bp.pcount, 1))) return rc;
    if (c->nb > 0) HIPCHK(c, hipMemcpy(bp.body_world, bw.data(), c->nb * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(bp.plane_base, base.data(), base.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    bp.plane_pairs = acc;
    bp.ready = true;
  }
  if (pcap > bp.pcap) {
    if ((rc = dalloc(c, &bp.pkey_in, (size_t)pcap))) return rc;
    if ((rc = dalloc(c, &bp.pkey, (size_t)pcap))) return rc;
    bp.pcap = pcap;
  }
  size_t need = 0;
  HIPCHK(c, bp_sort_bytes((int32_t)c->nb, bp.pcap, &need));
  if (need > bp.tmp_bytes) {
    dfree(bp.tmp);
    HIPCHK(c, hipMalloc(&bp.tmp, need));
    bp.tmp_bytes = need;
  }
  return CBZ_OK;
}

template <typename R>
BroadArgs<R> bp_args(cbz_ctx* c) {
  auto& bp = c->bp;
  BroadArgs<R> a;
  a.nb = (int32_t)c->nb;
  a.stride = c->nb;
  a.ctransform = (const R*)c->field[CBZ_CTRANSFORM];
  a.brad = c->brad;
  a.body_world = bp.body_world;
  a.body_off = c->d_body_off;
  a.plane_off = c->d_plane_off;
  a.lo = bp.lo;
  a.hi = bp.hi;
  a.hix = bp.hix;
  a.wild = bp.wild;
  a.key_in = bp.key_in;
  a.key_out = bp.key_out;
  a.val_in = bp.val_in;
  a.val_out = bp.val_out;
  a.pkey = bp.pkey_in;
  a.pcap = bp.pcap;
  a.plane_pairs = bp.plane_pairs;
  a.pcount = bp.pcount;
  return a;
}

// Broadphase into the explicit list: returns the number of listed pairs.
template <typename R>
int do_broadphase(cbz_ctx* c, int64_t* n_pairs) {
  auto& bp = c->bp;
  int rc;
  if ((rc = bp_ensure(c, 0))) return rc;  // scene tables and the body-plane pair count
  if (bp.plane_pairs > INT32_MAX / 2) return fail(c, CBZ_ERR_CAPACITY, "broadphase: too many body-plane pairs");
  // the body-plane keys always fit: the capacity is at least plane_pairs
  if ((rc = bp_ensure(c, std::min<int64_t>(bp.plane_pairs + 16 * c->nb + 1024, INT32_MAX / 2)))) return rc;
  if ((rc = refresh_brad(c))) return rc;
  HIPCHK(c, launch_bp_bounds_sort<R>(bp_args<R>(c), bp.tmp, bp.tmp_bytes, c->stream));
  BroadPlanes pp{(int32_t)c->nb, bp.body_world, c->d_body_off, c->d_plane_off, bp.plane_base, bp.pkey_in};
  int64_t found = 0;
  for (int round = 0; round < 2; ++round) {
    pp.pkey = bp.pkey_in;
    HIPCHK(c, launch_bp_sweep<R>(bp_args<R>(c), pp, c->stream));
    HIPCHK(c, hipMemcpyAsync(&found, bp.pcount, sizeof found, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (bp.plane_pairs + found > INT32_MAX / 2)
      return fail(c, CBZ_ERR_CAPACITY, "broadphase: more than 2^30 candidate pairs");
    if (bp.plane_pairs + found <= bp.pcap) break;
    if (round == 1) return fail(c, CBZ_ERR_HIP, "broadphase: pair count changed between sweeps");
    if ((rc = bp_ensure(c, std::min<int64_t>(bp.plane_pairs + found + found / 4 + 1024, INT32_MAX / 2)))) return rc;
  }
  const int64_t n = bp.plane_pairs + found;
  if ((rc = ensure_list(c, n))) return rc;
  BroadList L{n, c->nw, bp.pkey_in, bp.pkey, bp.body_world, c->d_body_off, c->d_plane_off, c->l_one, c->l_two,
              c->l_off};
  HIPCHK(c, launch_bp_list(L, bp.tmp, bp.tmp_bytes, c->stream));
  *n_pairs = n;
  return CBZ_OK;
}

int check_overflow(cbz_ctx* c) {
  int32_t of = 0, nc = 0;
  int64_t tot = 0;
  HIPCHK(c, hipMemcpyAsync(&of, c->overflow + 1, sizeof of, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&tot, c->total, sizeof tot, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&nc, c->n_cand, sizeof nc, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (of && nc > c->cand_cap)
    return fail(c, CBZ_ERR_CAPACITY,
                "narrowphase candidate list exceeded: " + std::to_string(nc) + " candidate pairs for " +
                    std::to_string(c->cand_cap) + " slots (raise cbz_config.contact_capacity)");
  if (of)
    return fail(c, CBZ_ERR_CAPACITY,
                "contact capacity exceeded: " + std::to_string(tot) + " contacts for " + std::to_string(c->cap) +
                    " slots (raise cbz_config.contact_capacity)");
  return CBZ_OK;
}

// The most bodies in one world, and the bound up to which the feature rows
// come from the per-world kernel (its per-body counts in LDS).
constexpr int32_t kFeatWorldBodiesMax = 4096;
int32_t max_world_bodies(const cbz_ctx* c) {
  int32_t m = 0;
  for (int32_t w = 0; w < c->nw; ++w) m = std::max(m, c->body_off[w + 1] - c->body_off[w]);
  return m;
}

// World order for the resolver: larger worlds first, so the longest
// serial chains start earliest.
void build_order(cbz_ctx* c, std::vector<int32_t>& ord) {
  ord.resize((size_t)c->nw);
  for (int32_t w = 0; w < c->nw; ++w) ord[w] = w;
  std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) {
    return (c->body_off[a + 1] - c->body_off[a]) > (c->body_off[b + 1] - c->body_off[b]);
  });
}


// ----------------------------------------------------------------------
// resolver surrogate: host orchestration
// ----------------------------------------------------------------------
struct LayerOff {
  int64_t w[kLayers], b[kLayers], total;
};
LayerOff layer_offsets() {
  LayerOff o{};
  int64_t off = 0;
  for (int l = 0; l < kLayers; ++l) {
    o.w[l] = off;
    off += int64_t(layer_out(l)) * layer_in(l);
    o.b[l] = off;
    off += layer_out(l);
  }
  o.total = off;
  return o;
}
constexpr int kNorm = kFeat + kFeat + kOut + kOut + kOut;
int64_t round32(int64_t n) { return (n + 31) / 32 * 32; }

// The training batch buffers for Mpad rows, each with its element count:
// the allocator and the shape check (cbz_sur_check_gemms) share this list.
template <typename F>
void sur_batch_buffers(cbz_context::Sur& u, int64_t Mpad, F&& f) {
  f(u.idx, Mpad);
  f(u.bX, Mpad * kFeatPad);
  f(u.bT, Mpad * kOut);
  f(u.At[0], Mpad * kFeatPad);
  for (int l = 0; l < kLayers; ++l) {
    const int64_t w = Mpad * layer_out_pad(l);
    f(u.Z[l], w);
    f(u.dZ[l], w);
    f(u.dZt[l], w);
    if (l < kLayers - 1) {
      f(u.Hb[l], w);
      f(u.At[l + 1], w);
    }
  }
  f(u.Yout, Mpad * kOut);
  f(u.dA, Mpad * kHidden);
  f(u.row_loss, Mpad);
}

// The model's device buffers (master weights, optimiser state,
// normalisation, packed BF16 weights) with their element counts.
template <typename F>
void sur_model_buffers(cbz_context::Sur& u, F&& f) {
  const int64_t total = layer_offsets().total;
  f(u.params, total);
  f(u.grad, total);
  f(u.adam_m, total);
  f(u.adam_v, total);
  f(u.norm, int64_t(kNorm));
  for (int l = 0; l < kLayers; ++l) {
    f(u.Wnk[l], int64_t(layer_out_pad(l)) * layer_in_pad(l));
    f(u.Wkn[l], int64_t(layer_out_pad(l)) * layer_in_pad(l));
  }
}

// The fused inference's features and predictions for Mpad rows.
template <typename F>
void sur_infer_buffers(cbz_context::Sur& u, int64_t Mpad, F&& f) {
  f(u.X, Mpad * kFeatPad);
  f(u.Yp, Mpad * kOut);
}

// The per-layer inference's activations for Mpad rows (cbz_sur_predict
// with CBZ_SUR_PREDICT_LAYERED only).
template <typename F>
void sur_layered_buffers(cbz_context::Sur& u, int64_t Mpad, F&& f) {
  f(u.H[0], Mpad * kHidden);
  f(u.H[1], Mpad * kHidden);
}


int sur_alloc_model(cbz_ctx* c) {
  auto& u = c->sur;
  if (u.params) return CBZ_OK;
  int rc = CBZ_OK;
  sur_model_buffers(u, [&](auto*& p, int64_t n) {
    if (rc == CBZ_OK) rc = dalloc(c, &p, (size_t)n);
    // zeroed: the optimiser state, and the padding of the packed weights
    if (rc == CBZ_OK && hipMemset(p, 0, (size_t)n * sizeof(*p)) != hipSuccess)
      rc = fail(c, CBZ_ERR_HIP, "sur: hipMemset failed");
  });
  return rc;
}

void sur_free_model(cbz_ctx* c) {
  auto& u = c->sur;
  dfree(u.params);
  dfree(u.grad);
  dfree(u.adam_m);
  dfree(u.adam_v);
  dfree(u.norm);
  for (int l = 0; l < kLayers; ++l) {
    dfree(u.Wnk[l]);
    dfree(u.Wkn[l]);
  }
  dfree(u.dsX);
  dfree(u.dsY);
  dfree(u.idx);
  dfree(u.bX);
  dfree(u.bT);
  for (auto& p : u.Z) dfree(p);
  for (auto& p : u.Hb) dfree(p);
  for (auto& p : u.At) dfree(p);
  for (auto& p : u.dZ) dfree(p);
  for (auto& p : u.dZt) dfree(p);
  dfree(u.Yout);
  dfree(u.dA);
  dfree(u.row_loss);
  dfree(u.losses);
  u = cbz_context::Sur{};
}

int sur_pack(cbz_ctx* c) {
  auto& u = c->sur;
  const LayerOff o = layer_offsets();
  for (int l = 0; l < kLayers; ++l)
    HIPCHK(c, launch_pack_weights(u.params + o.w[l], layer_out(l), layer_in(l), u.Wnk[l], layer_in_pad(l), u.Wkn[l],
                                  layer_out_pad(l), c->stream));
  return CBZ_OK;
}

// The bytes [p, p + bytes) lie inside one region of the shape check.
bool dry_inside(const cbz_ctx* c, const void* p, int64_t bytes) {
  const uintptr_t a = (uintptr_t)p;
  for (const auto& r : *c->dry_regions)
    if (a >= r.first && a < r.second) return bytes >= 0 && a + (uintptr_t)bytes <= r.second;
  return false;
}

// One GEMM of the surrogate: launched, or in the shape check compared with
// the extents of the buffers it reads and writes (k_gemm reads rows < M of A
// and rows < N of Bt, K elements each; writes Z/Y where row < m_valid and
// col < n_valid, H/Ht where row < M and col < N).
int run_gemm(cbz_ctx* c, const GemmArgs& g, const char* what) {
  if (!c->dry_regions) {
    HIPCHK(c, launch_gemm(g, c->stream));
    return CBZ_OK;
  }
  auto bad = [&](const char* op) {
    return fail(c, CBZ_ERR_STATE, std::string("GEMM ") + what + ": operand " + op + " exceeds its buffer (M " +
                                      std::to_string(g.M) + ", N " + std::to_string(g.N) + ", K " +
                                      std::to_string(g.K) + ")");
  };
  if (g.M <= 0 || g.N <= 0) return CBZ_OK;
  if (g.K <= 0 || g.K % 16 || g.lda < g.K || g.ldb < g.K) return bad("K");
  const int64_t b2 = 2, b4 = 4;
  if (!dry_inside(c, g.A, ((g.M - 1) * g.lda + g.K) * b2)) return bad("A");
  if (!dry_inside(c, g.Bt, ((g.N - 1) * g.ldb + g.K) * b2)) return bad("Bt");
  const int64_t mv = std::min(g.M, g.m_valid), nv = std::min(g.N, g.n_valid);
  if (g.bias && nv > 0 && !dry_inside(c, g.bias, nv * b4)) return bad("bias");
  if (mv > 0 && nv > 0) {
    if (g.Z && !dry_inside(c, g.Z, ((mv - 1) * g.ldz + nv) * b4)) return bad("Z");
    if (g.Y && !dry_inside(c, g.Y, ((mv - 1) * g.ldy + nv) * b4)) return bad("Y");
    if (g.y_scale && (!dry_inside(c, g.y_scale, nv * b4) || !dry_inside(c, g.y_shift, nv * b4))) return bad("y_scale");
  }
  if (g.H && !dry_inside(c, g.H, ((g.M - 1) * g.ldh + g.N) * b2)) return bad("H");
  if (g.Ht && !dry_inside(c, g.Ht, ((g.N - 1) * g.ldht + g.M) * b2)) return bad("Ht");
  return CBZ_OK;
}

// Forward pass over M rows (Mpad = M rounded up to 32) of BF16 inputs X.
// train: keep pre-activations, activations and their transposes for the
// backward pass and write normalised outputs to Yout; else write
// denormalised predictions to Yp.
int sur_forward(cbz_ctx* c, const __bf16* X, int64_t M, int64_t Mpad, bool train) {
  auto& u = c->sur;
  const LayerOff o = layer_offsets();
  const __bf16* in = X;
  for (int l = 0; l < kLayers; ++l) {
    GemmArgs g{};
    g.M = (int)Mpad;
    g.N = layer_out_pad(l);
    g.K = layer_in_pad(l);
    g.A = in;
    g.lda = layer_in_pad(l);
    g.Bt = u.Wnk[l];
    g.ldb = layer_in_pad(l);
    g.bias = u.params + o.b[l];
    g.act = l < kLayers - 1;
    g.m_valid = (int)M;
    g.n_valid = layer_out(l);
    if (l < kLayers - 1) {
      __bf16* out = train ? u.Hb[l] : u.H[l % 2];
      g.H = out;
      g.ldh = layer_out_pad(l);
      if (train) {
        g.Z = u.Z[l];
        g.ldz = layer_out_pad(l);
        g.Ht = u.At[l + 1];
        g.ldht = Mpad;
      }
      in = out;
    } else if (train) {
      g.Y = u.Yout;
      g.ldy = kOut;
    } else {
      g.Y = u.Yp;
      g.ldy = kOut;
      g.y_scale = u.norm + 2 * kFeat + kOut;  // out_std
      g.y_shift = u.norm + 2 * kFeat;         // out_mean
    }
    int rc = run_gemm(c, g, train ? "forward (training)" : "forward");
    if (rc) return rc;
  }
  return CBZ_OK;
}

// Inference forward: all four layers in one fused launch (k_mlp_forward),
// rows of X [M][kFeatPad] -> denormalised predictions in u.Yp; with M_dev,
// the rows are the device count clamped to M.
int sur_infer(cbz_ctx* c, const __bf16* X, int64_t M, const int64_t* M_dev = nullptr) {
  auto& u = c->sur;
  const LayerOff o = layer_offsets();
  MlpArgs f{};
  f.M = M;
  f.M_dev = M_dev;
  f.X = X;
  for (int l = 0; l < kLayers; ++l) {
    f.W[l] = u.Wnk[l];
    f.bias[l] = u.params + o.b[l];
  }
  f.y_scale = u.norm + 2 * kFeat + kOut;  // out_std
  f.y_shift = u.norm + 2 * kFeat;         // out_mean
  f.Y = u.Yp;
  HIPCHK(c, launch_mlp_forward(f, c->stream));
  return CBZ_OK;
}

// Backward pass of the mean-squared error against bT; gradients into grad.
int sur_backward(cbz_ctx* c, int64_t M, int64_t Mpad) {
  auto& u = c->sur;
  const LayerOff o = layer_offsets();
  const bool dry = c->dry_regions != nullptr;
  int rc;
  if (!dry)
    HIPCHK(c, launch_mse(u.Yout, u.bT, (int)Mpad, (int)M, u.row_loss, u.dZ[kLayers - 1], u.dZt[kLayers - 1],
                         (int)Mpad, c->stream));
  for (int l = kLayers - 1; l >= 0; --l) {
    const int N = layer_out(l), K = layer_in(l), Np = layer_out_pad(l), Kp = layer_in_pad(l);
    // dW = dZ^T * A_in : [Np x Mpad] * [Mpad x Kp]
    GemmArgs g{};
    g.M = Np;
    g.N = Kp;
    g.K = (int)Mpad;
    g.A = u.dZt[l];
    g.lda = Mpad;
    g.Bt = u.At[l];
    g.ldb = Mpad;
    g.Y = u.grad + o.w[l];
    g.ldy = K;
    g.m_valid = N;
    g.n_valid = K;
    if ((rc = run_gemm(c, g, "weight gradient"))) return rc;
    if (!dry) HIPCHK(c, launch_colsum(u.dZ[l], (int)Mpad, N, Np, u.grad + o.b[l], c->stream));
    if (l > 0) {
      // dA_in = dZ * W : [Mpad x Np] * [Np x Kp], B given as Wkn [Kp x Np]
      GemmArgs h{};
      h.M = (int)Mpad;
      h.N = Kp;
      h.K = Np;
      h.A = u.dZ[l];
      h.lda = Np;
      h.Bt = u.Wkn[l];
      h.ldb = Np;
      h.Y = u.dA;
      h.ldy = Kp;
      h.m_valid = (int)M;
      h.n_valid = Kp;
      if ((rc = run_gemm(c, h, "input gradient"))) return rc;
      if (!dry)
        HIPCHK(c, launch_silu_bwd(u.dA, u.Z[l - 1], (int)Mpad, (int)M, Kp, Kp, u.dZ[l - 1], u.dZt[l - 1], (int)Mpad,
                                c->stream));
    }
  }
  return CBZ_OK;
}

int sur_ensure_batch(cbz_ctx* c, int rows) {
  auto& u = c->sur;
  const int Mpad = (int)round32(rows);
  if (Mpad <= u.batch_cap) return CBZ_OK;
  int rc = CBZ_OK;
  sur_batch_buffers(u, Mpad, [&](auto*& p, int64_t n) {
    if (rc == CBZ_OK) rc = dalloc(c, &p, (size_t)n);
  });
  if (rc) return rc;
  u.batch_cap = Mpad;
  return CBZ_OK;
}

int sur_ensure_inference(cbz_ctx* c, int64_t rows) {
  auto& u = c->sur;
  int64_t Mpad = round32(std::max<int64_t>(rows, 32));
  if (Mpad <= u.inf_cap && u.mask) return CBZ_OK;
  Mpad = std::max(Mpad, u.inf_cap + u.inf_cap / 4);  // grow geometrically
  int rc = CBZ_OK;
  sur_infer_buffers(u, Mpad, [&](auto*& p, int64_t n) {
    if (rc == CBZ_OK) rc = dalloc(c, &p, (size_t)n);
  });
  if (rc) return rc;
  if ((rc = dalloc(c, &u.deg, (size_t)std::max<int64_t>(c->nb, 1)))) return rc;
  if ((rc = dalloc(c, &u.mask, (size_t)std::max(c->nw, 1)))) return rc;
  if ((rc = dalloc(c, &u.nfall, 1))) return rc;
  const size_t nb = (size_t)std::max<int64_t>(c->nb, 1);
  if ((rc = dalloc(c, &u.delta, 6 * nb))) return rc;
  if ((rc = dalloc_bytes(c, &u.vel0, 7 * nb * real_size(c)))) return rc;
  if ((rc = dalloc_bytes(c, &u.cvel0, 3 * (size_t)std::max<int64_t>(c->cap, 1) * real_size(c)))) return rc;
  if ((rc = dalloc(c, &u.pmask, (size_t)std::max(c->nw, 1)))) return rc;
  if ((rc = dalloc(c, &u.dpos, 6 * nb))) return rc;
  if ((rc = dalloc_bytes(c, &u.psnap, 22 * nb * real_size(c)))) return rc;
  u.inf_cap = Mpad;
  u.generation += 1;  // buffers moved: a captured surrogate step is stale
  return CBZ_OK;
}

int sur_ensure_layered(cbz_ctx* c, int64_t rows) {
  auto& u = c->sur;
  const int64_t Mpad = round32(std::max<int64_t>(rows, 32));
  if (Mpad <= u.lay_cap) return CBZ_OK;
  int rc = CBZ_OK;
  sur_layered_buffers(u, Mpad, [&](auto*& p, int64_t n) {
    if (rc == CBZ_OK) rc = dalloc(c, &p, (size_t)n);
  });
  if (rc) return rc;
  u.lay_cap = Mpad;
  return CBZ_OK;
}

// Profiling of one surrogate resolve (events ev[8..13] recorded, ev[14]
// recorded here): accumulates the phase times, the contact count, the worlds
// given a plain phase and the iterations of both phases.
int sur_profile(cbz_ctx* c, bool pos) {
  hipEvent_t* ev = c->ev + 8;
  HIPCHK(c, hipEventRecord(ev[6], c->stream));
  HIPCHK(c, hipEventSynchronize(ev[6]));
  int64_t contacts = 0;
  HIPCHK(c, hipMemcpy(&contacts, c->total, sizeof contacts, hipMemcpyDeviceToHost));
  contacts = std::min(contacts, c->cap);
  for (int k = 0; k < 6; ++k) {
    float t = 0;
    HIPCHK(c, hipEventElapsedTime(&t, ev[k], ev[k + 1]));
    c->sur_ms[k] += t;
  }
  std::vector<uint8_t> m((size_t)c->nw), pm((size_t)c->nw, 1);
  std::vector<int64_t> it(2 * (size_t)c->nw);
  HIPCHK(c, hipMemcpy(m.data(), c->sur.mask, m.size(), hipMemcpyDeviceToHost));
  if (pos) HIPCHK(c, hipMemcpy(pm.data(), c->sur.pmask, pm.size(), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(it.data(), c->world_iters, it.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
  int64_t fb = 0, pfb = 0, vi = 0, pi = 0;
  for (int32_t w = 0; w < c->nw; ++w) {
    fb += (m[w] || (pos && pm[w])) ? 1 : 0;
    pfb += pm[w] ? 1 : 0;
    pi += it[2 * (size_t)w];
    vi += it[2 * (size_t)w + 1];
  }
  c->sur_cnt[0] += 1;
  c->sur_cnt[1] += contacts;
  c->sur_cnt[2] += fb;
  c->sur_cnt[3] += vi;
  c->sur_cnt[4] += pfb;
  c->sur_cnt[5] += pi;
  return CBZ_OK;
}

// One resolve with the surrogate:
//   1. prepareContacts (a launch of the resolver without iterations);
//   2. the network on the prepared contacts: per contact the velocity phase's
//      changes of its bodies and the penetration the position phase resolves
//      there, summed per body (with that penetration's moves);
//   3. the position phase: per world the reference's loop polishes from the
//      predicted moves to positionEpsilon + pos_tol within pos_polish * C
//      iterations, or the world is restored and given the plain position
//      phase (every world plain when the position polish is off);
//   4. the velocity phase likewise, to velocityEpsilon + tol within
//      polish_mult * C iterations.
// tol < 0 gives every world both plain phases; CBZ_SUR_ZERO_PREDICTION
// polishes from zero predictions (the same method without the network).
// *nfall_dst counts the worlds that took a plain phase.
template <typename R>
int sur_resolve(cbz_ctx* c, double dt, int64_t mult, double tol, int64_t polish_mult, int32_t flags,
                int32_t* nfall_dst) {
  auto& u = c->sur;
  hipEvent_t* ev = c->ev + 8;  // prepare, features, network, sums, position phase, velocity phase
  const bool pos = tol >= 0 && u.pos_tol >= 0;
  int rc = sur_ensure_inference(c, c->cap);
  if (rc) return rc;
  u.last_pos = pos;
  if (c->prof) HIPCHK(c, hipEventRecord(ev[0], c->stream));
  if ((rc = do_resolve<R>(c, dt, 0, 0, RM_POS))) return rc;
  if (c->prof) HIPCHK(c, hipEventRecord(ev[1], c->stream));
  if (tol < 0 || (flags & CBZ_SUR_ZERO_PREDICTION)) {
    // every world plain (tol < 0), or polished from zero predictions
    if (c->prof) {
      HIPCHK(c, hipEventRecord(ev[2], c->stream));
      HIPCHK(c, hipEventRecord(ev[3], c->stream));
    }
    HIPCHK(c, launch_fill_u8(u.mask, c->nw, tol < 0 ? 1 : 0, c->stream));
    HIPCHK(c, launch_fill_f32(u.delta, 6 * c->nb, 0.f, c->stream));
    if (pos) {
      HIPCHK(c, launch_fill_u8(u.pmask, c->nw, 0, c->stream));
      HIPCHK(c, launch_fill_f32(u.dpos, 6 * c->nb, 0.f, c->stream));
    }
  } else {
    // The feature and network grids cover the contact capacity and read the
    // contact count on the device: no host round trip, so the step can be
    // captured in a graph.
    const int32_t max_nb = max_world_bodies(c);
    if (max_nb <= kFeatWorldBodiesMax)
      HIPCHK(c, launch_features_world<R>(bodies_view<R>(c), contacts_view<R>(c), c->world_cnt, c->world_coff,
                                         c->d_body_off, c->overflow, c->nw, max_nb, u.norm, u.norm + kFeat, u.X,
                                         nullptr, c->stream));
    else
      HIPCHK(c, launch_features<R>(bodies_view<R>(c), contacts_view<R>(c), c->cap, c->total, u.deg, u.norm,
                                   u.norm + kFeat, u.X, nullptr, c->stream));
    if (c->prof) HIPCHK(c, hipEventRecord(ev[2], c->stream));
    if ((rc = sur_infer(c, u.X, c->cap, c->total))) return rc;
    if (c->prof) HIPCHK(c, hipEventRecord(ev[3], c->stream));
    SurApply<R> a;
    a.b = bodies_view<R>(c);
    a.c = contacts_view<R>(c);
    a.world_cnt = c->world_cnt;
    a.world_coff = c->world_coff;
    a.body_off = c->d_body_off;
    a.overflow = c->overflow;
    a.Y = u.Yp;
    a.delta = u.delta;
    a.mask = u.mask;
    a.dpos = pos ? u.dpos : nullptr;
    a.pmask = pos ? u.pmask : nullptr;
    a.max_contacts = kSurContacts;
    HIPCHK(c, launch_sur_scatter<R>(a, c->nw, c->stream));
  }
  if (c->prof) HIPCHK(c, hipEventRecord(ev[4], c->stream));
  rc = pos ? do_resolve<R>(c, dt, mult, 0, RM_POS | RM_PREPARED | RM_POLISH, u.pmask, nullptr, u.pos_tol, u.pos_polish)
           : do_resolve<R>(c, dt, mult, 0, RM_POS | RM_PREPARED);
  if (rc) return rc;
  if (c->prof) HIPCHK(c, hipEventRecord(ev[5], c->stream));
  rc = do_resolve<R>(c, dt, mult, 0, RM_VEL | RM_POLISH, u.mask, nullptr, tol, polish_mult);
  if (rc) return rc;
  if (nfall_dst) HIPCHK(c, launch_count_u8(u.mask, c->nw, nfall_dst, c->stream, pos ? u.pmask : nullptr));
  return c->prof ? sur_profile(c, pos) : CBZ_OK;
}

}  // namespace

// ======================================================================
// C ABI
// ======================================================================
extern "C" {

int cbz_field_width(int32_t field) {
  if (field < 0 || field >= CBZ_NUM_FIELDS) return -1;
  return kFields[field].width;
}

int cbz_create(const cbz_config* cfg, cbz_ctx** out) {
  if (!out) return CBZ_ERR_ARG;
  *out = nullptr;
  cbz_ctx* c = new cbz_ctx();
  if (cfg) {
    if (cfg->precision != CBZ_F64 && cfg->precision != CBZ_F32) {
      delete c;
      return CBZ_ERR_ARG;
    }
    c->precision = cfg->precision;
    c->device = cfg->device;
    c->cap_cfg = cfg->contact_capacity;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || c->device < 0 || c->device >= ndev) {
    delete c;
    return CBZ_ERR_HIP;
  }
  bool ok = hipSetDevice(c->device) == hipSuccess &&
            hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
  for (auto& e : c->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) == hipSuccess &&
       hipEventCreateWithFlags(&c->mw_fork, hipEventDisableTiming) == hipSuccess &&
       hipEventCreateWithFlags(&c->mw_sel, hipEventDisableTiming) == hipSuccess &&
       hipEventCreateWithFlags(&c->mw_join, hipEventDisableTiming) == hipSuccess &&
       hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess;
  ok = ok && scratch_address(c->device, &c->scratch) == hipSuccess;
  if (!ok) {
    cbz_destroy(c);
    return CBZ_ERR_HIP;
  }
  *out = c;
  return CBZ_OK;
}

int cbz_destroy(cbz_ctx* c) {
  if (!c) return CBZ_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_scene(c);
  sur_free_model(c);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->pev)
    if (e) (void)hipEventDestroy(e);
  if (c->side) (void)hipStreamSynchronize(c->side);
  if (c->mw_fork) (void)hipEventDestroy(c->mw_fork);
  if (c->mw_sel) (void)hipEventDestroy(c->mw_sel);
  if (c->mw_join) (void)hipEventDestroy(c->mw_join);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CBZ_OK;
}

const char* cbz_last_error(const cbz_ctx* c) { return c ? c->err.c_str() : "null context"; }

int32_t cbz_num_worlds(const cbz_ctx* c) { return c ? c->nw : 0; }
int64_t cbz_num_bodies(const cbz_ctx* c) { return c ? c->nb : 0; }

int cbz_set_scene(cbz_ctx* c, int32_t n_worlds, const int32_t* body_off, const int32_t* plane_off,
                  const double* plane_normal, const double* plane_offset) {
  if (!c) return CBZ_ERR_ARG;
  if (n_worlds <= 0 || !body_off || !plane_off) return fail(c, CBZ_ERR_ARG, "set_scene: bad arguments");
  if (body_off[0] != 0 || plane_off[0] != 0) return fail(c, CBZ_ERR_ARG, "set_scene: offsets must start at 0");
  for (int32_t w = 0; w < n_worlds; ++w) {
    if (body_off[w + 1] < body_off[w] || plane_off[w + 1] < plane_off[w])
      return fail(c, CBZ_ERR_ARG, "set_scene: offsets must be non-decreasing");
    const int64_t n = body_off[w + 1] - body_off[w], np = plane_off[w + 1] - plane_off[w];
    if (n * (np + n) > (int64_t)INT32_MAX - 64)
      return fail(c, CBZ_ERR_ARG, "set_scene: world " + std::to_string(w) + " has too many ordered pairs");
  }
  const int32_t np = plane_off[n_worlds];
  if (np > 0 && (!plane_normal || !plane_offset)) return fail(c, CBZ_ERR_ARG, "set_scene: planes missing");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  free_scene(c);
  c->nw = n_worlds;
  c->nb = body_off[n_worlds];
  c->np = np;
  c->body_off.assign(body_off, body_off + n_worlds + 1);
  c->plane_off.assign(plane_off, plane_off + n_worlds + 1);
  const int64_t nb = c->nb;
  const size_t rs = real_size(c);
  int rc;
  for (int f = 0; f < CBZ_NUM_FIELDS; ++f) {
    const FieldInfo& fi = kFields[f];
    if (fi.kind == 3) continue;
    size_t el = fi.kind == 0 ? rs : fi.kind == 1 ? sizeof(int32_t) : sizeof(uint8_t);
    if ((rc = dalloc_bytes(c, &c->field[f], el * fi.width * (size_t)std::max<int64_t>(nb, 1)))) return rc;
  }
  if ((rc = dalloc_bytes(c, &c->pow_lin, rs * (size_t)std::max<int64_t>(nb, 1)))) return rc;
  if ((rc = dalloc_bytes(c, &c->pow_ang, rs * (size_t)std::max<int64_t>(nb, 1)))) return rc;
  if ((rc = dalloc(c, &c->brad, (size_t)nb))) return rc;
  if ((rc = dalloc_bytes(c, &c->plane_normal, rs * 3 * (size_t)std::max(np, 1)))) return rc;
  if ((rc = dalloc_bytes(c, &c->plane_offset, rs * (size_t)std::max(np, 1)))) return rc;
  if ((rc = dalloc_bytes(c, &c->plane_fric, rs * (size_t)std::max(np, 1)))) return rc;
  if ((rc = dalloc_bytes(c, &c->plane_rest, rs * (size_t)std::max(np, 1)))) return rc;
  if (np > 0) {
    if ((rc = upload_real(c, c->plane_normal, plane_normal, np, 3))) return rc;
    if ((rc = upload_real(c, c->plane_offset, plane_offset, np, 1))) return rc;
    const std::vector<double> pf((size_t)np, kFriction), pr((size_t)np, kRestitution);
    if ((rc = upload_real(c, c->plane_fric, pf.data(), np, 1))) return rc;
    if ((rc = upload_real(c, c->plane_rest, pr.data(), np, 1))) return rc;
  }
  if ((rc = dalloc(c, &c->d_body_off, (size_t)n_worlds + 1))) return rc;
  if ((rc = dalloc(c, &c->d_plane_off, (size_t)n_worlds + 1))) return rc;
  HIPCHK(c, hipMemcpy(c->d_body_off, body_off, (n_worlds + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_plane_off, plane_off, (n_worlds + 1) * sizeof(int32_t), hipMemcpyHostToDevice));

  // all-pairs tiles: row-aligned, nch = ceil((np + n) / 64) per row; count
  // tasks: nch column tiles x ceil(n / kRowBlock) row blocks per world
  std::vector<int32_t> wt0(n_worlds + 1), tw, wk0(n_worlds + 1), kw;
  int64_t nt = 0, nk = 0, max_wt = 0;
  for (int32_t w = 0; w < n_worlds; ++w) {
    const int64_t n = body_off[w + 1] - body_off[w], npw = plane_off[w + 1] - plane_off[w];
    const int64_t nch = (npw + n + 63) / 64;
    wt0[w] = (int32_t)nt;
    wk0[w] = (int32_t)nk;
    nt += n * nch;
    max_wt = std::max<int64_t>(max_wt, n * nch);
    nk += nch * ((n + kRowBlock - 1) / kRowBlock);
    if (nt > (int64_t)(INT32_MAX / 64) || nk > INT32_MAX / 2)
      return fail(c, CBZ_ERR_ARG, "set_scene: pair space too large");
  }
  wt0[n_worlds] = (int32_t)nt;
  wk0[n_worlds] = (int32_t)nk;
  tw.resize((size_t)nt);
  kw.resize((size_t)nk);
  for (int32_t w = 0; w < n_worlds; ++w) {
    for (int32_t t = wt0[w]; t < wt0[w + 1]; ++t) tw[t] = w;
    for (int32_t k = wk0[w]; k < wk0[w + 1]; ++k) kw[k] = w;
  }
  c->n_tiles = (int32_t)nt;
  c->n_tasks = (int32_t)nk;
  c->max_world_tiles = (int32_t)max_wt;
  if ((rc = dalloc(c, &c->d_tile_world, (size_t)std::max<int64_t>(nt, 1)))) return rc;
  if ((rc = dalloc(c, &c->d_world_tile0, (size_t)n_worlds + 1))) return rc;
  if ((rc = dalloc(c, &c->d_task_world, (size_t)std::max<int64_t>(nk, 1)))) return rc;
  if ((rc = dalloc(c, &c->d_world_task0, (size_t)n_worlds + 1))) return rc;
  if (nt > 0) HIPCHK(c, hipMemcpy(c->d_tile_world, tw.data(), nt * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_world_tile0, wt0.data(), wt0.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  if (nk > 0) HIPCHK(c, hipMemcpy(c->d_task_world, kw.data(), nk * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_world_task0, wk0.data(), wk0.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  if ((rc = ensure_tiles(c, (int32_t)std::max<int64_t>(nt, 1)))) return rc;

  // contacts
  c->cap = c->cap_cfg > 0 ? c->cap_cfg : 16 * nb + 4096;
  if ((rc = dalloc_bytes(c, &c->c_body0, sizeof(int32_t) * (size_t)c->cap))) return rc;
  if ((rc = dalloc_bytes(c, &c->c_body1, sizeof(int32_t) * (size_t)c->cap))) return rc;
  if ((rc = dalloc_bytes(c, &c->c_real, rs * kContactPlanes * (size_t)c->cap))) return rc;
  if ((rc = dalloc_bytes(c, &c->c_rec, rs * 16 * (size_t)c->cap))) return rc;
  if ((rc = dalloc(c, &c->hits, (size_t)c->cap))) return rc;
  if ((rc = dalloc(c, &c->n_hits, 1))) return rc;
  // candidate pairs of the cull: 4 per contact slot plus the scene's ordered
  // pairs up to 32 M (a dense pile keeps many pairs inside the cull's bound)
  int64_t pairs_total = 0;
  for (int32_t w = 0; w < n_worlds; ++w) {
    const int64_t n = body_off[w + 1] - body_off[w], npw = plane_off[w + 1] - plane_off[w];
    pairs_total += n * (n - 1 + npw);
  }
  c->cand_cap = std::min<int64_t>(4 * c->cap + std::min<int64_t>(pairs_total, int64_t(32) << 20), INT32_MAX);
  if ((rc = dalloc(c, &c->cand, (size_t)c->cand_cap))) return rc;
  if ((rc = dalloc(c, &c->n_cand, 1))) return rc;
  if ((rc = dalloc(c, &c->world_cnt, (size_t)n_worlds))) return rc;
  if ((rc = dalloc(c, &c->world_coff, (size_t)n_worlds))) return rc;
  if ((rc = dalloc(c, &c->total, 1))) return rc;
  if ((rc = dalloc(c, &c->overflow, 2))) return rc;
  if ((rc = dalloc(c, &c->world_iters, 2 * (size_t)n_worlds))) return rc;
  if ((rc = dalloc(c, &c->world_time, 4 * (size_t)n_worlds))) return rc;
  if ((rc = dalloc(c, &c->order, (size_t)n_worlds))) return rc;
  if ((rc = dalloc(c, &c->mw_list, (size_t)n_worlds))) return rc;
  if ((rc = dalloc(c, &c->mw_count, 2))) return rc;
  if ((rc = dalloc(c, &c->mw_flag, (size_t)n_worlds))) return rc;
  if ((rc = dalloc(c, &c->d_awake, kAwakeSlots))) return rc;
  if ((rc = dalloc(c, &c->d_prof, 4))) return rc;
  if ((rc = dalloc_bytes(c, &c->pos0, rs * 3 * (size_t)std::max<int64_t>(nb, 1)))) return rc;
  HIPCHK(c, hipMemset(c->world_cnt, 0, n_worlds * sizeof(int32_t)));
  HIPCHK(c, hipMemset(c->world_coff, 0, n_worlds * sizeof(int64_t)));
  HIPCHK(c, hipMemset(c->world_iters, 0, 2 * n_worlds * sizeof(int64_t)));
  HIPCHK(c, hipMemset(c->world_time, 0xFF, 4 * n_worlds * sizeof(int64_t)));  // -1: no launch yet
  HIPCHK(c, hipMemset(c->overflow, 0, 2 * sizeof(int32_t)));
  HIPCHK(c, hipMemset(c->total, 0, sizeof(int64_t)));
  std::vector<int32_t> ord;
  build_order(c, ord);
  HIPCHK(c, hipMemcpy(c->order, ord.data(), ord.size() * sizeof(int32_t), hipMemcpyHostToDevice));

  // NewRigidBody (rigidbody.go:104-114) inside NewCollisionCube(nil, 0)
  // (colliders.go:265-274).
  c->lin_damp.assign((size_t)nb, 0.95);
  c->ang_damp.assign((size_t)nb, 0.95);  // AngularDamping = defaultLinearDamping, rigidbody.go:108
  c->h_shape.assign((size_t)nb, CBZ_CUBE);
  c->h_half.assign((size_t)nb * 3, 0.0);
  c->h_offset_id.assign((size_t)nb, 1);
  std::vector<double> z((size_t)nb * 12, 0.0);
  const int zero_fields[] = {CBZ_HALF, CBZ_POS, CBZ_VEL, CBZ_ROT, CBZ_INV_MASS, CBZ_IIT_BODY, CBZ_TRANSFORM,
                             CBZ_CTRANSFORM, CBZ_LAST_ACC};
  for (int f : zero_fields)
    if ((rc = upload_real(c, c->field[f], z.data(), nb, kFields[f].width))) return rc;
  std::vector<double> t((size_t)nb * 12);
  for (int64_t i = 0; i < nb; ++i) {
    double* q = &t[(size_t)i * 4];
    q[0] = 1.0;
    q[1] = q[2] = q[3] = 0.0;
  }
  if ((rc = upload_real(c, c->field[CBZ_QUAT], t.data(), nb, 4))) return rc;
  for (int64_t i = 0; i < nb; ++i) {
    double* a = &t[(size_t)i * 3];
    a[0] = 0.0;
    a[1] = -9.78;  // defaultAcceleration, rigidbody.go:18
    a[2] = 0.0;
  }
  if ((rc = upload_real(c, c->field[CBZ_ACC], t.data(), nb, 3))) return rc;
  for (int64_t i = 0; i < nb; ++i) {
    double* m = &t[(size_t)i * 9];
    for (int k = 0; k < 9; ++k) m[k] = (k % 4 == 0) ? 1.0 : 0.0;
  }
  if ((rc = upload_real(c, c->field[CBZ_IIT_WORLD], t.data(), nb, 9))) return rc;
  for (int64_t i = 0; i < nb; ++i) {
    double* m = &t[(size_t)i * 12];
    for (int k = 0; k < 12; ++k) m[k] = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
  }
  if ((rc = upload_real(c, c->field[CBZ_OFFSET], t.data(), nb, 12))) return rc;
  for (int64_t i = 0; i < nb; ++i) t[i] = 0.6;  // SetAwake(true): motion = 2*sleepEpsilon
  if ((rc = upload_real(c, c->field[CBZ_MOTION], t.data(), nb, 1))) return rc;
  // contact materials: the reference's constants (colliders.go:201-202)
  for (int64_t i = 0; i < nb; ++i) t[i] = kFriction;
  if ((rc = upload_real(c, c->field[CBZ_FRICTION], t.data(), nb, 1))) return rc;
  for (int64_t i = 0; i < nb; ++i) t[i] = kRestitution;
  if ((rc = upload_real(c, c->field[CBZ_RESTITUTION], t.data(), nb, 1))) return rc;
  std::vector<int32_t> sh((size_t)nb, CBZ_CUBE);
  if (nb > 0) HIPCHK(c, hipMemcpy(c->field[CBZ_SHAPE], sh.data(), nb * sizeof(int32_t), hipMemcpyHostToDevice));
  std::vector<uint8_t> fl((size_t)nb, CBZ_FLAG_AWAKE | CBZ_FLAG_CAN_SLEEP | F_OFFSET_ID);
  if (nb > 0) HIPCHK(c, hipMemcpy(c->field[CBZ_FLAGS], fl.data(), nb, hipMemcpyHostToDevice));
  c->brad_dirty = true;
  c->pow_dirty = true;
  c->have_contacts = false;
  return CBZ_OK;
}

int cbz_set_field(cbz_ctx* c, int32_t field, const void* host) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (field < 0 || field >= CBZ_NUM_FIELDS || !host) return fail(c, CBZ_ERR_ARG, "set_field: bad field");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const FieldInfo& fi = kFields[field];
  const int64_t nb = c->nb;
  if (field == CBZ_LIN_DAMP || field == CBZ_ANG_DAMP) {
    auto& dst = field == CBZ_LIN_DAMP ? c->lin_damp : c->ang_damp;
    std::memcpy(dst.data(), host, nb * sizeof(double));
    c->pow_dirty = true;
    return CBZ_OK;
  }
  if (field == CBZ_SHAPE) {
    const int32_t* s = (const int32_t*)host;
    for (int64_t i = 0; i < nb; ++i)
      if (s[i] != CBZ_CUBE && s[i] != CBZ_SPHERE) return fail(c, CBZ_ERR_ARG, "set_field: unknown shape");
    std::memcpy(c->h_shape.data(), s, nb * sizeof(int32_t));
    if (nb > 0) HIPCHK(c, hipMemcpy(c->field[field], s, nb * sizeof(int32_t), hipMemcpyHostToDevice));
    c->brad_dirty = true;
    return CBZ_OK;
  }
  if (field == CBZ_FLAGS) {
    std::vector<uint8_t> fl((size_t)nb);
    const uint8_t* s = (const uint8_t*)host;
    for (int64_t i = 0; i < nb; ++i) fl[i] = uint8_t((s[i] & kUserFlags) | (c->h_offset_id[i] ? F_OFFSET_ID : 0));
    if (nb > 0) HIPCHK(c, hipMemcpy(c->field[field], fl.data(), nb, hipMemcpyHostToDevice));
    return CBZ_OK;
  }
  const double* d = (const double*)host;
  if ((rc = upload_real(c, c->field[field], d, nb, fi.width))) return rc;
  if (field == CBZ_HALF) {
    std::memcpy(c->h_half.data(), d, nb * 3 * sizeof(double));
    c->brad_dirty = true;
  }
  if (field == CBZ_OFFSET) {
    static const double id[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
    std::vector<uint8_t> fl((size_t)nb);
    if (nb > 0) HIPCHK(c, hipMemcpy(fl.data(), c->field[CBZ_FLAGS], nb, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nb; ++i) {
      bool same = std::memcmp(d + 12 * i, id, sizeof id) == 0;  // bitwise: +0.0 entries
      c->h_offset_id[i] = same ? 1 : 0;
      fl[i] = uint8_t((fl[i] & ~F_OFFSET_ID) | (same ? F_OFFSET_ID : 0));
    }
    if (nb > 0) HIPCHK(c, hipMemcpy(c->field[CBZ_FLAGS], fl.data(), nb, hipMemcpyHostToDevice));
  }
  return CBZ_OK;
}

int cbz_get_field(cbz_ctx* c, int32_t field, void* host) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (field < 0 || field >= CBZ_NUM_FIELDS || !host) return fail(c, CBZ_ERR_ARG, "get_field: bad field");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const FieldInfo& fi = kFields[field];
  const int64_t nb = c->nb;
  if (field == CBZ_LIN_DAMP || field == CBZ_ANG_DAMP) {
    const auto& src = field == CBZ_LIN_DAMP ? c->lin_damp : c->ang_damp;
    std::memcpy(host, src.data(), nb * sizeof(double));
    return CBZ_OK;
  }
  if (fi.kind == 1) {
    if (nb > 0) HIPCHK(c, hipMemcpy(host, c->field[field], nb * sizeof(int32_t), hipMemcpyDeviceToHost));
    return CBZ_OK;
  }
  if (fi.kind == 2) {
    if (nb > 0) HIPCHK(c, hipMemcpy(host, c->field[field], nb, hipMemcpyDeviceToHost));
    uint8_t* h = (uint8_t*)host;
    for (int64_t i = 0; i < nb; ++i) h[i] &= kUserFlags;
    return CBZ_OK;
  }
  return download_real(c, c->field[field], nb, 0, nb, fi.width, (double*)host);
}

int cbz_calculate_derived(cbz_ctx* c) {
  int rc = need_scene(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  if (c->precision == CBZ_F64)
    HIPCHK(c, launch_calculate_derived<double>(bodies_view<double>(c), c->stream));
  else
    HIPCHK(c, launch_calculate_derived<float>(bodies_view<float>(c), c->stream));
  return CBZ_OK;
}

int cbz_integrate(cbz_ctx* c, double dt) {
  int rc = need_scene(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  return c->precision == CBZ_F64 ? do_integrate<double>(c, dt) : do_integrate<float>(c, dt);
}

int cbz_narrowphase(cbz_ctx* c, int64_t* total) {
  int rc = need_scene(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, launch_fill_i32(c->overflow, 2, 0, c->stream));
  rc = c->precision == CBZ_F64 ? do_narrowphase<double>(c, -1) : do_narrowphase<float>(c, -1);
  if (rc) return rc;
  if ((rc = check_overflow(c))) return rc;
  if (total) HIPCHK(c, hipMemcpy(total, c->total, sizeof(int64_t), hipMemcpyDeviceToHost));
  return CBZ_OK;
}

int cbz_narrowphase_pairs(cbz_ctx* c, int64_t n_pairs, const int32_t* one, const int32_t* two, int64_t* total) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (n_pairs < 0 || (n_pairs > 0 && (!one || !two)) || n_pairs > INT32_MAX / 2)
    return fail(c, CBZ_ERR_ARG, "narrowphase_pairs: bad arguments");
  HIPCHK(c, hipSetDevice(c->device));
  // world of every operand; pairs grouped by non-decreasing world
  auto world_of_body = [&](int32_t b) -> int32_t {
    auto it = std::upper_bound(c->body_off.begin(), c->body_off.end(), b);
    return (int32_t)(it - c->body_off.begin()) - 1;
  };
  auto world_of_plane = [&](int32_t p) -> int32_t {
    auto it = std::upper_bound(c->plane_off.begin(), c->plane_off.end(), p);
    return (int32_t)(it - c->plane_off.begin()) - 1;
  };
  std::vector<int32_t> loff(c->nw + 1, 0), wq(c->nw, 0);
  int32_t prev = 0;
  for (int64_t k = 0; k < n_pairs; ++k) {
    int32_t wa = -1, wb = -1;
    if (one[k] >= 0) {
      if (one[k] >= c->nb) return fail(c, CBZ_ERR_ARG, "narrowphase_pairs: body index out of range");
      wa = world_of_body(one[k]);
    } else {
      if (-one[k] - 1 >= c->np) return fail(c, CBZ_ERR_ARG, "narrowphase_pairs: plane index out of range");
      wa = world_of_plane(-one[k] - 1);
    }
    if (two[k] >= 0) {
      if (two[k] >= c->nb) return fail(c, CBZ_ERR_ARG, "narrowphase_pairs: body index out of range");
      wb = world_of_body(two[k]);
    } else {
      if (-two[k] - 1 >= c->np) return fail(c, CBZ_ERR_ARG, "narrowphase_pairs: plane index out of range");
      wb = world_of_plane(-two[k] - 1);
    }
    if (wa != wb) return fail(c, CBZ_ERR_ARG, "narrowphase_pairs: pair spans two worlds");
    if (wa < prev) return fail(c, CBZ_ERR_ARG, "narrowphase_pairs: pairs not grouped by world");
    prev = wa;
    wq[wa]++;
  }
  for (int32_t w = 0; w < c->nw; ++w) loff[w + 1] = loff[w] + wq[w];
  if ((rc = ensure_list(c, n_pairs))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (n_pairs > 0) {
    HIPCHK(c, hipMemcpy(c->l_one, one, n_pairs * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->l_two, two, n_pairs * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  HIPCHK(c, hipMemcpy(c->l_off, loff.data(), loff.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  c->l_last = -1;
  if ((rc = run_list_narrowphase(c, wq, total))) return rc;
  c->l_last = n_pairs;
  return CBZ_OK;
}

int cbz_pair_contact_counts(cbz_ctx* c, int64_t n_pairs, int32_t* counts) {
  if (!c || n_pairs < 0 || (n_pairs > 0 && !counts)) return fail(c, CBZ_ERR_ARG, "pair_contact_counts: bad arguments");
  if (c->l_last < 0) return fail(c, CBZ_ERR_STATE, "pair_contact_counts: the last narrowphase had no pair list");
  if (n_pairs != c->l_last)
    return fail(c, CBZ_ERR_ARG, "pair_contact_counts: n_pairs differs from the last cbz_narrowphase_pairs");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (n_pairs > 0) HIPCHK(c, hipMemcpy(counts, c->l_cnt, n_pairs * sizeof(int32_t), hipMemcpyDeviceToHost));
  return CBZ_OK;
}

int cbz_narrowphase_broad(cbz_ctx* c, int64_t* total, int64_t* pairs_checked) {
  int rc = need_scene(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  int64_t n = 0;
  rc = c->precision == CBZ_F64 ? do_broadphase<double>(c, &n) : do_broadphase<float>(c, &n);
  if (rc) return rc;
  std::vector<int32_t> loff((size_t)c->nw + 1), wq((size_t)c->nw);
  HIPCHK(c, hipMemcpyAsync(loff.data(), c->l_off, loff.size() * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int32_t w = 0; w < c->nw; ++w) wq[w] = loff[w + 1] - loff[w];
  if (pairs_checked) *pairs_checked = n;
  return run_list_narrowphase(c, wq, total);
}

int cbz_contact_counts(cbz_ctx* c, int64_t* counts) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (!counts) return fail(c, CBZ_ERR_ARG, "contact_counts: null output");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<int32_t> t((size_t)c->nw);
  HIPCHK(c, hipMemcpy(t.data(), c->world_cnt, c->nw * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int32_t w = 0; w < c->nw; ++w) counts[w] = t[w];
  return CBZ_OK;
}

int cbz_get_contacts(cbz_ctx* c, int32_t world, cbz_contact_view* out) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (!out || world < -1 || world >= c->nw) return fail(c, CBZ_ERR_ARG, "get_contacts: bad arguments");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int64_t n = 0, off = 0;
  if (world < 0) {  // every world: the lists are packed in world order from slot 0
    if (!c->have_contacts) return CBZ_OK;
    HIPCHK(c, hipMemcpy(&n, c->total, sizeof n, hipMemcpyDeviceToHost));
    n = std::min(n, c->cap);
  } else {
    int32_t nw = 0;
    HIPCHK(c, hipMemcpy(&nw, c->world_cnt + world, sizeof nw, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(&off, c->world_coff + world, sizeof off, hipMemcpyDeviceToHost));
    n = nw;
  }
  if (n == 0) return CBZ_OK;
  if (off + n > c->cap) return fail(c, CBZ_ERR_CAPACITY, "get_contacts: world's contacts exceed the capacity");
  const size_t rs = real_size(c);
  const char* base = (const char*)c->c_real;
  const int64_t S = c->cap;
  if (out->body0) HIPCHK(c, hipMemcpy(out->body0, (int32_t*)c->c_body0 + off, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (out->body1) HIPCHK(c, hipMemcpy(out->body1, (int32_t*)c->c_body1 + off, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  struct Part {
    double* dst;
    int plane0;
    int k;
  } parts[] = {{out->point, 0, 3}, {out->normal, 3, 3}, {out->pen, 6, 1}, {out->c2w, 7, 9},
               {out->rel, 16, 6},  {out->cvel, 22, 3},  {out->ddv, 25, 1}};
  for (const Part& p : parts) {
    if (!p.dst) continue;
    if ((rc = download_real(c, base + (size_t)p.plane0 * S * rs, S, off, n, p.k, p.dst))) return rc;
  }
  // Contact.Friction / Contact.Restitution: per contact with materials on,
  // else the constants every contact carries
  const bool mat = (c->mat_flags & CBZ_MATERIALS) != 0;
  const Part mats[] = {{out->friction, 26, 1}, {out->restitution, 27, 1}};
  const double consts[] = {kFriction, kRestitution};
  for (int m = 0; m < 2; ++m) {
    if (!mats[m].dst) continue;
    if (mat) {
      if ((rc = download_real(c, base + (size_t)mats[m].plane0 * S * rs, S, off, n, 1, mats[m].dst))) return rc;
    } else {
      for (int64_t k = 0; k < n; ++k) mats[m].dst[k] = consts[m];
    }
  }
  return CBZ_OK;
}

int cbz_set_contacts(cbz_ctx* c, const int64_t* counts, const int32_t* body0, const int32_t* body1,
                     const double* point, const double* normal, const double* pen) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (!counts) return fail(c, CBZ_ERR_ARG, "set_contacts: null counts");
  int64_t tot = 0;
  std::vector<int32_t> cnt((size_t)c->nw);
  std::vector<int64_t> off((size_t)c->nw);
  for (int32_t w = 0; w < c->nw; ++w) {
    if (counts[w] < 0 || counts[w] > INT32_MAX) return fail(c, CBZ_ERR_ARG, "set_contacts: bad count");
    off[w] = tot;
    cnt[w] = (int32_t)counts[w];
    tot += counts[w];
  }
  if (tot > c->cap) return fail(c, CBZ_ERR_CAPACITY, "set_contacts: more contacts than slots");
  if (tot > 0 && (!body0 || !body1 || !point || !normal || !pen)) return fail(c, CBZ_ERR_ARG, "set_contacts: null record");
  for (int64_t k = 0; k < tot; ++k) {
    if (body0[k] >= c->nb || body1[k] >= c->nb || (body0[k] < 0 && body1[k] < 0))
      return fail(c, CBZ_ERR_ARG, "set_contacts: bad body index");
  }
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t rs = real_size(c);
  char* base = (char*)c->c_real;
  const int64_t S = c->cap;
  if (tot > 0) {
    HIPCHK(c, hipMemcpy(c->c_body0, body0, tot * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->c_body1, body1, tot * sizeof(int32_t), hipMemcpyHostToDevice));
    // upload_real writes [k][n] with stride n; write each component plane at stride S
    for (int k = 0; k < 3; ++k) {
      std::vector<double> t((size_t)tot);
      for (int64_t i = 0; i < tot; ++i) t[i] = point[3 * i + k];
      if ((rc = upload_real(c, base + ((size_t)0 + k) * S * rs, t.data(), tot, 1))) return rc;
      for (int64_t i = 0; i < tot; ++i) t[i] = normal[3 * i + k];
      if ((rc = upload_real(c, base + ((size_t)3 + k) * S * rs, t.data(), tot, 1))) return rc;
    }
    if ((rc = upload_real(c, base + (size_t)6 * S * rs, pen, tot, 1))) return rc;
    if (c->mat_flags & CBZ_MATERIALS) {  // the constants, until cbz_set_contact_materials
      const std::vector<double> f((size_t)tot, kFriction), e((size_t)tot, kRestitution);
      if ((rc = upload_real(c, base + (size_t)26 * S * rs, f.data(), tot, 1))) return rc;
      if ((rc = upload_real(c, base + (size_t)27 * S * rs, e.data(), tot, 1))) return rc;
    }
  }
  HIPCHK(c, hipMemcpy(c->world_cnt, cnt.data(), cnt.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->world_coff, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->total, &tot, sizeof tot, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemset(c->overflow, 0, 2 * sizeof(int32_t)));
  c->have_contacts = true;
  return CBZ_OK;
}

int cbz_set_materials(cbz_ctx* c, int32_t flags) {
  if (!c) return CBZ_ERR_ARG;
  if (flags & ~(CBZ_MATERIALS | CBZ_MATERIALS_FRICTIONLESS_TWO_BODY))
    return fail(c, CBZ_ERR_ARG, "set_materials: unknown flag");
  if ((flags & CBZ_MATERIALS_FRICTIONLESS_TWO_BODY) && !(flags & CBZ_MATERIALS))
    return fail(c, CBZ_ERR_ARG, "set_materials: CBZ_MATERIALS_FRICTIONLESS_TWO_BODY needs CBZ_MATERIALS");
  if (flags == c->mat_flags) return CBZ_OK;
  if (c->stream) {
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  c->mat_flags = flags;
  drop_graph(c);  // the captured kernels carry the contact view
  c->have_contacts = false;
  return CBZ_OK;
}

int cbz_set_plane_materials(cbz_ctx* c, const double* friction, const double* restitution) {
  int rc = need_scene(c);
  if (rc) return rc;
  for (int32_t k = 0; friction && k < c->np; ++k)
    if (!(friction[k] >= 0.0) || !std::isfinite(friction[k]))
      return fail(c, CBZ_ERR_ARG, "set_plane_materials: friction must be finite and >= 0");
  for (int32_t k = 0; restitution && k < c->np; ++k)
    if (!std::isfinite(restitution[k])) return fail(c, CBZ_ERR_ARG, "set_plane_materials: restitution not finite");
  if (c->np == 0) return CBZ_OK;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (friction && (rc = upload_real(c, c->plane_fric, friction, c->np, 1))) return rc;
  if (restitution && (rc = upload_real(c, c->plane_rest, restitution, c->np, 1))) return rc;
  return CBZ_OK;
}

int cbz_set_contact_materials(cbz_ctx* c, const double* friction, const double* restitution) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (!(c->mat_flags & CBZ_MATERIALS)) return fail(c, CBZ_ERR_STATE, "set_contact_materials: materials are off");
  if (!c->have_contacts) return fail(c, CBZ_ERR_STATE, "set_contact_materials: no contacts");
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = check_overflow(c))) return rc;
  int64_t tot = 0;
  HIPCHK(c, hipMemcpy(&tot, c->total, sizeof tot, hipMemcpyDeviceToHost));
  if (tot <= 0) return CBZ_OK;
  const size_t rs = real_size(c);
  char* base = (char*)c->c_real;
  const int64_t S = c->cap;
  if (friction && (rc = upload_real(c, base + (size_t)26 * S * rs, friction, tot, 1))) return rc;
  if (restitution && (rc = upload_real(c, base + (size_t)27 * S * rs, restitution, tot, 1))) return rc;
  return CBZ_OK;
}

int cbz_resolve(cbz_ctx* c, double dt, int64_t iter_per_contact, int64_t iter_fixed) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (!c->have_contacts) return fail(c, CBZ_ERR_STATE, "resolve: no contacts (run a narrowphase first)");
  if (iter_per_contact < 0 || iter_fixed < 0) return fail(c, CBZ_ERR_ARG, "resolve: negative iteration count");
  HIPCHK(c, hipSetDevice(c->device));
  rc = c->precision == CBZ_F64 ? do_resolve<double>(c, dt, iter_per_contact, iter_fixed)
                               : do_resolve<float>(c, dt, iter_per_contact, iter_fixed);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CBZ_OK;
}

int cbz_resolver_occupancy(cbz_ctx* c, int32_t* worlds_per_cu, int32_t* n_cu) {
  if (!c) return CBZ_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  int w = 0, cu = 0;
  HIPCHK(c, resolver_occupancy(c->precision == CBZ_F64, &w));
  HIPCHK(c, hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, c->device));
  if (worlds_per_cu) *worlds_per_cu = w;
  if (n_cu) *n_cu = cu;
  return CBZ_OK;
}

int cbz_set_resolver_split(cbz_ctx* c, int32_t min_contacts, int32_t max_worlds) {
  if (!c) return CBZ_ERR_ARG;
  if (min_contacts < 0 || max_worlds < 0)
    return fail(c, CBZ_ERR_ARG, "set_resolver_split: min_contacts and max_worlds must be >= 0");
  if (min_contacts != c->mw_min || max_worlds != c->mw_max) {
    drop_graph(c);  // captured launches carry the split
    c->mw_min = min_contacts;
    c->mw_max = max_worlds;
  }
  return CBZ_OK;
}

int cbz_get_resolver_split(cbz_ctx* c, int32_t* min_contacts, int32_t* max_worlds) {
  if (!c) return CBZ_ERR_ARG;
  if (min_contacts) *min_contacts = c->mw_min;
  if (max_worlds) *max_worlds = c->mw_max;
  return CBZ_OK;
}

int cbz_autotune_resolver(cbz_ctx* c, double dt, int64_t iter_per_contact, int32_t n_steps, int32_t surrogate,
                          double tol, int64_t polish_per_contact, int32_t n_cand, const int32_t* cand_min,
                          const int32_t* cand_max, double* ms_per_step, int32_t* best) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (n_steps <= 0 || n_cand <= 0 || !cand_min || !cand_max)
    return fail(c, CBZ_ERR_ARG, "autotune_resolver: bad arguments");
  if (c->prof) return fail(c, CBZ_ERR_STATE, "autotune_resolver: profiling is on");
  for (int32_t i = 0; i < n_cand; ++i)
    if (cand_min[i] < 0 || cand_max[i] < 0) return fail(c, CBZ_ERR_ARG, "autotune_resolver: negative split");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // the body state on the device, restored before every timed run
  std::vector<size_t> off(CBZ_NUM_FIELDS + 1, 0), len(CBZ_NUM_FIELDS, 0);
  for (int f = 0; f < CBZ_NUM_FIELDS; ++f) {
    const FieldInfo& fi = kFields[f];
    const size_t el = fi.kind == 0 ? real_size(c) : fi.kind == 1 ? sizeof(int32_t) : sizeof(uint8_t);
    len[f] = fi.kind == 3 ? 0 : el * fi.width * (size_t)c->nb;
    off[f + 1] = off[f] + ((len[f] + 7) & ~size_t(7));
  }
  void* snap = nullptr;
  if ((rc = dalloc_bytes(c, &snap, std::max<size_t>(off[CBZ_NUM_FIELDS], 8)))) return rc;
  auto copy = [&](bool save) -> hipError_t {
    for (int f = 0; f < CBZ_NUM_FIELDS; ++f) {
      const size_t n = len[f];
      if (!n || !c->field[f]) continue;
      char* d = (char*)snap + off[f];
      const hipError_t e = save ? hipMemcpyAsync(d, c->field[f], n, hipMemcpyDeviceToDevice, c->stream)
                                : hipMemcpyAsync(c->field[f], d, n, hipMemcpyDeviceToDevice, c->stream);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  };
  const int32_t old_min = c->mw_min, old_max = c->mw_max;
  std::vector<int64_t> fb((size_t)n_steps + 1);
  auto run = [&](int32_t steps) {
    return surrogate ? cbz_step_sur(c, dt, steps, iter_per_contact, tol, polish_per_contact, 0, nullptr, fb.data())
                     : cbz_step(c, dt, steps, iter_per_contact, nullptr);
  };
  int32_t pick = -1;
  double pick_ms = 0.0;
  hipError_t e = copy(true);
  for (int32_t i = 0; e == hipSuccess && !rc && i < n_cand; ++i) {
    if ((rc = cbz_set_resolver_split(c, cand_min[i], cand_max[i]))) break;
    // one untimed step captures the step for this split
    if ((e = copy(false)) != hipSuccess || (rc = run(1))) break;
    if ((e = copy(false)) != hipSuccess || (e = hipEventRecord(c->ev[0], c->stream)) != hipSuccess) break;
    if ((rc = run(n_steps))) break;
    if ((e = hipEventRecord(c->ev[1], c->stream)) != hipSuccess || (e = hipEventSynchronize(c->ev[1])) != hipSuccess)
      break;
    float ms = 0.f;
    if ((e = hipEventElapsedTime(&ms, c->ev[0], c->ev[1])) != hipSuccess) break;
    const double per = double(ms) / n_steps;
    if (ms_per_step) ms_per_step[i] = per;
    if (pick < 0 || per < pick_ms) {
      pick = i;
      pick_ms = per;
    }
  }
  if (e == hipSuccess) e = copy(false);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dfree(snap);
  if (!rc && e != hipSuccess) rc = fail(c, CBZ_ERR_HIP, std::string("autotune_resolver: ") + hipGetErrorString(e));
  if (rc) {
    cbz_set_resolver_split(c, old_min, old_max);
    return rc;
  }
  if (best) *best = pick;
  return cbz_set_resolver_split(c, cand_min[pick], cand_max[pick]);
}

int cbz_resolver_world_times(cbz_ctx* c, double* us) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (!us) return fail(c, CBZ_ERR_ARG, "resolver_world_times: null");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<int64_t> t(4 * (size_t)c->nw);
  if (c->nw > 0) HIPCHK(c, hipMemcpy(t.data(), c->world_time, t.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
  int khz = 0;
  HIPCHK(c, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
  if (khz <= 0) return fail(c, CBZ_ERR_HIP, "resolver_world_times: no wall-clock rate");
  int64_t t0 = INT64_MAX;
  for (int64_t v : t)
    if (v >= 0) t0 = std::min(t0, v);
  for (size_t k = 0; k < t.size(); ++k) us[k] = t[k] >= 0 ? double(t[k] - t0) * 1000.0 / khz : -1.0;
  return CBZ_OK;
}

int cbz_resolver_iterations(cbz_ctx* c, int64_t* iters) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (!iters) return fail(c, CBZ_ERR_ARG, "resolver_iterations: null output");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(iters, c->world_iters, 2 * c->nw * sizeof(int64_t), hipMemcpyDeviceToHost));
  return CBZ_OK;
}

int cbz_step(cbz_ctx* c, double dt, int32_t n_steps, int64_t iter_per_contact, int64_t* counts) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (n_steps < 0 || iter_per_contact < 0) return fail(c, CBZ_ERR_ARG, "step: bad arguments");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, launch_fill_i32(c->overflow, 2, 0, c->stream));
  if (!counts && !c->prof && n_steps > 0) {
    // Replay a captured graph of one step: the host-side inputs (the Pow
    // values, the bounding radii) are refreshed first, so the captured
    // launches do no host work and read them from device memory.
    rc = refresh_pow(c, dt);
    if (rc == CBZ_OK) rc = refresh_brad(c);
    if (rc) return rc;
    if (!c->gexec || c->g_dt != dt || c->g_mult != iter_per_contact) {
      drop_graph(c);
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int crc = c->precision == CBZ_F64 ? do_integrate<double>(c, dt) : do_integrate<float>(c, dt);
      if (crc == CBZ_OK)
        crc = c->precision == CBZ_F64 ? do_narrowphase<double>(c, -1) : do_narrowphase<float>(c, -1);
      if (crc == CBZ_OK)
        crc = c->precision == CBZ_F64 ? do_resolve<double>(c, dt, iter_per_contact, 0)
                                      : do_resolve<float>(c, dt, iter_per_contact, 0);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (crc) {
        if (g) (void)hipGraphDestroy(g);
        return crc;
      }
      if (e != hipSuccess) return fail(c, CBZ_ERR_HIP, std::string("step: capture: ") + hipGetErrorString(e));
      e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        c->gexec = nullptr;
        return fail(c, CBZ_ERR_HIP, std::string("step: instantiate: ") + hipGetErrorString(e));
      }
      c->g_dt = dt;
      c->g_mult = iter_per_contact;
    }
    for (int32_t s = 0; s < n_steps; ++s) HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    c->have_contacts = true;
    return check_overflow(c);
  }
  int32_t* hist = nullptr;
  if (counts && n_steps > 0) HIPCHK(c, hipMalloc((void**)&hist, (size_t)n_steps * c->nw * sizeof(int32_t)));
  // Profiling: per step s, events E[8s + k]: k = 0 step start, 7 after
  // integrate, 1..4 narrowphase (before count, after count, after scans,
  // after emit), 6 before resolve, 5 after resolve.  Nothing waits between
  // steps (the counters accumulate on the device), so the stream stays ahead
  // of the GPU and each event pair brackets its kernels alone.
  hipEvent_t* E = nullptr;
  if (c->prof && n_steps > 0) {
    const size_t need = 8 * (size_t)n_steps;
    while (c->pev.size() < need) {
      hipEvent_t e = nullptr;
      HIPCHK(c, hipEventCreate(&e));
      c->pev.push_back(e);
    }
    E = c->pev.data();
    HIPCHK(c, launch_fill_i32((int32_t*)c->d_prof, 8, 0, c->stream));
    HIPCHK(c, launch_fill_i32((int32_t*)c->d_awake, kAwakeSlots, 0, c->stream));
  }
  c->count_awake = c->prof;
  for (int32_t s = 0; s < n_steps && rc == CBZ_OK; ++s) {
    hipEvent_t* e = E ? E + 8 * (size_t)s : nullptr;
    if (e) HIPCHK(c, hipEventRecord(e[0], c->stream));
    rc = c->precision == CBZ_F64 ? do_integrate<double>(c, dt) : do_integrate<float>(c, dt);
    if (rc) break;
    if (e) HIPCHK(c, hipEventRecord(e[7], c->stream));
    rc = c->precision == CBZ_F64 ? do_narrowphase<double>(c, -1, e ? e + 1 : nullptr)
                                 : do_narrowphase<float>(c, -1, e ? e + 1 : nullptr);
    if (rc) break;
    if (hist)
      HIPCHK(c, hipMemcpyAsync(hist + (size_t)s * c->nw, c->world_cnt, c->nw * sizeof(int32_t),
                               hipMemcpyDeviceToDevice, c->stream));
    if (e) HIPCHK(c, hipEventRecord(e[6], c->stream));
    rc = c->precision == CBZ_F64 ? do_resolve<double>(c, dt, iter_per_contact, 0)
                                 : do_resolve<float>(c, dt, iter_per_contact, 0);
    if (rc) break;
    if (e) {
      HIPCHK(c, hipEventRecord(e[5], c->stream));
      HIPCHK(c, launch_prof_accum(c->d_awake, c->total, c->world_iters, 2 * c->nw, c->d_prof, c->stream));
    }
  }
  if (E && rc == CBZ_OK) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int32_t s = 0; s < n_steps; ++s) {
      hipEvent_t* e = E + 8 * (size_t)s;
      float t[5] = {0, 0, 0, 0, 0}, whole = 0;
      HIPCHK(c, hipEventElapsedTime(&t[0], e[0], e[7]));  // integrate
      HIPCHK(c, hipEventElapsedTime(&t[1], e[1], e[2]));  // count (cull + eval)
      HIPCHK(c, hipEventElapsedTime(&t[2], e[2], e[3]));  // scans
      HIPCHK(c, hipEventElapsedTime(&t[3], e[3], e[4]));  // emit
      HIPCHK(c, hipEventElapsedTime(&t[4], e[6], e[5]));  // resolve
      HIPCHK(c, hipEventElapsedTime(&whole, e[0], e[5]));
      for (int k = 0; k < 5; ++k) c->acc_ms[k] += t[k];
      c->acc_ms[5] += whole;
    }
    int64_t acc[4] = {0, 0, 0, 0};
    HIPCHK(c, hipMemcpy(acc, c->d_prof, sizeof acc, hipMemcpyDeviceToHost));
    for (int k = 0; k < 4; ++k) c->acc_cnt[k] += acc[k];
  }
  c->count_awake = false;
  if (rc == CBZ_OK) rc = check_overflow(c);
  if (rc == CBZ_OK && hist) {
    std::vector<int32_t> t((size_t)n_steps * c->nw);
    hipError_t e = hipMemcpy(t.data(), hist, t.size() * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(c, CBZ_ERR_HIP, std::string("step: count download: ") + hipGetErrorString(e));
    for (size_t k = 0; k < t.size(); ++k) counts[k] = t[k];
  }
  if (hist) (void)hipFree(hist);
  return rc;
}

int cbz_synchronize(cbz_ctx* c) {
  if (!c) return CBZ_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CBZ_OK;
}

int cbz_set_profiling(cbz_ctx* c, int32_t on) {
  if (!c) return CBZ_ERR_ARG;
  c->prof = on != 0;
  return CBZ_OK;
}

int cbz_phase_times(cbz_ctx* c, double* ms6, int64_t* counts4) {
  if (!c) return CBZ_ERR_ARG;
  if (ms6)
    for (int k = 0; k < 6; ++k) ms6[k] = c->acc_ms[k];
  if (counts4)
    for (int k = 0; k < 4; ++k) counts4[k] = c->acc_cnt[k];
  for (double& v : c->acc_ms) v = 0;
  for (int64_t& v : c->acc_cnt) v = 0;
  return CBZ_OK;
}

int cbz_integrate_timed(cbz_ctx* c, double dt, int32_t n, double* ms_per_launch, int64_t* integrated) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (n <= 0) return fail(c, CBZ_ERR_ARG, "integrate_timed: n must be positive");
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = refresh_pow(c, dt))) return rc;
  HIPCHK(c, launch_fill_i32((int32_t*)c->d_awake, kAwakeSlots, 0, c->stream));
  c->count_awake = true;
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  for (int32_t k = 0; k < n && rc == CBZ_OK; ++k)
    rc = c->precision == CBZ_F64 ? do_integrate<double>(c, dt) : do_integrate<float>(c, dt);
  c->count_awake = false;
  if (rc) return rc;
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev[1]));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  uint32_t aw[kAwakeSlots];
  HIPCHK(c, hipMemcpy(aw, c->d_awake, sizeof aw, hipMemcpyDeviceToHost));
  int64_t tot = 0;
  for (uint32_t v : aw) tot += v;
  if (ms_per_launch) *ms_per_launch = ms / n;
  if (integrated) *integrated = tot;
  return CBZ_OK;
}

int cbz_sur_phase_times(cbz_ctx* c, double* ms6, int64_t* counts6) {
  if (!c) return CBZ_ERR_ARG;
  if (ms6)
    for (int k = 0; k < 6; ++k) ms6[k] = c->sur_ms[k];
  if (counts6)
    for (int k = 0; k < 6; ++k) counts6[k] = c->sur_cnt[k];
  for (double& v : c->sur_ms) v = 0;
  for (int64_t& v : c->sur_cnt) v = 0;
  return CBZ_OK;
}

int cbz_selftest_f64(cbz_ctx* c, int32_t op, int64_t n, const double* a, const double* b, double* out) {
  if (!c || n < 0 || (n > 0 && (!a || !b || !out))) return fail(c, CBZ_ERR_ARG, "selftest: bad arguments");
  if (n == 0) return CBZ_OK;
  HIPCHK(c, hipSetDevice(c->device));
  double *da = nullptr, *db = nullptr, *dout = nullptr;
  HIPCHK(c, hipMalloc((void**)&da, n * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&db, n * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&dout, n * sizeof(double)));
  hipError_t e = hipMemcpy(da, a, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(db, b, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_selftest(op, n, da, db, dout, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(da);
  (void)hipFree(db);
  (void)hipFree(dout);
  if (e != hipSuccess) return fail(c, CBZ_ERR_HIP, std::string("selftest: ") + hipGetErrorString(e));
  return CBZ_OK;
}

int cbz_selftest_sort(cbz_ctx* c, int64_t n, const uint64_t* keys, const int32_t* vals, uint64_t* keys_out,
                      int32_t* vals_out) {
  if (!c || n < 0 || n >= (int64_t(1) << 31) || (n > 0 && (!keys || !keys_out)) || (!vals != !vals_out))
    return fail(c, CBZ_ERR_ARG, "selftest_sort: bad arguments");
  if (n == 0) return CBZ_OK;
  HIPCHK(c, hipSetDevice(c->device));
  const bool v = vals != nullptr;
  const size_t tb = radix_sort_bytes(n, v);
  void *dk = nullptr, *dko = nullptr, *dv = nullptr, *dvo = nullptr, *tmp = nullptr;
  hipError_t e = hipMalloc(&dk, n * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMalloc(&dko, n * sizeof(uint64_t));
  if (e == hipSuccess && v) e = hipMalloc(&dv, n * sizeof(int32_t));
  if (e == hipSuccess && v) e = hipMalloc(&dvo, n * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(&tmp, tb);
  if (e == hipSuccess) e = hipMemcpy(dk, keys, n * sizeof(uint64_t), hipMemcpyHostToDevice);
  if (e == hipSuccess && v) e = hipMemcpy(dv, vals, n * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = radix_sort_u64((const uint64_t*)dk, (uint64_t*)dko, (const int32_t*)dv, (int32_t*)dvo, n, tmp, tb, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(keys_out, dko, n * sizeof(uint64_t), hipMemcpyDeviceToHost);
  if (e == hipSuccess && v) e = hipMemcpy(vals_out, dvo, n * sizeof(int32_t), hipMemcpyDeviceToHost);
  for (void* q : {dk, dko, dv, dvo, tmp})
    if (q) (void)hipFree(q);
  if (e != hipSuccess) return fail(c, CBZ_ERR_HIP, std::string("selftest_sort: ") + hipGetErrorString(e));
  return CBZ_OK;
}

// --- resolver surrogate -------------------------------------------------
int64_t cbz_sur_num_params(void) { return layer_offsets().total; }

int cbz_sur_set_params(cbz_ctx* c, const float* p) {
  if (!c || !p) return fail(c, CBZ_ERR_ARG, "sur_set_params: null");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = sur_alloc_model(c);
  if (rc) return rc;
  const LayerOff o = layer_offsets();
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(c->sur.params, p, o.total * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemset(c->sur.adam_m, 0, o.total * sizeof(float)));
  HIPCHK(c, hipMemset(c->sur.adam_v, 0, o.total * sizeof(float)));
  c->sur.adam_t = 0;
  if ((rc = sur_pack(c))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->sur.have_model = true;
  return CBZ_OK;
}

int cbz_sur_get_params(cbz_ctx* c, float* p) {
  if (!c || !p) return fail(c, CBZ_ERR_ARG, "sur_get_params: null");
  if (!c->sur.params) return fail(c, CBZ_ERR_STATE, "sur_get_params: no model");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(p, c->sur.params, layer_offsets().total * sizeof(float), hipMemcpyDeviceToHost));
  return CBZ_OK;
}

int cbz_sur_set_norm(cbz_ctx* c, const float* n) {
  if (!c || !n) return fail(c, CBZ_ERR_ARG, "sur_set_norm: null");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = sur_alloc_model(c);
  if (rc) return rc;
  std::vector<float> h(kNorm);
  for (int k = 0; k < kFeat; ++k) {
    h[k] = n[k];
    h[kFeat + k] = 1.0f / std::max(n[kFeat + k], 1e-6f);
  }
  for (int k = 0; k < kOut; ++k) {
    h[2 * kFeat + k] = n[2 * kFeat + k];
    h[2 * kFeat + kOut + k] = std::max(n[2 * kFeat + kOut + k], 1e-6f);
    h[2 * kFeat + 2 * kOut + k] = 1.0f / h[2 * kFeat + kOut + k];
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(c->sur.norm, h.data(), kNorm * sizeof(float), hipMemcpyHostToDevice));
  return CBZ_OK;
}

int cbz_sur_get_norm(cbz_ctx* c, float* n) {
  if (!c || !n) return fail(c, CBZ_ERR_ARG, "sur_get_norm: null");
  if (!c->sur.norm) return fail(c, CBZ_ERR_STATE, "sur_get_norm: no model");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<float> h(kNorm);
  HIPCHK(c, hipMemcpy(h.data(), c->sur.norm, kNorm * sizeof(float), hipMemcpyDeviceToHost));
  for (int k = 0; k < kFeat; ++k) {
    n[k] = h[k];
    n[kFeat + k] = 1.0f / h[kFeat + k];
  }
  for (int k = 0; k < kOut; ++k) {
    n[2 * kFeat + k] = h[2 * kFeat + k];
    n[2 * kFeat + kOut + k] = h[2 * kFeat + kOut + k];
  }
  return CBZ_OK;
}

int cbz_sur_reserve(cbz_ctx* c, int64_t rows) {
  if (!c || rows < 0) return fail(c, CBZ_ERR_ARG, "sur_reserve: bad arguments");
  HIPCHK(c, hipSetDevice(c->device));
  auto& u = c->sur;
  if (rows <= u.ds_cap) return CBZ_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float *x = nullptr, *y = nullptr;
  HIPCHK(c, hipMalloc((void**)&x, (size_t)rows * kFeat * sizeof(float)));
  HIPCHK(c, hipMalloc((void**)&y, (size_t)rows * kOut * sizeof(float)));
  if (u.ds_rows > 0) {
    HIPCHK(c, hipMemcpy(x, u.dsX, u.ds_rows * kFeat * sizeof(float), hipMemcpyDeviceToDevice));
    HIPCHK(c, hipMemcpy(y, u.dsY, u.ds_rows * kOut * sizeof(float), hipMemcpyDeviceToDevice));
  }
  dfree(u.dsX);
  dfree(u.dsY);
  u.dsX = x;
  u.dsY = y;
  u.ds_cap = rows;
  return CBZ_OK;
}

int cbz_sur_collect(cbz_ctx* c, double dt, int64_t iter_per_contact, int64_t* added) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (c->mat_flags & CBZ_MATERIALS) return fail(c, CBZ_ERR_STATE, "sur_collect: contact materials are on");
  if (!c->have_contacts) return fail(c, CBZ_ERR_STATE, "sur_collect: no contacts (run a narrowphase first)");
  HIPCHK(c, hipSetDevice(c->device));
  auto& u = c->sur;
  int64_t tot = 0;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(&tot, c->total, sizeof tot, hipMemcpyDeviceToHost));
  tot = std::min(tot, c->cap);
  if (u.ds_rows + tot > u.ds_cap)
    return fail(c, CBZ_ERR_CAPACITY, "sur_collect: dataset full (cbz_sur_reserve more rows)");
  if ((rc = sur_ensure_inference(c, tot))) return rc;
  float* X = u.dsX + u.ds_rows * kFeat;
  float* Y = u.dsY + u.ds_rows * kOut;
  const bool f64 = c->precision == CBZ_F64;
  // the features of the prepared contacts (what the surrogate sees), then
  // the exact position and velocity phases with their labels
  rc = f64 ? do_resolve<double>(c, dt, 0, 0, RM_POS) : do_resolve<float>(c, dt, 0, 0, RM_POS);
  if (rc) return rc;
  const int32_t max_nb = max_world_bodies(c);
  if (max_nb <= kFeatWorldBodiesMax) {
    if (f64)
      HIPCHK(c, launch_features_world<double>(bodies_view<double>(c), contacts_view<double>(c), c->world_cnt,
                                              c->world_coff, c->d_body_off, c->overflow, c->nw, max_nb, nullptr,
                                              nullptr, nullptr, X, c->stream));
    else
      HIPCHK(c, launch_features_world<float>(bodies_view<float>(c), contacts_view<float>(c), c->world_cnt,
                                             c->world_coff, c->d_body_off, c->overflow, c->nw, max_nb, nullptr,
                                             nullptr, nullptr, X, c->stream));
  } else if (f64) {
    HIPCHK(c, launch_features<double>(bodies_view<double>(c), contacts_view<double>(c), tot, nullptr, u.deg, nullptr, nullptr,
                                      nullptr, X, c->stream));
  } else {
    HIPCHK(c, launch_features<float>(bodies_view<float>(c), contacts_view<float>(c), tot, nullptr, u.deg, nullptr, nullptr,
                                     nullptr, X, c->stream));
  }
  HIPCHK(c, launch_fill_f32(Y, tot * kOut, 0.f, c->stream));
  rc = f64 ? do_resolve<double>(c, dt, iter_per_contact, 0, RM_POS | RM_PREPARED | RM_LABEL, nullptr, Y)
           : do_resolve<float>(c, dt, iter_per_contact, 0, RM_POS | RM_PREPARED | RM_LABEL, nullptr, Y);
  if (rc) return rc;
  rc = f64 ? do_resolve<double>(c, dt, iter_per_contact, 0, RM_VEL | RM_LABEL, nullptr, Y)
           : do_resolve<float>(c, dt, iter_per_contact, 0, RM_VEL | RM_LABEL, nullptr, Y);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  u.ds_rows += tot;
  if (added) *added = tot;
  return CBZ_OK;
}

int cbz_sur_dataset(cbz_ctx* c, int64_t* rows, float* X, float* Y) {
  if (!c) return CBZ_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  auto& u = c->sur;
  if (rows) *rows = u.ds_rows;
  if (X && u.ds_rows) HIPCHK(c, hipMemcpy(X, u.dsX, u.ds_rows * kFeat * sizeof(float), hipMemcpyDeviceToHost));
  if (Y && u.ds_rows) HIPCHK(c, hipMemcpy(Y, u.dsY, u.ds_rows * kOut * sizeof(float), hipMemcpyDeviceToHost));
  return CBZ_OK;
}

int cbz_sur_load_dataset(cbz_ctx* c, int64_t rows, const float* X, const float* Y) {
  if (!c || rows < 0 || (rows > 0 && (!X || !Y))) return fail(c, CBZ_ERR_ARG, "sur_load_dataset: bad arguments");
  int rc = cbz_sur_reserve(c, rows);
  if (rc) return rc;
  auto& u = c->sur;
  if (rows) {
    HIPCHK(c, hipMemcpy(u.dsX, X, rows * kFeat * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(u.dsY, Y, rows * kOut * sizeof(float), hipMemcpyHostToDevice));
  }
  u.ds_rows = rows;
  return CBZ_OK;
}

int cbz_sur_init(cbz_ctx* c, uint64_t seed) {
  if (!c) return CBZ_ERR_ARG;
  auto& u = c->sur;
  if (u.ds_rows <= 0) return fail(c, CBZ_ERR_STATE, "sur_init: empty dataset (collect first)");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = sur_alloc_model(c);
  if (rc) return rc;
  float* st = nullptr;
  HIPCHK(c, hipMalloc((void**)&st, 2 * (kFeat + kOut) * sizeof(float)));
  HIPCHK(c, launch_col_stats(u.dsX, u.ds_rows, kFeat, st, st + kFeat, c->stream));
  HIPCHK(c, launch_col_stats(u.dsY, u.ds_rows, kOut, st + 2 * kFeat, st + 2 * kFeat + kOut, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<float> h(2 * (kFeat + kOut));
  HIPCHK(c, hipMemcpy(h.data(), st, h.size() * sizeof(float), hipMemcpyDeviceToHost));
  (void)hipFree(st);
  if ((rc = cbz_sur_set_norm(c, h.data()))) return rc;
  // He-uniform weights (a = sqrt(6 / fan_in)), the output layer scaled by
  // 0.1; zero biases.  SplitMix64 stream from the seed.
  const LayerOff o = layer_offsets();
  std::vector<float> p((size_t)o.total, 0.f);
  uint64_t x = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  auto next = [&]() {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  for (int l = 0; l < kLayers; ++l) {
    const double a = std::sqrt(6.0 / layer_in(l)) * (l == kLayers - 1 ? 0.1 : 1.0);
    for (int64_t i = 0; i < int64_t(layer_out(l)) * layer_in(l); ++i) {
      const double u01 = double(next() >> 11) * (1.0 / 9007199254740992.0);
      p[o.w[l] + i] = float((2.0 * u01 - 1.0) * a);
    }
  }
  return cbz_sur_set_params(c, p.data());
}

int cbz_sur_train(cbz_ctx* c, int32_t steps, int32_t batch, float lr, uint64_t seed, float* losses) {
  if (!c || steps < 0 || batch <= 0) return fail(c, CBZ_ERR_ARG, "sur_train: bad arguments");
  auto& u = c->sur;
  if (!u.have_model || u.ds_rows <= 0) return fail(c, CBZ_ERR_STATE, "sur_train: needs a model and a dataset");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = sur_ensure_batch(c, batch);
  if (rc) return rc;
  if (steps > u.losses_cap) {
    if ((rc = dalloc(c, &u.losses, (size_t)steps))) return rc;
    u.losses_cap = steps;
  }
  const int64_t Mpad = round32(batch);
  const LayerOff o = layer_offsets();
  const float* nm = u.norm;
  for (int32_t s = 0; s < steps; ++s) {
    HIPCHK(c, launch_sample_idx(u.idx, batch, seed, (uint64_t)s, u.ds_rows, c->stream));
    HIPCHK(c, launch_gather_batch(u.dsX, u.dsY, u.idx, (int)Mpad, batch, nm, nm + kFeat, nm + 2 * kFeat,
                                  nm + 2 * kFeat + 2 * kOut, u.bX, u.At[0], u.bT, c->stream));
    if ((rc = sur_forward(c, u.bX, batch, Mpad, true))) return rc;
    if ((rc = sur_backward(c, batch, Mpad))) return rc;
    HIPCHK(c, launch_sum_f32(u.row_loss, (int)Mpad, u.losses + s, c->stream));
    u.adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f / (1.0f - std::pow(b1, (float)u.adam_t));
    const float c2 = 1.0f / (1.0f - std::pow(b2, (float)u.adam_t));
    HIPCHK(c, launch_adam(u.params, u.grad, u.adam_m, u.adam_v, o.total, lr, b1, b2, 1e-8f, c1, c2, c->stream));
    if ((rc = sur_pack(c))) return rc;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (losses && steps > 0) HIPCHK(c, hipMemcpy(losses, u.losses, steps * sizeof(float), hipMemcpyDeviceToHost));
  return CBZ_OK;
}

int cbz_sur_check_gemms(int32_t rows, char* msg, int32_t msg_len) {
  if (rows <= 0) return CBZ_ERR_ARG;
  // A context with no device: every buffer the training step and the
  // inference pass use gets a disjoint fake address range of exactly the
  // size its allocator gives it, then both passes run with their GEMMs
  // checked instead of launched.
  cbz_context fake;
  std::vector<std::pair<uintptr_t, uintptr_t>> regions;
  uintptr_t next = uintptr_t(1) << 44;
  auto place = [&](auto*& p, int64_t n) {
    using T = std::remove_reference_t<decltype(*p)>;
    const uintptr_t bytes = uintptr_t(std::max<int64_t>(n, 1)) * sizeof(T);
    p = (decltype(p))next;
    regions.push_back({next, next + bytes});
    next += ((bytes + (uintptr_t(1) << 20)) >> 20 << 20) + (uintptr_t(1) << 20);
  };
  const int64_t Mpad = round32(rows);
  sur_model_buffers(fake.sur, place);
  sur_batch_buffers(fake.sur, Mpad, place);
  sur_infer_buffers(fake.sur, Mpad, place);
  sur_layered_buffers(fake.sur, Mpad, place);
  fake.dry_regions = &regions;
  int rc = sur_forward(&fake, fake.sur.bX, rows, Mpad, true);
  if (!rc) rc = sur_backward(&fake, rows, Mpad);
  if (!rc) rc = sur_forward(&fake, fake.sur.bX, rows, Mpad, false);
  if (msg && msg_len > 0) {
    std::snprintf(msg, (size_t)msg_len, "%s", rc ? fake.err.c_str() : "");
  }
  return rc;
}

// rows of raw features (and raw labels) -> the staging batch
static int sur_stage_rows(cbz_ctx* c, int32_t rows, const float* X, const float* Y) {
  auto& u = c->sur;
  int rc = sur_ensure_batch(c, rows);
  if (rc) return rc;
  const int64_t Mpad = round32(rows);
  float *dx = nullptr, *dy = nullptr;
  HIPCHK(c, hipMalloc((void**)&dx, (size_t)rows * kFeat * sizeof(float)));
  HIPCHK(c, hipMalloc((void**)&dy, (size_t)rows * kOut * sizeof(float)));
  std::vector<int64_t> idx((size_t)rows);
  for (int i = 0; i < rows; ++i) idx[i] = i;
  hipError_t e = hipMemcpy(dx, X, rows * kFeat * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess && Y) e = hipMemcpy(dy, Y, rows * kOut * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(u.idx, idx.data(), rows * sizeof(int64_t), hipMemcpyHostToDevice);
  const float* nm = u.norm;
  if (e == hipSuccess)
    e = launch_gather_batch(dx, Y ? dy : nullptr, u.idx, (int)Mpad, rows, nm, nm + kFeat, nm + 2 * kFeat,
                            nm + 2 * kFeat + 2 * kOut, u.bX, u.At[0], u.bT, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dx);
  (void)hipFree(dy);
  if (e != hipSuccess) return fail(c, CBZ_ERR_HIP, std::string("sur: staging: ") + hipGetErrorString(e));
  return CBZ_OK;
}

int cbz_sur_grad(cbz_ctx* c, int32_t rows, const float* X, const float* Y, float* grad, float* loss) {
  if (!c || rows <= 0 || !X || !Y) return fail(c, CBZ_ERR_ARG, "sur_grad: bad arguments");
  if (!c->sur.have_model) return fail(c, CBZ_ERR_STATE, "sur_grad: no model");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = sur_stage_rows(c, rows, X, Y);
  if (rc) return rc;
  auto& u = c->sur;
  const int64_t Mpad = round32(rows);
  if ((rc = sur_forward(c, u.bX, rows, Mpad, true))) return rc;
  if ((rc = sur_backward(c, rows, Mpad))) return rc;
  float* dl = nullptr;
  HIPCHK(c, hipMalloc((void**)&dl, sizeof(float)));
  HIPCHK(c, launch_sum_f32(u.row_loss, (int)Mpad, dl, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (loss) HIPCHK(c, hipMemcpy(loss, dl, sizeof(float), hipMemcpyDeviceToHost));
  (void)hipFree(dl);
  if (grad) HIPCHK(c, hipMemcpy(grad, u.grad, layer_offsets().total * sizeof(float), hipMemcpyDeviceToHost));
  return CBZ_OK;
}

int cbz_sur_predict(cbz_ctx* c, int32_t rows, const float* X, float* Y, int32_t flags) {
  if (!c || rows <= 0 || !X || !Y) return fail(c, CBZ_ERR_ARG, "sur_predict: bad arguments");
  if (!c->sur.have_model) return fail(c, CBZ_ERR_STATE, "sur_predict: no model");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = sur_stage_rows(c, rows, X, nullptr);
  if (rc) return rc;
  auto& u = c->sur;
  if ((rc = sur_ensure_inference(c, rows))) return rc;
  const int64_t Mpad = round32(rows);
  if (flags & CBZ_SUR_PREDICT_LAYERED) {
    if ((rc = sur_ensure_layered(c, rows))) return rc;
    rc = sur_forward(c, u.bX, rows, Mpad, false);
  } else
    rc = sur_infer(c, u.bX, rows);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(Y, u.Yp, (size_t)rows * kOut * sizeof(float), hipMemcpyDeviceToHost));
  return CBZ_OK;
}

int cbz_sur_predict_device(cbz_ctx* c, int32_t rows, const float* features, float* out) {
  if (!c || rows < 0 || (rows > 0 && (!features || !out)))
    return fail(c, CBZ_ERR_ARG, "sur_predict_device: bad arguments");
  if (rows > kNoAllocRows)
    return fail(c, CBZ_ERR_ARG, "sur_predict_device: " + std::to_string(rows) + " rows; at most " +
                                    std::to_string(kNoAllocRows) + " per call");
  if (!c->sur.have_model) return fail(c, CBZ_ERR_STATE, "sur_predict_device: no model");
  if (!c->scratch) return fail(c, CBZ_ERR_STATE, "sur_predict_device: no scratch address");
  if (rows == 0) return CBZ_OK;
  HIPCHK(c, hipSetDevice(c->device));
  auto& u = c->sur;
  const float* nm = u.norm;
  __bf16* X = (__bf16*)c->scratch;  // every row the network reads is written here first
  HIPCHK(c, launch_gather_batch(features, nullptr, nullptr, rows, rows, nm, nm + kFeat, nm + 2 * kFeat,
                                nm + 2 * kFeat + 2 * kOut, X, nullptr, nullptr, c->stream));
  const LayerOff o = layer_offsets();
  MlpArgs f{};
  f.M = rows;
  f.X = X;
  for (int l = 0; l < kLayers; ++l) {
    f.W[l] = u.Wnk[l];
    f.bias[l] = u.params + o.b[l];
  }
  f.y_scale = u.norm + 2 * kFeat + kOut;  // out_std
  f.y_shift = u.norm + 2 * kFeat;         // out_mean
  f.Y = out;
  HIPCHK(c, launch_mlp_forward(f, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CBZ_OK;
}

int cbz_sur_resolve(cbz_ctx* c, double dt, int64_t iter_per_contact, double tol, int64_t polish_per_contact,
                    int32_t flags, int64_t* n_fallback) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (c->mat_flags & CBZ_MATERIALS) return fail(c, CBZ_ERR_STATE, "sur_resolve: contact materials are on");
  if (iter_per_contact < 0 || polish_per_contact < 0 || !(tol == tol))
    return fail(c, CBZ_ERR_ARG, "sur_resolve: bad arguments");
  if (!c->have_contacts) return fail(c, CBZ_ERR_STATE, "sur_resolve: no contacts (run a narrowphase first)");
  if (!c->sur.have_model) return fail(c, CBZ_ERR_STATE, "sur_resolve: no model");
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = sur_ensure_inference(c, c->cap))) return rc;
  rc = c->precision == CBZ_F64
           ? sur_resolve<double>(c, dt, iter_per_contact, tol, polish_per_contact, flags, c->sur.nfall)
           : sur_resolve<float>(c, dt, iter_per_contact, tol, polish_per_contact, flags, c->sur.nfall);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (n_fallback) {
    int32_t v = 0;
    HIPCHK(c, hipMemcpy(&v, c->sur.nfall, sizeof v, hipMemcpyDeviceToHost));
    *n_fallback = v;
  }
  return CBZ_OK;
}

int cbz_sur_set_position(cbz_ctx* c, double tol, int64_t polish_per_contact) {
  if (!c || !(tol == tol) || polish_per_contact < 0) return fail(c, CBZ_ERR_ARG, "sur_set_position: bad arguments");
  if (tol != c->sur.pos_tol || polish_per_contact != c->sur.pos_polish) {
    c->sur.pos_tol = tol;
    c->sur.pos_polish = polish_per_contact;
    c->sur.generation += 1;  // a captured surrogate step is stale
  }
  return CBZ_OK;
}

int cbz_sur_get_position(cbz_ctx* c, double* tol, int64_t* polish_per_contact) {
  if (!c) return CBZ_ERR_ARG;
  if (tol) *tol = c->sur.pos_tol;
  if (polish_per_contact) *polish_per_contact = c->sur.pos_polish;
  return CBZ_OK;
}

int cbz_sur_world_mask(cbz_ctx* c, uint8_t* mask) {
  if (!c || !mask) return CBZ_ERR_ARG;
  if (!c->sur.mask) return fail(c, CBZ_ERR_STATE, "sur_world_mask: no surrogate resolve yet");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(mask, c->sur.mask, c->nw, hipMemcpyDeviceToHost));
  return CBZ_OK;
}

int cbz_sur_position_mask(cbz_ctx* c, uint8_t* mask) {
  if (!c || !mask) return CBZ_ERR_ARG;
  if (!c->sur.mask) return fail(c, CBZ_ERR_STATE, "sur_position_mask: no surrogate resolve yet");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->sur.last_pos)
    HIPCHK(c, hipMemcpy(mask, c->sur.pmask, c->nw, hipMemcpyDeviceToHost));
  else
    std::fill(mask, mask + c->nw, uint8_t(1));
  return CBZ_OK;
}

int cbz_step_sur(cbz_ctx* c, double dt, int32_t n_steps, int64_t iter_per_contact, double tol,
                 int64_t polish_per_contact, int32_t flags, int64_t* counts, int64_t* fallbacks) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (c->mat_flags & CBZ_MATERIALS) return fail(c, CBZ_ERR_STATE, "step_sur: contact materials are on");
  if (n_steps < 0 || iter_per_contact < 0 || polish_per_contact < 0 || !(tol == tol))
    return fail(c, CBZ_ERR_ARG, "step_sur: bad arguments");
  if (!c->sur.have_model) return fail(c, CBZ_ERR_STATE, "step_sur: no model");
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = sur_ensure_inference(c, c->cap))) return rc;
  const bool f64 = c->precision == CBZ_F64;
  if (!counts && !c->prof && n_steps > 0) {
    // Replay a captured graph of one step, as cbz_step does: the host-side
    // inputs are refreshed first, and every buffer the step uses exists.
    rc = refresh_pow(c, dt);
    if (rc == CBZ_OK) rc = refresh_brad(c);
    if (rc) return rc;
    if (n_steps > c->fb_cap) {
      const int64_t cap = std::max<int64_t>(n_steps, 2 * c->fb_cap);
      if ((rc = dalloc(c, &c->fb_hist, (size_t)cap))) return rc;
      if (!c->fb_ctr && (rc = dalloc(c, &c->fb_ctr, 1))) return rc;
      c->fb_cap = cap;
      if (c->sgexec) (void)hipGraphExecDestroy(c->sgexec);
      c->sgexec = nullptr;
    }
    if (!c->sgexec || c->sg_dt != dt || c->sg_mult != iter_per_contact || c->sg_tol != tol ||
        c->sg_polish != polish_per_contact || c->sg_flags != flags || c->sg_gen != c->sur.generation) {
      if (c->sgexec) (void)hipGraphExecDestroy(c->sgexec);
      c->sgexec = nullptr;
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int crc = f64 ? do_integrate<double>(c, dt) : do_integrate<float>(c, dt);
      if (crc == CBZ_OK) crc = f64 ? do_narrowphase<double>(c, -1) : do_narrowphase<float>(c, -1);
      if (crc == CBZ_OK)
        crc = f64 ? sur_resolve<double>(c, dt, iter_per_contact, tol, polish_per_contact, flags, c->sur.nfall)
                  : sur_resolve<float>(c, dt, iter_per_contact, tol, polish_per_contact, flags, c->sur.nfall);
      if (crc == CBZ_OK && launch_record_count(c->sur.nfall, c->fb_hist, c->fb_ctr, c->fb_cap, c->stream) != hipSuccess)
        crc = fail(c, CBZ_ERR_HIP, "step_sur: record launch");
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (crc) {
        if (g) (void)hipGraphDestroy(g);
        return crc;
      }
      if (e != hipSuccess) return fail(c, CBZ_ERR_HIP, std::string("step_sur: capture: ") + hipGetErrorString(e));
      e = hipGraphInstantiate(&c->sgexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        c->sgexec = nullptr;
        return fail(c, CBZ_ERR_HIP, std::string("step_sur: instantiate: ") + hipGetErrorString(e));
      }
      c->sg_dt = dt;
      c->sg_mult = iter_per_contact;
      c->sg_tol = tol;
      c->sg_polish = polish_per_contact;
      c->sg_flags = flags;
      c->sg_gen = c->sur.generation;
    }
    HIPCHK(c, launch_fill_i32(c->overflow, 2, 0, c->stream));
    HIPCHK(c, launch_fill_i32(c->fb_ctr, 1, 0, c->stream));
    for (int32_t s = 0; s < n_steps; ++s) HIPCHK(c, hipGraphLaunch(c->sgexec, c->stream));
    c->have_contacts = true;
    if ((rc = check_overflow(c))) return rc;
    if (fallbacks) {
      std::vector<int32_t> f((size_t)n_steps);
      HIPCHK(c, hipMemcpy(f.data(), c->fb_hist, f.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
      for (size_t k = 0; k < f.size(); ++k) fallbacks[k] = f[k];
    }
    return CBZ_OK;
  }
  int32_t* hist = nullptr;
  int32_t* fb = nullptr;
  if (n_steps > 0) {
    if (counts) HIPCHK(c, hipMalloc((void**)&hist, (size_t)n_steps * c->nw * sizeof(int32_t)));
    HIPCHK(c, hipMalloc((void**)&fb, (size_t)n_steps * sizeof(int32_t)));
  }
  HIPCHK(c, launch_fill_i32(c->overflow, 2, 0, c->stream));
  for (int32_t s = 0; s < n_steps && rc == CBZ_OK; ++s) {
    rc = f64 ? do_integrate<double>(c, dt) : do_integrate<float>(c, dt);
    if (rc) break;
    rc = f64 ? do_narrowphase<double>(c, -1) : do_narrowphase<float>(c, -1);
    if (rc) break;
    if (hist)
      HIPCHK(c, hipMemcpyAsync(hist + (size_t)s * c->nw, c->world_cnt, c->nw * sizeof(int32_t),
                               hipMemcpyDeviceToDevice, c->stream));
    rc = f64 ? sur_resolve<double>(c, dt, iter_per_contact, tol, polish_per_contact, flags, fb + s)
             : sur_resolve<float>(c, dt, iter_per_contact, tol, polish_per_contact, flags, fb + s);
  }
  if (rc == CBZ_OK) rc = check_overflow(c);
  if (rc == CBZ_OK && n_steps > 0) {
    std::vector<int32_t> t((size_t)n_steps * c->nw), f((size_t)n_steps);
    hipError_t e = hipSuccess;
    if (hist) e = hipMemcpy(t.data(), hist, t.size() * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(f.data(), fb, f.size() * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(c, CBZ_ERR_HIP, std::string("step_sur: download: ") + hipGetErrorString(e));
    if (counts)
      for (size_t k = 0; k < t.size(); ++k) counts[k] = t[k];
    if (fallbacks)
      for (size_t k = 0; k < f.size(); ++k) fallbacks[k] = f[k];
  }
  if (hist) (void)hipFree(hist);
  if (fb) (void)hipFree(fb);
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// State checkpoints: header, then the payload (little-endian).
//   header (64 bytes): magic "CBZSTAT1", u32 version, u32 header bytes,
//     i32 n_worlds, i32 n_planes, i64 n_bodies, i32 material flags,
//     i32 field count, i32 precision of the saving context, i32 0,
//     u64 payload bytes, u64 FNV-1a hash of the payload
//   payload: i32 body_off[n_worlds + 1], i32 plane_off[n_worlds + 1]
//     (zero-padded to 8 bytes), f64 plane normal [n_planes][3], offset,
//     friction, restitution [n_planes]; then per field: i32 id, i32 width,
//     i32 kind (0 f64, 1 i32, 2 u8), i32 0, i64 data bytes, data in the host
//     layout of cbz_get_field (zero-padded to 8 bytes)
// ---------------------------------------------------------------------------
namespace {

constexpr char kCkptMagic[8] = {'C', 'B', 'Z', 'S', 'T', 'A', 'T', '1'};
constexpr uint32_t kCkptVersion = 1;
constexpr size_t kCkptHeader = 64;

struct CkptHeader {
  char magic[8];
  uint32_t version;
  uint32_t header_bytes;
  int32_t n_worlds;
  int32_t n_planes;
  int64_t n_bodies;
  int32_t mat_flags;
  int32_t n_fields;
  int32_t precision;
  int32_t zero;
  uint64_t payload_bytes;
  uint64_t hash;
};
static_assert(sizeof(CkptHeader) == kCkptHeader, "checkpoint header layout");

size_t pad8(size_t n) { return (n + 7) & ~size_t(7); }

uint64_t fnv1a(const uint8_t* p, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

size_t field_bytes(int f, int64_t nb) {
  const FieldInfo& fi = kFields[f];
  const size_t el = (fi.kind == 1) ? sizeof(int32_t) : (fi.kind == 2) ? 1 : sizeof(double);
  return el * (size_t)fi.width * (size_t)nb;
}

int field_kind(int f) { return kFields[f].kind == 1 ? 1 : kFields[f].kind == 2 ? 2 : 0; }

size_t ckpt_payload_bytes(const cbz_ctx* c) {
  size_t n = pad8(2 * sizeof(int32_t) * (size_t)(c->nw + 1));
  n += sizeof(double) * 6 * (size_t)c->np;
  for (int f = 0; f < CBZ_NUM_FIELDS; ++f) n += 24 + pad8(field_bytes(f, c->nb));
  return n;
}

}  // namespace

extern "C" {

int cbz_checkpoint_size(cbz_ctx* c, int64_t* bytes) {
  int rc = need_scene(c);
  if (rc) return rc;
  if (!bytes) return fail(c, CBZ_ERR_ARG, "checkpoint_size: null output");
  *bytes = (int64_t)(kCkptHeader + ckpt_payload_bytes(c));
  return CBZ_OK;
}

int cbz_checkpoint_save(cbz_ctx* c, void* buf, int64_t bytes) {
  int rc = need_scene(c);
  if (rc) return rc;
  const size_t payload = ckpt_payload_bytes(c);
  if (!buf || bytes < (int64_t)(kCkptHeader + payload))
    return fail(c, CBZ_ERR_ARG, "checkpoint_save: buffer smaller than cbz_checkpoint_size");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  uint8_t* out = (uint8_t*)buf;
  std::memset(out, 0, kCkptHeader + payload);
  uint8_t* p = out + kCkptHeader;
  std::memcpy(p, c->body_off.data(), sizeof(int32_t) * (c->nw + 1));
  std::memcpy(p + sizeof(int32_t) * (c->nw + 1), c->plane_off.data(), sizeof(int32_t) * (c->nw + 1));
  p += pad8(2 * sizeof(int32_t) * (size_t)(c->nw + 1));
  const int64_t np = c->np;
  if (np > 0) {
    std::vector<double> t((size_t)np * 3);
    if ((rc = download_real(c, c->plane_normal, np, 0, np, 3, t.data()))) return rc;
    std::memcpy(p, t.data(), sizeof(double) * 3 * np);
    p += sizeof(double) * 3 * np;
    void* const one[3] = {c->plane_offset, c->plane_fric, c->plane_rest};
    for (void* src : one) {
      if ((rc = download_real(c, src, np, 0, np, 1, t.data()))) return rc;
      std::memcpy(p, t.data(), sizeof(double) * np);
      p += sizeof(double) * np;
    }
  }
  std::vector<uint8_t> tmp;
  for (int f = 0; f < CBZ_NUM_FIELDS; ++f) {
    const size_t fb = field_bytes(f, c->nb);
    const int32_t hdr[4] = {f, kFields[f].width, field_kind(f), 0};
    const int64_t len = (int64_t)fb;
    std::memcpy(p, hdr, sizeof hdr);
    std::memcpy(p + 16, &len, sizeof len);
    p += 24;
    if (fb > 0) {
      tmp.resize(fb);
      if ((rc = cbz_get_field(c, f, tmp.data()))) return rc;
      std::memcpy(p, tmp.data(), fb);
    }
    p += pad8(fb);
  }
  CkptHeader h{};
  std::memcpy(h.magic, kCkptMagic, sizeof kCkptMagic);
  h.version = kCkptVersion;
  h.header_bytes = (uint32_t)kCkptHeader;
  h.n_worlds = c->nw;
  h.n_planes = c->np;
  h.n_bodies = c->nb;
  h.mat_flags = c->mat_flags;
  h.n_fields = CBZ_NUM_FIELDS;
  h.precision = c->precision;
  h.payload_bytes = payload;
  h.hash = fnv1a(out + kCkptHeader, payload);
  std::memcpy(out, &h, sizeof h);
  return CBZ_OK;
}

int cbz_checkpoint_load(cbz_ctx* c, const void* buf, int64_t bytes) {
  if (!c) return CBZ_ERR_ARG;
  if (!buf || bytes < (int64_t)kCkptHeader) return fail(c, CBZ_ERR_ARG, "checkpoint_load: truncated header");
  const uint8_t* in = (const uint8_t*)buf;
  CkptHeader h;
  std::memcpy(&h, in, sizeof h);
  if (std::memcmp(h.magic, kCkptMagic, sizeof kCkptMagic) != 0)
    return fail(c, CBZ_ERR_ARG, "checkpoint_load: not a cubez_amd state checkpoint");
  if (h.version != kCkptVersion || h.header_bytes != kCkptHeader)
    return fail(c, CBZ_ERR_ARG, "checkpoint_load: unsupported version " + std::to_string(h.version));
  if (h.n_worlds <= 0 || h.n_planes < 0 || h.n_bodies < 0 || h.n_fields != CBZ_NUM_FIELDS ||
      bytes - (int64_t)kCkptHeader < (int64_t)h.payload_bytes)
    return fail(c, CBZ_ERR_ARG, "checkpoint_load: truncated or inconsistent header");
  const uint8_t* p = in + kCkptHeader;
  if (fnv1a(p, h.payload_bytes) != h.hash) return fail(c, CBZ_ERR_ARG, "checkpoint_load: payload hash mismatch");
  const int32_t nw = h.n_worlds;
  const size_t off_bytes = pad8(2 * sizeof(int32_t) * (size_t)(nw + 1));
  if (h.payload_bytes < off_bytes + sizeof(double) * 6 * (size_t)h.n_planes)
    return fail(c, CBZ_ERR_ARG, "checkpoint_load: payload too short");
  std::vector<int32_t> bo((size_t)nw + 1), po((size_t)nw + 1);
  std::memcpy(bo.data(), p, sizeof(int32_t) * (nw + 1));
  std::memcpy(po.data(), p + sizeof(int32_t) * (nw + 1), sizeof(int32_t) * (nw + 1));
  if (bo[nw] != h.n_bodies || po[nw] != h.n_planes)
    return fail(c, CBZ_ERR_ARG, "checkpoint_load: offsets disagree with the header");
  const uint8_t* q = p + off_bytes;
  const int64_t np = h.n_planes;
  std::vector<double> pn((size_t)np * 3), pofs((size_t)np), pf((size_t)np), pr((size_t)np);
  std::memcpy(pn.data(), q, sizeof(double) * 3 * np);
  q += sizeof(double) * 3 * np;
  std::memcpy(pofs.data(), q, sizeof(double) * np);
  q += sizeof(double) * np;
  std::memcpy(pf.data(), q, sizeof(double) * np);
  q += sizeof(double) * np;
  std::memcpy(pr.data(), q, sizeof(double) * np);
  q += sizeof(double) * np;
  // every field record must be present and whole before the context changes
  const uint8_t* end = p + h.payload_bytes;
  const uint8_t* fields[CBZ_NUM_FIELDS] = {};
  for (int f = 0; f < CBZ_NUM_FIELDS; ++f) {
    if (end - q < 24) return fail(c, CBZ_ERR_ARG, "checkpoint_load: truncated field record");
    int32_t hdr[4];
    int64_t len;
    std::memcpy(hdr, q, sizeof hdr);
    std::memcpy(&len, q + 16, sizeof len);
    if (hdr[0] != f || hdr[1] != kFields[f].width || hdr[2] != field_kind(f) ||
        len != (int64_t)field_bytes(f, h.n_bodies) || end - (q + 24) < (ptrdiff_t)pad8((size_t)len))
      return fail(c, CBZ_ERR_ARG, "checkpoint_load: field record " + std::to_string(f) + " does not match");
    fields[f] = q + 24;
    q += 24 + pad8((size_t)len);
  }
  for (int64_t i = 0; i < h.n_bodies; ++i) {
    int32_t sh;
    std::memcpy(&sh, fields[CBZ_SHAPE] + 4 * i, sizeof sh);
    if (sh != CBZ_CUBE && sh != CBZ_SPHERE) return fail(c, CBZ_ERR_ARG, "checkpoint_load: unknown shape");
  }
  if (h.mat_flags & ~(CBZ_MATERIALS | CBZ_MATERIALS_FRICTIONLESS_TWO_BODY))
    return fail(c, CBZ_ERR_ARG, "checkpoint_load: unknown material flags");
  int rc = cbz_set_scene(c, nw, bo.data(), po.data(), pn.data(), pofs.data());
  if (rc) return rc;
  if ((rc = cbz_set_materials(c, h.mat_flags))) return rc;
  if ((rc = cbz_set_plane_materials(c, pf.data(), pr.data()))) return rc;
  // OFFSET before FLAGS: the flags upload keeps the identity-offset bit the
  // offset upload derives
  std::vector<int> order;
  for (int f = 0; f < CBZ_NUM_FIELDS; ++f)
    if (f != CBZ_FLAGS) order.push_back(f);
  order.push_back(CBZ_FLAGS);
  for (int f : order) {
    if (h.n_bodies == 0) break;
    if ((rc = cbz_set_field(c, f, fields[f]))) return rc;
  }
  return CBZ_OK;
}

}  // extern "C"
