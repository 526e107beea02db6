// cbz_mlp.h — shapes and launch interfaces of the resolver surrogate MLP.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cbz_types.h"

namespace cbz {

constexpr int kFeat = 40;      // input features per contact
constexpr int kFeatPad = 48;   // padded to the MFMA K step (16)
constexpr int kHidden = 256;   // hidden width
constexpr int kOutVel = 12;    // outputs 0-11: dv0, dw0, dv1, dw1 (the velocity phase's changes)
constexpr int kOutPos = 12;    // output 12: the penetration the position phase resolves at the contact
constexpr int kOut = 13;       // outputs per contact
constexpr int kOutPad = 32;    // padded to the MFMA tile (32)
constexpr int kLayers = 4;
constexpr int kSurBodies = 256;     // bodies per world the surrogate takes
constexpr int kSurContacts = 2048;  // contacts per world the surrogate takes
// contacts per world the position polish takes (the one-wavefront
// kernel's widest register class); larger worlds get the plain position phase
constexpr int kSurPosContacts = 2048;

// layer l maps kIn[l] -> kOutL[l] features (unpadded)
__host__ __device__ constexpr int layer_in(int l) { return l == 0 ? kFeat : kHidden; }
__host__ __device__ constexpr int layer_out(int l) { return l == kLayers - 1 ? kOut : kHidden; }
__host__ __device__ constexpr int layer_in_pad(int l) { return l == 0 ? kFeatPad : kHidden; }
__host__ __device__ constexpr int layer_out_pad(int l) { return l == kLayers - 1 ? kOutPad : kHidden; }

struct GemmArgs {
  int M, N, K;                  // padded sizes: M, N multiples of 32, K of 16
  const __bf16* A;              // [M x K] row-major
  int64_t lda;
  const __bf16* Bt;             // [N x K] row-major (B transposed)
  int64_t ldb;
  const float* bias;            // [N] or null
  int act;                      // 1: SiLU
  float* Z;                     // pre-activation FP32 [M x N] (or null)
  int64_t ldz;
  float* Y;                     // activation FP32 (or null)
  int64_t ldy;
  const float* y_scale;         // Y = act * scale + shift per column (or null)
  const float* y_shift;
  __bf16* H;                    // activation BF16 row-major (or null)
  int64_t ldh;
  __bf16* Ht;                   // activation BF16 transposed [N x M] (or null)
  int64_t ldht;
  int m_valid, n_valid;         // rows / columns that hold data
};

// fused inference forward: all four layers per block of rows
// Grid sizes of the grid-stride kernels of the surrogate step: whatever the
// contact count, one launch of at most this many workgroups covers it.
constexpr int kGridStrideBlocks = 8192;  // 256-thread workgroups (features)
constexpr int kMlpGridBlocks = 1024;     // fused-forward workgroups (2 per CU fit its LDS)

struct MlpArgs {
  int64_t M;                      // rows (the capacity when M_dev is set)
  const int64_t* M_dev = nullptr; // rows counted on the device (clamped to M), or null
  const __bf16* X;                // [M][kFeatPad] normalised features
  const __bf16* W[kLayers];       // per layer [Npad][Kpad] row-major (the k_gemm Bt operand)
  const float* bias[kLayers];
  const float* y_scale;           // output = y * scale + shift (denormalisation)
  const float* y_shift;
  float* Y;                       // [M][kOut]
};

template <typename R>
struct SurApply {
  Bodies<R> b;
  Contacts<R> c;
  const int32_t* world_cnt;
  const int64_t* world_coff;
  const int32_t* body_off;
  const int32_t* overflow;
  const float* Y;               // [contacts][kOut] predictions
  float* delta;                 // out: [6][n_bodies] summed velocity and rotation change per body
  uint8_t* mask;                // out: 1 = plain velocity phase, 0 = polish
  float* dpos;                  // out: [6][n_bodies] summed position and orientation change per body
  uint8_t* pmask;               // out: 1 = plain position phase, 0 = polish (null: no position polish)
  int max_contacts;
};

hipError_t launch_gemm(const GemmArgs& g, hipStream_t st);
hipError_t launch_mlp_forward(const MlpArgs& f, hipStream_t st);
template <typename R>
hipError_t launch_features(const Bodies<R>& b, const Contacts<R>& c, int64_t n_cap, const int64_t* n_dev,
                           int32_t* deg, const float* mean, const float* inv_std, __bf16* X, float* raw,
                           hipStream_t st);
// The feature rows of the step's contacts (normalised BF16 rows into X and/or
// raw FP32 rows into raw, either may be null), one workgroup per world, the
// per-body contact counts in LDS; worlds of at most 4096 bodies (else
// hipErrorInvalidValue: use launch_features).
template <typename R>
hipError_t launch_features_world(const Bodies<R>& b, const Contacts<R>& c, const int32_t* world_cnt,
                                 const int64_t* world_coff, const int32_t* body_off, const int32_t* overflow,
                                 int n_worlds, int max_world_bodies, const float* mean, const float* inv_std,
                                 __bf16* X, float* raw, hipStream_t st);
template <typename R>
hipError_t launch_sur_scatter(const SurApply<R>& a, int n_worlds, hipStream_t st);
hipError_t launch_gather_batch(const float* Xraw, const float* Yraw, const int64_t* idx, int rows, int valid_rows,
                               const float* in_mean, const float* in_inv_std, const float* out_mean,
                               const float* out_inv_std, __bf16* X, __bf16* Xt, float* T, hipStream_t st);
hipError_t launch_silu_bwd(const float* dA, const float* Z, int rows, int valid_rows, int cols, int ld, __bf16* dZ,
                           __bf16* dZt, int ldt, hipStream_t st);
hipError_t launch_sample_idx(int64_t* idx, int n, uint64_t seed, uint64_t step, int64_t rows, hipStream_t st);
hipError_t launch_col_stats(const float* A, int64_t rows, int cols, float* mean, float* stdv, hipStream_t st);
// count of worlds whose byte is set in m or (when not null) in m2
hipError_t launch_count_u8(const uint8_t* m, int n, int32_t* out, hipStream_t st, const uint8_t* m2 = nullptr);
hipError_t launch_fill_f32(float* p, int64_t n, float v, hipStream_t st);
hipError_t launch_record_count(const int32_t* value, int32_t* hist, int32_t* ctr, int64_t cap, hipStream_t st);
hipError_t launch_mse(const float* Y, const float* T, int rows, int valid_rows, float* row_loss, __bf16* dY,
                      __bf16* dYt, int ldt, hipStream_t st);
hipError_t launch_colsum(const __bf16* A, int rows, int cols, int ld, float* out, hipStream_t st);
hipError_t launch_sum_f32(const float* a, int n, float* out, hipStream_t st);
hipError_t launch_adam(float* p, const float* grad, float* m, float* v, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2, hipStream_t st);
hipError_t launch_pack_weights(const float* W, int N, int K, __bf16* Wnk, int kpad, __bf16* Wkn, int npad,
                               hipStream_t st);

}  // namespace cbz
