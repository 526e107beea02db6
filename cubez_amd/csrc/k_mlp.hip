// k_mlp.hip — the learned resolver surrogate: a 40 -> 256 -> 256 -> 256 -> 12
// MLP with SiLU, BF16 weights and activations, FP32 accumulation, run on
// the matrix cores (v_mfma_f32_32x32x16_bf16) for inference and training.
//
// Per contact it predicts the velocity and rotation changes that the exact
// resolver's velocity phase (adjustVelocities, contact.go:390-445) gives the
// contact's two bodies, summed over the iterations that chose the contact.
// Summing the predictions of a body's contacts gives the body's change.
//
// Kernels
//   k_degree      contacts per body (an input feature)
//   k_features    the 40 input features of every contact, normalised, BF16
//   k_gemm        C = A * B (+bias, SiLU) on MFMA 32x32x16, FP32 accumulate
//   k_silu_bwd, k_mse, k_colsum, k_adam   backward pass and optimiser
//   k_sur_scatter per world: sum the predicted changes per body (the
//                 resolver then polishes or re-solves the world)
#include <algorithm>
#include <climits>

#include "cbz_devmath.h"
#include "cbz_launch.h"
#include "cbz_mlp.h"
#include "cbz_types.h"
#include "cbz_wave.h"

namespace cbz {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <typename R>
__device__ __forceinline__ V3<R> ld3c(const R* a, int64_t s, int64_t i) {
  return V3<R>{a[i], a[s + i], a[2 * s + i]};
}

// SiLU with the hardware reciprocal (about 1 ulp): its result is rounded to
// BF16 for the next layer
__device__ __forceinline__ float silu(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
// Rows handled by a grid-stride kernel: n_cap, or the device count n_dev
// (read when the kernel runs) clamped to n_cap.  The grids of the surrogate
// step are sized from capacities, so the step needs no host round trip and
// can be captured in a graph.
__device__ __forceinline__ int64_t rows_of(int64_t n_cap, const int64_t* n_dev) {
  if (!n_dev) return n_cap;
  const int64_t n = *n_dev;
  return n < n_cap ? n : n_cap;
}

__global__ void k_degree(const int32_t* body0, const int32_t* body1, int64_t n_cap, const int64_t* n_dev,
                         int32_t* deg) {
  const int64_t n = rows_of(n_cap, n_dev);
  for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < n; g += int64_t(gridDim.x) * blockDim.x) {
    atomicAdd(&deg[body0[g]], 1);
    if (body1[g] >= 0) atomicAdd(&deg[body1[g]], 1);
  }
}

// Feature layout (kFeat = 40): normal 3, rel0 3, rel1 3, v0-v1 3, w0-w1 3,
// invMass0, invMass1, iitWorld0 9, iitWorld1 9, penetration,
// desiredDeltaVelocity, contactVelocity.x, 1/deg0, 1/deg1.
// deg0, deg1: the contact counts of the contact's bodies (deg1 unused when
// there is no second body).
template <typename R>
__device__ __forceinline__ void contact_features(const Bodies<R>& b, const Contacts<R>& c, int64_t g, int deg0,
                                                 int deg1, const float* mean, const float* inv_std, __bf16* X,
                                                 float* raw) {
  const int64_t S = b.stride, CS = c.cap;
  const int b0 = c.body0[g], b1 = c.body1[g];
  float f[kFeat];
  int k = 0;
  for (int d = 0; d < 3; ++d) f[k++] = float(c.normal[d * CS + g]);
  for (int d = 0; d < 6; ++d) f[k++] = float(c.rel[d * CS + g]);
  for (int d = 0; d < 3; ++d) f[k++] = float(b.vel[d * S + b0]) - (b1 >= 0 ? float(b.vel[d * S + b1]) : 0.f);
  for (int d = 0; d < 3; ++d) f[k++] = float(b.rot[d * S + b0]) - (b1 >= 0 ? float(b.rot[d * S + b1]) : 0.f);
  f[k++] = float(b.inv_mass[b0]);
  f[k++] = b1 >= 0 ? float(b.inv_mass[b1]) : 0.f;
  for (int d = 0; d < 9; ++d) f[k++] = float(b.iit_world[d * S + b0]);
  for (int d = 0; d < 9; ++d) f[k++] = b1 >= 0 ? float(b.iit_world[d * S + b1]) : 0.f;
  f[k++] = float(c.pen[g]);
  f[k++] = float(c.ddv[g]);
  f[k++] = float(c.cvel[g]);
  f[k++] = 1.0f / float(deg0);
  f[k++] = b1 >= 0 ? 1.0f / float(deg1) : 0.f;
  if (raw)
    for (int i = 0; i < kFeat; ++i) raw[g * kFeat + i] = f[i];
  if (X) {
    // the row's 48 BF16 values as six 16-byte stores (kFeatPad * 2 bytes
    // per row keeps every row 16-byte aligned)
    bf16x8* x = (bf16x8*)(X + g * kFeatPad);
#pragma unroll
    for (int q = 0; q < kFeatPad / 8; ++q) {
      bf16x8 v;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int i = 8 * q + j;
        v[j] = __bf16(i < kFeat ? (f[i] - mean[i]) * inv_std[i] : 0.f);
      }
      x[q] = v;
    }
  }
}

template <typename R>
__global__ void k_features(Bodies<R> b, Contacts<R> c, int64_t n_cap, const int64_t* n_dev, const int32_t* deg,
                           const float* mean, const float* inv_std, __bf16* X, float* raw) {
  const int64_t n = rows_of(n_cap, n_dev);
  for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < n; g += int64_t(gridDim.x) * blockDim.x) {
    const int b0 = c.body0[g], b1 = c.body1[g];
    contact_features(b, c, g, deg[b0], b1 >= 0 ? deg[b1] : 0, mean, inv_std, X, raw);
  }
}

// The same rows for the step's contacts, one workgroup per world: the
// world's contacts are counted per body in LDS (integer counts, so the same
// degrees as k_degree's), then its rows written.  Worlds of at most
// kFeatWorldBodies bodies; the launcher takes the per-contact kernels for
// larger ones.
constexpr int kFeatWorldBodies = 4096;
template <typename R>
__global__ void __launch_bounds__(256) k_features_world(Bodies<R> b, Contacts<R> c, const int32_t* world_cnt,
                                                        const int64_t* world_coff, const int32_t* body_off,
                                                        const int32_t* overflow, const float* mean,
                                                        const float* inv_std, __bf16* X, float* raw) {
  __shared__ int32_t s_deg[kFeatWorldBodies];
  const int w = blockIdx.x, t = threadIdx.x;
  if (overflow[0] != 0) return;  // the step reports the overflow; no row is read
  const int C = world_cnt[w];
  const int b0w = body_off[w], n = body_off[w + 1] - b0w;
  const int64_t base = world_coff[w];
  for (int i = t; i < n; i += 256) s_deg[i] = 0;
  __syncthreads();
  // prepare has made body0 a body of the world for every contact; the
  // index tests keep a malformed contact from touching other LDS
  for (int j = t; j < C; j += 256) {
    const int64_t g = base + j;
    const unsigned l0 = unsigned(c.body0[g] - b0w), l1 = unsigned(c.body1[g] - b0w);
    if (l0 < unsigned(n)) atomicAdd(&s_deg[l0], 1);
    if (c.body1[g] >= 0 && l1 < unsigned(n)) atomicAdd(&s_deg[l1], 1);
  }
  __syncthreads();
  for (int j = t; j < C; j += 256) {
    const int64_t g = base + j;
    const unsigned l0 = unsigned(c.body0[g] - b0w), l1 = unsigned(c.body1[g] - b0w);
    const int d0 = l0 < unsigned(n) ? s_deg[l0] : 0, d1 = l1 < unsigned(n) ? s_deg[l1] : 0;
    contact_features(b, c, g, d0, d1, mean, inv_std, X, raw);
  }
}

// Batch sampling on the device: row i of training step `step` is
// SplitMix64(seed, step * n + i) mod rows (uniform with replacement).
__global__ void k_sample_idx(int64_t* idx, int n, uint64_t seed, uint64_t step, int64_t rows) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t z = seed * 0xD1B54A32D192ED03ull + (step * uint64_t(n) + uint64_t(i) + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  idx[i] = int64_t(z % uint64_t(rows));
}

// raw features / labels of a sampled batch -> normalised BF16 inputs and
// FP32 targets (training).  idx null takes rows 0..valid_rows-1 in order;
// Xt (the transposed copy) and T (the targets) may be null.
__global__ void k_gather_batch(const float* Xraw, const float* Yraw, const int64_t* idx, int rows, int valid_rows,
                               const float* in_mean, const float* in_inv_std, const float* out_mean,
                               const float* out_inv_std, __bf16* X, __bf16* Xt, float* T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const bool v = i < valid_rows;
  const int64_t s = v ? (idx ? idx[i] : int64_t(i)) : 0;  // idx null: row i itself
  for (int k = 0; k < kFeatPad; ++k) {
    const __bf16 x = __bf16(v && k < kFeat ? (Xraw[s * kFeat + k] - in_mean[k]) * in_inv_std[k] : 0.f);
    X[int64_t(i) * kFeatPad + k] = x;
    if (Xt) Xt[int64_t(k) * rows + i] = x;
  }
  if (T)
    for (int k = 0; k < kOut; ++k)
      T[int64_t(i) * kOut + k] = v && Yraw ? (Yraw[s * kOut + k] - out_mean[k]) * out_inv_std[k] : 0.f;
}

// column mean and standard deviation of a row-major FP32 table (one block
// per column, FP64 accumulation); std floored at 1e-6.
__global__ void k_col_stats(const float* A, int64_t rows, int cols, float* mean, float* stdv) {
  __shared__ double s1[256], s2[256];
  const int c = blockIdx.x;
  double a = 0, b = 0;
  for (int64_t r = threadIdx.x; r < rows; r += 256) {
    const double v = A[r * cols + c];
    a += v;
    b += v * v;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double m = rows > 0 ? s1[0] / double(rows) : 0.0;
    const double var = rows > 0 ? s2[0] / double(rows) - m * m : 0.0;
    mean[c] = float(m);
    stdv[c] = float(var > 1e-12 ? sqrt(var) : 1e-6);
  }
}

__global__ void k_count_u8(const uint8_t* m, const uint8_t* m2, int n, int32_t* out) {
  __shared__ int part[256];
  int s = 0;
  for (int i = threadIdx.x; i < n; i += 256) s += (m[i] || (m2 && m2[i])) ? 1 : 0;
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

// Appends *value to hist at *ctr (while below cap) and advances *ctr: one
// entry per replay of a captured step.
__global__ void k_record_count(const int32_t* value, int32_t* hist, int32_t* ctr, int64_t cap) {
  if (threadIdx.x != 0) return;
  const int32_t i = *ctr;
  if (i < cap) hist[i] = *value;
  *ctr = i + 1;
}

__global__ void k_fill_f32(float* p, int64_t n, float v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// GEMM on the matrix cores
// ---------------------------------------------------------------------------
// C[M x N] = A[M x K] * B[K x N] with A row-major (K contiguous) and B given
// as Bt = [N x K] row-major (K contiguous): both operands load their MFMA
// fragments as 16-byte vectors.  Fragment layout of v_mfma_f32_32x32x16_bf16
// for lane l (r = l % 32, h = l / 32): A[r][8h .. 8h+7], B[8h .. 8h+7][r],
// and accumulator j holds D[8 (j/4) + 4h + j%4][r].  Each wave computes one
// 32x32 tile; a block holds WM x WN waves.  K is a multiple of 16.  A tile
// that extends past row M of A or row N of Bt reads zeros there and writes
// nothing there, so an operand needs only M (N) rows.
template <int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN) k_gemm(GemmArgs g) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int m0 = (blockIdx.x * WM + wave % WM) * 32;
  const int n0 = (blockIdx.y * WN + wave / WM) * 32;
  if (m0 >= g.M || n0 >= g.N) return;
  const int r = lane & 31, h = lane >> 5;
  const bool a_in = m0 + r < g.M, b_in = n0 + r < g.N;
  const __bf16* ap = g.A + int64_t(a_in ? m0 + r : 0) * g.lda + 8 * h;
  const __bf16* bp = g.Bt + int64_t(b_in ? n0 + r : 0) * g.ldb + 8 * h;
  f32x16 acc = {};
  const bf16x8 zero = {};
  for (int k = 0; k < g.K; k += 16) {
    const bf16x8 a = a_in ? *(const bf16x8*)(ap + k) : zero;
    const bf16x8 bb = b_in ? *(const bf16x8*)(bp + k) : zero;
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bb, acc, 0, 0, 0);
  }
  const int col = n0 + r;
  if (col >= g.N) return;
  // bias has n_valid entries (padding columns produce zeros)
  const float bias = g.bias && col < g.n_valid ? g.bias[col] : 0.f;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int row = m0 + 8 * (j >> 2) + 4 * h + (j & 3);
    if (row >= g.M) continue;
    const float z = acc[j] + bias;
    const float y = g.act ? silu(z) : z;
    if (row < g.m_valid && col < g.n_valid) {
      if (g.Z) g.Z[int64_t(row) * g.ldz + col] = z;
      if (g.Y) g.Y[int64_t(row) * g.ldy + col] = g.y_scale ? y * g.y_scale[col] + g.y_shift[col] : y;
    }
    // BF16 activations keep their zero padding (rows >= m_valid, cols >= n_valid)
    const float yv = (row < g.m_valid && col < g.n_valid) ? y : 0.f;
    if (g.H) g.H[int64_t(row) * g.ldh + col] = __bf16(yv);
    if (g.Ht) g.Ht[int64_t(col) * g.ldht + row] = __bf16(yv);
  }
}

// dZ = dA * SiLU'(Z), written as BF16 row-major and transposed.
__global__ void k_silu_bwd(const float* dA, const float* Z, int rows, int valid_rows, int cols, int ld, __bf16* dZ,
                           __bf16* dZt, int ldt) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(rows) * cols) return;
  const int r = int(i / cols), c = int(i % cols);
  const float v = r < valid_rows ? dA[int64_t(r) * ld + c] * silu_grad(Z[int64_t(r) * ld + c]) : 0.f;
  dZ[int64_t(r) * ld + c] = __bf16(v);
  dZt[int64_t(c) * ldt + r] = __bf16(v);
}

// Mean-squared error over rows x kOut; writes dL/dY as BF16 (row-major,
// width kOutPad, and transposed) and the per-row squared error.
__global__ void k_mse(const float* Y, const float* T, int rows, int valid_rows, float* row_loss, __bf16* dY,
                      __bf16* dYt, int ldt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  float acc = 0.f;
  for (int k = 0; k < kOutPad; ++k) {
    float d = 0.f;
    if (k < kOut && i < valid_rows) {
      const float e = Y[int64_t(i) * kOut + k] - T[int64_t(i) * kOut + k];
      acc += e * e;
      d = 2.0f * e / float(valid_rows * kOut);
    }
    dY[int64_t(i) * kOutPad + k] = __bf16(d);
    dYt[int64_t(k) * ldt + i] = __bf16(d);
  }
  row_loss[i] = i < valid_rows ? acc / float(valid_rows * kOut) : 0.f;
}

// column sums of a row-major BF16 matrix: bias gradient, and (FP32 input)
// the loss.  One thread per column, rows summed in order.
__global__ void k_colsum(const __bf16* A, int rows, int cols, int ld, float* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.f;
  for (int r = 0; r < rows; ++r) s += float(A[int64_t(r) * ld + c]);
  out[c] = s;
}
__global__ void k_sum_f32(const float* a, int n, float* out) {
  __shared__ float part[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) s += a[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

// Adam on the FP32 master parameters, then refresh the BF16 copies of one
// layer's weights: Wnk [N x Kpad] (forward B-operand layout) and Wkn
// [K x Npad] (backward), padding left at zero.
__global__ void k_adam(float* p, const float* grad, float* m, float* v, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gr = grad[i];
  const float mi = b1 * m[i] + (1.f - b1) * gr;
  const float vi = b2 * v[i] + (1.f - b2) * gr * gr;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}
__global__ void k_pack_weights(const float* W, int N, int K, __bf16* Wnk, int kpad, __bf16* Wkn, int npad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N * K) return;
  const int n = i / K, k = i % K;
  const __bf16 w = __bf16(W[i]);
  Wnk[int64_t(n) * kpad + k] = w;
  Wkn[int64_t(k) * npad + n] = w;
}

// ---------------------------------------------------------------------------
// fused forward (inference)
// ---------------------------------------------------------------------------
// One workgroup (4 waves) per block of kFuseRows rows runs all four layers:
// the block's activations stay in LDS (one 128 x 256 BF16 tile, updated in
// place), so HBM sees only the BF16 features in and the 13 outputs out; the
// weights stream from the L2, once per 128 rows.  Each wave computes a
// 128-row x 64-column slab of a hidden layer as 4 x 2 MFMA 32x32x16 tiles
// (eight independent accumulators).  The k order of every dot product is
// that of k_gemm, so the results are the same bits.  LDS element (row, k) sits at
// row * 256 + (((k / 8) ^ (row % 32)) * 8) + k % 8: the 16-byte fragments
// that the 32 rows of a tile read at one k land in different banks.
constexpr int kFuseRows = 128;
constexpr int kRowTiles = kFuseRows / 32;  // 32-row MFMA tiles per block (one per wave in layer 3)
constexpr int kFuseWaves = 4;
constexpr int kColTiles = kHidden / 32 / kFuseWaves;  // 32-column tiles per wave

__device__ __forceinline__ int fswz(int row, int k) { return row * kHidden + ((((k >> 3) ^ (row & 31))) << 3) + (k & 7); }

// One block of kFuseRows rows starting at m0 (of M rows), all four layers.
__device__ __forceinline__ void mlp_block(const MlpArgs& f, int64_t M, int64_t m0, __bf16* H) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 31, h = lane >> 5;
  // The weights are the MFMA's first operand and the rows its second, so
  // acc[t][u][j] is feature 32 (wave kColTiles + u) + 8(j/4) + 4h + j%4 of tile row
  // 32t + r: a lane holds runs of 4 consecutive features of one row, stored
  // to the LDS tile 8 bytes at a time.
  f32x16 acc[kRowTiles][kColTiles];
  // hidden-layer epilogue: bias, SiLU, BF16 into the LDS tile
  auto store_hidden = [&](const float* bias) {
#pragma unroll
    for (int u = 0; u < kColTiles; ++u) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int col = wave * 32 * kColTiles + 32 * u + 8 * q + 4 * h;
        const float4 bv{bias[col], bias[col + 1], bias[col + 2], bias[col + 3]};
#pragma unroll
        for (int t = 0; t < kRowTiles; ++t) {
          const int row = 32 * t + r;
          bf16x4 v;
          v[0] = __bf16(silu(acc[t][u][4 * q + 0] + bv.x));
          v[1] = __bf16(silu(acc[t][u][4 * q + 1] + bv.y));
          v[2] = __bf16(silu(acc[t][u][4 * q + 2] + bv.z));
          v[3] = __bf16(silu(acc[t][u][4 * q + 3] + bv.w));
          *(bf16x4*)(H + fswz(row, col)) = v;
        }
      }
    }
  };
  // layer 0: rows from the features in HBM
#pragma unroll
  for (int t = 0; t < kRowTiles; ++t)
#pragma unroll
    for (int u = 0; u < kColTiles; ++u) acc[t][u] = f32x16{};
#pragma unroll
  for (int k = 0; k < kFeatPad; k += 16) {
    bf16x8 a[kRowTiles], b[kColTiles];
#pragma unroll
    for (int t = 0; t < kRowTiles; ++t) {
      const int64_t row = m0 + 32 * t + r;
      a[t] = row < M ? *(const bf16x8*)(f.X + row * kFeatPad + k + 8 * h) : bf16x8{};
    }
#pragma unroll
    for (int u = 0; u < kColTiles; ++u) b[u] = *(const bf16x8*)(f.W[0] + (wave * 32 * kColTiles + 32 * u + r) * kFeatPad + k + 8 * h);
#pragma unroll
    for (int t = 0; t < kRowTiles; ++t)
#pragma unroll
      for (int u = 0; u < kColTiles; ++u) acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[u], a[t], acc[t][u], 0, 0, 0);
  }
  store_hidden(f.bias[0]);
  __syncthreads();
  // layers 1 and 2: rows from the LDS tile, results back into it
  for (int l = 1; l < kLayers - 1; ++l) {
#pragma unroll
    for (int t = 0; t < kRowTiles; ++t)
#pragma unroll
      for (int u = 0; u < kColTiles; ++u) acc[t][u] = f32x16{};
    const __bf16* W = f.W[l];
#pragma unroll 4
    for (int k = 0; k < kHidden; k += 16) {
      bf16x8 a[kRowTiles], b[kColTiles];
#pragma unroll
      for (int t = 0; t < kRowTiles; ++t) a[t] = *(const bf16x8*)(H + fswz(32 * t + r, k + 8 * h));
#pragma unroll
      for (int u = 0; u < kColTiles; ++u) b[u] = *(const bf16x8*)(W + (wave * 32 * kColTiles + 32 * u + r) * kHidden + k + 8 * h);
#pragma unroll
      for (int t = 0; t < kRowTiles; ++t)
#pragma unroll
        for (int u = 0; u < kColTiles; ++u) acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[u], a[t], acc[t][u], 0, 0, 0);
    }
    __syncthreads();  // every wave has read the tile
    store_hidden(f.bias[l]);
    __syncthreads();
  }
  // layer 3: kOut outputs (one 32-feature tile); wave t takes row tile t;
  // lane (r, h) holds features 8q + 4h .. +3 of row 32*wave + r
  if (wave < kRowTiles) {
    f32x16 y = {};
    const __bf16* W = f.W[kLayers - 1];
#pragma unroll 4
    for (int k = 0; k < kHidden; k += 16) {
      const bf16x8 a = *(const bf16x8*)(H + fswz(32 * wave + r, k + 8 * h));
      const bf16x8 b = *(const bf16x8*)(W + r * kHidden + k + 8 * h);
      y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, a, y, 0, 0, 0);
    }
    const int64_t row = m0 + 32 * wave + r;
    if (row < M) {
      // rows of kOut (odd) floats: scalar stores
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int col = 8 * q + 4 * h + j;
          if (col < kOut)
            f.Y[row * kOut + col] = (y[4 * q + j] + f.bias[kLayers - 1][col]) * f.y_scale[col] + f.y_shift[col];
        }
    }
  }
}

__global__ void __launch_bounds__(64 * kFuseWaves) __attribute__((amdgpu_waves_per_eu(2, 2))) k_mlp_forward(MlpArgs f) {
  __shared__ __align__(16) __bf16 H[kFuseRows * kHidden];
  const int64_t M = rows_of(f.M, f.M_dev);
  // grid-stride over blocks of kFuseRows rows
  for (int64_t m0 = int64_t(blockIdx.x) * kFuseRows; m0 < M; m0 += int64_t(gridDim.x) * kFuseRows) {
    mlp_block(f, M, m0, H);
    __syncthreads();  // every wave is done with the tile before the next block's rows
  }
}

hipError_t launch_mlp_forward(const MlpArgs& f, hipStream_t st) {
  if (f.M <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((f.M + kFuseRows - 1) / kFuseRows, kMlpGridBlocks);
  k_mlp_forward<<<int(blocks), 64 * kFuseWaves, 0, st>>>(f);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// per-body sums of the predictions
// ---------------------------------------------------------------------------
// One workgroup per world.  Each contact's predicted velocity changes (dv0,
// dw0 for its first body, dv1, dw1 for its second) are added to its bodies
// in 32.32 fixed point with LDS integer atomics: integer addition is
// associative, so the sums do not depend on the order the atomics land in
// and a replay gives the same bits.  The predicted position output is the
// penetration the position phase resolves at the contact over all its
// selections; the contact's move for that amount is applyPositionChange's
// (contact.go:286-386: the inertia-weighted split between the bodies'
// linear and angular moves, the angular limit of 0.2 |r x n|-projection),
// summed per body the same way.  Sleeping bodies get no change (the exact
// loop wakes them when a contact needs it).  A world with a non-finite or
// implausibly large prediction, or too large for the surrogate, is marked for
// the plain phases (mask bytes 1); the others are marked for polishing (0).
constexpr float kMaxPredicted = 1.0e4f;  // |change| in m/s or rad/s, or the resolved penetration in m
constexpr double kFix = 4294967296.0;     // 2^32

// applyPositionChange's moves of a contact's two bodies for `amount` of
// penetration resolved (contact.go:286-386), with the bodies' world inverse
// inertia iw and inverse masses im (zeros for an absent second body).
template <typename R>
__device__ __forceinline__ void position_move(const V3<R>& n, const V3<R> rel[2], bool has1, const R (&iw)[2][9],
                                              const R im[2], R amount, V3<R> lin[2], V3<R> ang[2]) {
  R angI[2] = {R(0), R(0)}, linI[2] = {R(0), R(0)};
  R total = R(0);
  for (int i = 0; i < 2; ++i) {
    lin[i] = ang[i] = V3<R>{R(0), R(0), R(0)};
    if (i == 1 && !has1) continue;
    V3<R> aiw = cross(rel[i], n);
    aiw = m3v(iw[i], aiw);
    aiw = cross(aiw, rel[i]);
    angI[i] = dot(aiw, n);
    linI[i] = im[i];
    total += linI[i] + angI[i];
  }
  if (!(total > R(0))) return;  // two bodies of infinite mass: nothing moves
  for (int i = 0; i < 2; ++i) {
    if (i == 1 && !has1) continue;
    const R sign = i == 0 ? R(1) : R(-1);
    R angM = sign * amount * (angI[i] / total);
    R linM = sign * amount * (linI[i] / total);
    const V3<R> projection = add_scaled(rel[i], n, -dot(rel[i], n));
    const R maxMagnitude = R(0.2) * mag(projection);
    if (angM < -maxMagnitude) {
      const R totalMove = angM + linM;
      angM = -maxMagnitude;
      linM = totalMove - angM;
    } else if (angM > maxMagnitude) {
      const R totalMove = angM + linM;
      angM = maxMagnitude;
      linM = totalMove - angM;
    }
    if (angM != R(0)) ang[i] = mul(m3v(iw[i], cross(rel[i], n)), angM / angI[i]);
    lin[i] = mul(n, linM);
  }
}

template <typename R>
__global__ void __launch_bounds__(256) k_sur_scatter(SurApply<R> a) {
  constexpr int NA = 12;  // accumulators per body: velocity 3, rotation 3, position 3, orientation 3
  __shared__ unsigned long long acc[NA][kSurBodies];
  __shared__ int bad;
  const int w = blockIdx.x, t = threadIdx.x;
  const int C = a.world_cnt[w];
  const int b0w = a.body_off[w], n = a.body_off[w + 1] - b0w;
  const bool pos = a.pmask != nullptr;
  if (a.overflow[0] != 0 || C <= 0) {  // nothing to resolve (the resolver's guard)
    if (t == 0) {
      a.mask[w] = 0;
      if (pos) a.pmask[w] = 0;
    }
    return;
  }
  if (n > kSurBodies || C > a.max_contacts) {
    if (t == 0) {
      a.mask[w] = 1;
      if (pos) a.pmask[w] = 1;
    }
    return;
  }
  const int na = pos ? NA : 6;
  for (int i = t; i < na * kSurBodies; i += 256) acc[i / kSurBodies][i % kSurBodies] = 0ull;
  if (t == 0) bad = 0;
  __syncthreads();
  const int64_t base = a.world_coff[w];
  const int64_t S = a.b.stride, CS = a.c.cap;
  const uint8_t* fl = a.b.flags;
  for (int j = t; j < C; j += 256) {
    const int64_t g = base + j;
    const int gb[2] = {a.c.body0[g], a.c.body1[g]};
    const float* y = a.Y + g * kOut;
    bool ok = true;
    for (int k = 0; k < kOut; ++k) ok &= fabsf(y[k]) <= kMaxPredicted;  // false for NaN
    if (!ok) {
      bad = 1;
      continue;
    }
    bool aw[2];
    for (int bb = 0; bb < 2; ++bb) {
      aw[bb] = gb[bb] >= 0 && (fl[gb[bb]] & F_AWAKE);
      if (!aw[bb]) continue;
      const int l = gb[bb] - b0w;
      for (int k = 0; k < 6; ++k)
        atomicAdd(&acc[k][l], (unsigned long long)__double2ll_rn(double(y[6 * bb + k]) * kFix));
    }
    const R amount = R(y[kOutPos]);
    if (!pos || !(amount > R(0)) || !(aw[0] || aw[1])) continue;
    const bool has1 = gb[1] >= 0;
    R iw[2][9] = {}, im[2] = {R(0), R(0)};
    for (int bb = 0; bb < (has1 ? 2 : 1); ++bb) {
      for (int k = 0; k < 9; ++k) iw[bb][k] = a.b.iit_world[k * S + gb[bb]];
      im[bb] = a.b.inv_mass[gb[bb]];
    }
    const V3<R> nrm = ld3c(a.c.normal, CS, g);
    const V3<R> rel[2] = {ld3c(a.c.rel, CS, g), ld3c(a.c.rel + 3 * CS, CS, g)};
    V3<R> lin[2], ang[2];
    position_move(nrm, rel, has1, iw, im, amount, lin, ang);
    for (int bb = 0; bb < 2; ++bb) {
      if (!aw[bb]) continue;
      const int l = gb[bb] - b0w;
      const R v[6] = {lin[bb].x, lin[bb].y, lin[bb].z, ang[bb].x, ang[bb].y, ang[bb].z};
      for (int k = 0; k < 6; ++k)
        atomicAdd(&acc[6 + k][l], (unsigned long long)__double2ll_rn(double(v[k]) * kFix));
    }
  }
  __syncthreads();
  for (int i = t; i < n; i += 256) {
    for (int k = 0; k < 6; ++k) a.delta[k * S + b0w + i] = float(double((long long)acc[k][i]) / kFix);
    if (pos)
      for (int k = 0; k < 6; ++k) a.dpos[k * S + b0w + i] = float(double((long long)acc[6 + k][i]) / kFix);
  }
  if (t == 0) {
    a.mask[w] = bad ? 1 : 0;
    if (pos) a.pmask[w] = (bad || C > kSurPosContacts) ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------
hipError_t launch_gemm(const GemmArgs& g, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  // the fragment loads step K by 16 and read K contiguous elements per row
  if (g.K <= 0 || g.K % 16 != 0 || g.lda < g.K || g.ldb < g.K) return hipErrorInvalidValue;
  if (g.N >= 128) {
    dim3 grid((g.M + 63) / 64, (g.N + 63) / 64);
    k_gemm<2, 2><<<grid, 256, 0, st>>>(g);
  } else {
    dim3 grid((g.M + 127) / 128, (g.N + 31) / 32);
    k_gemm<4, 1><<<grid, 256, 0, st>>>(g);
  }
  return hipGetLastError();
}

template <typename R>
hipError_t launch_features(const Bodies<R>& b, const Contacts<R>& c, int64_t n_cap, const int64_t* n_dev,
                           int32_t* deg, const float* mean, const float* inv_std, __bf16* X, float* raw,
                           hipStream_t st) {
  if (n_cap <= 0) return hipSuccess;
  hipError_t e = launch_fill_i32(deg, b.n, 0, st);
  if (e != hipSuccess) return e;
  const int blocks = int(std::min<int64_t>((n_cap + 255) / 256, kGridStrideBlocks));
  k_degree<<<blocks, 256, 0, st>>>(c.body0, c.body1, n_cap, n_dev, deg);
  k_features<R><<<blocks, 256, 0, st>>>(b, c, n_cap, n_dev, deg, mean, inv_std, X, raw);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_features_world(const Bodies<R>& b, const Contacts<R>& c, const int32_t* world_cnt,
                                 const int64_t* world_coff, const int32_t* body_off, const int32_t* overflow,
                                 int n_worlds, int max_world_bodies, const float* mean, const float* inv_std,
                                 __bf16* X, float* raw, hipStream_t st) {
  if (n_worlds <= 0) return hipSuccess;
  if (max_world_bodies > kFeatWorldBodies) return hipErrorInvalidValue;
  k_features_world<R><<<n_worlds, 256, 0, st>>>(b, c, world_cnt, world_coff, body_off, overflow, mean, inv_std, X,
                                                raw);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_sur_scatter(const SurApply<R>& a, int n_worlds, hipStream_t st) {
  if (n_worlds <= 0) return hipSuccess;
  k_sur_scatter<R><<<n_worlds, 256, 0, st>>>(a);
  return hipGetLastError();
}

hipError_t launch_gather_batch(const float* Xraw, const float* Yraw, const int64_t* idx, int rows, int valid_rows,
                               const float* in_mean, const float* in_inv_std, const float* out_mean,
                               const float* out_inv_std, __bf16* X, __bf16* Xt, float* T, hipStream_t st) {
  k_gather_batch<<<(rows + 255) / 256, 256, 0, st>>>(Xraw, Yraw, idx, rows, valid_rows, in_mean, in_inv_std,
                                                   out_mean, out_inv_std, X, Xt, T);
  return hipGetLastError();
}
hipError_t launch_sample_idx(int64_t* idx, int n, uint64_t seed, uint64_t step, int64_t rows, hipStream_t st) {
  if (n <= 0 || rows <= 0) return hipSuccess;
  k_sample_idx<<<(n + 255) / 256, 256, 0, st>>>(idx, n, seed, step, rows);
  return hipGetLastError();
}
hipError_t launch_col_stats(const float* A, int64_t rows, int cols, float* mean, float* stdv, hipStream_t st) {
  k_col_stats<<<cols, 256, 0, st>>>(A, rows, cols, mean, stdv);
  return hipGetLastError();
}
hipError_t launch_count_u8(const uint8_t* m, int n, int32_t* out, hipStream_t st, const uint8_t* m2) {
  k_count_u8<<<1, 256, 0, st>>>(m, m2, n, out);
  return hipGetLastError();
}
hipError_t launch_record_count(const int32_t* value, int32_t* hist, int32_t* ctr, int64_t cap, hipStream_t st) {
  k_record_count<<<1, 64, 0, st>>>(value, hist, ctr, cap);
  return hipGetLastError();
}
hipError_t launch_fill_f32(float* p, int64_t n, float v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_fill_f32<<<int((n + 255) / 256), 256, 0, st>>>(p, n, v);
  return hipGetLastError();
}
hipError_t launch_silu_bwd(const float* dA, const float* Z, int rows, int valid_rows, int cols, int ld, __bf16* dZ,
                           __bf16* dZt, int ldt, hipStream_t st) {
  const int64_t n = int64_t(rows) * cols;
  k_silu_bwd<<<int((n + 255) / 256), 256, 0, st>>>(dA, Z, rows, valid_rows, cols, ld, dZ, dZt, ldt);
  return hipGetLastError();
}
hipError_t launch_mse(const float* Y, const float* T, int rows, int valid_rows, float* row_loss, __bf16* dY,
                      __bf16* dYt, int ldt, hipStream_t st) {
  k_mse<<<(rows + 255) / 256, 256, 0, st>>>(Y, T, rows, valid_rows, row_loss, dY, dYt, ldt);
  return hipGetLastError();
}
hipError_t launch_colsum(const __bf16* A, int rows, int cols, int ld, float* out, hipStream_t st) {
  k_colsum<<<(cols + 63) / 64, 64, 0, st>>>(A, rows, cols, ld, out);
  return hipGetLastError();
}
hipError_t launch_sum_f32(const float* a, int n, float* out, hipStream_t st) {
  k_sum_f32<<<1, 256, 0, st>>>(a, n, out);
  return hipGetLastError();
}
hipError_t launch_adam(float* p, const float* grad, float* m, float* v, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2, hipStream_t st) {
  k_adam<<<int((n + 255) / 256), 256, 0, st>>>(p, grad, m, v, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}
hipError_t launch_pack_weights(const float* W, int N, int K, __bf16* Wnk, int kpad, __bf16* Wkn, int npad,
                               hipStream_t st) {
  k_pack_weights<<<(N * K + 255) / 256, 256, 0, st>>>(W, N, K, Wnk, kpad, Wkn, npad);
  return hipGetLastError();
}

template hipError_t launch_features<double>(const Bodies<double>&, const Contacts<double>&, int64_t, const int64_t*,
                                            int32_t*,
                                            const float*, const float*, __bf16*, float*, hipStream_t);
template hipError_t launch_features<float>(const Bodies<float>&, const Contacts<float>&, int64_t, const int64_t*,
                                           int32_t*,
                                           const float*, const float*, __bf16*, float*, hipStream_t);
template hipError_t launch_features_world<double>(const Bodies<double>&, const Contacts<double>&, const int32_t*,
                                                  const int64_t*, const int32_t*, const int32_t*, int, int,
                                                  const float*, const float*, __bf16*, float*, hipStream_t);
template hipError_t launch_features_world<float>(const Bodies<float>&, const Contacts<float>&, const int32_t*,
                                                 const int64_t*, const int32_t*, const int32_t*, int, int,
                                                 const float*, const float*, __bf16*, float*, hipStream_t);
template hipError_t launch_sur_scatter<double>(const SurApply<double>&, int, hipStream_t);
template hipError_t launch_sur_scatter<float>(const SurApply<float>&, int, hipStream_t);

}  // namespace cbz
