// k_sort.hip — stable LSD radix sort of 64-bit keys for wave64: eight passes
// of 8-bit digits, each three kernels:
//   k_radix_hist     per tile of 4096 keys, the digit histogram, written
//                    digit-major ([digit][tile]) so that one exclusive scan
//                    gives every (digit, tile) its output offset;
//   k_radix_scan     that exclusive scan, one workgroup;
//   k_radix_scatter  per tile, each key's rank among the tile's earlier keys
//                    of its digit.  A tile is four waves, each a contiguous
//                    run of 16 rounds of 64 keys; within a round the lanes
//                    with the same digit are found with eight ballots, so a
//                    key's rank is the popcount of its group below its lane
//                    plus the running count of the digit in its wave, and
//                    a wave starts at its tile's offset plus the earlier
//                    waves' counts.  Equal digits keep the input order, so
//                    the sort is stable.
#include "cbz_sort.h"
#include "cbz_wave.h"

namespace cbz {

constexpr int kSortWaves = 4, kSortThreads = 64 * kSortWaves, kSortRounds = 16;
constexpr int64_t kSortTile = int64_t(kSortThreads) * kSortRounds;  // 4096 keys
constexpr int kRadix = 256;
static_assert(kSortThreads == kRadix, "one thread per digit in the per-tile steps");

// The lanes of the wave whose digit equals this lane's (valid lanes only).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bb = __ballot(valid && bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

__global__ void __launch_bounds__(kSortThreads)
    k_radix_hist(const uint64_t* keys, int64_t n, int shift, uint32_t* hist, int32_t tiles) {
  __shared__ uint32_t s_cnt[kRadix];
  const int t = threadIdx.x;
  s_cnt[t] = 0;
  __syncthreads();
  const int64_t base = int64_t(blockIdx.x) * kSortTile;
  for (int r = 0; r < kSortRounds; ++r) {
    const int64_t i = base + int64_t(r) * kSortThreads + t;
    if (i < n) atomicAdd(&s_cnt[uint32_t(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[int64_t(t) * tiles + blockIdx.x] = s_cnt[t];
}

// Exclusive scan of hist[0..len) in place, one workgroup of 1024 threads,
// thread t over a contiguous chunk.
__global__ void __launch_bounds__(1024) k_radix_scan(uint32_t* hist, int64_t len) {
  __shared__ int32_t s_wave[16];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int64_t per = (len + 1023) / 1024;
  const int64_t b0 = int64_t(t) * per, b1 = b0 + per < len ? b0 + per : len;
  uint32_t sum = 0;
  for (int64_t i = b0; i < b1; ++i) sum += hist[i];
  const int inc = wave_incl_scan_i32(int(sum));
  if (lane == 63) s_wave[wv] = inc;
  __syncthreads();
  uint32_t pre = uint32_t(inc) - sum;
  for (int v = 0; v < wv; ++v) pre += uint32_t(s_wave[v]);
  for (int64_t i = b0; i < b1; ++i) {
    const uint32_t h = hist[i];
    hist[i] = pre;
    pre += h;
  }
}

template <bool VALS>
__global__ void __launch_bounds__(kSortThreads)
    k_radix_scatter(const uint64_t* kin, uint64_t* kout, const int32_t* vin, int32_t* vout, int64_t n, int shift,
                    const uint32_t* off, int32_t tiles) {
  __shared__ uint32_t s_cnt[kSortWaves][kRadix];  // each wave's digit counts
  __shared__ uint32_t s_run[kSortWaves][kRadix];  // each wave's next position per digit
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int v = 0; v < kSortWaves; ++v) s_cnt[v][t] = 0;
  __syncthreads();
  const int64_t wbase = int64_t(blockIdx.x) * kSortTile + int64_t(wv) * kSortRounds * 64;
  const uint64_t below = (uint64_t(1) << lane) - 1;
  for (int r = 0; r < kSortRounds; ++r) {
    const int64_t i = wbase + int64_t(r) * 64 + lane;
    const bool valid = i < n;
    const uint32_t d = valid ? uint32_t(kin[i] >> shift) & 255u : 0u;
    const uint64_t m = match_digit(d, valid);
    if (valid && (m & below) == 0) s_cnt[wv][d] += uint32_t(__popcll(m));
  }
  __syncthreads();
  {
    uint32_t p = off[int64_t(t) * tiles + blockIdx.x];
    for (int v = 0; v < kSortWaves; ++v) {
      s_run[v][t] = p;
      p += s_cnt[v][t];
    }
  }
  __syncthreads();
  for (int r = 0; r < kSortRounds; ++r) {
    const int64_t i = wbase + int64_t(r) * 64 + lane;
    const bool valid = i < n;
    const uint64_t k = valid ? kin[i] : 0;
    const uint32_t d = uint32_t(k >> shift) & 255u;
    const uint64_t m = match_digit(d, valid);
    // every lane of the group reads the running position before its
    // lowest lane advances it (one wave: the read precedes the write)
    const uint32_t pos = s_run[wv][d] + uint32_t(__popcll(m & below));
    if (valid) {
      kout[pos] = k;
      if constexpr (VALS) vout[pos] = vin[i];
    }
    if (valid && (m & below) == 0) s_run[wv][d] += uint32_t(__popcll(m));
  }
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

size_t radix_sort_bytes(int64_t n, bool with_values) {
  if (n <= 0) return 256;
  const int64_t tiles = (n + kSortTile - 1) / kSortTile;
  size_t b = align256(size_t(n) * sizeof(uint64_t));
  if (with_values) b += align256(size_t(n) * sizeof(int32_t));
  return b + align256(size_t(kRadix) * size_t(tiles) * sizeof(uint32_t));
}

hipError_t radix_sort_u64(const uint64_t* keys_in, uint64_t* keys_out, const int32_t* vals_in, int32_t* vals_out,
                          int64_t n, void* tmp, size_t tmp_bytes, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const bool vals = vals_in != nullptr;
  if (n >= (int64_t(1) << 31) || (vals != (vals_out != nullptr)) || !tmp ||
      tmp_bytes < radix_sort_bytes(n, vals))
    return hipErrorInvalidValue;
  const int32_t tiles = int32_t((n + kSortTile - 1) / kSortTile);
  char* p = (char*)tmp;
  uint64_t* ktmp = (uint64_t*)p;
  p += align256(size_t(n) * sizeof(uint64_t));
  int32_t* vtmp = nullptr;
  if (vals) {
    vtmp = (int32_t*)p;
    p += align256(size_t(n) * sizeof(int32_t));
  }
  uint32_t* hist = (uint32_t*)p;
  // eight passes: in -> tmp, then tmp -> out and out -> tmp alternately,
  // so that the last (odd) pass writes out
  for (int pass = 0; pass < 8; ++pass) {
    const int shift = 8 * pass;
    const uint64_t* ks = pass == 0 ? keys_in : (pass & 1) ? ktmp : keys_out;
    uint64_t* kd = (pass & 1) ? keys_out : ktmp;
    const int32_t* vs = pass == 0 ? vals_in : (pass & 1) ? vtmp : vals_out;
    int32_t* vd = (pass & 1) ? vals_out : vtmp;
    k_radix_hist<<<tiles, kSortThreads, 0, st>>>(ks, n, shift, hist, tiles);
    k_radix_scan<<<1, 1024, 0, st>>>(hist, int64_t(kRadix) * tiles);
    if (vals)
      k_radix_scatter<true><<<tiles, kSortThreads, 0, st>>>(ks, kd, vs, vd, n, shift, hist, tiles);
    else
      k_radix_scatter<false><<<tiles, kSortThreads, 0, st>>>(ks, kd, nullptr, nullptr, n, shift, hist, tiles);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace cbz
