// cbz_sort.h — the device sort of the broadphase (k_sort.hip): a stable
// least-significant-digit radix sort of 64-bit keys, with optional 32-bit
// values carried along.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cbz {

// Bytes of temporary storage radix_sort_u64 needs for n keys.
size_t radix_sort_bytes(int64_t n, bool with_values);

// keys_out[0..n) = keys_in sorted ascending, equal keys in input order;
// vals_out the values in the same permutation (both value pointers null:
// keys only).  keys_in/vals_in are only read.  n < 2^31.
hipError_t radix_sort_u64(const uint64_t* keys_in, uint64_t* keys_out, const int32_t* vals_in, int32_t* vals_out,
                          int64_t n, void* tmp, size_t tmp_bytes, hipStream_t st);

}  // namespace cbz
