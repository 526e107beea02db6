// cbz_broad.h — the sort-and-sweep broadphase (k_broad.hip) and its buffers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cbz {

template <typename R>
struct BroadArgs {
  int32_t nb;
  int64_t stride;
  const R* ctransform;       // [12][stride]: collider centre in planes 9..11
  const double* brad;        // [nb] bounding radius
  const int32_t* body_world; // [nb]
  const int32_t* body_off;   // [nw + 1]
  const int32_t* plane_off;  // [nw + 1]
  double* lo;                // [2][nb] lower y, z bounds
  double* hi;                // [2][nb] upper y, z bounds
  float* hix;                // [nb] upper x bound, rounded up
  uint8_t* wild;             // [nb] centre or radius not finite
  uint64_t* key_in;          // [nb] world << 32 | ordered lower x bound (rounded down)
  uint64_t* key_out;         // [nb] sorted
  int32_t* val_in;           // [nb] body index
  int32_t* val_out;          // [nb] sorted
  uint64_t* pkey;            // [pcap] pair keys, body << 32 | slot in its row
  int64_t pcap;
  int64_t plane_pairs;       // body-plane pairs, written first
  int64_t* pcount;           // body-pair keys appended (may exceed pcap)
};

struct BroadPlanes {
  int32_t nb;
  const int32_t* body_world;
  const int32_t* body_off;
  const int32_t* plane_off;
  const int64_t* plane_pair_base;  // [nw] first body-plane key of each world
  uint64_t* pkey;
};

struct BroadList {
  int64_t n;                 // pair keys
  int32_t nw;
  const uint64_t* pkey_in;   // unsorted keys
  uint64_t* pkey;            // sorted keys
  const int32_t* body_world;
  const int32_t* body_off;
  const int32_t* plane_off;
  int32_t* one;              // [n] explicit pair list
  int32_t* two;
  int32_t* list_off;         // [nw + 1] first pair of each world
};

template <typename R>
hipError_t launch_bp_bounds_sort(const BroadArgs<R>& a, void* tmp, size_t tmp_bytes, hipStream_t st);
template <typename R>
hipError_t launch_bp_sweep(const BroadArgs<R>& a, const BroadPlanes& p, hipStream_t st);
hipError_t launch_bp_list(const BroadList& L, void* tmp, size_t tmp_bytes, hipStream_t st);
hipError_t bp_sort_bytes(int32_t n, int64_t m, size_t* bytes);

}  // namespace cbz
