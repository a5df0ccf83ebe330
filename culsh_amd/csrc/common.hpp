// Shared device/host helpers for libculsh (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace culsh {

constexpr int kWave = 64;

// 64-bit finaliser of SplitMix64 (Steele, Lea, Flood, "Fast splittable
// pseudorandom number generators", OOPSLA'14).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Hash term of dimension m holding code c; a bucket key is the sum over m of
// these terms (mod 2^64), so a key can be updated one dimension at a time.
__host__ __device__ inline uint64_t code_term(uint64_t table_seed, int m, int c) {
  return mix64(table_seed ^ ((uint64_t)(uint32_t)m << 32) ^ (uint64_t)(uint32_t)c);
}

// Bucket key layout: the top `tb` bits hold the table index, the rest the
// top bits of the hash sum.  Sorting all tables' keys at once therefore
// groups them table by table.
__host__ __device__ inline uint64_t make_key(uint64_t hsum, int t, int tb) {
  return ((uint64_t)t << (64 - tb)) | (hsum >> tb);
}

__device__ inline uint64_t shfl_u64(uint64_t v, int src) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  lo = __shfl(lo, src);
  hi = __shfl(hi, src);
  return ((uint64_t)hi << 32) | lo;
}

__device__ inline uint64_t shfl_xor_u64(uint64_t v, int m) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  lo = __shfl_xor(lo, m);
  hi = __shfl_xor(hi, m);
  return ((uint64_t)hi << 32) | lo;
}

__device__ inline double shfl_f64(double v, int src) {
  return __longlong_as_double((long long)shfl_u64((uint64_t)__double_as_longlong(v), src));
}

__device__ inline double shfl_xor_f64(double v, int m) {
  return __longlong_as_double((long long)shfl_xor_u64((uint64_t)__double_as_longlong(v), m));
}

__device__ inline uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += shfl_xor_u64(v, off);
  return v;
}

// Orders LDS accesses of the lanes of one wave (no workgroup barrier).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Lower bound of `key` in the sorted unique key array, through the prefix
// directory: dir[v] = first bucket whose key >> shift >= v.  Returns the
// bucket index, or -1 when the key is absent.
__device__ inline long long lookup_key(uint64_t key, const uint32_t* __restrict__ dir, int shift,
                                       const uint64_t* __restrict__ ukeys) {
  uint64_t v = key >> shift;
  uint32_t lo = dir[v], hi = dir[v + 1];
  const uint32_t end = hi;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (ukeys[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < end && ukeys[lo] == key) ? (long long)lo : -1;
}

}  // namespace culsh
