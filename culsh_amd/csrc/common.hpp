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

// ---- DPP cross-lane moves inside groups of 8 lanes (gfx9 DPP controls) ----
// quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror (i <-> 7-i).
constexpr int kDppXor1 = 0xB1;
constexpr int kDppXor2 = 0x4E;
constexpr int kDppHalfMirror = 0x141;

// Every source lane of these patterns is inside the octet, so no lane takes
// the bound_ctrl zero and the result needs no "old" operand (no copy).
template <int CTRL>
__device__ inline uint32_t dpp_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}

// Reductions over the 8 lanes of an octet; every lane gets the result.  All
// 64 lanes of the wave must execute them.
__device__ inline uint64_t octet_min_u64(uint64_t v) {
  uint32_t hi = (uint32_t)(v >> 32), lo = (uint32_t)v;
#define CULSH_OCT_MIN64(C)                                \
  {                                                       \
    const uint32_t h2 = dpp_u32<C>(hi), l2 = dpp_u32<C>(lo); \
    const bool lt = (h2 < hi) || (h2 == hi && l2 < lo);   \
    hi = lt ? h2 : hi;                                    \
    lo = lt ? l2 : lo;                                    \
  }
  CULSH_OCT_MIN64(kDppXor1)
  CULSH_OCT_MIN64(kDppXor2)
  CULSH_OCT_MIN64(kDppHalfMirror)
#undef CULSH_OCT_MIN64
  return ((uint64_t)hi << 32) | lo;
}

// Lexicographic minimum of (v, slot) over the octet, carrying payload from
// the winning lane; slots must be distinct within the octet.
__device__ inline void octet_argmin_u64(uint64_t& v, uint32_t& slot, uint64_t& payload) {
  uint32_t hi = (uint32_t)(v >> 32), lo = (uint32_t)v, ph = (uint32_t)(payload >> 32),
           pl = (uint32_t)payload;
#define CULSH_OCT_ARGMIN(C)                                                              \
  {                                                                                      \
    const uint32_t h2 = dpp_u32<C>(hi), l2 = dpp_u32<C>(lo), s2 = dpp_u32<C>(slot);      \
    const uint32_t p2h = dpp_u32<C>(ph), p2l = dpp_u32<C>(pl);                           \
    const uint64_t a = ((uint64_t)hi << 32) | lo, b = ((uint64_t)h2 << 32) | l2;         \
    const bool lt = b < a || (b == a && s2 < slot);                                      \
    hi = lt ? h2 : hi;                                                                   \
    lo = lt ? l2 : lo;                                                                   \
    slot = lt ? s2 : slot;                                                               \
    ph = lt ? p2h : ph;                                                                  \
    pl = lt ? p2l : pl;                                                                  \
  }
  CULSH_OCT_ARGMIN(kDppXor1)
  CULSH_OCT_ARGMIN(kDppXor2)
  CULSH_OCT_ARGMIN(kDppHalfMirror)
#undef CULSH_OCT_ARGMIN
  v = ((uint64_t)hi << 32) | lo;
  payload = ((uint64_t)ph << 32) | pl;
}

__device__ inline void octet_argmin_u64(uint64_t& v, uint32_t& slot, uint32_t& payload) {
  uint32_t hi = (uint32_t)(v >> 32), lo = (uint32_t)v;
#define CULSH_OCT_ARGMIN(C)                                                              \
  {                                                                                      \
    const uint32_t h2 = dpp_u32<C>(hi), l2 = dpp_u32<C>(lo), s2 = dpp_u32<C>(slot);      \
    const uint32_t p2 = dpp_u32<C>(payload);                                             \
    const uint64_t a = ((uint64_t)hi << 32) | lo, b = ((uint64_t)h2 << 32) | l2;         \
    const bool lt = b < a || (b == a && s2 < slot);                                      \
    hi = lt ? h2 : hi;                                                                   \
    lo = lt ? l2 : lo;                                                                   \
    slot = lt ? s2 : slot;                                                               \
    payload = lt ? p2 : payload;                                                         \
  }
  CULSH_OCT_ARGMIN(kDppXor1)
  CULSH_OCT_ARGMIN(kDppXor2)
  CULSH_OCT_ARGMIN(kDppHalfMirror)
#undef CULSH_OCT_ARGMIN
  v = ((uint64_t)hi << 32) | lo;
}

__device__ inline uint32_t octet_min_u32(uint32_t v) {
  v = min(v, dpp_u32<kDppXor1>(v));
  v = min(v, dpp_u32<kDppXor2>(v));
  v = min(v, dpp_u32<kDppHalfMirror>(v));
  return v;
}

__device__ inline uint64_t octet_or_u64(uint64_t v) {
  uint32_t hi = (uint32_t)(v >> 32), lo = (uint32_t)v;
  hi |= dpp_u32<kDppXor1>(hi);
  lo |= dpp_u32<kDppXor1>(lo);
  hi |= dpp_u32<kDppXor2>(hi);
  lo |= dpp_u32<kDppXor2>(lo);
  hi |= dpp_u32<kDppHalfMirror>(hi);
  lo |= dpp_u32<kDppHalfMirror>(lo);
  return ((uint64_t)hi << 32) | lo;
}

__device__ inline uint64_t octet_sum_u64(uint64_t v) {
  uint32_t hi = (uint32_t)(v >> 32), lo = (uint32_t)v;
#define CULSH_OCT_ADD64(C)                                  \
  {                                                         \
    const uint32_t h2 = dpp_u32<C>(hi), l2 = dpp_u32<C>(lo); \
    const uint64_t s = (((uint64_t)hi << 32) | lo) + (((uint64_t)h2 << 32) | l2); \
    hi = (uint32_t)(s >> 32);                               \
    lo = (uint32_t)s;                                       \
  }
  CULSH_OCT_ADD64(kDppXor1)
  CULSH_OCT_ADD64(kDppXor2)
  CULSH_OCT_ADD64(kDppHalfMirror)
#undef CULSH_OCT_ADD64
  return ((uint64_t)hi << 32) | lo;
}

// ---- Bucket records and the probe index ----
// One record per bucket: its key, the start of its members in the id array
// and its size.  Records are sorted by key (table by table).
struct __align__(16) BucketRec {
  uint64_t key;
  uint32_t start;
  uint32_t size;
};

// Probe index: the structures a multiprobe lookup reads.  Most probes ask
// for a bucket that does not exist, so an absent key must be rejected from
// cache, and a present one resolved with as few HBM lines as possible.
//  * Key filter: a register-blocked Bloom filter per table, one 64-bit word
//    per key (word mulhi(key << tb, nw_t) of the table's nw_t =
//    ceil(B_t * kFilterBits / 64) words, kFilterHashes bits chosen by a
//    second hash of the key).  At kFilterBits bits per bucket it holds the
//    table being probed in the L2 and passes about 1% of absent keys.
//  * Key directory: slot v covers keys with key >> shift == v, 2^P slots per
//    table with 2^P >= 2 * (largest table's bucket count), so most slots
//    hold at most one key.  Packed u64 entry:
//      bits  0-31  index of the slot's first record
//      bits 32-55  low 24 bits of that record's key (fingerprint)
//      bits 56-63  number of records in the slot, saturated at 255
//  * Bucket lines: one line of lstride (a multiple of 64) bytes per record,
//    u32 start, u32 size, i16 codes[M]: a directory hit is confirmed by
//    comparing the probe's full code with the line's and resolved from the
//    same line.
constexpr int kFilterBits = 10;
constexpr int kFilterHashes = 5;
constexpr uint64_t kFpMask = 0xFFFFFFull;

struct ProbeIndex {
  const uint64_t* filter;      // key filter words of all tables
  const uint2* tfw;            // [L]: first filter word, word count of table t
  const uint64_t* dir;         // [nslots + 1] key directory
  const unsigned char* lines;  // [B][lstride]: bucket lines in record order
  const BucketRec* recs;       // key-sorted records (multi-key slots)
  int tb;                      // table bits of a key
  int shift;                   // directory slot = key >> shift
  int lstride;                 // bytes per bucket line
};

__host__ __device__ inline int bucket_line_bytes(int M) { return (8 + 2 * M + 63) / 64 * 64; }

__device__ __host__ inline uint64_t dir_entry(uint32_t first, uint64_t first_key, uint32_t count) {
  const uint64_t c = count > 255u ? 255u : count;
  return (uint64_t)first | ((first_key & kFpMask) << 32) | (c << 56);
}

// Filter word of `key` in its table (tfw = the table's first word and count).
__device__ inline uint32_t filter_word(uint64_t key, int tb, uint2 tfw) {
  return tfw.x + (uint32_t)__umul64hi(key << tb, (uint64_t)tfw.y);
}

// The kFilterHashes bits of `key` inside its filter word.
__device__ inline uint64_t filter_bits(uint64_t key) {
  const uint64_t h = key * 0x9E3779B97F4A7C15ull;
  uint64_t m = 0;
#pragma unroll
  for (int i = 0; i < kFilterHashes; ++i) m |= 1ull << ((h >> (58 - 6 * i)) & 63);
  return m;
}

// Record index of `key`, or -1, given its directory entry e = dir[key >> shift].
// Most absent keys are rejected by the entry alone (empty slot or
// fingerprint mismatch).
__device__ inline long long lookup_resolve(uint64_t key, uint64_t e,
                                           const uint64_t* __restrict__ dir, int shift,
                                           const BucketRec* __restrict__ recs) {
  const uint64_t v = key >> shift;
  const uint32_t cnt = (uint32_t)(e >> 56);
  if (cnt == 0) return -1;
  uint32_t lo = (uint32_t)e;
  if (cnt == 1) {
    if (((e >> 32) & kFpMask) != (key & kFpMask)) return -1;
    return recs[lo].key == key ? (long long)lo : -1;
  }
  uint32_t hi = cnt < 255 ? lo + cnt : (uint32_t)dir[v + 1];
  const uint32_t end = hi;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (recs[mid].key < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < end && recs[lo].key == key) ? (long long)lo : -1;
}

// As lookup_resolve, for a caller that then verifies the bucket's full code:
// a single-entry slot whose fingerprint matches needs no key load (the
// slot's keys are all of this table, and equal codes of one table have equal
// keys).
__device__ inline long long lookup_unverified(uint64_t key, uint64_t e,
                                              const uint64_t* __restrict__ dir, int shift,
                                              const BucketRec* __restrict__ recs) {
  const uint32_t cnt = (uint32_t)(e >> 56);
  if (cnt == 0) return -1;
  if (cnt == 1) return ((e >> 32) & kFpMask) == (key & kFpMask) ? (long long)(uint32_t)e : -1;
  return lookup_resolve(key, e, dir, shift, recs);
}

__device__ inline const unsigned char* bucket_line(const ProbeIndex& P, uint32_t g) {
  return P.lines + (size_t)g * (size_t)P.lstride;
}
__device__ inline const int16_t* line_codes(const unsigned char* line) {
  return reinterpret_cast<const int16_t*>(line + 8);
}
__device__ inline uint32_t line_size(const unsigned char* line) {
  return reinterpret_cast<const uint32_t*>(line)[1];
}
// Member range of a bucket line: (start, size).
__device__ inline uint2 line_range(const unsigned char* line) {
  return *reinterpret_cast<const uint2*>(line);
}

}  // namespace culsh
