// Device-wide exclusive scan and stable LSD radix sort of (u64 key, u32 value)
// pairs, for gfx950 wave64.
//
// Radix sort: 8-bit digits, reduce-then-scan per pass:
//   k_rs_hist     per-tile digit histogram -> hist[digit][tile]
//   scan          exclusive scan of hist in digit-major order -> offsets
//   k_rs_scatter  per-tile stable ranking (wave ballot multi-split, Merrill &
//                 Garland style); the tile is put in digit order in LDS and
//                 each digit's run is written out contiguously
// Stability: inside a tile, element order is (wave, item, lane); a wave's
// running per-digit counters rank its items in that order, waves are
// prefixed in index order and tiles by the scanned histogram.
#pragma once
#include "common.hpp"

namespace culsh {

constexpr int kScanTile = 2048;  // 256 threads x 8
constexpr int kRsItems = 16;
constexpr int kRsTile = 256 * kRsItems;

// Exclusive scan across the 256 threads of a block; `total` gets the sum.
__device__ inline uint32_t block_exclusive_scan_256(uint32_t v, uint32_t* smem4, uint32_t& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) smem4[w] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint32_t s = smem4[i];
    if (i < w) pre += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return pre + x - v;
}

// Exclusive scan of one 2048-element tile; tile totals go to sums (if any).
// In-place (in == out) is allowed.
__global__ __launch_bounds__(256) void k_scan_tiles(const uint32_t* in, uint32_t* out, long long n,
                                                    uint32_t* sums) {
  __shared__ uint32_t sm[4];
  __shared__ uint32_t tile[kScanTile];
  const long long base = (long long)blockIdx.x * kScanTile;
  for (int i = threadIdx.x; i < kScanTile; i += 256) {
    long long idx = base + i;
    tile[i] = idx < n ? in[idx] : 0u;
  }
  __syncthreads();
  uint32_t v[8], s = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    uint32_t t = tile[threadIdx.x * 8 + i];
    v[i] = s;
    s += t;
  }
  uint32_t tot;
  uint32_t pre = block_exclusive_scan_256(s, sm, tot);
#pragma unroll
  for (int i = 0; i < 8; ++i) tile[threadIdx.x * 8 + i] = pre + v[i];
  __syncthreads();
  for (int i = threadIdx.x; i < kScanTile; i += 256) {
    long long idx = base + i;
    if (idx < n) out[idx] = tile[i];
  }
  if (threadIdx.x == 0 && sums) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t* out, long long n,
                                                  const uint32_t* tile_offsets) {
  const uint32_t add = tile_offsets[blockIdx.x];
  const long long base = (long long)blockIdx.x * kScanTile;
  for (int i = threadIdx.x; i < kScanTile; i += 256) {
    long long idx = base + i;
    if (idx < n) out[idx] += add;
  }
}

// Scratch (in u32 elements) needed by scan_exclusive for n elements.
inline long long scan_scratch_elems(long long n) {
  long long total = 0;
  while (n > kScanTile) {
    n = (n + kScanTile - 1) / kScanTile;
    total += n;
  }
  return total;
}

// Exclusive scan of n u32 values (in-place allowed).  scratch must hold
// scan_scratch_elems(n) elements.
inline void scan_exclusive(const uint32_t* in, uint32_t* out, long long n, uint32_t* scratch,
                           hipStream_t s) {
  if (n <= 0) return;
  long long nb = (n + kScanTile - 1) / kScanTile;
  if (nb == 1) {
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(256), 0, s, in, out, n, (uint32_t*)nullptr);
    return;
  }
  uint32_t* sums = scratch;
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nb), dim3(256), 0, s, in, out, n, sums);
  scan_exclusive(sums, sums, nb, scratch + nb, s);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(256), 0, s, out, n, (const uint32_t*)sums);
}

__global__ __launch_bounds__(256) void k_rs_hist(const uint64_t* __restrict__ keys, long long n,
                                                 int shift, uint32_t* __restrict__ hist,
                                                 long long ntiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const long long base = (long long)blockIdx.x * kRsTile;
#pragma unroll
  for (int i = 0; i < kRsItems; ++i) {
    long long idx = base + (long long)i * 256 + threadIdx.x;
    if (idx < n) atomicAdd(&h[(uint32_t)(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(long long)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_rs_scatter(const uint64_t* __restrict__ kin,
                                                    const uint32_t* __restrict__ vin,
                                                    uint64_t* __restrict__ kout,
                                                    uint32_t* __restrict__ vout, long long n,
                                                    int shift, const uint32_t* __restrict__ offs,
                                                    long long ntiles) {
  __shared__ uint32_t whist[4][256];
  __shared__ uint32_t s_lstart[256], s_gstart[256], s_sm4[4];
  __shared__ uint64_t s_k[kRsTile];
  __shared__ uint32_t s_v[kRsTile];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < 1024; i += 256) (&whist[0][0])[i] = 0;
  __syncthreads();
  const long long tile0 = (long long)blockIdx.x * kRsTile;
  const long long base = tile0 + (long long)w * 64 * kRsItems;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  uint64_t k[kRsItems];
  uint32_t v[kRsItems], rank[kRsItems];
#pragma unroll
  for (int it = 0; it < kRsItems; ++it) {
    const long long idx = base + it * 64 + lane;
    const bool valid = idx < n;
    k[it] = valid ? kin[idx] : 0ull;
    v[it] = valid ? vin[idx] : 0u;
  }
  // Rank of each item among the wave's items of its digit, in input order.
#pragma unroll
  for (int it = 0; it < kRsItems; ++it) {
    const long long idx = base + it * 64 + lane;
    const bool valid = idx < n;
    const uint32_t dg = (uint32_t)(k[it] >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (dg >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t before = whist[w][dg];
    rank[it] = before + (uint32_t)__popcll(peers & lt);
    wave_lds_sync();
    if (valid && lane == __ffsll((unsigned long long)peers) - 1)
      whist[w][dg] = before + (uint32_t)__popcll(peers);
    wave_lds_sync();
  }
  __syncthreads();
  // Per digit: offsets of the waves' runs, the digit's start in the tile and
  // its start in the output.
  {
    const int d = threadIdx.x;
    uint32_t run = 0;
#pragma unroll
    for (int ww = 0; ww < 4; ++ww) {
      const uint32_t c = whist[ww][d];
      whist[ww][d] = run;
      run += c;
    }
    uint32_t total;
    s_lstart[d] = block_exclusive_scan_256(run, s_sm4, total);
    s_gstart[d] = offs[(long long)d * ntiles + blockIdx.x];
  }
  __syncthreads();
  // The tile in digit order in LDS (stable), then written out as one
  // contiguous run per digit.
#pragma unroll
  for (int it = 0; it < kRsItems; ++it) {
    const long long idx = base + it * 64 + lane;
    if (idx < n) {
      const uint32_t dg = (uint32_t)(k[it] >> shift) & 255u;
      const uint32_t loc = s_lstart[dg] + whist[w][dg] + rank[it];
      s_k[loc] = k[it];
      s_v[loc] = v[it];
    }
  }
  __syncthreads();
  const int cnt = (int)(n - tile0 < kRsTile ? n - tile0 : kRsTile);
  for (int j = threadIdx.x; j < cnt; j += 256) {
    const uint64_t kk = s_k[j];
    const uint32_t dg = (uint32_t)(kk >> shift) & 255u;
    const uint32_t pos = s_gstart[dg] + (uint32_t)j - s_lstart[dg];
    kout[pos] = kk;
    vout[pos] = s_v[j];
  }
}

inline long long rs_tiles(long long n) { return (n + kRsTile - 1) / kRsTile; }

// Completes a stable sort by key bits [shift, 64) into a stable sort by the
// whole key.  A run of equal key >> shift needs work only when it holds an
// inversion (a key above its successor): k_sort_inversions marks the first
// element of each such run in a bitmap, then k_sort_runs insertion-sorts
// each marked run by the full key, stably, one thread per run.  Runs are
// disjoint, and permuting a run leaves key >> shift unchanged at every
// position of it (a thread that reads the first element of the next run, to
// find where its own run ends, sees the same key >> shift whatever that
// run's thread has moved there).

// Thread i at an inversion (keys[i-1] > keys[i], equal key >> shift) finds
// its run's first element h -- a doubling search backwards, then a binary
// search (the keys are sorted by key >> shift) -- and sets bit h of `mark`
// (ceil(n / 32) words, zero on entry).
__global__ __launch_bounds__(256) void k_sort_inversions(const uint64_t* __restrict__ keys,
                                                         long long n, int shift,
                                                         uint32_t* __restrict__ mark) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i <= 0 || i >= n) return;
  const uint64_t a = keys[i - 1], b = keys[i];
  const uint64_t hi = b >> shift;
  if ((a >> shift) != hi || a <= b) return;
  long long l = i - 1, lo = 0;  // keys[l] >> shift == hi; the run starts in [lo, l]
  for (long long step = 1;; step <<= 1) {
    const long long c = l - step;
    if (c < 0) break;
    if ((keys[c] >> shift) != hi) {
      lo = c + 1;
      break;
    }
    l = c;
  }
  while (lo < l) {
    const long long mid = lo + ((l - lo) >> 1);
    if ((keys[mid] >> shift) < hi)
      lo = mid + 1;
    else
      l = mid;
  }
  atomicOr(&mark[lo >> 5], 1u << (lo & 31));
}

// Thread j takes word j of the bitmap and sorts the runs whose first
// elements it marks.
__global__ __launch_bounds__(256) void k_sort_runs(uint64_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals, long long n,
                                                   int shift, const uint32_t* __restrict__ mark) {
  const long long j = (long long)blockIdx.x * 256 + threadIdx.x;
  if (j >= (n + 31) / 32) return;
  uint32_t w = mark[j];
  while (w != 0u) {
    const long long h = j * 32 + (__ffs(w) - 1);
    w &= w - 1u;
    const uint64_t hi = keys[h] >> shift;
    for (long long a = h + 1; a < n && (keys[a] >> shift) == hi; ++a) {
      const uint64_t k = keys[a];
      if (keys[a - 1] <= k) continue;
      const uint32_t v = vals[a];
      long long b = a;
      while (b > h && keys[b - 1] > k) {
        keys[b] = keys[b - 1];
        vals[b] = vals[b - 1];
        --b;
      }
      keys[b] = k;
      vals[b] = v;
    }
  }
}

// Scratch (u32 elements) for radix_sort_pairs on n elements.
inline long long rs_scratch_elems(long long n) {
  long long h = 256 * rs_tiles(n);
  return h + scan_scratch_elems(h);
}

// Stable sort of (keys, vals) by key bits [begin_bit, end_bit) (multiples of
// 8).  Ping-pongs between (*k, *v) and (k_alt, v_alt); on return *k / *v
// point at the sorted arrays (pointers may be swapped).
inline void radix_sort_pairs(uint64_t** k, uint32_t** v, uint64_t* k_alt, uint32_t* v_alt,
                             long long n, int begin_bit, int end_bit, uint32_t* scratch,
                             hipStream_t s) {
  if (n <= 1) return;
  const long long nt = rs_tiles(n);
  uint32_t* hist = scratch;
  uint32_t* sscr = scratch + 256 * nt;
  uint64_t* ka = *k;
  uint32_t* va = *v;
  uint64_t* kb = k_alt;
  uint32_t* vb = v_alt;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)nt), dim3(256), 0, s, (const uint64_t*)ka, n,
                       shift, hist, nt);
    scan_exclusive(hist, hist, 256 * nt, sscr, s);
    hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)nt), dim3(256), 0, s, (const uint64_t*)ka,
                       (const uint32_t*)va, kb, vb, n, shift, (const uint32_t*)hist, nt);
    uint64_t* tk = ka;
    ka = kb;
    kb = tk;
    uint32_t* tv = va;
    va = vb;
    vb = tv;
  }
  *k = ka;
  *v = va;
}

}  // namespace culsh
