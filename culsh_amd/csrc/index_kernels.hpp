// Index-side kernels: projection -> floor -> code, FP64 boundary recheck,
// bucket keys, bucket heads, bucket arrays, probe index (key filter, key
// directory, bucket lines).
//
// Restates cuLSH_Indexing.cu:181-274 (indexTableData) for all L tables at
// once: the reference's per-table SGEMM + subtract + divide + floor
// (Indexing.cu:27-66, Kernels.cu:40-101), M-pass LSD row sort
// (Indexing.cu:71-93) and head-flag bucketing (Indexing.cu:98-144).
#pragma once
#include "common.hpp"

namespace culsh {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr float kU32 = 5.9604644775390625e-08f;  // 2^-24, FP32 unit roundoff
constexpr int kCodeMax = 32767;
constexpr int kPjRows = 64, kPjCols = 64, kPjKc = 32;  // projection tile and d-step

// Upper bound of the Euclidean norm of each row (16 lanes per row).
__global__ __launch_bounds__(256) void k_row_norms(const float* __restrict__ X, long long N, int D,
                                                   float* __restrict__ nrm) {
  const int g = threadIdx.x >> 4, lig = threadIdx.x & 15;
  const long long n = (long long)blockIdx.x * 16 + g;
  double acc = 0.0;
  if (n < N) {
    const float* x = X + n * D;
    for (int d = lig; d < D; d += 16) acc += (double)x[d] * (double)x[d];
  }
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 16);
  if (n < N && lig == 0) nrm[n] = (float)(sqrt(acc) * (1.0 + 0x1.0p-20));
}

// FP16 copy of the rows (round to nearest even) and the largest squared row
// error ||x - x16||^2 over the rows (FP64 sum, rounded up), kept as the bits
// of a non-negative double, whose integer order is its numeric order; a
// value that leaves the half range makes the error infinite or NaN, and the
// caller then keeps to the FP32 rows.  16 lanes per row, rows grid-strided;
// one atomic per workgroup.
__global__ __launch_bounds__(256) void k_rows_f16(const float* __restrict__ X, long long N, int D,
                                                  uint16_t* __restrict__ x16,
                                                  unsigned long long* __restrict__ e2max_bits) {
  __shared__ double s_max[4];
  const int g = threadIdx.x >> 4, lig = threadIdx.x & 15;
  double mx = 0.0;
  for (long long n = (long long)blockIdx.x * 16 + g; n < N; n += (long long)gridDim.x * 16) {
    const float* x = X + n * D;
    uint16_t* h = x16 + n * D;
    double acc = 0.0;
    for (int d = lig; d < D; d += 16) {
      const _Float16 v = (_Float16)x[d];
      h[d] = __builtin_bit_cast(uint16_t, v);
      const double e = (double)x[d] - (double)(float)v;
      acc = fma(e, e, acc);
    }
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 16);
    const double up = acc * (1.0 + 0x1.0p-30);  // the sum's rounding, D < 2^22
    mx = (up != up || mx != mx) ? (double)NAN : fmax(mx, up);
  }
  // Largest over the workgroup (NaN wins), then one atomic.
#pragma unroll
  for (int off = 16; off < 64; off <<= 1) {
    const double o = __shfl_xor(mx, off);
    mx = (o != o || mx != mx) ? (double)NAN : fmax(mx, o);
  }
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = s_max[0];
    for (int w = 1; w < 4; ++w) m = (s_max[w] != s_max[w] || m != m) ? (double)NAN : fmax(m, s_max[w]);
    atomicMax(e2max_bits, (unsigned long long)__double_as_longlong(m != m ? INFINITY : m));
  }
}

// FP32 squared norm of each row (the norm expansion's |x|^2) and the largest
// row norm (as float bits; norms are non-negative so integer max orders
// them).
__global__ __launch_bounds__(256) void k_row_sqnorms(const float* __restrict__ X, long long N,
                                                     int D, float* __restrict__ xn2,
                                                     unsigned* __restrict__ max_norm_bits) {
  const int g = threadIdx.x >> 4, lig = threadIdx.x & 15;
  const long long n = (long long)blockIdx.x * 16 + g;
  float acc = 0.f;
  if (n < N) {
    const float* x = X + n * D;
    for (int d = lig; d < D; d += 16) acc = fmaf(x[d], x[d], acc);
  }
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 16);
  if (n < N && lig == 0) {
    xn2[n] = acc;
    atomicMax(max_norm_bits, __float_as_uint(sqrtf(acc) * (1.0f + 1e-4f)));
  }
}

// Epilogue of a projection tile element: v = (s - b) / W, code = floor(v),
// and whether the element needs FP64 re-evaluation, all in FP32.  `errb`
// bounds |s - x.a| (the dot product error of the fast path, already rounded
// up); rW = fl(1 / W).  With t = (x.a - b) / W exact (b is a float):
//   |v - t| <= (errb + 2.01 u (|s| + |b|)) rW (1 + 1.01 u) + 1.01 u |v|
// (u = 2^-24; the subtraction, the rounding of rW and of the product).  eps
// below exceeds that bound with room for its own six roundings; fr = v -
// floor(v) is exact except for v in (-1, 0) (error <= 2^-25), hence the
// extra 2^-24 in the comparison.  An element not flagged has no integer
// within |v - t| of v, so floor(v) = floor(t); and the FP64 oracle's own
// error (about D 2^-53 ||x|| ||a_j||) is far below errb, so its floor is the
// same.  Writes the code to *code; returns the recheck decision (invalid
// elements: false, code untouched).
__device__ inline bool project_epilogue(bool valid, float s, float bb, float rW, float errb,
                                        int16_t* code, int* __restrict__ err) {
  if (!valid) return false;
  const float v = (s - bb) * rW;
  const float f = floorf(v);
  const float eps =
      ((errb + kU32 * 3.0f * (fabsf(s) + fabsf(bb))) * rW + kU32 * 2.0f * fabsf(v)) *
          (1.0f + 0x1.0p-18f) + 1e-30f + kU32;
  if (f < -(float)(kCodeMax - 1) || f > (float)(kCodeMax - 1)) {
    atomicOr(err, 1);
    return false;
  }
  *code = (int16_t)(int)f;
  const float fr = v - f;
  return fr <= eps || 1.0f - fr <= eps;
}

// Writes a 64 x 64 tile of codes staged in LDS (sc[r * 64 + c] for point
// n0 + r, column j0 + c) to codes [L][N][M]: table by table, the tile's
// rows of one table are one run of M-code rows, so consecutive threads
// write consecutive code pairs (4-byte stores when M and the table's first
// column in the tile are even, else one code per thread).
__device__ inline void store_code_tile(const int16_t* sc, long long n0, int j0, long long N,
                                       int J, int M, int16_t* __restrict__ codes) {
  const int tid = threadIdx.x;
  const int rows = (int)min((long long)kPjRows, N - n0);
  const int jend = min(j0 + kPjCols, J);
  for (int t = j0 / M; t * M < jend; ++t) {
    const int ms = max(j0 - t * M, 0), me = min(jend - t * M, M);  // m range in the tile
    const int cs = t * M + ms - j0;                                // its first tile column
    int16_t* dst = codes + ((long long)t * N + n0) * M;
    // e / n for e < 2^12, n <= 64 as floor((e + 0.5) / n) in FP32: the
    // quotient is at least 1/128 away from an integer, far beyond the
    // rounding error.
    if ((M & 1) == 0 && (ms & 1) == 0 && (me & 1) == 0) {
      const int np = (me - ms) >> 1;
      const float rn = 1.0f / (float)np;
      for (int e = tid; e < rows * np; e += 256) {
        const int r = (int)(((float)e + 0.5f) * rn), q = e - r * np;
        const int c = cs + 2 * q;
        const uint32_t v = (uint32_t)(uint16_t)sc[r * kPjCols + c] |
                           ((uint32_t)(uint16_t)sc[r * kPjCols + c + 1] << 16);
        *reinterpret_cast<uint32_t*>(dst + (long long)r * M + ms + 2 * q) = v;
      }
    } else {
      const int nm = me - ms;
      const float rn = 1.0f / (float)nm;
      for (int e = tid; e < rows * nm; e += 256) {
        const int r = (int)(((float)e + 0.5f) * rn), q = e - r * nm;
        dst[(long long)r * M + ms + q] = sc[r * kPjCols + cs + q];
      }
    }
  }
}

constexpr int kRcShards = 8;  // recheck list shards (one counter each)

// Epilogue of a 64 x 64 projection tile held as MFMA accumulators (16 rows
// per wave, lane layout of v_mfma_f32_16x16x*): codes (staged in the 8 KB
// of LDS at sc, which the caller no longer needs, then written by
// store_code_tile), then the flagged (n, j) appended to recheck-list shard
// blockIdx % kRcShards with one atomic per workgroup.  Shard k holds entries
// list[k * seg, k * seg + count[k]); a count above seg means the shard
// overflowed.  `kerr` is the fast path's dot product error factor (bound =
// kerr ||x|| ||a_j||).  Every thread of the workgroup must call it, after a
// barrier that ends every read of sc's previous contents.
__device__ inline void projection_tile_epilogue(
    const floatx4 (&acc)[4], long long n0, int j0, long long N, int J, int M,
    const float* __restrict__ bj, const float* __restrict__ naj, const float* __restrict__ nx,
    float W, double kerr, int16_t* sc, int16_t* __restrict__ codes,
    unsigned long long* __restrict__ list, unsigned long long* __restrict__ count,
    unsigned long long seg, int* __restrict__ err) {
  __shared__ unsigned s_wcnt[4];
  __shared__ unsigned long long s_base;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const float rW = 1.0f / W;
  // kerr rounded up to FP32; the two FP32 products of errb stay within
  // (1 + 2^-20) of their exact value.
  const float kf = (float)kerr * (1.0f + 0x1.0p-20f);
  float xn[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const long long n = n0 + w * 16 + (lane >> 4) * 4 + i;
    xn[i] = n < N ? kf * nx[n] : 0.f;
  }
  unsigned bits = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int j = j0 + c * 16 + (lane & 15);
    const bool jv = j < J;
    const float bb = jv ? bj[j] : 0.f;
    const float na = jv ? naj[j] * (1.0f + 0x1.0p-20f) : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = w * 16 + (lane >> 4) * 4 + i;
      const bool valid = jv && n0 + r < N;
      if (project_epilogue(valid, acc[c][i], bb, rW, xn[i] * na, &sc[r * kPjCols + c * 16 + (lane & 15)],
                           err))
        bits |= 1u << (c * 4 + i);
    }
  }
  __syncthreads();
  store_code_tile(sc, n0, j0, N, J, M, codes);
  // Offsets: inclusive scan of the per-lane counts within the wave, then
  // per-wave totals, then one atomic for the workgroup.
  const unsigned lc = (unsigned)__popc(bits);
  unsigned incl = lc;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned y = __shfl_up(incl, off);
    if (lane >= off) incl += y;
  }
  if (lane == 63) s_wcnt[w] = incl;
  __syncthreads();
  if (tid == 0) {
    const unsigned tot = s_wcnt[0] + s_wcnt[1] + s_wcnt[2] + s_wcnt[3];
    s_base = tot ? atomicAdd(&count[blockIdx.x % kRcShards], (unsigned long long)tot) : 0ull;
  }
  __syncthreads();
  unsigned long long off = s_base + (incl - lc);
  for (int ww = 0; ww < w; ++ww) off += s_wcnt[ww];
  unsigned long long* shard = list + (unsigned long long)(blockIdx.x % kRcShards) * seg;
  while (bits) {
    const int b = __ffs(bits) - 1;
    bits &= bits - 1;
    const int c = b >> 2, i = b & 3;
    const long long n = n0 + w * 16 + (lane >> 4) * 4 + i;
    const int j = j0 + c * 16 + (lane & 15);
    if (off < seg) shard[off] = ((unsigned long long)n << 24) | (unsigned long long)j;
    ++off;
  }
}

// Maps a linear workgroup id to (row tile, column tile) so that the column
// tiles of one row tile run on the same XCD (workgroups are dealt to the 8
// XCDs round-robin): the X tile is then fetched from HBM once per XCD L2.
__device__ inline bool xcd_tile(int nct, long long ntiles, long long& rt, int& ct) {
  const long long b = blockIdx.x;
  const long long per = (long long)gridDim.x / 8;
  const long long logical = (b & 7) * per + (b >> 3);
  if (logical >= ntiles) return false;
  rt = logical / nct;
  ct = (int)(logical - rt * nct);
  return true;
}


// Loads the 64 x 32 X tile and the 32 x 64 A tile of one projection step
// into registers as 16-byte vectors (thread e of 512 takes row e >> 3 of X,
// columns 4 (e & 7) .. +3, and row e >> 4 of A, columns 4 (e & 15) .. +3).
// Every load is issued from a clamped, always-valid address before any is
// consumed; elements past D or N are zeroed when stored.  X's rows are read
// as vectors when D % 4 == 0 and X is 16-byte aligned (then a vector either
// lies inside the row or wholly past it), else element by element.
__device__ inline void load_tile(const float* __restrict__ X, long long N, int D,
                                 const float* __restrict__ Ad, int Jp, long long n0, int j0,
                                 int d0, float4 (&xv)[2], float4 (&av)[2]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int e = tid + 256 * k;
    const int r = e >> 3, c = (e & 7) * 4;
    const long long n = min(n0 + r, N - 1);
    if ((D & 3) == 0 && (reinterpret_cast<uintptr_t>(X) & 15) == 0) {
      const int d = min(d0 + c, D - 4);
      xv[k] = *reinterpret_cast<const float4*>(X + n * D + d);
    } else {
      const float* xr = X + n * D;
      xv[k] = make_float4(xr[min(d0 + c, D - 1)], xr[min(d0 + c + 1, D - 1)],
                          xr[min(d0 + c + 2, D - 1)], xr[min(d0 + c + 3, D - 1)]);
    }
    const int ra = e >> 4, ca = (e & 15) * 4;
    av[k] = *reinterpret_cast<const float4*>(Ad + (long long)min(d0 + ra, D - 1) * Jp + j0 + ca);
  }
}

// Stores a step's registers into LDS, zeroing out-of-range elements.
__device__ inline void store_tile(long long N, int D, long long n0, int d0, const float4 (&xv)[2],
                                  const float4 (&av)[2], float (*Xs)[kPjKc + 1],
                                  float (*As)[kPjCols + 1]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int e = tid + 256 * k;
    const int r = e >> 3, c = (e & 7) * 4;
    const bool rv = n0 + r < N;
    Xs[r][c] = (rv && d0 + c < D) ? xv[k].x : 0.f;
    Xs[r][c + 1] = (rv && d0 + c + 1 < D) ? xv[k].y : 0.f;
    Xs[r][c + 2] = (rv && d0 + c + 2 < D) ? xv[k].z : 0.f;
    Xs[r][c + 3] = (rv && d0 + c + 3 < D) ? xv[k].w : 0.f;
    const int ra = e >> 4, ca = (e & 15) * 4;
    const bool av_ok = d0 + ra < D;
    As[ra][ca] = av_ok ? av[k].x : 0.f;
    As[ra][ca + 1] = av_ok ? av[k].y : 0.f;
    As[ra][ca + 2] = av_ok ? av[k].z : 0.f;
    As[ra][ca + 3] = av_ok ? av[k].w : 0.f;
  }
}

// FP32 projection on MFMA (v_mfma_f32_16x16x4_f32): a 64-point x 64-column
// tile per workgroup, 16 rows per wave, D in steps of 32 through two LDS
// buffers: the next step's global loads are in flight while the MFMAs of
// the current one run, and each step takes one barrier.  Ad is [Dp][Jp]
// (column j = t*M + m of table t), zero padded.  Plain-form error bound:
// |s - x.a| <= 2 (D+2) u ||x|| ||a_j||.
__global__ __launch_bounds__(256) void k_project_f32(
    const float* __restrict__ X, long long N, int D, const float* __restrict__ Ad, int Jp, int J,
    int M, const float* __restrict__ bj, const float* __restrict__ naj,
    const float* __restrict__ nx, float W, int16_t* __restrict__ codes,
    unsigned long long* __restrict__ list, unsigned long long* __restrict__ count,
    unsigned long long seg, int* __restrict__ err) {
  __shared__ float Xs[2][kPjRows][kPjKc + 1];
  __shared__ float As[2][kPjKc][kPjCols + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int nct = Jp / kPjCols;
  const long long nrt = (N + kPjRows - 1) / kPjRows;
  long long rt;
  int ct;
  if (!xcd_tile(nct, nrt * nct, rt, ct)) return;
  const long long n0 = rt * kPjRows;
  const int j0 = ct * kPjCols;
  floatx4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
  float4 xv[2], av[2];
  load_tile(X, N, D, Ad, Jp, n0, j0, 0, xv, av);
  store_tile(N, D, n0, 0, xv, av, Xs[0], As[0]);
  __syncthreads();
  const int nsteps = (D + kPjKc - 1) / kPjKc;
  for (int st = 0; st < nsteps; ++st) {
    const int cur = st & 1;
    const bool more = st + 1 < nsteps;
    if (more) load_tile(X, N, D, Ad, Jp, n0, j0, (st + 1) * kPjKc, xv, av);
#pragma unroll
    for (int kk = 0; kk < kPjKc; kk += 4) {
      const float a = Xs[cur][w * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float b = As[cur][kk + (lane >> 4)][c * 16 + (lane & 15)];
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[c], 0, 0, 0);
      }
    }
    if (more) store_tile(N, D, n0, (st + 1) * kPjKc, xv, av, Xs[cur ^ 1], As[cur ^ 1]);
    __syncthreads();
  }
  const double ka = 2.0 * (double)(D + 2) * (double)kU32;
  projection_tile_epilogue(acc, n0, j0, N, J, M, bj, naj, nx, W, ka,
                           reinterpret_cast<int16_t*>(&Xs[0][0][0]), codes, list, count, seg, err);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// float -> bf16, round to nearest even (finite inputs).
__device__ inline uint16_t f2bf_bits(float x) {
  const uint32_t b = __float_as_uint(x);
  return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}
__device__ inline float bf_bits2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Splits x into hi + lo bf16 parts: |x - hi - lo| <= 2^-16 |x|.
__device__ inline void split_bf16(float x, uint16_t& hi, uint16_t& lo) {
  hi = f2bf_bits(x);
  lo = f2bf_bits(x - bf_bits2f(hi));
}

constexpr int kBfPad = 40;  // LDS row of 32 bf16 plus 16 bytes of padding

// Precision form: BF16x3 split-precision projection on MFMA
// (v_mfma_f32_16x16x32_bf16), x.a ~ xh.ah + xh.al + xl.ah with FP32
// accumulation.  Same tiling and epilogue as k_project_f32.  A arrives
// split once at upload (AhT / AlT: [Jp][Dp] bf16, d contiguous per column,
// Dp = D rounded up to 32, zero padded); each X tile element is split once
// while it is staged in LDS; every MFMA operand (8 consecutive k of one row
// or column) is one 16-byte LDS read.  Error bound: each term is within
// 3.01 * 2^-16 |x_d a_d| of x_d a_d (dropped xl.al and the split residuals)
// and the 3D-term FP32 accumulation adds 3.1 D 2^-24 sum |x_d a_d|; with
// sum |x_d a_d| <= ||x|| ||a_j|| and a factor 2 margin.
__global__ __launch_bounds__(256) void k_project_bf16x3(
    const float* __restrict__ X, long long N, int D, const uint16_t* __restrict__ AhT,
    const uint16_t* __restrict__ AlT, int Dp, int Jp, int J, int M,
    const float* __restrict__ bj, const float* __restrict__ naj, const float* __restrict__ nx,
    float W, int16_t* __restrict__ codes, unsigned long long* __restrict__ list,
    unsigned long long* __restrict__ count, unsigned long long seg, int* __restrict__ err) {
  __shared__ __align__(16) uint16_t s_bf[4][kPjRows][kBfPad];  // Xh, Xl, Ah, Al
  uint16_t(*Xh)[kBfPad] = s_bf[0];
  uint16_t(*Xl)[kBfPad] = s_bf[1];
  uint16_t(*Ah)[kBfPad] = s_bf[2];
  uint16_t(*Al)[kBfPad] = s_bf[3];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int nct = Jp / kPjCols;
  const long long nrt = (N + kPjRows - 1) / kPjRows;
  long long rt;
  int ct;
  if (!xcd_tile(nct, nrt * nct, rt, ct)) return;
  const long long n0 = rt * kPjRows;
  const int j0 = ct * kPjCols;
  floatx4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
  // Staging roles: row / column sr, eight consecutive k from sk.
  const int sr = tid >> 2, sk = (tid & 3) * 8;
  const bool vecx = (D & 3) == 0;
  const long long xn = n0 + sr;
  const int kb = (lane >> 4) * 8;  // this lane's 8 consecutive k of the 32
  for (int d0 = 0; d0 < D; d0 += kPjKc) {
    float xv[8];
    const int dd = d0 + sk;
    if (xn < N && vecx && dd + 8 <= D) {
      const float4* xp = reinterpret_cast<const float4*>(X + xn * D + dd);
      const float4 a = xp[0], b = xp[1];
      xv[0] = a.x; xv[1] = a.y; xv[2] = a.z; xv[3] = a.w;
      xv[4] = b.x; xv[5] = b.y; xv[6] = b.z; xv[7] = b.w;
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) xv[i] = (xn < N && dd + i < D) ? X[xn * D + dd + i] : 0.f;
    }
    const uint4 ah4 = *reinterpret_cast<const uint4*>(AhT + (long long)(j0 + sr) * Dp + dd);
    const uint4 al4 = *reinterpret_cast<const uint4*>(AlT + (long long)(j0 + sr) * Dp + dd);
    uint32_t hw[4], lw[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      uint16_t h0, l0, h1, l1;
      split_bf16(xv[2 * i], h0, l0);
      split_bf16(xv[2 * i + 1], h1, l1);
      hw[i] = (uint32_t)h0 | ((uint32_t)h1 << 16);
      lw[i] = (uint32_t)l0 | ((uint32_t)l1 << 16);
    }
    *reinterpret_cast<uint4*>(&Xh[sr][sk]) = make_uint4(hw[0], hw[1], hw[2], hw[3]);
    *reinterpret_cast<uint4*>(&Xl[sr][sk]) = make_uint4(lw[0], lw[1], lw[2], lw[3]);
    *reinterpret_cast<uint4*>(&Ah[sr][sk]) = ah4;
    *reinterpret_cast<uint4*>(&Al[sr][sk]) = al4;
    __syncthreads();
    const bf16x8 xh = *reinterpret_cast<const bf16x8*>(&Xh[w * 16 + (lane & 15)][kb]);
    const bf16x8 xl = *reinterpret_cast<const bf16x8*>(&Xl[w * 16 + (lane & 15)][kb]);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const bf16x8 ah = *reinterpret_cast<const bf16x8*>(&Ah[c * 16 + (lane & 15)][kb]);
      const bf16x8 al = *reinterpret_cast<const bf16x8*>(&Al[c * 16 + (lane & 15)][kb]);
      acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, ah, acc[c], 0, 0, 0);
      acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, al, acc[c], 0, 0, 0);
      acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl, ah, acc[c], 0, 0, 0);
    }
    __syncthreads();
  }
  const double kb3 = 2.0 * (3.01 * 0x1.0p-16 + 3.1 * (double)D * 0x1.0p-24);
  projection_tile_epilogue(acc, n0, j0, N, J, M, bj, naj, nx, W, kb3,
                           reinterpret_cast<int16_t*>(&s_bf[0][0][0]), codes, list, count, seg,
                           err);
}

// Re-evaluates the recorded projections in FP64 in the oracle's operation
// order -- sequential FMA over d ascending, then (acc - b) / W
// (cuLSH_Indexing.cu:43-62 in FP64) -- and patches their codes.
// Aorig is [L*M][D]: column j of the projection contiguous over d.  A
// workgroup takes 256 recorded (n, j); for each 32-dimension chunk, 32
// lanes per entry load its x and a segments (128-byte runs) into LDS, then
// thread i extends entry i's sequential FMA chain (the oracle's order).
constexpr int kRcDc = 32;
__global__ __launch_bounds__(256) void k_recheck(const float* __restrict__ X, long long N, int D,
                                                 const float* __restrict__ Aorig,
                                                 const float* __restrict__ bj, float W, int M,
                                                 const unsigned long long* __restrict__ list,
                                                 unsigned long long count,
                                                 int16_t* __restrict__ codes,
                                                 int* __restrict__ err) {
  __shared__ float xs[256][kRcDc + 1];
  __shared__ float as[256][kRcDc + 1];
  __shared__ unsigned long long s_e[256];
  const int tid = threadIdx.x;
  const unsigned long long base = (unsigned long long)blockIdx.x * 256;
  const unsigned long long i = base + tid;
  s_e[tid] = i < count ? list[i] : ~0ull;
  const int lane32 = tid & 31, grp = tid >> 5;  // 8 groups of 32 lanes
  double acc = 0.0;
  for (int d0 = 0; d0 < D; d0 += kRcDc) {
    const int dl = min(kRcDc, D - d0);
    __syncthreads();
    // 32 rows per group of 32 lanes; every load in flight before any store.
    float xv[32], av[32];
    const int dd = min(d0 + lane32, D - 1);
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      const unsigned long long e = s_e[grp + 8 * k];
      const unsigned long long ee = e == ~0ull ? 0ull : e;
      const long long n = (long long)(ee >> 24);
      const int j = (int)(ee & 0xFFFFFFull);
      xv[k] = X[n * D + dd];
      av[k] = Aorig[(long long)j * D + dd];
    }
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      const int r = grp + 8 * k;
      const bool ok = s_e[r] != ~0ull && lane32 < dl;
      xs[r][lane32] = ok ? xv[k] : 0.f;
      as[r][lane32] = ok ? av[k] : 0.f;
    }
    __syncthreads();
    for (int dd = 0; dd < dl; ++dd) acc = fma((double)xs[tid][dd], (double)as[tid][dd], acc);
  }
  if (i >= count) return;
  const unsigned long long e = s_e[tid];
  const long long n = (long long)(e >> 24);
  const int j = (int)(e & 0xFFFFFFull);
  const double f = floor((acc - (double)bj[j]) / (double)W);
  if (f < -(double)(kCodeMax - 1) || f > (double)(kCodeMax - 1)) {
    atomicOr(err, 1);
    return;
  }
  const int t = j / M, m = j - t * M;
  codes[((long long)t * N + n) * M + m] = (int16_t)(int)f;
}

// FP64 projection of every (n, j) in the oracle's order, tiled: a workgroup
// computes 64 points x 64 columns, staging 32-dimension chunks of both as
// doubles in LDS (coalesced row reads); thread (r, c) keeps 4 x 4 sums, each
// a sequential FMA over d ascending -- bit-identical to k_recheck's, so no
// recheck follows.  The path for wide rows (large D), where the FP32 error
// bound flags too many projections (the recheck list overflows).
constexpr int kP64Tile = 64, kP64Dc = 32;
__global__ __launch_bounds__(256) void k_project_f64_tiled(const float* __restrict__ X, long long N,
                                                           int D, const float* __restrict__ Aorig,
                                                           const float* __restrict__ bj, float W,
                                                           int M, int J,
                                                           int16_t* __restrict__ codes,
                                                           int* __restrict__ err) {
  __shared__ double xs[kP64Tile][kP64Dc + 1];
  __shared__ double as[kP64Tile][kP64Dc + 1];
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  const int ctiles = (J + kP64Tile - 1) / kP64Tile;
  const long long n0 = (long long)(blockIdx.x / ctiles) * kP64Tile;
  const int j0 = (int)(blockIdx.x % ctiles) * kP64Tile;
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;
  for (int d0 = 0; d0 < D; d0 += kP64Dc) {
    const int dl = min(kP64Dc, D - d0);
    __syncthreads();
    float xv[8], av[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int e = tid + 256 * k;
      const int row = e >> 5, dd = e & 31;
      const int d = min(d0 + dd, D - 1);
      xv[k] = X[min(n0 + row, N - 1) * D + d];
      av[k] = Aorig[(long long)min(j0 + row, J - 1) * D + d];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int e = tid + 256 * k;
      const int row = e >> 5, dd = e & 31;
      xs[row][dd] = (n0 + row < N && dd < dl) ? (double)xv[k] : 0.0;
      as[row][dd] = (j0 + row < J && dd < dl) ? (double)av[k] : 0.0;
    }
    __syncthreads();
    for (int dd = 0; dd < dl; ++dd) {
      double xv[4], av[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) xv[r] = xs[tr * 4 + r][dd];
#pragma unroll
      for (int c = 0; c < 4; ++c) av[c] = as[tc + 16 * c][dd];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = fma(xv[r], av[c], acc[r][c]);
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const long long n = n0 + tr * 4 + r;
    if (n >= N) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int j = j0 + tc + 16 * c;
      if (j >= J) continue;
      const double f = floor((acc[r][c] - (double)bj[j]) / (double)W);
      if (f < -(double)(kCodeMax - 1) || f > (double)(kCodeMax - 1)) {
        atomicOr(err, 1);
        continue;
      }
      const int t = j / M, m = j - t * M;
      codes[((long long)t * N + n) * M + m] = (int16_t)(int)f;
    }
  }
}

// Bucket key of every (table, point): keys[t*N + n], vals = n.  The
// workgroup's 256 code rows are contiguous (row i = t*N + n at codes + i*M;
// 512 M bytes from a 16-byte aligned start) and are staged in LDS with
// 16-byte loads, so the row reads coalesce; each thread then sums its row's
// terms from LDS.  M <= 64.
__global__ __launch_bounds__(256) void k_keys(const int16_t* __restrict__ codes, long long N,
                                              int L, int M, const uint64_t* __restrict__ seeds,
                                              int tb, uint64_t* __restrict__ keys,
                                              uint32_t* __restrict__ vals) {
  __shared__ uint4 s_rows[256 * 64 * 2 / 16];
  int16_t* s16 = reinterpret_cast<int16_t*>(s_rows);
  const long long n_all = (long long)L * N;
  const long long i0 = (long long)blockIdx.x * 256;
  const int rows = (int)(n_all - i0 < 256 ? n_all - i0 : 256);
  const int nel = rows * M, n16 = nel >> 3;
  const uint4* src = reinterpret_cast<const uint4*>(codes + i0 * M);
  for (int c = threadIdx.x; c < n16; c += 256) s_rows[c] = src[c];
  for (int e = 8 * n16 + threadIdx.x; e < nel; e += 256) s16[e] = codes[i0 * M + e];
  __syncthreads();
  if ((int)threadIdx.x >= rows) return;
  const long long i = i0 + threadIdx.x;
  const int t = (int)(i / N);
  const long long n = i - (long long)t * N;
  const uint64_t seed = seeds[t];
  const int16_t* c = s16 + threadIdx.x * M;
  uint64_t h = 0;
  for (int m = 0; m < M; ++m) h += code_term(seed, m, (int)c[m]);
  keys[i] = make_key(h, t, tb);
  vals[i] = (uint32_t)n;
}

// Head flags over the sorted (key, id) array: a new bucket starts where the
// key changes.
__global__ __launch_bounds__(256) void k_heads(const uint64_t* __restrict__ keys, long long n_all,
                                               uint32_t* __restrict__ flag) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n_all) return;
  flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// Bucket records, bucket codes and the exactness check of the hashed keys in
// one pass over the sorted (key, id) array.  Every element's code row is
// gathered once into LDS; heads write their bucket's record (key, start;
// the size follows in k_rec_sizes) and code row (row-major [B][M], M <= 64)
// as contiguous runs; every other element compares its codes with its
// predecessor's (same key), so all members of a bucket equal its head by
// induction -- two different code vectors with one 64-bit key set
// *collision and the host rebuilds with another hash seed.  The first
// element of a workgroup also gathers the row before it.
__global__ __launch_bounds__(256) void k_bucket_fill(
    const uint64_t* __restrict__ keys, const uint32_t* __restrict__ ids,
    const uint32_t* __restrict__ flag, const uint32_t* __restrict__ excl, long long n_all,
    long long N, int M, int tb, const int16_t* __restrict__ codes, BucketRec* __restrict__ recs,
    int16_t* __restrict__ bcodes, int* __restrict__ collision) {
  __shared__ uint4 s_rec[256];
  __shared__ unsigned char s_hloc[256];
  __shared__ int16_t s_code[257 * 64];  // row r + 1: element i0 + r; row 0: element i0 - 1
  const int tid = threadIdx.x;
  const long long i0 = (long long)blockIdx.x * 256, i = i0 + tid;
  const bool in = i < n_all;
  const bool head = in && flag[i] != 0u;
  const uint32_t kb = excl[i0];
  const int nh = __syncthreads_count(head);
  auto gather = [&](long long e, int row) {
    const int t = (int)(keys[e] >> (64 - tb));
    const int16_t* src = codes + ((long long)t * N + ids[e]) * M;
    int16_t* dst = s_code + row * M;
    if ((M & 1) == 0) {
      const uint32_t* s32 = reinterpret_cast<const uint32_t*>(src);
      for (int m = 0; m < M / 2; ++m) {
        const uint32_t v = s32[m];
        dst[2 * m] = (int16_t)(v & 0xFFFFu);
        dst[2 * m + 1] = (int16_t)(v >> 16);
      }
    } else {
      for (int m = 0; m < M; ++m) dst[m] = src[m];
    }
  };
  if (in) gather(i, tid + 1);
  if (tid == 0 && in && !head) gather(i - 1, 0);  // i > 0: element 0 is a head
  if (head) {
    const int loc = (int)(excl[i] - kb);
    const uint64_t key = keys[i];
    s_rec[loc] = make_uint4((uint32_t)key, (uint32_t)(key >> 32), (uint32_t)i, 0u);
    s_hloc[loc] = (unsigned char)tid;
  }
  __syncthreads();
  if (in && !head) {
    const int16_t* a = s_code + (tid + 1) * M;
    const int16_t* p = s_code + tid * M;
    for (int m = 0; m < M; ++m)
      if (a[m] != p[m]) {
        atomicOr(collision, 1);
        break;
      }
  }
  uint4* rdst = reinterpret_cast<uint4*>(recs + kb);
  for (int j = tid; j < nh; j += 256) rdst[j] = s_rec[j];
  int16_t* cdst = bcodes + (long long)kb * M;
  for (int e = tid; e < nh * M; e += 256) {
    const int j = e / M, m = e - j * M;
    cdst[e] = s_code[((int)s_hloc[j] + 1) * M + m];
  }
}

// Bucket sizes from consecutive starts.
__global__ __launch_bounds__(256) void k_rec_sizes(BucketRec* __restrict__ recs, long long B,
                                                   long long n_all) {
  const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
  if (k >= B) return;
  const uint32_t end = k + 1 < B ? recs[k + 1].start : (uint32_t)n_all;
  recs[k].size = end - recs[k].start;
}

// First bucket of each table (positions t*N in the sorted order) and the
// total bucket count at out[L].
__global__ void k_table_bases(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ excl,
                              long long N, int L, uint32_t* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < L) out[t] = excl[(long long)t * N];
  if (t == L) {
    const long long last = (long long)L * N - 1;
    out[L] = excl[last] + flag[last];
  }
}

// Probe index (see ProbeIndex), built from the key-sorted records.
//
// Key filter: the records of one filter word are consecutive (the word is
// monotone in the key), so a wave ORs the bits of each run of equal words
// among its lanes (a segmented scan: lanes o apart with equal words hold
// only that word's records between them) and the last lane of each run ORs
// them into the word; the words start zeroed.
__global__ __launch_bounds__(256) void k_filter(const BucketRec* __restrict__ recs, long long B,
                                                int tb, const uint2* __restrict__ tfw,
                                                uint64_t* __restrict__ filter) {
  const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
  const bool in = g < B;
  const uint64_t key = in ? recs[g].key : 0ull;
  const uint32_t w = in ? filter_word(key, tb, tfw[(int)(key >> (64 - tb))]) : 0xffffffffu;
  unsigned long long bits = in ? filter_bits(key) : 0ull;
  const int lane = (int)(threadIdx.x & 63);
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t wo = __shfl_up(w, o, 64);
    const unsigned long long bo = __shfl_up(bits, o, 64);
    if (lane >= o && wo == w) bits |= bo;
  }
  const uint32_t wn = __shfl_down(w, 1, 64);
  if (in && (lane == 63 || wn != w))
    atomicOr(reinterpret_cast<unsigned long long*>(filter) + w, bits);
}

// Key directory (see ProbeIndex) in one pass over the key-sorted records:
// the first record of each occupied slot writes the slot's entry (first
// record, fingerprint, count up to 255), and the last record of a slot whose
// next slot is empty writes that slot's start (so a saturated slot's end is
// its successor's first record).  Every other slot keeps the empty entry of
// the fill; slot nslots is the end sentinel (first = B).
__global__ __launch_bounds__(256) void k_dir_clear(uint64_t* __restrict__ dir, long long n) {
  const long long v = (long long)blockIdx.x * 256 + threadIdx.x;
  if (v < n) dir[v] = 0ull;
}

__global__ __launch_bounds__(256) void k_directory(const BucketRec* __restrict__ recs, long long B,
                                                   int shift, long long nslots,
                                                   uint64_t* __restrict__ dir) {
  const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
  if (r == 0) dir[nslots] = dir_entry((uint32_t)B, 0, 0);
  if (r >= B) return;
  const uint64_t key = recs[r].key;
  const long long v = (long long)(key >> shift);
  if (r == 0 || (long long)(recs[r - 1].key >> shift) != v) {
    uint32_t cnt = 1;
    while (cnt < 255 && r + cnt < B && (long long)(recs[r + cnt].key >> shift) == v) ++cnt;
    dir[v] = dir_entry((uint32_t)r, key, cnt);
  }
  const long long vn = r + 1 < B ? (long long)(recs[r + 1].key >> shift) : nslots;
  if (vn != v && vn != v + 1 && v + 1 < nslots) dir[v + 1] = dir_entry((uint32_t)(r + 1), 0, 0);
}

// Bucket lines (start, size, the codes in pairs, zero padding): a
// workgroup copies the code rows of its 256 records into LDS with
// coalesced loads, then writes their lines as consecutive 16-byte chunks.
__global__ __launch_bounds__(256) void k_bucket_lines(const BucketRec* __restrict__ recs,
                                                      const int16_t* __restrict__ bcodes,
                                                      long long B, int M, int lstride,
                                                      uint4* __restrict__ lines) {
  __shared__ int16_t s_code[256 * 64];
  const int tid = threadIdx.x;
  const long long g0 = (long long)blockIdx.x * 256;
  const int nb = (int)min(256ll, B - g0);
  const int16_t* src = bcodes + g0 * M;
  for (int e = tid; e < nb * M; e += 256) s_code[e] = src[e];
  __syncthreads();
  const int cpl = lstride / 16;
  uint4* dst = lines + g0 * cpl;
  for (int e = tid; e < nb * cpl; e += 256) {
    const int r = e / cpl, ch = e - r * cpl;
    const BucketRec rec = recs[g0 + r];
    uint32_t w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int wi = ch * 4 + i;
      uint32_t v = 0;
      if (wi == 0) {
        v = rec.start;
      } else if (wi == 1) {
        v = rec.size;
      } else {
        const int m = 2 * (wi - 2);
        if (m < M) v = (uint32_t)(uint16_t)s_code[r * M + m];
        if (m + 1 < M) v |= (uint32_t)(uint16_t)s_code[r * M + m + 1] << 16;
      }
      w[i] = v;
    }
    dst[e] = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

}  // namespace culsh
