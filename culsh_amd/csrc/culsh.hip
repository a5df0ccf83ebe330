// libculsh — MI355X-native LSH index/query path behind a C ABI.
//
// Host orchestration for cuLSH::HashTables (cuLSH_HashTables.cu:51-520) and
// cuLSH::SearchTables (cuLSH_SearchTables.cu:42-463).  The index lives in
// HBM: dataset, bucket keys, offsets, member ids, codes and the probe index
// (key filter, key directory and bucket lines);
// both index() and search() run as a short chain of device kernels on one
// stream per object.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/culsh_c.h"
#include "batch_kernels.hpp"
#include "canon.hpp"
#include "common.hpp"
#include "index_file.hpp"
#include "index_kernels.hpp"
#include "philox.hpp"
#include "search_kernels.hpp"
#include "sort.hpp"

using namespace culsh;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(CULSH_E_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                     \
  } while (0)

#define TRY(expr)               \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != CULSH_OK) return rc_; \
  } while (0)

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // Grow-only allocation of at least `count` elements.
  int ensure(size_t count) {
    if (count <= n && p) return CULSH_OK;
    release();
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc((void**)&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(CULSH_E_MEMORY, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    n = std::max<size_t>(count, 1);
    return CULSH_OK;
  }
  int exact(size_t count) {
    release();
    return ensure(count);
  }
};

int ceil_log2_host(long long x) {
  int l = 0;
  while ((1ll << l) < x) ++l;
  return l;
}

int floor_log2_host(long long x) {
  int l = 0;
  while ((2ll << l) <= x) ++l;
  return l;
}

// Generation stamps of dataset contents on the device: every index build or
// load and every upload of a search dataset takes a new one, so a cache of
// per-row data keyed on (pointer, generation) never outlives its dataset
// even when an allocation is reused at the same address.
std::atomic<unsigned long long> g_data_gen{0};
unsigned long long next_data_gen() { return ++g_data_gen; }

inline unsigned grid_for(long long n, int block = 256) {
  return (unsigned)((n + block - 1) / block);
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

}  // namespace

// ---------------------------------------------------------------------------
// HashTables
// ---------------------------------------------------------------------------
struct culsh_ht {
  int N = 0, D = 0, L = 0, M = 0;
  float W = 0.f;
  int form = CULSH_FORM_PLAIN;
  bool has_params = false;
  std::vector<float> A, b;  // [L][D*M] (column-major per table), [L][M]
  hipStream_t stream = nullptr;

  // Projection parameters on the device.
  int J = 0, Jp = 0;
  DevBuf<float> dAd;      // [D][Jp]   column j = t*M + m
  DevBuf<float> dAorig;   // [J][D]    column j contiguous over d
  DevBuf<float> dbj;      // [Jp]
  DevBuf<float> dnaj;     // [Jp]      upper bound of ||a_j||_2
  DevBuf<double> dAt64;   // [L][D][M]
  DevBuf<uint16_t> dAhT, dAlT;  // [Jp][Dp] bf16 hi / lo parts of column j over d
  int Dp = 0;                   // D rounded up to 32
  DevBuf<double> db64;    // [L][M]
  DevBuf<uint64_t> dseeds;
  uint64_t hash_seed = 0x243F6A8885A308D3ull;

  // Index.
  bool indexed = false;
  DevBuf<float> dX_own;
  const float* dX = nullptr;          // resident dataset [N][D]
  const float* hX_indexed = nullptr;  // host pointer index() was given
  unsigned long long data_gen = 0;    // generation of the resident dataset
  // FP16 copy of the resident rows for the batching form's distance pass
  // (built on first use per dataset generation; rows16_ok false when D % 8
  // != 0 or a value leaves the half range) and its bound on ||x - x16||^2.
  DevBuf<uint16_t> dX16;
  unsigned long long rows16_gen = 0;
  bool rows16_ok = false;
  double rows16_e2 = 0.0;
  DevBuf<BucketRec> recs;             // [B_total] sorted by key: key, start, size
  DevBuf<uint32_t> sids;              // [L*N] members, bucket by bucket
  DevBuf<int16_t> bcodes;             // [B_total][M]
  DevBuf<uint64_t> filter;            // key filter words of all tables (ProbeIndex)
  DevBuf<uint2> tfw;                  // [L] first filter word, word count per table
  DevBuf<uint64_t> dir;               // [nslots + 1] packed directory entries
  DevBuf<unsigned char> lines;        // [B_total][lstride] bucket lines
  int lstride = 64;
  std::vector<uint32_t> table_base;   // [L + 1] first global bucket of each table
  int tb = 1, P = 1, shift = 0;
  long long nslots = 0;
  unsigned long long rechecked = 0;

  // Lexicographic numbering (renum_ok) and host mirrors (canon), built on
  // demand.
  bool renum_ok = false;
  bool canon = false;
  DevBuf<int> renumber;  // global bucket -> table-local lexicographic index
  std::vector<unsigned> hB;
  std::vector<std::vector<unsigned>> hTotal, hStart, hSize;
  std::vector<std::vector<float>> hCodes;

  ~culsh_ht() {
    if (stream) (void)hipStreamDestroy(stream);
  }
  long long B_total() const { return table_base.empty() ? 0 : (long long)table_base[L]; }
  void drop_index() {
    indexed = false;
    renum_ok = false;
    canon = false;
    data_gen = next_data_gen();
    dX_own.release();
    dX = nullptr;
    dX16.release();
    rows16_ok = false;
    hX_indexed = nullptr;
    recs.release();
    sids.release();
    bcodes.release();
    filter.release();
    tfw.release();
    dir.release();
    lines.release();
    renumber.release();
    table_base.clear();
    hB.clear();
    hTotal.clear();
    hStart.clear();
    hSize.clear();
    hCodes.clear();
    rechecked = 0;
  }
};

namespace {

int ht_stream(culsh_ht* ht) {
  if (!ht->stream) HIP_TRY(hipStreamCreateWithFlags(&ht->stream, hipStreamNonBlocking));
  return CULSH_OK;
}

bool params_ok(int N, int D, int L, int M, float W) {
  return N > 0 && D > 0 && L > 0 && M > 0 && W > 0.0f;
}

int check_limits(int N, int D, int L, int M) {
  const std::string m = culsh_file::limits_message(N, D, L, M);
  if (!m.empty()) return fail(CULSH_E_LIMIT, "%s", m.c_str());
  return CULSH_OK;
}

// Uploads the projection parameters in the layouts the kernels read.
// float -> bf16 bits, round to nearest even, and the split x ~ hi + lo used
// by the precision form's projection (|x - hi - lo| <= 2^-16 |x|).
uint16_t host_f2bf(float x) {
  uint32_t b;
  std::memcpy(&b, &x, 4);
  return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}
float host_bf2f(uint16_t h) {
  const uint32_t b = (uint32_t)h << 16;
  float x;
  std::memcpy(&x, &b, 4);
  return x;
}

int upload_params(culsh_ht* ht) {
  TRY(ht_stream(ht));
  const int D = ht->D, L = ht->L, M = ht->M;
  ht->J = L * M;
  ht->Jp = (ht->J + kPjCols - 1) / kPjCols * kPjCols;
  const int Jp = ht->Jp;
  // Ad carries 64 floats of padding after its last row (the probe kernel's
  // lanes past M read it and discard it).
  std::vector<float> Ad((size_t)D * Jp + 64, 0.f), Aorig((size_t)ht->J * D), bj(Jp, 0.f),
      naj(Jp, 0.f);
  std::vector<double> At64((size_t)L * D * M), b64((size_t)L * M);
  std::vector<uint64_t> seeds(L);
  ht->Dp = (D + kPjKc - 1) / kPjKc * kPjKc;
  const int Dp = ht->Dp;
  std::vector<uint16_t> AhT((size_t)Jp * Dp, 0), AlT((size_t)Jp * Dp, 0);
  for (int t = 0; t < L; ++t) {
    const float* At = ht->A.data() + (size_t)t * D * M;
    for (int m = 0; m < M; ++m) {
      const int j = t * M + m;
      double nrm = 0.0;
      for (int d = 0; d < D; ++d) {
        const float a = At[(size_t)m * D + d];
        Ad[(size_t)d * Jp + j] = a;
        Aorig[(size_t)j * D + d] = a;
        const uint16_t hi = host_f2bf(a);
        AhT[(size_t)j * Dp + d] = hi;
        AlT[(size_t)j * Dp + d] = host_f2bf(a - host_bf2f(hi));
        At64[((size_t)t * D + d) * M + m] = (double)a;
        nrm += (double)a * (double)a;
      }
      bj[j] = ht->b[(size_t)t * M + m];
      naj[j] = (float)(std::sqrt(nrm) * (1.0 + 0x1.0p-20));
      b64[(size_t)t * M + m] = (double)ht->b[(size_t)t * M + m];
    }
    seeds[t] = mix64(ht->hash_seed + (uint64_t)t);
  }
  TRY(ht->dAd.exact(Ad.size()));
  TRY(ht->dAorig.exact(Aorig.size()));
  TRY(ht->dbj.exact(Jp));
  TRY(ht->dnaj.exact(Jp));
  TRY(ht->dAt64.exact(At64.size()));
  TRY(ht->db64.exact(b64.size()));
  TRY(ht->dseeds.exact(L));
  TRY(ht->dAhT.exact(AhT.size()));
  TRY(ht->dAlT.exact(AlT.size()));
  HIP_TRY(hipMemcpy(ht->dAhT.p, AhT.data(), AhT.size() * 2, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ht->dAlT.p, AlT.data(), AlT.size() * 2, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ht->dAd.p, Ad.data(), Ad.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ht->dAorig.p, Aorig.data(), Aorig.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ht->dbj.p, bj.data(), Jp * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ht->dnaj.p, naj.data(), Jp * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ht->dAt64.p, At64.data(), At64.size() * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ht->db64.p, b64.data(), b64.size() * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ht->dseeds.p, seeds.data(), L * 8, hipMemcpyHostToDevice));
  return CULSH_OK;
}

int set_params(culsh_ht* ht, int N, int D, int L, int M, float W) {
  ht->drop_index();
  ht->has_params = false;
  ht->N = N;
  ht->D = D;
  ht->L = L;
  ht->M = M;
  ht->W = W;
  if (!params_ok(N, D, L, M, W))
    return fail(CULSH_E_ARG,
                "Projection parameters must be positive! Have you reset the index structure "
                "properly???");
  TRY(check_limits(N, D, L, M));
  return CULSH_OK;
}

// Probe index over the key-sorted bucket records (see ProbeIndex): the key
// filter (ceil(B_t * kFilterBits / 64) words per table), the key directory
// (2^P slots per table with 2^P >= 2 * (largest table's bucket count)) and
// the bucket lines.
int build_probe_index(culsh_ht* ht) {
  hipStream_t s = ht->stream;
  const int L = ht->L, M = ht->M;
  const long long B = ht->B_total();
  std::vector<uint2> tfw(L);
  long long nw = 0, maxB = 1;
  for (int t = 0; t < L; ++t) {
    const long long Bt = (long long)ht->table_base[t + 1] - ht->table_base[t];
    const long long w = std::max<long long>(1, (Bt * kFilterBits + 63) / 64);
    tfw[t] = make_uint2((uint32_t)nw, (uint32_t)w);
    nw += w;
    maxB = std::max(maxB, Bt);
  }
  ht->P = std::max(1, ceil_log2_host(maxB) + 1);
  ht->P = std::min(ht->P, 30 - ht->tb);  // at most 2^30 slots (8 GiB)
  ht->shift = 64 - ht->tb - ht->P;
  ht->nslots = 1ll << (ht->tb + ht->P);
  ht->lstride = bucket_line_bytes(M);
  TRY(ht->tfw.exact(L));
  TRY(ht->filter.exact((size_t)nw));
  TRY(ht->dir.exact(ht->nslots + 1));
  TRY(ht->lines.exact((size_t)std::max<long long>(B, 1) * ht->lstride));
  HIP_TRY(hipMemcpyAsync(ht->tfw.p, tfw.data(), L * sizeof(uint2), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemsetAsync(ht->filter.p, 0, (size_t)nw * 8, s));
  hipLaunchKernelGGL(k_dir_clear, dim3(grid_for(ht->nslots + 1)), dim3(256), 0, s, ht->dir.p,
                     ht->nslots + 1);
  hipLaunchKernelGGL(k_directory, dim3(grid_for(std::max<long long>(B, 1))), dim3(256), 0, s,
                     (const BucketRec*)ht->recs.p, B, ht->shift, ht->nslots, ht->dir.p);
  if (B > 0) {
    hipLaunchKernelGGL(k_filter, dim3(grid_for(B)), dim3(256), 0, s, (const BucketRec*)ht->recs.p,
                       B, ht->tb, (const uint2*)ht->tfw.p, ht->filter.p);
    hipLaunchKernelGGL(k_bucket_lines, dim3(grid_for(B)), dim3(256), 0, s,
                       (const BucketRec*)ht->recs.p, (const int16_t*)ht->bcodes.p, B, M,
                       ht->lstride, (uint4*)ht->lines.p);
  }
  HIP_TRY(hipStreamSynchronize(s));  // tfw's host staging goes out of scope
  return CULSH_OK;
}

ProbeIndex probe_index(const culsh_ht* ht) {
  ProbeIndex P;
  P.filter = ht->filter.p;
  P.tfw = ht->tfw.p;
  P.dir = ht->dir.p;
  P.lines = ht->lines.p;
  P.recs = ht->recs.p;
  P.tb = ht->tb;
  P.shift = ht->shift;
  P.lstride = ht->lstride;
  return P;
}

// Keys -> sort -> buckets -> probe index, from codes [L][N][M] on the device.
int build_buckets(culsh_ht* ht, const DevBuf<int16_t>& codes, FILE* dbg) {
  hipStream_t s = ht->stream;
  const long long N = ht->N, L = ht->L, M = ht->M, n_all = L * N;
  ht->tb = std::max(1, ceil_log2_host(L));

  DevBuf<uint64_t> k0, k1;
  DevBuf<uint32_t> v0, v1, flag, excl, scratch;
  DevBuf<int> dflags;
  TRY(k0.ensure(n_all));
  TRY(k1.ensure(n_all));
  TRY(v0.ensure(n_all));
  TRY(v1.ensure(n_all));
  TRY(flag.ensure(n_all));
  TRY(excl.ensure(n_all));
  TRY(scratch.ensure(std::max(rs_scratch_elems(n_all), scan_scratch_elems(n_all)) + 1));
  TRY(dflags.ensure(1));
  DevBuf<uint32_t> bases;
  TRY(bases.ensure(L + 1));
  for (int attempt = 0;; ++attempt) {
    if (attempt > 0) {
      // 64-bit hash collision between two distinct codes: new seeds.
      ht->hash_seed = mix64(ht->hash_seed + 0x9E3779B97F4A7C15ull);
      std::vector<uint64_t> seeds(L);
      for (int t = 0; t < L; ++t) seeds[t] = mix64(ht->hash_seed + (uint64_t)t);
      HIP_TRY(hipMemcpy(ht->dseeds.p, seeds.data(), L * 8, hipMemcpyHostToDevice));
    }
    if (attempt > 4) return fail(CULSH_E_DEVICE, "bucket key collisions persist after reseeding");
    hipLaunchKernelGGL(k_keys, dim3(grid_for(n_all)), dim3(256), 0, s, codes.p, N, (int)L,
                       (int)M, (const uint64_t*)ht->dseeds.p, ht->tb, k0.p, v0.p);
    uint64_t* kp = k0.p;
    uint32_t* vp = v0.p;
    uint64_t* ka = k1.p;
    uint32_t* va = v1.p;
    // Radix passes over the key's table bits and its top hash bits only:
    // with at least log2(N) + 2 hash bits above the lowest sorted bit, runs
    // of equal sorted bits are a bucket's members plus rare collisions, and
    // k_sort_inversions / k_sort_runs finish the order of the full key.
    const int hbits = ceil_log2_host(N) + 2;
    const int sb = std::max(0, (64 - ht->tb - hbits) / 8 * 8);
    radix_sort_pairs(&kp, &vp, ka, va, n_all, sb, 64, scratch.p, s);
    if (sb > 0) {
      // excl (free until the head scan) holds the bitmap of runs to sort.
      const long long words = (n_all + 31) / 32;
      HIP_TRY(hipMemsetAsync(excl.p, 0, (size_t)words * sizeof(uint32_t), s));
      hipLaunchKernelGGL(k_sort_inversions, dim3(grid_for(n_all)), dim3(256), 0, s,
                         (const uint64_t*)kp, n_all, sb, excl.p);
      hipLaunchKernelGGL(k_sort_runs, dim3(grid_for(words)), dim3(256), 0, s, kp, vp, n_all, sb,
                         (const uint32_t*)excl.p);
    }
    if (kp != k0.p) {
      std::swap(k0.p, k1.p);
      std::swap(v0.p, v1.p);
    }
    hipLaunchKernelGGL(k_heads, dim3(grid_for(n_all)), dim3(256), 0, s, (const uint64_t*)k0.p,
                       n_all, flag.p);
    scan_exclusive(flag.p, excl.p, n_all, scratch.p, s);
    hipLaunchKernelGGL(k_table_bases, dim3(1), dim3(std::max(64, (int)((L + 64) / 64 * 64))), 0,
                       s, (const uint32_t*)flag.p, (const uint32_t*)excl.p, N, (int)L, bases.p);
    ht->table_base.assign(L + 1, 0);
    HIP_TRY(hipMemcpyAsync(ht->table_base.data(), bases.p, (L + 1) * 4, hipMemcpyDeviceToHost,
                           s));
    HIP_TRY(hipStreamSynchronize(s));
    const long long B = ht->table_base[L];
    TRY(ht->recs.exact(B));
    TRY(ht->bcodes.exact(B * M));
    HIP_TRY(hipMemsetAsync(dflags.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_bucket_fill, dim3(grid_for(n_all)), dim3(256), 0, s,
                       (const uint64_t*)k0.p, (const uint32_t*)v0.p, (const uint32_t*)flag.p,
                       (const uint32_t*)excl.p, n_all, N, (int)M, ht->tb, codes.p, ht->recs.p,
                       ht->bcodes.p, dflags.p);
    hipLaunchKernelGGL(k_rec_sizes, dim3(grid_for(B)), dim3(256), 0, s, ht->recs.p, B, n_all);
    int collision = 0;
    HIP_TRY(hipMemcpyAsync(&collision, dflags.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (!collision) break;
    if (dbg) fprintf(dbg, "[HashTables::index]\tbucket key collision, reseeding\n");
  }
  // The sorted member ids become the index's id array.
  ht->sids.release();
  ht->sids.p = v0.p;
  ht->sids.n = v0.n;
  v0.p = nullptr;
  v0.n = 0;
  TRY(build_probe_index(ht));
  HIP_TRY(hipStreamSynchronize(s));
  HIP_TRY(hipGetLastError());
  return CULSH_OK;
}

// Projection of the resident dataset into codes [L][N][M].
int project_dataset(culsh_ht* ht, DevBuf<int16_t>& codes, FILE* dbg) {
  hipStream_t s = ht->stream;
  const long long N = ht->N;
  const int D = ht->D, M = ht->M, J = ht->J, Jp = ht->Jp;
  TRY(codes.exact((size_t)N * J));
  DevBuf<float> nx;
  TRY(nx.ensure(N));
  hipLaunchKernelGGL(k_row_norms, dim3(grid_for(N, 16)), dim3(256), 0, s, ht->dX, N, D, nx.p);
  // Recheck list: kRcShards shards of `seg` entries, one counter each.
  const unsigned long long seg =
      std::max<unsigned long long>(1ull << 20, (unsigned long long)(N * J / 64)) / kRcShards;
  const unsigned long long cap = seg * kRcShards;
  DevBuf<unsigned long long> list, cnt;
  DevBuf<int> err;
  TRY(list.ensure(cap));
  TRY(cnt.ensure(kRcShards));
  TRY(err.ensure(1));
  HIP_TRY(hipMemsetAsync(cnt.p, 0, 8 * kRcShards, s));
  HIP_TRY(hipMemsetAsync(err.p, 0, 4, s));
  const long long nrt = (N + kPjRows - 1) / kPjRows;
  const long long tiles = nrt * (Jp / kPjCols);
  const long long grid = (tiles + 7) / 8 * 8;
  if (grid >= (1ll << 31)) return fail(CULSH_E_LIMIT, "projection grid too large");
  if (ht->form == CULSH_FORM_PRECISION)
    hipLaunchKernelGGL(k_project_bf16x3, dim3((unsigned)grid), dim3(256), 0, s, ht->dX, N, D,
                       (const uint16_t*)ht->dAhT.p, (const uint16_t*)ht->dAlT.p, ht->Dp, Jp, J,
                       M, (const float*)ht->dbj.p, (const float*)ht->dnaj.p, (const float*)nx.p,
                       ht->W, codes.p, list.p, cnt.p, seg, err.p);
  else
    hipLaunchKernelGGL(k_project_f32, dim3((unsigned)grid), dim3(256), 0, s, ht->dX, N, D,
                       (const float*)ht->dAd.p, Jp, J, M, (const float*)ht->dbj.p,
                       (const float*)ht->dnaj.p, (const float*)nx.p, ht->W, codes.p, list.p,
                       cnt.p, seg, err.p);
  unsigned long long counts[kRcShards];
  HIP_TRY(hipMemcpyAsync(counts, cnt.p, sizeof(counts), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  HIP_TRY(hipGetLastError());
  unsigned long long count = 0;
  bool overflow = false;
  for (int k = 0; k < kRcShards; ++k) {
    count += counts[k];
    overflow |= counts[k] > seg;
  }
  ht->rechecked = count;
  if (overflow) {
    if (dbg) fprintf(dbg, "[HashTables::index]\t%llu boundary projections; full FP64 pass\n", count);
    const long long t64 = ((N + kP64Tile - 1) / kP64Tile) * ((J + kP64Tile - 1) / kP64Tile);
    if (t64 >= (1ll << 31)) return fail(CULSH_E_LIMIT, "FP64 projection grid too large");
    hipLaunchKernelGGL(k_project_f64_tiled, dim3((unsigned)t64), dim3(256), 0, s, ht->dX, N, D,
                       (const float*)ht->dAorig.p, (const float*)ht->dbj.p, ht->W, M, J, codes.p,
                       err.p);
  } else {
    for (int k = 0; k < kRcShards; ++k)
      if (counts[k] > 0)
        hipLaunchKernelGGL(k_recheck, dim3(grid_for((long long)counts[k])), dim3(256), 0, s,
                           ht->dX, N, D, (const float*)ht->dAorig.p, (const float*)ht->dbj.p,
                           ht->W, M, (const unsigned long long*)list.p + k * seg, counts[k],
                           codes.p, err.p);
  }
  int herr = 0;
  HIP_TRY(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  HIP_TRY(hipGetLastError());
  if (herr)
    return fail(CULSH_E_RANGE, "a hash code falls outside [-32766, 32766] (W too small for the data?)");
  if (dbg) fprintf(dbg, "[HashTables::index]\tprojected %lld x %d, %llu FP64 rechecks\n", N, J, count);
  return CULSH_OK;
}

// The FP16 copy of the resident rows for the current dataset generation
// (one conversion pass and one host synchronisation on first use).
int ensure_rows16(culsh_ht* ht) {
  if (ht->rows16_gen == ht->data_gen) return CULSH_OK;
  ht->rows16_ok = false;
  ht->dX16.release();
  ht->rows16_gen = ht->data_gen;
  if (!ht->dX || ht->D % 8 != 0) return CULSH_OK;
  hipStream_t s = ht->stream;
  const long long N = ht->N;
  DevBuf<unsigned long long> e2;
  if (ht->dX16.exact((size_t)N * ht->D) != CULSH_OK || e2.ensure(1) != CULSH_OK) {
    // The copy only saves bytes: without the memory for it the batches
    // read the FP32 rows.
    ht->dX16.release();
    (void)hipGetLastError();
    return CULSH_OK;
  }
  HIP_TRY(hipMemsetAsync(e2.p, 0, 8, s));
  hipLaunchKernelGGL(k_rows_f16, dim3(std::min<unsigned>(grid_for(N, 16), 4096u)), dim3(256), 0,
                     s, ht->dX, N, ht->D, ht->dX16.p, e2.p);
  unsigned long long bits = 0;
  HIP_TRY(hipMemcpyAsync(&bits, e2.p, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  HIP_TRY(hipGetLastError());
  double v;
  std::memcpy(&v, &bits, 8);
  if (std::isfinite(v)) {
    ht->rows16_ok = true;
    ht->rows16_e2 = v;
  } else {
    ht->dX16.release();
  }
  return CULSH_OK;
}

int index_resident(culsh_ht* ht, FILE* dbg) {
  DevBuf<int16_t> codes;
  TRY(project_dataset(ht, codes, dbg));
  TRY(build_buckets(ht, codes, dbg));
  ht->indexed = true;
  ht->renum_ok = false;
  ht->canon = false;
  if (dbg)
    fprintf(dbg, "[HashTables::index]\t%d tables indexed, %lld buckets in total\n", ht->L,
            ht->B_total());
  return CULSH_OK;
}

// Lexicographic order of every table's buckets (global bucket indices,
// table by table) into perm [B]: the reference's bucket order
// (cuLSH_Indexing.cu:71-93), computed on the device by stable LSD passes
// over keys that each pack as many code columns as fit 64 bits, last
// columns first, then a pass on the table index.
int lex_order(culsh_ht* ht, DevBuf<uint32_t>& perm) {
  hipStream_t s = ht->stream;
  const long long B = ht->B_total();
  const int M = ht->M;
  DevBuf<uint64_t> k0, k1;
  DevBuf<uint32_t> p1, scr;
  DevBuf<int> rng;
  TRY(perm.exact(B));
  TRY(p1.exact(B));
  TRY(k0.exact(B));
  TRY(k1.exact(B));
  TRY(scr.exact(rs_scratch_elems(B) + 1));
  TRY(rng.exact(2));
  hipLaunchKernelGGL(k_lex_init, dim3(1), dim3(64), 0, s, rng.p);
  const long long nc = B * M;
  hipLaunchKernelGGL(k_code_range, dim3((unsigned)std::min<long long>(grid_for(nc), 4096)),
                     dim3(256), 0, s, (const int16_t*)ht->bcodes.p, nc, rng.p);
  int range[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(range, rng.p, sizeof(range), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const int w = std::max(1, ceil_log2_host((long long)range[1] - range[0] + 1));
  const int per_key = std::max(1, 64 / w);
  uint64_t *ka = k0.p, *kb = k1.p;
  uint32_t *pa = perm.p, *pb = p1.p;
  auto pass = [&](int bits) {
    uint64_t* kk = ka;
    uint32_t* vv = pa;
    radix_sort_pairs(&kk, &vv, kb, pb, B, 0, (bits + 7) / 8 * 8, scr.p, s);
    if (kk != ka) {
      std::swap(ka, kb);
      std::swap(pa, pb);
    }
  };
  hipLaunchKernelGGL(k_iota, dim3(grid_for(B)), dim3(256), 0, s, pa, B);
  for (int hi = M; hi > 0; hi -= per_key) {
    const int m0 = std::max(0, hi - per_key), ncols = hi - m0;
    hipLaunchKernelGGL(k_lex_key, dim3(grid_for(B)), dim3(256), 0, s, (const uint32_t*)pa, B,
                       (const int16_t*)ht->bcodes.p, M, m0, ncols, w, range[0], ka);
    pass(ncols * w);
  }
  hipLaunchKernelGGL(k_lex_table_key, dim3(grid_for(B)), dim3(256), 0, s, (const uint32_t*)pa, B,
                     (const BucketRec*)ht->recs.p, ht->tb, ka);
  pass(ht->tb);
  if (pa != perm.p) std::swap(perm.p, p1.p);
  HIP_TRY(hipGetLastError());
  return CULSH_OK;
}

// The table-local lexicographic index of every bucket (for the reference's
// query-bucket numbering), on the device.
int ensure_renumber(culsh_ht* ht) {
  if (ht->renum_ok) return CULSH_OK;
  if (!ht->indexed) return fail(CULSH_E_STATE, "no index: call index() or load() first");
  const long long B = ht->B_total();
  DevBuf<uint32_t> perm, tbase;
  TRY(lex_order(ht, perm));
  TRY(tbase.exact(ht->L + 1));
  HIP_TRY(hipMemcpyAsync(tbase.p, ht->table_base.data(), (ht->L + 1) * 4, hipMemcpyHostToDevice,
                         ht->stream));
  TRY(ht->renumber.exact(std::max<long long>(B, 1)));
  hipLaunchKernelGGL(k_lex_renumber, dim3(grid_for(B)), dim3(256), 0, ht->stream,
                     (const uint32_t*)perm.p, B, (const BucketRec*)ht->recs.p, ht->tb,
                     (const uint32_t*)tbase.p, ht->renumber.p);
  HIP_TRY(hipStreamSynchronize(ht->stream));
  HIP_TRY(hipGetLastError());
  ht->renum_ok = true;
  return CULSH_OK;
}

// Lexicographic numbering and the host mirrors of the bucket arrays that the
// reference's getters and save() return: built on the device in the
// reference's layout, then copied to host memory.
int ensure_canonical(culsh_ht* ht) {
  if (ht->canon) return CULSH_OK;
  if (!ht->indexed) return fail(CULSH_E_STATE, "no index: call index() or load() first");
  hipStream_t s = ht->stream;
  const long long N = ht->N, L = ht->L, M = ht->M, B = ht->B_total();
  DevBuf<uint32_t> perm, tbase, sz, gst, lst, tot, scr;
  DevBuf<float> codes;
  TRY(lex_order(ht, perm));
  TRY(tbase.exact(L + 1));
  TRY(sz.exact(B));
  TRY(gst.exact(B));
  TRY(lst.exact(B));
  TRY(tot.exact((size_t)L * N));
  TRY(codes.exact((size_t)B * M));
  TRY(scr.exact(scan_scratch_elems(B) + 1));
  HIP_TRY(hipMemcpyAsync(tbase.p, ht->table_base.data(), (L + 1) * 4, hipMemcpyHostToDevice, s));
  TRY(ht->renumber.exact(std::max<long long>(B, 1)));
  const uint32_t* pp = perm.p;
  const BucketRec* rp = ht->recs.p;
  hipLaunchKernelGGL(k_lex_renumber, dim3(grid_for(B)), dim3(256), 0, s, pp, B, rp, ht->tb,
                     (const uint32_t*)tbase.p, ht->renumber.p);
  hipLaunchKernelGGL(k_lex_sizes, dim3(grid_for(B)), dim3(256), 0, s, pp, B, rp, sz.p);
  scan_exclusive(sz.p, gst.p, B, scr.p, s);
  hipLaunchKernelGGL(k_lex_starts, dim3(grid_for(B)), dim3(256), 0, s, pp, B, rp, ht->tb, N,
                     (const uint32_t*)gst.p, lst.p);
  hipLaunchKernelGGL(k_lex_members, dim3(grid_for(B)), dim3(256), 0, s, pp, B, rp,
                     (const uint32_t*)ht->sids.p, (const uint32_t*)gst.p, tot.p);
  hipLaunchKernelGGL(k_lex_codes, dim3(grid_for(B * M)), dim3(256), 0, s, pp, B,
                     (const int16_t*)ht->bcodes.p, (int)M, rp, ht->tb, (const uint32_t*)tbase.p,
                     codes.p);
  HIP_TRY(hipGetLastError());
  ht->hB.assign(L, 0);
  ht->hTotal.assign(L, {});
  ht->hStart.assign(L, {});
  ht->hSize.assign(L, {});
  ht->hCodes.assign(L, {});
  for (int t = 0; t < L; ++t) {
    const long long b0 = ht->table_base[t], Bt = (long long)ht->table_base[t + 1] - b0;
    ht->hB[t] = (unsigned)Bt;
    ht->hTotal[t].resize(N);
    ht->hStart[t].resize(Bt);
    ht->hSize[t].resize(Bt);
    ht->hCodes[t].resize((size_t)Bt * M);
    HIP_TRY(hipMemcpyAsync(ht->hTotal[t].data(), tot.p + (size_t)t * N, N * 4,
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(ht->hStart[t].data(), lst.p + b0, Bt * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(ht->hSize[t].data(), sz.p + b0, Bt * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(ht->hCodes[t].data(), codes.p + (size_t)b0 * M, (size_t)Bt * M * 4,
                           hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(hipStreamSynchronize(s));
  ht->renum_ok = true;
  ht->canon = true;
  return CULSH_OK;
}

}  // namespace

extern "C" {

const char* culsh_last_error(void) { return g_err.c_str(); }
const char* culsh_version(void) { return "culsh-mi355x 0.1 (gfx950)"; }

int culsh_ht_create(culsh_ht** out) {
  if (!out) return fail(CULSH_E_ARG, "null output pointer");
  *out = new culsh_ht;
  return CULSH_OK;
}

void culsh_ht_destroy(culsh_ht* ht) { delete ht; }

int culsh_ht_reset(culsh_ht* ht, int N, int D, int L, int M, float W, const float* const* A,
                   const float* const* b, FILE* dbg) {
  if (!ht) return fail(CULSH_E_ARG, "null HashTables");
  TRY(set_params(ht, N, D, L, M, W));
  if (!A || !b) return fail(CULSH_E_ARG, "null A or b");
  ht->A.resize((size_t)L * D * M);
  ht->b.resize((size_t)L * M);
  for (int t = 0; t < L; ++t) {
    if (!A[t] || !b[t]) return fail(CULSH_E_ARG, "null A[%d] or b[%d]", t, t);
    std::memcpy(ht->A.data() + (size_t)t * D * M, A[t], sizeof(float) * D * M);
    std::memcpy(ht->b.data() + (size_t)t * M, b[t], sizeof(float) * M);
  }
  TRY(upload_params(ht));
  ht->has_params = true;
  if (dbg)
    fprintf(dbg,
            "[HashTables::reset]\tIndexing structure reset with parameters (N, D, L, M, W) = (%d, "
            "%d, %d, %d, %.2f), matrices [A] & vectors [b] have been manually set.\n",
            N, D, L, M, W);
  return CULSH_OK;
}

int culsh_ht_reset_random(culsh_ht* ht, int N, int D, int L, int M, float W,
                          unsigned long long seed, FILE* dbg) {
  if (!ht) return fail(CULSH_E_ARG, "null HashTables");
  TRY(set_params(ht, N, D, L, M, W));
  // cuLSH_HashTables.cu:122-187: A ~ N(0,1); b = (U - min U) * W, U ~ (0,1].
  ht->A.resize((size_t)L * D * M);
  ht->b.resize((size_t)L * M);
  for (size_t i = 0; i < ht->A.size(); ++i) ht->A[i] = (float)culsh_rng::normal(seed, 101, i);
  for (int t = 0; t < L; ++t) {
    float* bt = ht->b.data() + (size_t)t * M;
    for (int m = 0; m < M; ++m)
      bt[m] = (float)culsh_rng::uniform_open0(seed, 102, (uint64_t)t * M + m);
    float mn = bt[0];
    for (int m = 1; m < M; ++m) mn = std::min(mn, bt[m]);
    for (int m = 0; m < M; ++m) bt[m] = (bt[m] - mn) * W;
  }
  TRY(upload_params(ht));
  ht->has_params = true;
  if (dbg)
    fprintf(dbg,
            "[HashTables::reset]\tIndexing structure reset with parameters (N, D, L, M, W) = (%d, "
            "%d, %d, %d, %.2f), and random matrices [A] & vectors [b].\n",
            N, D, L, M, W);
  return CULSH_OK;
}

static int index_common(culsh_ht* ht, const float* matrix, bool device, FILE* dbg) {
  if (!ht) return fail(CULSH_E_ARG, "null HashTables");
  if (!params_ok(ht->N, ht->D, ht->L, ht->M, ht->W))
    return fail(CULSH_E_STATE,
                "Parameters must be positive. Have you reset the index structure properly???");
  if (!ht->has_params)
    return fail(CULSH_E_STATE,
                "Pointers to projection matrices [A] and vectors [b] must be valid. Have you "
                "reset the index structure properly???");
  if (!matrix) return fail(CULSH_E_ARG, "null dataset");
  TRY(ht_stream(ht));
  ht->drop_index();
  const size_t n = (size_t)ht->N * ht->D;
  if (device) {
    ht->dX = matrix;
  } else {
    TRY(ht->dX_own.exact(n));
    HIP_TRY(hipMemcpy(ht->dX_own.p, matrix, n * 4, hipMemcpyHostToDevice));
    ht->dX = ht->dX_own.p;
    ht->hX_indexed = matrix;
  }
  ht->data_gen = next_data_gen();
  int rc = index_resident(ht, dbg);
  if (rc != CULSH_OK) {
    std::string keep = g_err;
    ht->drop_index();
    g_err = keep;
  }
  return rc;
}

int culsh_ht_index(culsh_ht* ht, const float* matrix, FILE* dbg) {
  return index_common(ht, matrix, false, dbg);
}

int culsh_ht_index_device(culsh_ht* ht, const float* d_matrix, FILE* dbg) {
  return index_common(ht, d_matrix, true, dbg);
}

int culsh_ht_N(const culsh_ht* ht) { return ht ? ht->N : 0; }
int culsh_ht_D(const culsh_ht* ht) { return ht ? ht->D : 0; }
int culsh_ht_L(const culsh_ht* ht) { return ht ? ht->L : 0; }
int culsh_ht_M(const culsh_ht* ht) { return ht ? ht->M : 0; }
float culsh_ht_W(const culsh_ht* ht) { return ht ? ht->W : 0.f; }

const float* culsh_ht_A(culsh_ht* ht, int t) {
  if (!ht || !ht->has_params || t < 0 || t >= ht->L) return nullptr;
  return ht->A.data() + (size_t)t * ht->D * ht->M;
}
const float* culsh_ht_b(culsh_ht* ht, int t) {
  if (!ht || !ht->has_params || t < 0 || t >= ht->L) return nullptr;
  return ht->b.data() + (size_t)t * ht->M;
}

unsigned culsh_ht_buckets(culsh_ht* ht, int t) {
  if (!ht || !ht->indexed || t < 0 || t >= ht->L) return 0;
  return ht->table_base[t + 1] - ht->table_base[t];
}
const unsigned* culsh_ht_buckets_all(culsh_ht* ht) {
  if (!ht || ensure_canonical(ht) != CULSH_OK) return nullptr;
  return ht->hB.data();
}
const unsigned* culsh_ht_buckets_total_indices(culsh_ht* ht, int t) {
  if (!ht || t < 0 || t >= ht->L || ensure_canonical(ht) != CULSH_OK) return nullptr;
  return ht->hTotal[t].data();
}
const unsigned* culsh_ht_buckets_starting_indices(culsh_ht* ht, int t) {
  if (!ht || t < 0 || t >= ht->L || ensure_canonical(ht) != CULSH_OK) return nullptr;
  return ht->hStart[t].data();
}
const unsigned* culsh_ht_buckets_sizes(culsh_ht* ht, int t) {
  if (!ht || t < 0 || t >= ht->L || ensure_canonical(ht) != CULSH_OK) return nullptr;
  return ht->hSize[t].data();
}
const float* culsh_ht_buckets_codes(culsh_ht* ht, int t) {
  if (!ht || t < 0 || t >= ht->L || ensure_canonical(ht) != CULSH_OK) return nullptr;
  return ht->hCodes[t].data();
}
unsigned long long culsh_ht_rechecked(const culsh_ht* ht) { return ht ? ht->rechecked : 0; }

int culsh_ht_set_form(culsh_ht* ht, int form) {
  if (!ht) return fail(CULSH_E_ARG, "null HashTables");
  if (form != CULSH_FORM_PLAIN && form != CULSH_FORM_PRECISION)
    return fail(CULSH_E_ARG, "unknown form %d", form);
  ht->form = form;
  return CULSH_OK;
}

// File layout of HashTables::save / load (cuLSH_HashTables.cu:395-483),
// native endianness, no header:
//   int32 N, D, L, M; float W
//   per table: float A_t[D*M], float b_t[M]
//   uint32 B[L]
//   per table: uint32 totalIndices[N], startingIndices[B_t], sizes[B_t],
//              float codes[B_t*M] (column-major [B_t x M])
int culsh_ht_save(culsh_ht* ht, const char* filename, FILE* dbg) {
  if (!ht || !filename) return fail(CULSH_E_ARG, "null argument");
  if (dbg)
    fprintf(dbg, "[HashTables::save]\tSaving indexing structure to file \"%s\"...\n", filename);
  if (!ht->has_params)
    return fail(CULSH_E_STATE,
                "Pointers to matrices [A] and/or vectors [b] are null. Have you reset the index "
                "structure properly???");
  if (!ht->indexed)
    return fail(CULSH_E_STATE, "Pointers to bucket objects are null. Have you performed indexing???");
  TRY(ensure_canonical(ht));
  FILE* fp = fopen(filename, "wb");
  if (!fp)
    return fail(CULSH_E_IO, "Could not create file \"%s\" in order to store the hash tables...",
                filename);
  const int L = ht->L, M = ht->M, D = ht->D, N = ht->N;
  bool ok = fwrite(&ht->N, 4, 1, fp) == 1 && fwrite(&ht->D, 4, 1, fp) == 1 &&
            fwrite(&ht->L, 4, 1, fp) == 1 && fwrite(&ht->M, 4, 1, fp) == 1 &&
            fwrite(&ht->W, 4, 1, fp) == 1;
  for (int t = 0; ok && t < L; ++t)
    ok = fwrite(ht->A.data() + (size_t)t * D * M, 4, (size_t)D * M, fp) == (size_t)D * M &&
         fwrite(ht->b.data() + (size_t)t * M, 4, M, fp) == (size_t)M;
  ok = ok && fwrite(ht->hB.data(), 4, L, fp) == (size_t)L;
  for (int t = 0; ok && t < L; ++t) {
    const size_t B = ht->hB[t];
    ok = fwrite(ht->hTotal[t].data(), 4, N, fp) == (size_t)N &&
         fwrite(ht->hStart[t].data(), 4, B, fp) == B &&
         fwrite(ht->hSize[t].data(), 4, B, fp) == B &&
         fwrite(ht->hCodes[t].data(), 4, B * M, fp) == B * M;
  }
  ok = ok && !ferror(fp);
  if (fclose(fp) != 0) ok = false;
  if (!ok) return fail(CULSH_E_IO, "writing \"%s\" failed", filename);
  return CULSH_OK;
}

namespace {

// Device index from lexicographic bucket arrays (the saved form): bucket
// keys, records sorted by key, member ids in key order, codes, directory,
// and the lexicographic numbering for the host mirrors.
int index_from_buckets(culsh_ht* ht, FILE* dbg) {
  const long long N = ht->N, L = ht->L, M = ht->M;
  ht->tb = std::max(1, ceil_log2_host(L));
  long long Btot = 0;
  std::vector<long long> lexbase(L + 1, 0);
  for (int t = 0; t < L; ++t) lexbase[t + 1] = lexbase[t] + ht->hB[t];
  Btot = lexbase[L];
  std::vector<int16_t> codes16((size_t)Btot * M);
  for (int t = 0; t < L; ++t) {
    const long long Bt = ht->hB[t];
    for (long long k = 0; k < Bt; ++k)
      for (int m = 0; m < M; ++m) {
        const float c = ht->hCodes[t][(size_t)m * Bt + k];
        if (!(c == std::floor(c)) || std::fabs(c) > (float)(kCodeMax - 1))
          return fail(CULSH_E_RANGE, "table %d bucket %lld: code %g is not a 16-bit integer", t,
                      k, (double)c);
        codes16[(size_t)(lexbase[t] + k) * M + m] = (int16_t)c;
      }
  }
  std::vector<BucketRec> recs(Btot);
  std::vector<uint32_t> sids((size_t)L * N);
  std::vector<int> renum(Btot);
  std::vector<int16_t> bcodes((size_t)Btot * M);
  for (int attempt = 0;; ++attempt) {
    if (attempt > 4) return fail(CULSH_E_DEVICE, "bucket key collisions persist after reseeding");
    std::vector<uint64_t> seeds(L);
    for (int t = 0; t < L; ++t) seeds[t] = mix64(ht->hash_seed + (uint64_t)t);
    std::vector<std::pair<uint64_t, uint32_t>> order(Btot);  // (key, lexicographic global index)
    for (int t = 0; t < L; ++t)
      for (long long k = lexbase[t]; k < lexbase[t + 1]; ++k) {
        uint64_t h = 0;
        for (int m = 0; m < M; ++m) h += code_term(seeds[t], m, codes16[(size_t)k * M + m]);
        order[k] = {make_key(h, t, ht->tb), (uint32_t)k};
      }
    std::sort(order.begin(), order.end());
    bool collision = false;
    for (long long i = 1; i < Btot; ++i)
      if (order[i].first == order[i - 1].first) collision = true;
    if (collision) {
      if (dbg) fprintf(dbg, "[HashTables::load]\tbucket key collision, reseeding\n");
      ht->hash_seed = mix64(ht->hash_seed + 0x9E3779B97F4A7C15ull);
      continue;
    }
    uint32_t pos = 0;
    for (long long g = 0; g < Btot; ++g) {
      const uint32_t lk = order[g].second;
      const int t = (int)(order[g].first >> (64 - ht->tb));
      const long long k = lk - lexbase[t];
      const uint32_t st = ht->hStart[t][k], sz = ht->hSize[t][k];
      recs[g] = BucketRec{order[g].first, pos, sz};
      for (uint32_t r = 0; r < sz; ++r) sids[pos + r] = ht->hTotal[t][st + r];
      pos += sz;
      std::memcpy(&bcodes[(size_t)g * M], &codes16[(size_t)lk * M], M * sizeof(int16_t));
      renum[g] = (int)k;
    }
    HIP_TRY(hipMemcpy(ht->dseeds.p, seeds.data(), L * 8, hipMemcpyHostToDevice));
    break;
  }
  ht->table_base.assign(L + 1, 0);
  for (int t = 0; t <= L; ++t) ht->table_base[t] = (uint32_t)lexbase[t];
  TRY(ht->recs.exact(Btot));
  TRY(ht->bcodes.exact(Btot * M));
  TRY(ht->sids.exact((size_t)L * N));
  TRY(ht->renumber.exact(std::max<long long>(Btot, 1)));
  HIP_TRY(hipMemcpy(ht->recs.p, recs.data(), Btot * sizeof(BucketRec), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ht->bcodes.p, bcodes.data(), bcodes.size() * 2, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ht->sids.p, sids.data(), sids.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ht->renumber.p, renum.data(), Btot * 4, hipMemcpyHostToDevice));
  TRY(build_probe_index(ht));
  HIP_TRY(hipStreamSynchronize(ht->stream));
  HIP_TRY(hipGetLastError());
  ht->indexed = true;
  ht->renum_ok = true;
  ht->canon = true;
  return CULSH_OK;
}

int load_file(culsh_ht* ht, FILE* fp, const char* filename, FILE* dbg) {
  // The whole file is read and checked on the host first.
  culsh_file::IndexFile f;
  std::string err;
  const int rc = culsh_file::read_index_file(fp, filename, f, err);
  if (rc != CULSH_OK) return fail(rc, "%s", err.c_str());
  if (dbg)
    fprintf(dbg,
            "[HashTables::load]\tIndexing structure loaded from file \"%s\", (N, D, L, M, W) = "
            "(%d, %d, %d, %d, %.2f)\n",
            filename, f.N, f.D, f.L, f.M, f.W);
  TRY(set_params(ht, f.N, f.D, f.L, f.M, f.W));
  ht->A = std::move(f.A);
  ht->b = std::move(f.b);
  TRY(upload_params(ht));
  ht->has_params = true;
  ht->hB = std::move(f.B);
  ht->hTotal = std::move(f.total);
  ht->hStart = std::move(f.start);
  ht->hSize = std::move(f.size);
  ht->hCodes = std::move(f.codes);
  return index_from_buckets(ht, dbg);
}

}  // namespace

int culsh_ht_load(culsh_ht* ht, const char* filename, FILE* dbg) {
  if (!ht || !filename) return fail(CULSH_E_ARG, "null argument");
  if (dbg)
    fprintf(dbg, "[HashTables::reset]\tLoading indexing structure from file \"%s\"...\n",
            filename);
  FILE* fp = fopen(filename, "rb");
  if (!fp)
    return fail(CULSH_E_IO,
                "Could not open file \"%s\" in order to load the hash tables, or file does not "
                "exist...",
                filename);
  int rc = load_file(ht, fp, filename, dbg);
  fclose(fp);
  if (rc != CULSH_OK) {
    std::string keep = g_err;
    ht->drop_index();
    g_err = keep;
  }
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// SearchTables
// ---------------------------------------------------------------------------
struct culsh_st {
  culsh_ht* ht = nullptr;
  int K = 0, T = 1, Q = 0;
  int form = CULSH_FORM_PLAIN;
  hipStream_t stream = nullptr;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  // Host search(): per-phase shader-clock totals of the candidate stage
  // (union, FP32 distances, top-K), gathered while `stamp` is set.
  DevBuf<unsigned long long> dphase;
  bool stamp = false;
  int max_pass = 0;  // queries per pass cap (0: from the free HBM)
  DevBuf<float> dQ, dX_alt, ddists;
  DevBuf<int> dqb, dids, dov, dqb_local;
  DevBuf<uint2> dqr;  // probe member ranges (start, size) for the candidate stage
  DevBuf<unsigned long long> dcnu_qt, dcnu_q, dcu_q, dcnt, dmax;
  DevBuf<uint32_t> dscratch;  // working sets of overflowed queries
  DevBuf<int> dnov;
  // Asynchronous device searches: fixed overflow working sets, a sticky
  // device error flag, its pinned host copy and the completion event.
  DevBuf<uint32_t> dovfix;
  DevBuf<int> derr;
  int* herr = nullptr;
  hipEvent_t evdone = nullptr;
  bool pending = false;
  // Precision form: locality order of the queries and group working sets.
  DevBuf<uint64_t> dlk0, dlk1;
  DevBuf<uint32_t> dlv0, dlv1, dlscr, dgscratch;
  bool gscratch_clean = false;  // group hash slots hold their empty state
  size_t g_region = 0;           // ... for this per-workgroup region size
  int g_hbits = 0;               // ... and hash size
  // A search dataset uploaded by search() (when it is not the index's
  // resident one): the host pointer, index and index generation it was
  // uploaded for, and its own data generation.
  const float* alt_host = nullptr;
  const culsh_ht* alt_ht = nullptr;
  unsigned long long alt_ht_gen = 0, alt_gen = 0;
  // |x|^2 and max |x| of the dataset the last precision search used, keyed
  // on (device pointer, data generation).
  DevBuf<float> dxn2;
  DevBuf<unsigned> dxmax;
  const float* xn2_for = nullptr;
  unsigned long long xn2_gen = 0;
  float xmax = 0.f;
  std::vector<int> ids;
  std::vector<float> dists;
  std::vector<std::vector<int>> qbuckets;
  std::vector<unsigned long long> cnu, cu;
  unsigned long long tot_u = 0, tot_nu = 0, rechecked = 0;
  culsh_times times{};
  ~culsh_st() {
    if (pending && evdone) (void)hipEventSynchronize(evdone);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (evdone) (void)hipEventDestroy(evdone);
    if (herr) (void)hipHostFree(herr);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void clear_results() {
    ids.clear();
    dists.clear();
    qbuckets.clear();
    cnu.clear();
    cu.clear();
    tot_u = tot_nu = rechecked = 0;
    Q = 0;
  }
};

namespace {

int st_setup(culsh_st* st) {
  if (!st->stream) HIP_TRY(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));
  for (auto& e : st->ev)
    if (!e) HIP_TRY(hipEventCreate(&e));
  if (!st->evdone) HIP_TRY(hipEventCreate(&st->evdone));
  if (!st->herr) {
    HIP_TRY(hipHostMalloc((void**)&st->herr, sizeof(int), hipHostMallocDefault));
    *st->herr = 0;
  }
  if (!st->derr.p) {
    TRY(st->derr.ensure(1));
    HIP_TRY(hipMemset(st->derr.p, 0, sizeof(int)));
  }
  return CULSH_OK;
}

// Reports (once) an error of an earlier asynchronous search whose work has
// completed; with `block`, first waits for the last one.
int st_check_async(culsh_st* st, bool block) {
  if (!st->pending) return CULSH_OK;
  if (block) {
    HIP_TRY(hipEventSynchronize(st->evdone));
  } else {
    const hipError_t q = hipEventQuery(st->evdone);
    if (q == hipErrorNotReady) return CULSH_OK;
    HIP_TRY(q);
  }
  st->pending = false;
  if (*st->herr) {
    *st->herr = 0;
    HIP_TRY(hipMemset(st->derr.p, 0, sizeof(int)));
    return fail(CULSH_E_LIMIT,
                "an earlier asynchronous search had a query with more than %u non-unique "
                "candidates or more than %d candidates within the FP32 error bound of its "
                "K-th distance (its results were not written)",
                kOvAsyncUcap, kScapG);
  }
  return CULSH_OK;
}

// Queries whose candidates overflowed the LDS kernel: global-memory working
// sets sized from their largest C_nu, as many workgroups as fit in a 4 GiB
// scratch budget.
int run_overflow(culsh_st* st, const float* dQ, const float* dX, int Q, int* d_ids,
                 float* d_dists, hipStream_t s, int nov) {
  culsh_ht* ht = st->ht;
  const int L = ht->L, D = ht->D, T = st->T, K = st->K;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, st->dmax.p, 1);
  hipLaunchKernelGGL(k_max_cnu, dim3(grid_for(nov)), dim3(256), 0, s, (const int*)st->dov.p, nov,
                     (const unsigned long long*)st->dcnu_qt.p, L, st->dmax.p);
  unsigned long long maxcnu = 0;
  HIP_TRY(hipMemcpyAsync(&maxcnu, st->dmax.p, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (maxcnu >= (1ull << 30))
    return fail(CULSH_E_LIMIT, "a query has %llu candidates (limit 2^30)", maxcnu);
  const int hbits = std::max(13, ceil_log2_host((long long)(2 * maxcnu)));
  const unsigned ucap = (unsigned)std::max<unsigned long long>(maxcnu, 1);
  const size_t region = (1ull << hbits) + 2ull * ucap;  // u32 words per workgroup
  const size_t budget = (4ull << 30) / 4;
  const int G = (int)std::max<size_t>(1, std::min<size_t>((size_t)nov, budget / region));
  TRY(st->dscratch.ensure(region * G));
  HIP_TRY(hipMemsetAsync(st->dnov.p, 0, 4, s));
  hipLaunchKernelGGL(k_candidates_global, dim3(G), dim3(256), 0, s, dX, dQ, Q, D, L, T, K,
                     (const uint2*)st->dqr.p, (const uint32_t*)ht->sids.p, (const unsigned long long*)st->dcnu_qt.p,
                     d_ids, d_dists, st->dcnu_q.p, st->dcu_q.p, (const int*)st->dov.p, nov,
                     st->dscratch.p, hbits, ucap, st->dcnt.p, st->dnov.p,
                     st->stamp ? st->dphase.p : nullptr);
  HIP_TRY(hipGetLastError());
  int err = 0;
  HIP_TRY(hipMemcpyAsync(&err, st->dnov.p, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (err)
    return fail(CULSH_E_LIMIT,
                "a query has more than %d candidates within the FP32 error bound of its K-th "
                "distance",
                kScapG);
  return CULSH_OK;
}

// Locality order of the queries: a stable radix sort of coarse hash codes of
// table 0, so queries likely to share buckets become neighbours.
int locality_order(culsh_st* st, const float* dQ, int Q, hipStream_t s, const uint32_t** order) {
  culsh_ht* ht = st->ht;
  TRY(st->dlk0.ensure(Q));
  TRY(st->dlk1.ensure(Q));
  TRY(st->dlv0.ensure(Q));
  TRY(st->dlv1.ensure(Q));
  TRY(st->dlscr.ensure(rs_scratch_elems(Q) + 1));
  hipLaunchKernelGGL(k_locality_key, dim3(grid_for(Q)), dim3(256), 0, s, dQ, Q, ht->D, ht->M,
                     (const float*)ht->dAorig.p, (const float*)ht->dbj.p, ht->W, st->dlk0.p,
                     st->dlv0.p);
  uint64_t* kp = st->dlk0.p;
  uint32_t* vp = st->dlv0.p;
  radix_sort_pairs(&kp, &vp, st->dlk1.p, st->dlv1.p, Q, 0, 64, st->dlscr.p, s);
  *order = vp;
  return CULSH_OK;
}

// Precision form candidate stage: |x|^2 of the dataset (cached per dataset
// pointer), locality order of the queries, group working sets sized from the
// largest group's summed C_nu (one host synchronisation), then
// k_group_candidates.  Queries it cannot finish go to the overflow list.
int run_group_candidates(culsh_st* st, const float* dQ, const float* dX, int Q, int* d_ids,
                         float* d_dists, hipStream_t s) {
  culsh_ht* ht = st->ht;
  const int L = ht->L, D = ht->D, T = st->T, K = st->K;
  const long long N = ht->N;
  const unsigned long long gen = dX == ht->dX ? ht->data_gen : st->alt_gen;
  if (st->xn2_for != dX || st->xn2_gen != gen) {
    TRY(st->dxn2.ensure(N));
    TRY(st->dxmax.ensure(1));
    HIP_TRY(hipMemsetAsync(st->dxmax.p, 0, 4, s));
    hipLaunchKernelGGL(k_row_sqnorms, dim3(grid_for(N, 16)), dim3(256), 0, s, dX, N, D,
                       st->dxn2.p, st->dxmax.p);
    unsigned bits = 0;
    HIP_TRY(hipMemcpyAsync(&bits, st->dxmax.p, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::memcpy(&st->xmax, &bits, 4);
    st->xn2_for = dX;
    st->xn2_gen = gen;
  }
  const uint32_t* order = nullptr;
  TRY(locality_order(st, dQ, Q, s, &order));
  const int ng = (Q + kGroup - 1) / kGroup;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, st->dmax.p, 1);
  hipLaunchKernelGGL(k_group_cnu, dim3(grid_for(ng)), dim3(256), 0, s, order, Q, L,
                     (const unsigned long long*)st->dcnu_qt.p, st->dmax.p);
  unsigned long long gmax = 0;
  HIP_TRY(hipMemcpyAsync(&gmax, st->dmax.p, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (gmax >= (1ull << 29))
    return fail(CULSH_E_LIMIT, "a query group has %llu candidates (limit 2^29)", gmax);
  // Capacity rounded to a power of two, so that the scratch layout (which
  // the clean state below belongs to) changes rarely between searches.
  const unsigned qcap = 1u << ceil_log2_host((long long)std::max<unsigned long long>(gmax, 1));
  const int hbits = std::max(8, ceil_log2_host(2ll * qcap));
  const size_t region = 2 * (1ull << hbits) + 5ull * qcap;
  const size_t budget = (4ull << 30) / 4;
  const int G = (int)std::max<size_t>(
      1, std::min<size_t>(std::min<size_t>((size_t)ng, budget / region), 4096));
  if (st->dgscratch.n < region * G) {
    TRY(st->dgscratch.exact(region * G));
    st->gscratch_clean = false;
  }
  // Each group leaves the hash ids and masks of its workgroup's region
  // all-ones, but only for the layout (region, hbits) it ran with: another
  // layout puts hash slots over words of the other arrays.
  if (st->g_region != region || st->g_hbits != hbits) st->gscratch_clean = false;
  if (!st->gscratch_clean) {
    // Empty hash ids and complemented masks are all-ones words.
    HIP_TRY(hipMemsetAsync(st->dgscratch.p, 0xFF, st->dgscratch.n * 4, s));
    st->gscratch_clean = true;
    st->g_region = region;
    st->g_hbits = hbits;
  }
  hipLaunchKernelGGL(k_group_candidates, dim3(G), dim3(256), 0, s, dX,
                     (const float*)st->dxn2.p, st->xmax, dQ, Q, D, L, T, K,
                     (const uint2*)st->dqr.p, (const uint32_t*)ht->sids.p,
                     (const unsigned long long*)st->dcnu_qt.p, order, d_ids, d_dists, st->dcnu_q.p, st->dcu_q.p, st->dgscratch.p, region,
                     hbits, qcap, st->dov.p, st->dnov.p, st->dcnt.p);
  HIP_TRY(hipGetLastError());
  return CULSH_OK;
}

// Working buffers of one device search.
struct SearchBufs {
  int* qb;                          // [L][Q*T] probe buckets, or null
  uint2* qr;                        // [L][Q*T] their member ranges (start, size)
  unsigned long long* cnu_qt;       // [Q*L]
  unsigned long long *cnu_q, *cu_q; // [Q]
  int* ov;                          // [Q] overflowed queries
  int* nov;                         // their count (zeroed before the launch)
  unsigned long long* cnt;          // FP64 recheck counter (zeroed)
  bool order;                       // may run large batches in locality order
  unsigned long long* phase;        // [3] phase clocks of the candidate stage, or null
};

// Probe stage then candidate stage on `s` (events ev[0..2] around them when
// `timed`).  The caller has zeroed *b.nov and *b.cnt on `s`.
int launch_probe_and_candidates(culsh_st* st, const float* dQ, const float* dX, int Q,
                                int* d_ids, float* d_dists, hipStream_t s, const SearchBufs& b,
                                bool timed) {
  culsh_ht* ht = st->ht;
  const int L = ht->L, M = ht->M, D = ht->D, T = st->T, K = st->K;
  if (timed) HIP_TRY(hipEventRecord(st->ev[0], s));
  const ProbeIndex pix = probe_index(ht);
  const long long pairs = (long long)Q * L;
  const int lds8 = probe8_lds_bytes(D, M);
  if (T <= 128 && lds8 <= 64 * 1024) {
    // Octet kernel: 32 queries of one table per workgroup.
    const dim3 g8((unsigned)(((Q + 31) / 32) * (long long)L));
#define CULSH_PROBE8(JJ, MM)                                                                  \
  hipLaunchKernelGGL((k_query_probe8<JJ, MM>), g8, dim3(256), lds8, s, dQ, Q, D, L, M, T,     \
                     (const float*)ht->dAd.p, ht->Jp, (const double*)ht->db64.p, (double)ht->W, \
                     (const uint64_t*)ht->dseeds.p, pix, b.qb, b.qr, b.cnu_qt)
#define CULSH_PROBE8_M(JJ)   \
  if (M <= 8)                \
    CULSH_PROBE8(JJ, 1);     \
  else if (M <= 16)          \
    CULSH_PROBE8(JJ, 2);     \
  else if (M <= 24)          \
    CULSH_PROBE8(JJ, 3);     \
  else if (M <= 32)          \
    CULSH_PROBE8(JJ, 4);     \
  else                       \
    CULSH_PROBE8(JJ, 8);
    if (T <= 8) {
      CULSH_PROBE8_M(1)
    } else if (T <= 16) {
      CULSH_PROBE8_M(2)
    } else if (T <= 32) {
      CULSH_PROBE8_M(4)
    } else if (T <= 64) {
      CULSH_PROBE8_M(8)
    } else {
      CULSH_PROBE8_M(16)
    }
#undef CULSH_PROBE8_M
#undef CULSH_PROBE8
  } else {
    // Wave kernel: one wave per (query, table).
    const dim3 pgrid((unsigned)((pairs + 3) / 4));
#define CULSH_PROBE(JJ)                                                                       \
  hipLaunchKernelGGL(k_query_probe<JJ>, pgrid, dim3(256), 0, s, dQ, Q, D, L, M, T,           \
                     (const double*)ht->dAt64.p, (const double*)ht->db64.p, (double)ht->W,   \
                     (const uint64_t*)ht->dseeds.p, pix, b.qb, b.qr, b.cnu_qt)
    if (T <= 64)
      CULSH_PROBE(1);
    else if (T <= 128)
      CULSH_PROBE(2);
    else if (T <= 256)
      CULSH_PROBE(4);
    else if (T <= 512)
      CULSH_PROBE(8);
    else if (T <= 1024)
      CULSH_PROBE(16);
    else
      return fail(CULSH_E_LIMIT, "T = %d exceeds the supported maximum of 1024", T);
#undef CULSH_PROBE
  }
  if (timed) HIP_TRY(hipEventRecord(st->ev[1], s));
  if (st->form == CULSH_FORM_PRECISION_GROUPED) {
    if (timed) HIP_TRY(hipEventRecord(st->ev[3], s));
    TRY(run_group_candidates(st, dQ, dX, Q, d_ids, d_dists, s));
  } else {
    // Large batches run in locality order, each XCD taking a contiguous run
    // of it, so queries that share candidate rows share an L2.
    const uint32_t* order = nullptr;
    unsigned grid = (unsigned)Q;
    if (b.order && Q >= kOrderMinQ) {
      TRY(locality_order(st, dQ, Q, s, &order));
      grid = (unsigned)((Q + 7) / 8 * 8);
    }
    // The precision form's distance pass reads the resident rows' FP16 copy
    // when it exists (results unchanged; the FP64 pass reads the FP32 rows).
    const bool half = st->form == CULSH_FORM_PRECISION && dX == ht->dX && ht->rows16_ok &&
                      ht->rows16_gen == ht->data_gen;
    if (timed) HIP_TRY(hipEventRecord(st->ev[3], s));
    hipLaunchKernelGGL(k_candidates, dim3(grid), dim3(256), 0, s, dX, dQ, Q, D, L, T, K,
                       (const uint2*)b.qr, (const uint32_t*)ht->sids.p, (const unsigned long long*)b.cnu_qt, d_ids,
                       d_dists, b.cnu_q, b.cu_q, b.ov, b.nov, b.cnt, order, b.phase,
                       half ? (const uint16_t*)ht->dX16.p : nullptr, ht->rows16_e2);
  }
  if (timed) HIP_TRY(hipEventRecord(st->ev[2], s));
  HIP_TRY(hipGetLastError());
  return CULSH_OK;
}

// The device search pipeline on `s`: probe kernel, candidate kernel (and the
// global-memory kernel for queries that overflow it).  Synchronises to check
// the per-query capacity flags, unless `wait` is false and the distances are
// direct (plain and precision forms).
// `buckets`: keep each probe's bucket index (the host search's query
// buckets); device searches expose none and skip those writes.
int run_search(culsh_st* st, const float* dQ, const float* dX, int Q, int* d_ids, float* d_dists,
               hipStream_t s, bool timed, bool wait = true, bool buckets = true) {
  TRY(st_check_async(st, wait));
  // A device search still running on another stream uses the same working
  // buffers: order this one after it.
  if (st->pending) HIP_TRY(hipStreamWaitEvent(s, st->evdone, 0));
  // Direct distances can finish without the host: overflowed queries go
  // to a fixed-size device kernel.  The grouped form sizes its group
  // working sets on the host.
  const bool async = !wait && st->form != CULSH_FORM_PRECISION_GROUPED;
  culsh_ht* ht = st->ht;
  if (st->form == CULSH_FORM_PRECISION && dX == ht->dX) TRY(ensure_rows16(ht));
  const int L = ht->L, D = ht->D, T = st->T, K = st->K;
  const long long QT = (long long)Q * T;
  if ((long long)L * QT >= (1ll << 31)) return fail(CULSH_E_LIMIT, "L*Q*T exceeds 2^31 - 1");
  if (buckets) TRY(st->dqb.ensure((size_t)L * QT));
  TRY(st->dqr.ensure((size_t)L * QT));
  TRY(st->dcnu_qt.ensure((size_t)Q * L));
  TRY(st->dcnu_q.ensure(Q));
  TRY(st->dcu_q.ensure(Q));
  TRY(st->dov.ensure(Q));
  TRY(st->dnov.ensure(1));
  TRY(st->dcnt.ensure(1));
  TRY(st->dmax.ensure(1));
  HIP_TRY(hipMemsetAsync(st->dnov.p, 0, 4, s));
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, st->dcnt.p, 1);
  SearchBufs b{buckets ? st->dqb.p : nullptr, st->dqr.p, st->dcnu_qt.p, st->dcnu_q.p,
               st->dcu_q.p, st->dov.p, st->dnov.p,   st->dcnt.p,
               true,       st->stamp ? st->dphase.p : nullptr};
  TRY(launch_probe_and_candidates(st, dQ, dX, Q, d_ids, d_dists, s, b, timed));
  if (async) {
    TRY(st->dovfix.ensure((size_t)kOvAsyncGrid *
                          ((1ull << kOvAsyncHbits) + 2ull * kOvAsyncUcap)));
    hipLaunchKernelGGL(k_candidates_overflow, dim3(kOvAsyncGrid), dim3(256), 0, s, dX, dQ, Q, D,
                       L, T, K, (const uint2*)st->dqr.p, (const uint32_t*)ht->sids.p,
                       (const unsigned long long*)st->dcnu_qt.p,
                       d_ids, d_dists, st->dcnu_q.p, st->dcu_q.p, (const int*)st->dov.p,
                       (const int*)st->dnov.p, st->dovfix.p, kOvAsyncHbits, kOvAsyncUcap,
                       st->dcnt.p, st->derr.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(st->herr, st->derr.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipEventRecord(st->evdone, s));
    st->pending = true;
    return CULSH_OK;
  }
  int nov = 0;
  HIP_TRY(hipMemcpyAsync(&nov, st->dnov.p, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (nov > 0) TRY(run_overflow(st, dQ, dX, Q, d_ids, d_dists, s, nov));
  return CULSH_OK;
}

}  // namespace

extern "C" {

int culsh_st_create(culsh_st** out) {
  if (!out) return fail(CULSH_E_ARG, "null output pointer");
  *out = new culsh_st;
  return CULSH_OK;
}

void culsh_st_destroy(culsh_st* st) { delete st; }

int culsh_st_reset(culsh_st* st, culsh_ht* ht, int K, int T) {
  if (!st) return fail(CULSH_E_ARG, "null SearchTables");
  if (!ht) return fail(CULSH_E_ARG, "null HashTables");
  st->ht = ht;
  st->K = K;
  st->T = T;
  st->clear_results();
  return CULSH_OK;
}

int culsh_st_set_form(culsh_st* st, int form) {
  if (!st) return fail(CULSH_E_ARG, "null SearchTables");
  if (form != CULSH_FORM_PLAIN && form != CULSH_FORM_PRECISION &&
      form != CULSH_FORM_PRECISION_GROUPED)
    return fail(CULSH_E_ARG, "unknown form %d", form);
  st->form = form;
  return CULSH_OK;
}

static int st_check(culsh_st* st, int Q) {
  if (!st) return fail(CULSH_E_ARG, "null SearchTables");
  if (!st->ht)
    return fail(CULSH_E_STATE,
                "No indexing structure has been associated with searching structure...");
  if (Q <= 0 || st->K <= 0 || st->T <= 0 || st->ht->L <= 0)
    return fail(CULSH_E_ARG,
                "Integer parameters must be all positive. Have you reset the indexing structure "
                "properly before assigning it to this searching structure???");
  if (!st->ht->indexed) return fail(CULSH_E_STATE, "the hash tables hold no index");
  return st_setup(st);
}

}  // extern "C"

namespace {

// Queries per pass of a search (cuLSH_SearchTables.cu:283-311 sizes its
// batches from the free device memory the same way): the probe buffers
// address L*Q*T slots with 31-bit offsets, and a pass's working set must
// fit half the HBM free at this moment (plus what the object already holds,
// which the pass reuses).  `host`: the pass also stages queries, results and
// lexicographic query buckets.
int search_pass(culsh_st* st, int Q, bool host, int* out) {
  culsh_ht* ht = st->ht;
  const long long L = ht->L, T = st->T, K = st->K, D = ht->D;
  long long q = std::min<long long>(Q, ((1ll << 31) - 1) / (L * T));
  // Per query: member ranges (8 L T), C_nu per table, counters; the host
  // search also keeps bucket indices and stages queries, results and
  // lexicographic query buckets.
  const long long per = 8 * L * T + 8 * L + 16 + 4 + 48 + (host ? 8 * L * T + 4 * D + 8 * K : 0);
  size_t freeb = 0, totalb = 0;
  HIP_TRY(hipMemGetInfo(&freeb, &totalb));
  long long held =
      (long long)(st->dqb.n * 4 + st->dqr.n * 8 + st->dcnu_qt.n * 8 + st->dcnu_q.n * 16);
  if (host) held += (long long)(st->dqb_local.n * 4 + st->dQ.n * 4 + st->dids.n * 4 + st->ddists.n * 4);
  q = std::min<long long>(q, ((long long)(freeb / 2) + held) / per);
  if (st->max_pass > 0) q = std::min<long long>(q, st->max_pass);
  if (q < 1)
    return fail(CULSH_E_MEMORY, "free HBM (%zu bytes) holds no query of this search", freeb);
  *out = (int)q;
  return CULSH_OK;
}

}  // namespace

extern "C" {

int culsh_st_search(culsh_st* st, const float* queries, int Q, const float* dataset, FILE* dbg) {
  TRY(st_check(st, Q));
  if (!queries || !dataset) return fail(CULSH_E_ARG, "null queries or dataset");
  culsh_ht* ht = st->ht;
  const int D = ht->D, K = st->K, T = st->T, L = ht->L;
  st->clear_results();
  TRY(ensure_renumber(ht));
  const float* dX = ht->dX;
  if (!dX || dataset != ht->hX_indexed) {
    // A dataset other than the resident one (an index built from a device
    // pointer or loaded from a file, or another host array): uploaded for
    // the distances on the first search that passes this host pointer for
    // this index, then reused while the pointer and the index stay the same.
    if (!st->dX_alt.p || st->alt_host != dataset || st->alt_ht != ht ||
        st->alt_ht_gen != ht->data_gen) {
      st->alt_host = nullptr;
      TRY(st->dX_alt.ensure((size_t)ht->N * D));
      HIP_TRY(hipMemcpy(st->dX_alt.p, dataset, (size_t)ht->N * D * 4, hipMemcpyHostToDevice));
      st->alt_host = dataset;
      st->alt_ht = ht;
      st->alt_ht_gen = ht->data_gen;
      st->alt_gen = next_data_gen();
    }
    dX = st->dX_alt.p;
  }
  int pass = 0;
  TRY(search_pass(st, Q, true, &pass));
  if (dbg)
    fprintf(dbg, "[SearchTables::search]\tPerforming searching. (Q, K, T) = (%d, %d, %d), %d "
                 "queries per pass\n", Q, K, T, pass);
  const long long QT = (long long)Q * T;
  st->ids.resize((size_t)Q * K);
  st->dists.resize((size_t)Q * K);
  st->cnu.resize(Q);
  st->cu.resize(Q);
  st->qbuckets.assign(L, std::vector<int>(QT));
  TRY(st->dphase.ensure(3));
  double t_buckets = 0, t_sort = 0, t_cand = 0;
  unsigned long long cyc[3] = {0, 0, 0}, nrc_all = 0;
  for (int q0 = 0; q0 < Q; q0 += pass) {
    const int qn = std::min(pass, Q - q0);
    const long long qnT = (long long)qn * T;
    TRY(st->dQ.ensure((size_t)qn * D));
    TRY(st->dids.ensure((size_t)qn * K));
    TRY(st->ddists.ensure((size_t)qn * K));
    HIP_TRY(hipMemcpyAsync(st->dQ.p, queries + (size_t)q0 * D, (size_t)qn * D * 4,
                           hipMemcpyHostToDevice, st->stream));
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st->stream, st->dphase.p, 3);
    st->stamp = true;
    const int rc = run_search(st, st->dQ.p, dX, qn, st->dids.p, st->ddists.p, st->stream, true);
    st->stamp = false;
    TRY(rc);
    HIP_TRY(hipEventRecord(st->ev[4], st->stream));
    TRY(st->dqb_local.ensure((size_t)L * qnT));
    hipLaunchKernelGGL(k_local_buckets, dim3(grid_for(L * qnT)), dim3(256), 0, st->stream,
                       (const int*)st->dqb.p, qnT, L, (const int*)ht->renumber.p,
                       st->dqb_local.p);
    HIP_TRY(hipMemcpyAsync(st->ids.data() + (size_t)q0 * K, st->dids.p, (size_t)qn * K * 4,
                           hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipMemcpyAsync(st->dists.data() + (size_t)q0 * K, st->ddists.p, (size_t)qn * K * 4,
                           hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipMemcpyAsync(st->cnu.data() + q0, st->dcnu_q.p, (size_t)qn * 8,
                           hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipMemcpyAsync(st->cu.data() + q0, st->dcu_q.p, (size_t)qn * 8, hipMemcpyDeviceToHost,
                           st->stream));
    for (int t = 0; t < L; ++t)
      HIP_TRY(hipMemcpyAsync(st->qbuckets[t].data() + (size_t)q0 * T,
                             st->dqb_local.p + (size_t)t * qnT, qnT * 4, hipMemcpyDeviceToHost,
                             st->stream));
    unsigned long long nrc = 0, pc[3] = {0, 0, 0};
    HIP_TRY(hipMemcpyAsync(&nrc, st->dcnt.p, 8, hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipMemcpyAsync(pc, st->dphase.p, 24, hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    nrc_all += nrc;
    for (int i = 0; i < 3; ++i) cyc[i] += pc[i];
    float a = 0.f, b = 0.f, c = 0.f;
    HIP_TRY(hipEventElapsedTime(&a, st->ev[0], st->ev[1]));
    HIP_TRY(hipEventElapsedTime(&b, st->ev[1], st->ev[3]));
    HIP_TRY(hipEventElapsedTime(&c, st->ev[3], st->ev[4]));
    t_buckets += a;
    t_sort += b;
    t_cand += c;
  }
  st->Q = Q;
  st->rechecked = nrc_all;
  for (int q = 0; q < Q; ++q) {
    st->tot_nu += st->cnu[q];
    st->tot_u += st->cu[q];
  }
  // cuLSH_SearchTables.cu:342-362,410,440 times its steps separately; here
  // the candidate stage is one fused pass, whose measured time (candidate
  // kernels, overflow kernel included) is split by the shader-clock shares
  // of its phases: union (collect; the dedup happens in the same hash
  // insert, so unique stays 0), FP32 distances, top-K (knn).  `sort` is the
  // locality ordering of large query sets.
  const double ctot = (double)cyc[0] + (double)cyc[1] + (double)cyc[2];
  const double f0 = ctot > 0 ? cyc[0] / ctot : 0.0, f1 = ctot > 0 ? cyc[1] / ctot : 0.0,
               f2 = ctot > 0 ? cyc[2] / ctot : 0.0;
  st->times = culsh_times{};
  st->times.buckets = (float)(t_buckets / 1000.0);
  st->times.sort = (float)(t_sort / 1000.0);
  st->times.collect = (float)(t_cand * f0 / 1000.0);
  st->times.unique = 0.f;
  st->times.distances = (float)(t_cand * f1 / 1000.0);
  st->times.knn = (float)(t_cand * f2 / 1000.0);
  if (st->form == CULSH_FORM_PRECISION_GROUPED) {
    // The grouped kernel is not phase-stamped: its whole time is reported
    // as distance time.
    st->times.collect = st->times.knn = 0.f;
    st->times.distances = (float)(t_cand / 1000.0);
  }
  if (dbg)
    fprintf(dbg, "Non-unique: %llu, unique: %llu, selectivity: %.2f%%\n", st->tot_nu, st->tot_u,
            culsh_st_selectivity(st) * 100.0);
  return CULSH_OK;
}

int culsh_st_set_pass_limit(culsh_st* st, int max_queries) {
  if (!st) return fail(CULSH_E_ARG, "null SearchTables");
  if (max_queries < 0) return fail(CULSH_E_ARG, "max_queries must be >= 0");
  st->max_pass = max_queries;
  return CULSH_OK;
}

int culsh_st_search_device(culsh_st* st, const float* d_queries, int Q, int* d_knn_ids,
                           float* d_knn_dists, void* stream) {
  TRY(st_check(st, Q));
  if (!d_queries || !d_knn_ids || !d_knn_dists) return fail(CULSH_E_ARG, "null device pointer");
  if (!st->ht->dX) return fail(CULSH_E_STATE, "the index holds no resident dataset");
  int pass = 0;
  TRY(search_pass(st, Q, false, &pass));
  const int D = st->ht->D, K = st->K;
  for (int q0 = 0; q0 < Q; q0 += pass) {
    const int qn = std::min(pass, Q - q0);
    TRY(run_search(st, d_queries + (size_t)q0 * D, st->ht->dX, qn, d_knn_ids + (size_t)q0 * K,
                   d_knn_dists + (size_t)q0 * K, (hipStream_t)stream, true, false, false));
  }
  return CULSH_OK;
}

int culsh_st_wait(culsh_st* st) {
  if (!st) return fail(CULSH_E_ARG, "null SearchTables");
  return st_check_async(st, true);
}

int culsh_st_kernel_ms(culsh_st* st, float* probe_ms, float* candidates_ms) {
  if (!st || !st->ev[2]) return fail(CULSH_E_STATE, "no search has run");
  TRY(st_check_async(st, true));
  float a = 0.f, b = 0.f;
  HIP_TRY(hipEventElapsedTime(&a, st->ev[0], st->ev[1]));
  HIP_TRY(hipEventElapsedTime(&b, st->ev[1], st->ev[2]));
  if (probe_ms) *probe_ms = a;
  if (candidates_ms) *candidates_ms = b;
  return CULSH_OK;
}

const int* culsh_st_knn_ids(culsh_st* st) { return st && !st->ids.empty() ? st->ids.data() : nullptr; }
const float* culsh_st_knn_dists(culsh_st* st) {
  return st && !st->dists.empty() ? st->dists.data() : nullptr;
}
const int* culsh_st_query_buckets(culsh_st* st, int t) {
  if (!st || t < 0 || t >= (int)st->qbuckets.size()) return nullptr;
  return st->qbuckets[t].data();
}
float culsh_st_selectivity(culsh_st* st) {
  if (!st || !st->ht || !st->ht->N || !st->Q) return 0.f;
  return st->tot_u / (float)st->ht->N / (float)st->Q;
}
float culsh_st_nonunique_pct(culsh_st* st) {
  if (!st || !st->tot_u) return 0.f;
  return (st->tot_nu - st->tot_u) / (float)st->tot_u;
}
unsigned long long culsh_st_total_unique(culsh_st* st) { return st ? st->tot_u : 0; }
unsigned long long culsh_st_total_nonunique(culsh_st* st) { return st ? st->tot_nu : 0; }
void culsh_st_times(culsh_st* st, culsh_times* out) {
  if (st && out) *out = st->times;
}
const unsigned long long* culsh_st_query_cnu(culsh_st* st) {
  return st && !st->cnu.empty() ? st->cnu.data() : nullptr;
}
const unsigned long long* culsh_st_query_cu(culsh_st* st) {
  return st && !st->cu.empty() ? st->cu.data() : nullptr;
}
unsigned long long culsh_st_rechecked(culsh_st* st) { return st ? st->rechecked : 0; }

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free device search (culsh_st_search_device_noalloc)
// ---------------------------------------------------------------------------
// Working memory of the allocation-free search: one copy per device, held
// until the process ends.  Host code never passes this name to a kernel or a
// copy (on the host it is only a stand-in for the device array); it takes
// the device address from hipGetSymbolAddress, once per device.
static constexpr size_t kNoAllocBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

// Fixed overflow working sets of the allocation-free search: workgroups,
// hash bits and the largest candidate list of one query (2^17 non-unique
// candidates; a query with more sets the caller's status word).
constexpr int kNaOvGrid = 4;
constexpr int kNaOvHbits = 18;
constexpr unsigned kNaOvUcap = 1u << 17;
constexpr size_t kNaOvBytes = (size_t)kNaOvGrid * ((1ull << kNaOvHbits) + 2ull * kNaOvUcap) * 4;

std::mutex g_na_mu;
constexpr int kNaMaxDevices = 64;
void* g_na_ptr[kNaMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once).
int noalloc_scratch(void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNaMaxDevices)
    return fail(CULSH_E_DEVICE, "device %d beyond the %d the library tracks", dev, kNaMaxDevices);
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CULSH_E_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_na_ptr[dev] = p;
  }
  *out = g_na_ptr[dev];
  return CULSH_OK;
}

// Clears the counters of one allocation-free search and the caller's status.
__global__ void k_na_clear(int* nov, unsigned long long* cnt, int* status) {
  if (threadIdx.x == 0) {
    *nov = 0;
    *cnt = 0ull;
    *status = 0;
  }
}

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Byte offsets of the search buffers in the scratch array.
struct NaLayout {
  size_t ovfix, ctr, cnu_q, cu_q, ov, cnu_qt, qr, total;
};

NaLayout na_layout(long long Q, int L, int T) {
  NaLayout o;
  o.ovfix = 0;
  o.ctr = align256(kNaOvBytes);
  o.cnu_q = o.ctr + 256;
  o.cu_q = align256(o.cnu_q + 8ull * Q);
  o.ov = align256(o.cu_q + 8ull * Q);
  o.cnu_qt = align256(o.ov + 4ull * Q);
  o.qr = align256(o.cnu_qt + 8ull * Q * L);
  o.total = o.qr + 8ull * Q * T * L;
  return o;
}

}  // namespace

extern "C" int culsh_st_search_device_noalloc(culsh_st* st, const float* d_queries, int Q,
                                              int* d_knn_ids, float* d_knn_dists, int* d_status,
                                              void* stream) {
  if (!st) return fail(CULSH_E_ARG, "null SearchTables");
  if (!st->ht)
    return fail(CULSH_E_STATE,
                "No indexing structure has been associated with searching structure...");
  culsh_ht* ht = st->ht;
  if (Q <= 0 || st->K <= 0 || st->T <= 0)
    return fail(CULSH_E_ARG, "Q, K and T must be positive");
  if (!ht->indexed || !ht->dX) return fail(CULSH_E_STATE, "the index holds no resident dataset");
  if (!d_queries || !d_knn_ids || !d_knn_dists || !d_status)
    return fail(CULSH_E_ARG, "null device pointer");
  if (st->form == CULSH_FORM_PRECISION_GROUPED)
    return fail(CULSH_E_ARG,
                "the allocation-free search runs direct distances only (not the grouped form)");
  const int L = ht->L, T = st->T, D = ht->D, K = st->K;
  if (T > 1024) return fail(CULSH_E_LIMIT, "T = %d exceeds the supported maximum of 1024", T);
  if ((long long)L * Q * T >= (1ll << 31)) return fail(CULSH_E_LIMIT, "L*Q*T exceeds 2^31 - 1");
  const NaLayout lay = na_layout(Q, L, T);
  if (lay.total > kNoAllocBytes) {
    const NaLayout one = na_layout(1, L, T);
    const long long per = (long long)(one.total - one.cnu_q);
    return fail(CULSH_E_LIMIT,
                "allocation-free search: Q = %d with L = %d, T = %d needs %zu bytes of working "
                "memory; its fixed scratch holds %zu bytes (about %lld queries at this L and T)",
                Q, L, T, lay.total, kNoAllocBytes,
                (long long)((kNoAllocBytes - lay.cnu_q) / (size_t)(per + 64)));
  }
  void* basev = nullptr;
  TRY(noalloc_scratch(&basev));
  unsigned char* base = static_cast<unsigned char*>(basev);
  hipStream_t s = (hipStream_t)stream;
  int* nov = reinterpret_cast<int*>(base + lay.ctr);
  unsigned long long* cnt = reinterpret_cast<unsigned long long*>(base + lay.ctr + 8);
  hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(64), 0, s, nov, cnt, d_status);
  SearchBufs b{nullptr, reinterpret_cast<uint2*>(base + lay.qr),
               reinterpret_cast<unsigned long long*>(base + lay.cnu_qt),
               reinterpret_cast<unsigned long long*>(base + lay.cnu_q),
               reinterpret_cast<unsigned long long*>(base + lay.cu_q),
               reinterpret_cast<int*>(base + lay.ov), nov, cnt, false, nullptr};
  TRY(launch_probe_and_candidates(st, d_queries, ht->dX, Q, d_knn_ids, d_knn_dists, s, b, false));
  hipLaunchKernelGGL(k_candidates_overflow, dim3(kNaOvGrid), dim3(256), 0, s, ht->dX, d_queries,
                     Q, D, L, T, K, (const uint2*)b.qr, (const uint32_t*)ht->sids.p, (const unsigned long long*)b.cnu_qt,
                     d_knn_ids, d_knn_dists, b.cnu_q, b.cu_q, (const int*)b.ov,
                     (const int*)b.nov, reinterpret_cast<uint32_t*>(base + lay.ovfix),
                     kNaOvHbits, kNaOvUcap, b.cnt, d_status);
  HIP_TRY(hipGetLastError());
  return CULSH_OK;
}

// Batching form (culsh_bt_*): shares this translation unit's helpers and
// index state.
#include "batch_host.hpp"

// Serving form (culsh_srv_*): drives the batching form's plans.
#include "serve_host.hpp"
