// ORACLE — test infrastructure only.
//
// A CPU (OpenMP, FP64) restatement of the cuLSH indexing and query algorithm.
// It is the correctness checker for the HIP path and the CPU baseline that
// bench.py times.  Nothing in the shipped library links, loads or calls it;
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
//
// Each function cites the reference file:line it restates.  Arithmetic that
// the reference does in FP32 is done here in FP64 with a fixed operation
// order, so that a device kernel that performs the same IEEE operations in
// the same order reproduces these results bit for bit:
//   * projection  v = (fma-chain_d(x_d * a_d) - b) / W, d = 0..D-1 ascending
//   * deltas      dn = v - floor(v), up = 1 - dn
//   * probe score built incrementally exactly as the reference queue does
//   * distance    fma-chain_d((q_d - x_d)^2), d = 0..D-1 ascending
//
// Pins (tests/test_oracle.py): closed-form projections on exactly
// representable inputs, brute-force group-by for buckets, brute-force subset
// enumeration for the multiprobe sequence, linear-scan lookup, set union for
// candidates, exact integer distances, and a closed-form top-K tie rule
// checked against this file's literal heap simulation.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

// ---------------------------------------------------------------------------
// Projection: cuLSH_Indexing.cu:27-66 (SGEMM X'*A, then -b, then /W) and
// cuLSH_Kernels.cu:91-101 (floorf).  FP64, sequential FMA over d.
// A_t column-major [D x M]: (d, m) at m*D + d (cuLSH_Indexing.cu:43-51).
// ---------------------------------------------------------------------------
inline double project1(const float* x, const float* a_col, int D, float b, float W) {
  double acc = 0.0;
  for (int d = 0; d < D; ++d) acc = std::fma((double)x[d], (double)a_col[d], acc);
  return (acc - (double)b) / (double)W;
}

// All M projections of one vector for one table, vectorised across m with
// the same per-(n, m) operation order as project1.
void project_row(const float* x, const double* At /*[D][M]*/, const double* b, double W, int D,
                 int M, double* out) {
  double acc[256];
  for (int m = 0; m < M; ++m) acc[m] = 0.0;
  for (int d = 0; d < D; ++d) {
    const double xd = (double)x[d];
    const double* row = At + (size_t)d * M;
    for (int m = 0; m < M; ++m) acc[m] = std::fma(xd, row[m], acc[m]);
  }
  for (int m = 0; m < M; ++m) out[m] = (acc[m] - b[m]) / W;
}

struct Table {
  int B = 0;
  std::vector<uint32_t> total;   // [N] bucket members, bucket by bucket
  std::vector<uint32_t> start;   // [B]
  std::vector<uint32_t> size;    // [B]
  std::vector<int32_t> codes;    // [B][M] row-major, lexicographically increasing
};

int lex_cmp(const int32_t* a, const int32_t* b, int M) {
  for (int m = 0; m < M; ++m) {
    if (a[m] < b[m]) return -1;
    if (a[m] > b[m]) return 1;
  }
  return 0;
}

// Heap helpers restating cuLSH_Kernels.cu:144-175 (balanceHeap_deltas).
// perm holds indices into d; -1 marks an empty slot and acts as +infinity.
void sift_deltas(int* perm, const double* d, int total, int levels) {
  int parent = 0;
  for (int level = 0; level < levels - 1; ++level) {
    int left = parent * 2 + 1, right = left + 1, mx = parent;
    if (left < total && perm[left] == -1)
      mx = left;
    else if (right < total && perm[right] == -1)
      mx = right;
    else {
      if (left < total && d[perm[left]] > d[perm[mx]]) mx = left;
      if (right < total && d[perm[right]] > d[perm[mx]]) mx = right;
    }
    if (mx == parent) break;
    std::swap(perm[parent], perm[mx]);
    parent = mx;
  }
}

// ceil(log2(x)) for x >= 1: the heap depth the reference derives from
// __float2int_ru(log2f(x)) (cuLSH_Kernels.cu:202,419,452).
int ceil_log2(int x) {
  int l = 0;
  while ((1 << l) < x) ++l;
  return l;
}

// Heap helpers restating cuLSH_Kernels.cu:356-398 (balanceHeap_knnIds): a
// max-heap on (distance, id), the larger id being "bigger" on equal distance.
void sift_knn(double* heap, int* ids, int size, int levels) {
  int parent = 0;
  for (int level = 0; level < levels - 1; ++level) {
    int left = parent * 2 + 1, right = left + 1, mx = parent;
    double diff;
    if (left < size && (diff = heap[left] - heap[parent]) >= 0) {
      if (diff > 0 || ids[left] > ids[parent]) mx = left;
    }
    if (right < size && (diff = heap[right] - heap[mx]) >= 0) {
      if (diff > 0 || ids[right] > ids[mx]) mx = right;
    }
    if (mx == parent) break;
    std::swap(heap[parent], heap[mx]);
    std::swap(ids[parent], ids[mx]);
    parent = mx;
  }
}

}  // namespace

extern "C" {

struct orc_index {
  int N, D, L, M;
  float W;
  std::vector<float> A;      // [L][D*M] as given (column-major per table)
  std::vector<float> b;      // [L][M]
  std::vector<double> At;    // [L][D][M] transposed, FP64
  std::vector<double> b64;   // [L][M]
  std::vector<Table> tables;
};

// Projection of N vectors for one table (building block for tests).
// v is [N x M] row-major: v[n*M + m].
void orc_project(const float* X, long long N, int D, const float* A_t, const float* b_t, int M,
                 float W, double* v) {
#pragma omp parallel for schedule(static)
  for (long long n = 0; n < N; ++n)
    for (int m = 0; m < M; ++m)
      v[n * M + m] = project1(X + (size_t)n * D, A_t + (size_t)m * D, D, b_t[m], W);
}

// HashTables::reset(N,D,L,M,W,A,b) + HashTables::index (cuLSH_HashTables.cu:213-241,
// 486-520) -> indexTableData per table (cuLSH_Indexing.cu:181-274):
//   codes = floor(v)                                  (Kernels.cu:91-101)
//   stable LSD sort of rows, column M-1 first          (Indexing.cu:71-93)
//     == lexicographic order of code vectors, ties by ascending id
//   head flags, starts, sizes                          (Indexing.cu:98-144)
//   codes of bucket heads                              (Indexing.cu:237-247)
// Returns NULL on invalid parameters (cuLSH_HashTables.cu:70).
orc_index* orc_index_build(int N, int D, int L, int M, float W, const float* A, const float* b,
                           const float* X) {
  if (!(N > 0 && D > 0 && L > 0 && M > 0 && W > 0.0f) || M > 256) return nullptr;
  orc_index* ix = new orc_index;
  ix->N = N;
  ix->D = D;
  ix->L = L;
  ix->M = M;
  ix->W = W;
  ix->A.assign(A, A + (size_t)L * D * M);
  ix->b.assign(b, b + (size_t)L * M);
  ix->At.resize((size_t)L * D * M);
  ix->b64.resize((size_t)L * M);
  for (int t = 0; t < L; ++t) {
    for (int d = 0; d < D; ++d)
      for (int m = 0; m < M; ++m)
        ix->At[((size_t)t * D + d) * M + m] = (double)A[(size_t)t * D * M + (size_t)m * D + d];
    for (int m = 0; m < M; ++m) ix->b64[(size_t)t * M + m] = (double)b[(size_t)t * M + m];
  }
  ix->tables.resize(L);
  // Tables are independent: they are processed in chunks of G (the thread
  // count), projecting the chunk's tables over all points in parallel, then
  // sorting and bucketing one table per thread.  Every table's result is the
  // same as when processed alone.
  const int G = std::max(1, std::min(L, omp_get_max_threads()));
  std::vector<std::vector<int32_t>> codes(G);
  for (int t0 = 0; t0 < L; t0 += G) {
    const int nt = std::min(G, L - t0);
    for (int g = 0; g < nt; ++g) codes[g].resize((size_t)N * M);
#pragma omp parallel for schedule(static)
    for (long long n = 0; n < N; ++n) {
      double v[256];
      for (int g = 0; g < nt; ++g) {
        const int t = t0 + g;
        project_row(X + (size_t)n * D, ix->At.data() + (size_t)t * D * M,
                    ix->b64.data() + (size_t)t * M, (double)W, D, M, v);
        for (int m = 0; m < M; ++m) codes[g][(size_t)n * M + m] = (int32_t)std::floor(v[m]);
      }
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int g = 0; g < nt; ++g) {
      const int32_t* cp = codes[g].data();
      std::vector<uint32_t> order(N);
      for (int n = 0; n < N; ++n) order[n] = (uint32_t)n;
      std::stable_sort(order.begin(), order.end(), [cp, M](uint32_t a, uint32_t c) {
        return lex_cmp(cp + (size_t)a * M, cp + (size_t)c * M, M) < 0;
      });
      Table& tb = ix->tables[t0 + g];
      tb.start.clear();
      tb.size.clear();
      tb.codes.clear();
      for (int i = 0; i < N; ++i) {
        if (i == 0 || lex_cmp(cp + (size_t)order[i] * M, cp + (size_t)order[i - 1] * M, M) != 0) {
          tb.start.push_back((uint32_t)i);
          tb.codes.insert(tb.codes.end(), cp + (size_t)order[i] * M,
                          cp + (size_t)order[i] * M + M);
        }
      }
      tb.total = std::move(order);
      tb.B = (int)tb.start.size();
      tb.size.resize(tb.B);
      for (int k = 0; k < tb.B; ++k)
        tb.size[k] = (k + 1 < tb.B ? tb.start[k + 1] : (uint32_t)N) - tb.start[k];
    }
  }
  return ix;
}

void orc_index_free(orc_index* ix) { delete ix; }

unsigned orc_index_buckets(const orc_index* ix, int t) { return (unsigned)ix->tables[t].B; }

// Bucket arrays in the reference's host layout (cuLSH_HashTables.cu:57-61):
// codes as float column-major [B x M], (k, m) at m*B + k.
void orc_index_table(const orc_index* ix, int t, unsigned* total, unsigned* start, unsigned* size,
                     float* codes) {
  const Table& tb = ix->tables[t];
  if (total) std::memcpy(total, tb.total.data(), sizeof(uint32_t) * ix->N);
  if (start) std::memcpy(start, tb.start.data(), sizeof(uint32_t) * tb.B);
  if (size) std::memcpy(size, tb.size.data(), sizeof(uint32_t) * tb.B);
  if (codes)
    for (int k = 0; k < tb.B; ++k)
      for (int m = 0; m < ix->M; ++m)
        codes[(size_t)m * tb.B + k] = (float)tb.codes[(size_t)k * ix->M + m];
}

// Multiprobe sequence for one (query, table): cuLSH_Kernels.cu:188-257
// (kernel_calculateProbingCodes) with deltas from cuLSH_Kernels.cu:266-289
// (kernel_calculateDeltas), in FP64.
//   v       [M] unfloored projections
//   codes   [T][M] out: probe t's code (probe 0 = floor(v))
//   masks   [T] out: perturbation set of probe t as bits over perm positions
//   perm    [min(2M,64)] out: delta indices in ascending delta order
// 64-bit shifts replace the reference's 32-bit `1 << i` (undefined for
// i >= 31, cuLSH_Kernels.cu:237,242,247); a set whose largest position is
// maxT-1 is re-emitted unchanged, as the reference does (:241-253).
// Returns maxT.
int orc_probe_sequence(const double* v, int M, int T, int* codes, unsigned long long* masks,
                       int* perm) {
  const int twoM = 2 * M;
  const int maxT = twoM < 64 ? twoM : 64;
  double d[512];
  int base[256];
  for (int m = 0; m < M; ++m) {
    double f = std::floor(v[m]);
    base[m] = (int)f;
    double dn = v[m] - f;
    d[m] = dn;
    d[M + m] = 1.0 - dn;
  }
  for (int m = 0; m < M; ++m) codes[m] = base[m];
  if (masks) masks[0] = 0;
  if (T <= 1) return maxT;
  int pm[64];
  const int levels = ceil_log2(twoM + 1);
  for (int i = 0; i < maxT; ++i) pm[i] = -1;
  for (int i = 0; i < twoM; ++i)
    if (pm[0] == -1 || d[i] < d[pm[0]]) {
      pm[0] = i;
      sift_deltas(pm, d, maxT, levels);
    }
  for (int i = maxT - 1; i; --i) {
    std::swap(pm[0], pm[i]);
    sift_deltas(pm, d, i, levels);
  }
  if (perm) std::memcpy(perm, pm, sizeof(int) * maxT);
  std::vector<uint64_t> comb(T);
  std::vector<double> score(T);
  comb[0] = 1;
  score[0] = d[pm[0]];
  int size = 1;
  for (int t = 1; t < T; ++t) {
    double minValue = DBL_MAX;
    int minIndex = 0;
    for (int i = 0; i < size; ++i)
      if (score[i] < minValue) {
        minIndex = i;
        minValue = score[i];
      }
    uint64_t probe = comb[minIndex];
    int* c = codes + (size_t)t * M;
    for (int m = 0; m < M; ++m) c[m] = base[m];
    int maxPos = 0;
    for (int i = 0; i < 64; ++i)
      if (probe & (1ull << i)) {
        c[pm[i] % M] += (pm[i] >= M) ? 1 : -1;
        maxPos = i;
      }
    if (masks) masks[t] = probe;
    if (maxPos < maxT - 1) {
      probe |= (1ull << (maxPos + 1));
      minValue += d[pm[maxPos + 1]];
      comb[minIndex] = probe;
      score[minIndex] = minValue;
      probe -= (1ull << maxPos);
      minValue -= d[pm[maxPos]];
      comb[size] = probe;
      score[size] = minValue;
      ++size;
    }
  }
  return maxT;
}

// Top-K by the reference's heap (cuLSH_Kernels.cu:406-456, driven by
// cuLSH_Querying.cu:231-277), literally, in FP64.  ids[] must be ascending
// (the sorted-unique candidate list, cuLSH_SearchTables.cu:344-360); the scan
// runs from the last candidate to the first.  Outputs are ascending by
// (distance, id), padded with (FLT_MAX, -1).
void orc_topk_heap(const double* dist, const unsigned* ids, int C, int K, int* out_ids,
                   double* out_d) {
  std::vector<double> heap(K, (double)FLT_MAX);
  std::vector<int> hid(K, -1);
  const int levels = ceil_log2(K + 1);
  for (int i = C - 1; i >= 0; --i)
    if (dist[i] < heap[0]) {
      heap[0] = dist[i];
      hid[0] = (int)ids[i];
      sift_knn(heap.data(), hid.data(), K, levels);
    }
  for (int i = K - 1; i >= 1; --i) {
    std::swap(heap[0], heap[i]);
    std::swap(hid[0], hid[i]);
    sift_knn(heap.data(), hid.data(), i, ceil_log2(i + 1));
  }
  for (int k = 0; k < K; ++k) {
    out_ids[k] = hid[k];
    out_d[k] = heap[k];
  }
}

// Squared L2 (cuLSH_Kernels.cu:301-348) in FP64: sequential FMA over d of
// the exact differences.
double orc_dist(const float* q, const float* x, int D) {
  double acc = 0.0;
  for (int d = 0; d < D; ++d) {
    double df = (double)q[d] - (double)x[d];
    acc = std::fma(df, df, acc);
  }
  return acc;
}

// Lookup of one probe code in a table: exact match over the lexicographically
// sorted bucket codes, else -1.  This is what cuLSH_Querying.cu:124-165
// (joint stable sort, predecessor scan, verify kernel Kernels.cu:124-138)
// computes.
int orc_lookup(const orc_index* ix, int t, const int* code) {
  const Table& tb = ix->tables[t];
  const int M = ix->M;
  int lo = 0, hi = tb.B;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    int c = lex_cmp(tb.codes.data() + (size_t)mid * M, code, M);
    if (c < 0)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo < tb.B && lex_cmp(tb.codes.data() + (size_t)lo * M, code, M) == 0) return lo;
  return -1;
}

// SearchTables::search (cuLSH_SearchTables.cu:197-463) for all queries:
//   per table: project, multiprobe or floor, lookup       (Querying.cu:171-225)
//   C_nu(q) = sum of matched bucket sizes                  (SearchTables.cu:249-255)
//   U(q) = sorted unique union of matched bucket members   (SearchTables.cu:321-364)
//   squared L2 over U(q), heap top-K                       (Querying.cu:231-343)
// Outputs (any may be NULL):
//   knn_ids [Q*K] int, knn_dists [Q*K] float ((float) of the FP64 distance)
//   query_buckets [L][Q*T] (entry q*T+p), lexicographic bucket numbering
//   cnu, cu [Q] per-query candidate counts
// Returns 0 on success, -1 on invalid arguments.
int orc_search(const orc_index* ix, const float* queries, int Q, const float* X, int K, int T,
               int* knn_ids, float* knn_dists, int* query_buckets, unsigned long long* cnu,
               unsigned long long* cu) {
  if (!ix || Q <= 0 || K <= 0 || T <= 0) return -1;
  const int D = ix->D, L = ix->L, M = ix->M;
#pragma omp parallel
  {
    std::vector<double> v(M);
    std::vector<int> pc((size_t)T * M);
    std::vector<uint32_t> cand;
    std::vector<double> dist;
    std::vector<int> oid(K);
    std::vector<double> od(K);
#pragma omp for schedule(dynamic, 4)
    for (int q = 0; q < Q; ++q) {
      const float* y = queries + (size_t)q * D;
      cand.clear();
      unsigned long long nonu = 0;
      for (int t = 0; t < L; ++t) {
        project_row(y, ix->At.data() + (size_t)t * D * M, ix->b64.data() + (size_t)t * M,
                    (double)ix->W, D, M, v.data());
        orc_probe_sequence(v.data(), M, T, pc.data(), nullptr, nullptr);
        const Table& tb = ix->tables[t];
        for (int p = 0; p < T; ++p) {
          int k = orc_lookup(ix, t, pc.data() + (size_t)p * M);
          if (query_buckets) query_buckets[(size_t)t * Q * T + (size_t)q * T + p] = k;
          if (k >= 0) {
            nonu += tb.size[k];
            cand.insert(cand.end(), tb.total.begin() + tb.start[k],
                        tb.total.begin() + tb.start[k] + tb.size[k]);
          }
        }
      }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      if (cnu) cnu[q] = nonu;
      if (cu) cu[q] = cand.size();
      dist.resize(cand.size());
      for (size_t i = 0; i < cand.size(); ++i)
        dist[i] = orc_dist(y, X + (size_t)cand[i] * D, D);
      orc_topk_heap(dist.data(), cand.data(), (int)cand.size(), K, oid.data(), od.data());
      for (int k = 0; k < K; ++k) {
        if (knn_ids) knn_ids[(size_t)q * K + k] = oid[k];
        if (knn_dists) knn_dists[(size_t)q * K + k] = (float)od[k];
      }
    }
  }
  return 0;
}

// Exact brute-force kNN in FP64 (ground truth for recall): for each query the
// K smallest distances over all N vectors, ties by ascending id.
void orc_bruteforce_knn(const float* X, long long N, int D, const float* queries, int Q, int K,
                        int* ids, double* dists) {
#pragma omp parallel
  {
    std::vector<std::pair<double, int>> best;
#pragma omp for schedule(dynamic, 1)
    for (int q = 0; q < Q; ++q) {
      best.clear();
      const float* y = queries + (size_t)q * D;
      for (long long n = 0; n < N; ++n) {
        double dd = orc_dist(y, X + (size_t)n * D, D);
        if ((int)best.size() < K) {
          best.emplace_back(dd, (int)n);
          std::push_heap(best.begin(), best.end());
        } else if (std::make_pair(dd, (int)n) < best.front()) {
          std::pop_heap(best.begin(), best.end());
          best.back() = std::make_pair(dd, (int)n);
          std::push_heap(best.begin(), best.end());
        }
      }
      std::sort_heap(best.begin(), best.end());
      for (int k = 0; k < K; ++k) {
        ids[(size_t)q * K + k] = k < (int)best.size() ? best[k].second : -1;
        dists[(size_t)q * K + k] = k < (int)best.size() ? best[k].first : (double)FLT_MAX;
      }
    }
  }
}

int orc_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
