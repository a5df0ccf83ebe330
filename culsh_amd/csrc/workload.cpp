// Synthetic workload generator for the LSH index/query path.
//
// Produces datasets, queries and projection parameters in the layout the
// cuLSH API consumes: column-major float [D x N] (vector n is contiguous at
// n*D, cuLSH README.md:9), A_t as column-major [D x M] ((d,m) at m*D+d,
// cuLSH_Indexing.cu:43-51) and b_t as [M].
//
// Randomness: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11), a counter-based generator, so every
// element is a pure function of (seed, stream, index) and the fill is
// parallel and reproducible.  Normals come from the Box-Muller transform;
// Gamma(2, theta) is -theta * (ln U1 + ln U2) (sum of two exponentials).
//
// Host-only code (g++), built into libculsh_workload.so.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "philox.hpp"

namespace {

using culsh_rng::draw4;
using culsh_rng::normal;
using culsh_rng::u01_open0;
using culsh_rng::u01_open1;
using culsh_rng::uniform;

enum : uint32_t {
  S_CENTER = 1,
  S_ASSIGN = 2,
  S_NOISE = 3,
  S_ZERO = 4,
  S_GAMMA = 5,
  S_PICK = 6,
  S_A = 7,
  S_B = 8,
};

}  // namespace

extern "C" {

// Gaussian mixture: n_centers centres drawn U[lo, hi)^D, each point picks a
// centre uniformly and adds N(0, sigma^2) per dimension.
void culsh_gen_mixture(float* X, long long N, int D, int n_centers, double sigma,
                       double lo, double hi, unsigned long long seed) {
  double* centers = new double[(size_t)n_centers * D];
  for (long long i = 0; i < (long long)n_centers * D; ++i)
    centers[i] = lo + (hi - lo) * uniform(seed, S_CENTER, (uint64_t)i);
#pragma omp parallel for schedule(static)
  for (long long n = 0; n < N; ++n) {
    int c = (int)(uniform(seed, S_ASSIGN, (uint64_t)n) * n_centers);
    if (c >= n_centers) c = n_centers - 1;
    const double* cc = centers + (size_t)c * D;
    float* x = X + (size_t)n * D;
    for (int d = 0; d < D; ++d)
      x[d] = (float)(cc[d] + sigma * normal(seed, S_NOISE, (uint64_t)n * D + d));
  }
  delete[] centers;
}

// Sparse non-negative integers: per dimension zero with probability p_zero,
// otherwise min(maxv, round(Gamma(2, theta))).
void culsh_gen_gamma_ints(float* X, long long N, int D, double p_zero, double theta,
                          double maxv, unsigned long long seed) {
#pragma omp parallel for schedule(static)
  for (long long n = 0; n < N; ++n) {
    float* x = X + (size_t)n * D;
    for (int d = 0; d < D; ++d) {
      uint64_t i = (uint64_t)n * D + d;
      if (uniform(seed, S_ZERO, i) < p_zero) {
        x[d] = 0.0f;
        continue;
      }
      uint32_t w[4];
      draw4(seed, S_GAMMA, i, w);
      double g = -theta * (std::log(u01_open0(w[0], w[1])) + std::log(u01_open0(w[2], w[3])));
      double r = std::round(g);
      x[d] = (float)(r > maxv ? maxv : r);
    }
  }
}

// Queries near the data: a uniformly picked data vector plus N(0, noise^2).
// noise == 0 copies the picked vector.
void culsh_gen_queries_near(float* Qm, int Q, const float* X, long long N, int D,
                            double noise, unsigned long long seed) {
#pragma omp parallel for schedule(static)
  for (int q = 0; q < Q; ++q) {
    long long j = (long long)(uniform(seed, S_PICK, (uint64_t)q) * (double)N);
    if (j >= N) j = N - 1;
    const float* x = X + (size_t)j * D;
    float* y = Qm + (size_t)q * D;
    for (int d = 0; d < D; ++d)
      y[d] = (float)((double)x[d] + noise * normal(seed, S_NOISE, (uint64_t)q * D + d));
  }
}

// Projection parameters for L tables: A_t entries N(0,1) (column-major
// [D x M], table t at offset t*D*M) and b_t entries U[0, W) (table t at t*M).
void culsh_gen_projection(float* A, float* b, int L, int D, int M, float W,
                          unsigned long long seed) {
  long long na = (long long)L * D * M;
  for (long long i = 0; i < na; ++i) A[i] = (float)normal(seed, S_A, (uint64_t)i);
  for (long long i = 0; i < (long long)L * M; ++i) {
    float v = (float)(uniform(seed, S_B, (uint64_t)i) * (double)W);
    if (v >= W) v = std::nextafter(W, 0.0f);
    b[i] = v;
  }
}

// Raw Philox words, for known-answer tests of the generator itself.
void culsh_philox4x32_10(const unsigned int ctr[4], const unsigned int key[2], unsigned int out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  culsh_rng::philox_block(c, key[0], key[1]);
  std::memcpy(out, c, sizeof(c));
}

}  // extern "C"
