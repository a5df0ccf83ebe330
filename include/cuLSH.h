/*
 * cuLSH.h — drop-in C++ interface for the MI355X-native LSH library.
 *
 * Re-creates cuLSH::HashTables and cuLSH::SearchTables with the reference's
 * signatures (cuLSH_HashTables.cu:77-107, cuLSH_SearchTables.cu:44-105) as
 * inline forwarders to the C ABI in culsh_c.h.  Plain C++: callers need no
 * HIP headers; link with -lculsh.  Failures return false and print
 * "[Error @ file, l.N] message" to stdout, as the reference does (cuLSH.h:79).
 */
#ifndef CULSH_DROPIN_H
#define CULSH_DROPIN_H

#include <cstdio>

#include "culsh_c.h"

namespace cuLSH {

namespace detail {
inline bool check(int rc, const char* file, int line) {
  if (rc == CULSH_OK) return true;
  printf("[Error @ %s, l.%d] %s\n", file, line, culsh_last_error());
  return false;
}
}  // namespace detail

#define CULSH_CHECK(expr) ::cuLSH::detail::check((expr), __FILE__, __LINE__)

class HashTables {
 public:
  HashTables() { culsh_ht_create(&h_); }
  ~HashTables() { culsh_ht_destroy(h_); }
  HashTables(const HashTables&) = delete;
  HashTables& operator=(const HashTables&) = delete;

  // Random A ~ N(0,1), b = (U - min U) W (cuLSH_HashTables.cu:190-210).  The
  // reference seeds with time(0); `seed` here defaults to the same.
  bool reset(int N, int D, int L, int M, float W, FILE* debugStream = 0) {
    return CULSH_CHECK(culsh_ht_reset_random(h_, N, D, L, M, W,
                                             (unsigned long long)time_seed(), debugStream));
  }
  bool reset(int N, int D, int L, int M, float W, float** A, float** b, FILE* debugStream = 0) {
    return CULSH_CHECK(culsh_ht_reset(h_, N, D, L, M, W, (const float* const*)A,
                                      (const float* const*)b, debugStream));
  }
  bool reset(const char* filename, FILE* debugStream = 0) {
    return CULSH_CHECK(culsh_ht_load(h_, filename, debugStream));
  }
  bool index(const float* matrix, FILE* debugStream = 0) {
    return CULSH_CHECK(culsh_ht_index(h_, matrix, debugStream));
  }

  int getN(void) { return culsh_ht_N(h_); }
  int getD(void) { return culsh_ht_D(h_); }
  int getL(void) { return culsh_ht_L(h_); }
  int getM(void) { return culsh_ht_M(h_); }
  float getW(void) { return culsh_ht_W(h_); }
  const float* getA(int table) { return culsh_ht_A(h_, table); }
  const float* getb(int table) { return culsh_ht_b(h_, table); }

  unsigned getBuckets(int table) { return culsh_ht_buckets(h_, table); }
  const unsigned* getBuckets(void) { return culsh_ht_buckets_all(h_); }
  const unsigned* getBuckets_totalIndices(int table) {
    return culsh_ht_buckets_total_indices(h_, table);
  }
  const unsigned* getBuckets_startingIndices(int table) {
    return culsh_ht_buckets_starting_indices(h_, table);
  }
  const unsigned* getBuckets_sizes(int table) { return culsh_ht_buckets_sizes(h_, table); }
  const float* getBuckets_codes(int table) { return culsh_ht_buckets_codes(h_, table); }

  bool save(const char* filename, FILE* debugStream = 0) {
    return CULSH_CHECK(culsh_ht_save(h_, filename, debugStream));
  }
  bool load(const char* filename, FILE* debugStream = 0) {
    return CULSH_CHECK(culsh_ht_load(h_, filename, debugStream));
  }

  culsh_ht* handle() { return h_; }

 private:
  static unsigned long long time_seed();
  culsh_ht* h_ = nullptr;
};

class SearchTables {
 public:
  typedef struct {
    float buckets;    // finding the queries' matching buckets (probe kernel)
    float collect;    // gathering candidates from matched buckets, deduplicated
                      // in the same pass (the candidate stage's measured share)
    float sort;       // locality ordering of large query sets
    float unique;     // 0: deduplication happens while collecting
    float distances;  // FP32 distances (share of the candidate stage)
    float knn;        // K-th selection, FP64 recheck and top-K (share)
    void reset(void) { buckets = collect = sort = unique = distances = knn = 0.0f; }
  } timesStruct;

  SearchTables() { culsh_st_create(&h_); }
  ~SearchTables() { culsh_st_destroy(h_); }
  SearchTables(const SearchTables&) = delete;
  SearchTables& operator=(const SearchTables&) = delete;

  void reset(HashTables* hashtables, int K, int T = 1) {
    CULSH_CHECK(culsh_st_reset(h_, hashtables ? hashtables->handle() : 0, K, T));
  }
  bool search(const float* queries, const int Q, const float* dataset, FILE* debugStream = 0) {
    return CULSH_CHECK(culsh_st_search(h_, queries, Q, dataset, debugStream));
  }

  float getSelectivity(void) { return culsh_st_selectivity(h_); }
  float getNonUniquePercentage(void) { return culsh_st_nonunique_pct(h_); }
  const int* getQueryBuckets(int table) { return culsh_st_query_buckets(h_, table); }
  unsigned getTotalCandidates_unique(void) { return (unsigned)culsh_st_total_unique(h_); }
  unsigned getTotalCandidates_nonUnique(void) { return (unsigned)culsh_st_total_nonunique(h_); }
  const int* getKnnIds(void) { return culsh_st_knn_ids(h_); }
  const float* getKnnDistances(void) { return culsh_st_knn_dists(h_); }
  const float* getKnnDists(void) { return culsh_st_knn_dists(h_); }
  void getTimes(timesStruct* times) {
    culsh_times t;
    culsh_st_times(h_, &t);
    times->buckets = t.buckets;
    times->collect = t.collect;
    times->sort = t.sort;
    times->unique = t.unique;
    times->distances = t.distances;
    times->knn = t.knn;
  }

 private:
  culsh_st* h_ = nullptr;
};

}  // namespace cuLSH

#include <ctime>
inline unsigned long long cuLSH::HashTables::time_seed() { return (unsigned long long)time(0); }

#undef CULSH_CHECK
#endif /* CULSH_DROPIN_H */
