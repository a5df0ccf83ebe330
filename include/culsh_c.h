/*
 * culsh_c.h — C ABI of the MI355X-native LSH index/query library (libculsh.so).
 *
 * Every entry point takes plain pointers and sizes.  Host pointers are read
 * only during the call.  Status codes: 0 = success, non-zero = failure (see
 * CULSH_E_*); culsh_last_error() returns the message of the calling thread's
 * last failure.  Objects are not thread-safe; use one per thread.
 *
 * Layouts follow the reference (cuLSH README.md:9, cuLSH_Indexing.cu:43-51):
 *   dataset / queries : column-major float [D x N] (vector n at n*D)
 *   A[t]              : column-major float [D x M] ((d,m) at m*D + d)
 *   b[t]              : float [M]
 *   knn ids / dists   : [K x Q] ((k, q) at q*K + k), ascending (dist, id),
 *                       squared L2, padded with (FLT_MAX, -1)
 *   query buckets     : per table int [Q*T], entry q*T + p, -1 = no bucket
 *   bucket codes      : float column-major [B x M] ((k, m) at m*B + k)
 * Bucket numbering and bucket arrays are in lexicographic code order, as the
 * reference produces them (cuLSH_Indexing.cu:71-93).
 */
#ifndef CULSH_C_H
#define CULSH_C_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct culsh_ht culsh_ht; /* cuLSH::HashTables  (cuLSH_HashTables.cu:51-108) */
typedef struct culsh_st culsh_st; /* cuLSH::SearchTables (cuLSH_SearchTables.cu:42-107) */

/* cuLSH::SearchTables::timesStruct (cuLSH_SearchTables.cu:44-52), seconds,
 * of the last culsh_st_search: buckets = projection, multiprobe and lookup
 * (probe kernel); sort = locality ordering of large query sets; collect,
 * distances, knn = the fused candidate stage's time (overflow kernel
 * included) split by the measured shader-clock shares of its phases --
 * member gathering with hash dedup (so unique, a separate step in the
 * reference, is 0), FP32 distances, and selection + FP64 recheck + top-K.
 * The precision form's grouped kernel reports its whole time as distances. */
typedef struct {
  float buckets, collect, sort, unique, distances, knn;
} culsh_times;

enum {
  CULSH_OK = 0,
  CULSH_E_ARG = 1,     /* invalid argument or parameters not positive */
  CULSH_E_STATE = 2,   /* call out of order (e.g. index before reset) */
  CULSH_E_DEVICE = 3,  /* HIP runtime error */
  CULSH_E_MEMORY = 4,  /* allocation failed / does not fit */
  CULSH_E_IO = 5,      /* file open/read/write failed */
  CULSH_E_RANGE = 6,   /* a hash code does not fit the 16-bit code store */
  CULSH_E_LIMIT = 7    /* a size limit of this implementation (M > 64, ...) */
};

/* Forms (culsh_ht_set_form for the index projection, culsh_st_set_form for
 * the search).  Every form returns the same ids and distances (the FP64
 * oracle's): FP64 rechecks decide every value near a bucket edge or the
 * K-th distance. */
enum {
  CULSH_FORM_PLAIN = 0,      /* FP32 projection (MFMA), direct FP32 distances */
  CULSH_FORM_PRECISION = 1,  /* BF16x3 split-precision projection (MFMA);
                                searches of the resident dataset compute the
                                distance pass on an FP16 copy of its rows
                                (2 N D bytes, built with one synchronization by
                                the first such search when D % 8 == 0 and the
                                values fit a half), others direct FP32 */
  CULSH_FORM_PRECISION_GROUPED = 2  /* search only: distances by the norm
                                expansion |q|^2 + |x|^2 - 2 q.x on MFMA for groups
                                of 16 locality-ordered queries (pays only when
                                the queries of a group share most candidates) */
};

const char* culsh_last_error(void);
const char* culsh_version(void);

/* ---- HashTables -------------------------------------------------------- */
int culsh_ht_create(culsh_ht** out);
void culsh_ht_destroy(culsh_ht* ht);

/* HashTables::reset(N,D,L,M,W,dbg) (cuLSH_HashTables.cu:190-210): random
 * A ~ N(0,1), b = (U - min U) * W per table.  The reference seeds with
 * time(0); this entry point takes the seed explicitly. */
int culsh_ht_reset_random(culsh_ht* ht, int N, int D, int L, int M, float W,
                          unsigned long long seed, FILE* dbg);
/* HashTables::reset(N,D,L,M,W,A,b,dbg) (cuLSH_HashTables.cu:213-241):
 * deep-copies A[t] ([D x M]) and b[t] ([M]) for t < L. */
int culsh_ht_reset(culsh_ht* ht, int N, int D, int L, int M, float W,
                   const float* const* A, const float* const* b, FILE* dbg);
/* HashTables::reset(filename, dbg) / load (cuLSH_HashTables.cu:243-249,441-483). */
int culsh_ht_load(culsh_ht* ht, const char* filename, FILE* dbg);
/* HashTables::save (cuLSH_HashTables.cu:395-438), byte-compatible format. */
int culsh_ht_save(culsh_ht* ht, const char* filename, FILE* dbg);

/* HashTables::index(matrix, dbg) (cuLSH_HashTables.cu:486-520): matrix is a
 * host [D x N] dataset.  The dataset stays resident on the device. */
int culsh_ht_index(culsh_ht* ht, const float* matrix, FILE* dbg);
/* Same, from a device pointer (dataset already in HBM, not copied: the
 * caller keeps it alive while the index is used). */
int culsh_ht_index_device(culsh_ht* ht, const float* d_matrix, FILE* dbg);

/* Getters (cuLSH_HashTables.cu:89-103).  Bucket arrays are host mirrors,
 * materialised on first use, valid until the next reset/index/load. */
int culsh_ht_N(const culsh_ht* ht);
int culsh_ht_D(const culsh_ht* ht);
int culsh_ht_L(const culsh_ht* ht);
int culsh_ht_M(const culsh_ht* ht);
float culsh_ht_W(const culsh_ht* ht);
const float* culsh_ht_A(culsh_ht* ht, int table);
const float* culsh_ht_b(culsh_ht* ht, int table);
unsigned culsh_ht_buckets(culsh_ht* ht, int table);
const unsigned* culsh_ht_buckets_all(culsh_ht* ht);
const unsigned* culsh_ht_buckets_total_indices(culsh_ht* ht, int table);
const unsigned* culsh_ht_buckets_starting_indices(culsh_ht* ht, int table);
const unsigned* culsh_ht_buckets_sizes(culsh_ht* ht, int table);
const float* culsh_ht_buckets_codes(culsh_ht* ht, int table);
/* Number of data projections re-evaluated in FP64 by the last index() (their
 * FP32/BF16x3 value was within the rounding-error bound of a bucket edge). */
unsigned long long culsh_ht_rechecked(const culsh_ht* ht);
/* Projection form of index(): CULSH_FORM_PLAIN (default) or _PRECISION. */
int culsh_ht_set_form(culsh_ht* ht, int form);

/* ---- SearchTables ------------------------------------------------------ */
int culsh_st_create(culsh_st** out);
void culsh_st_destroy(culsh_st* st);
/* SearchTables::reset(ht, K, T) (cuLSH_SearchTables.cu:127-139). */
int culsh_st_reset(culsh_st* st, culsh_ht* ht, int K, int T);
/* SearchTables::search(queries, Q, dataset, dbg) (cuLSH_SearchTables.cu:197-463):
 * host [D x Q] queries; distances are taken against `dataset` (when it is
 * the host pointer index() was given, the resident copy is used; another
 * host array is uploaded once and reused while the same pointer is passed
 * for the same index).  Like the reference (cuLSH_SearchTables.cu:283-311),
 * the queries run in passes sized from the free device memory at the call;
 * a pass also holds at most (2^31 - 1) / (L T) queries. */
int culsh_st_search(culsh_st* st, const float* queries, int Q, const float* dataset, FILE* dbg);
/* Caps the queries per pass of culsh_st_search / culsh_st_search_device
 * (0, the default: the free device memory decides).  No counterpart in the
 * reference; results do not depend on it. */
int culsh_st_set_pass_limit(culsh_st* st, int max_queries);
/* Device-resident search: queries, ids and dists are device pointers
 * ([D x Q], [K x Q], [K x Q]); runs on `stream` (a hipStream_t, 0 = null
 * stream) and returns without synchronising.  Query buckets and candidate
 * totals of this call are not copied to the host.  In the plain form the
 * call never waits for the device: a query whose candidates exceed the
 * device path's limits (2^20 non-unique candidates, 4096 candidates within
 * the FP32 error bound of the K-th distance) is reported by culsh_st_wait,
 * culsh_st_kernel_ms or a later call on the object that finds the search
 * completed; its results are not written. */
int culsh_st_search_device(culsh_st* st, const float* d_queries, int Q, int* d_knn_ids,
                           float* d_knn_dists, void* stream);
/* Waits for the last device search and reports its (or an earlier one's)
 * failure. */
int culsh_st_wait(culsh_st* st);
/* Allocation-free variant of culsh_st_search_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory -- code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Same kernels and results as culsh_st_search_device (not the
 * grouped form; the precision form reads the FP16 row copy only when an
 * earlier search built it; no locality reordering of large query sets,
 * which changes no result), but every working buffer lives in one fixed
 * 64 MiB device array
 * of the library (one per GPU, kept until the process ends), so the call
 * allocates nothing and never synchronises.  Limits: the working set of
 * the call (about 20 + 8 L + 8 L T bytes per query, plus 8 MiB) must fit the
 * array, else CULSH_E_LIMIT before any launch (c3's L = 32, T = 64: about
 * 3,500 queries); a query with more than 2^17 non-unique candidates or more
 * than 4096 candidates within the FP32 error bound of its K-th distance
 * gets no results and sets *d_status (a device int the caller provides,
 * cleared by the call) to non-zero.  All calls on one GPU share the array:
 * calls that overlap on the device race (order them on one stream). */
int culsh_st_search_device_noalloc(culsh_st* st, const float* d_queries, int Q, int* d_knn_ids,
                                   float* d_knn_dists, int* d_status, void* stream);
/* Distance form of search(): CULSH_FORM_PLAIN (default), _PRECISION (the
 * FP16 distance pass on the resident dataset) or _PRECISION_GROUPED (grouped
 * norm expansion). */
int culsh_st_set_form(culsh_st* st, int form);

/* Results and metrics (cuLSH_SearchTables.cu:83-105). */
const int* culsh_st_knn_ids(culsh_st* st);
const float* culsh_st_knn_dists(culsh_st* st);
const int* culsh_st_query_buckets(culsh_st* st, int table);
float culsh_st_selectivity(culsh_st* st);
float culsh_st_nonunique_pct(culsh_st* st);
unsigned long long culsh_st_total_unique(culsh_st* st);
unsigned long long culsh_st_total_nonunique(culsh_st* st);
void culsh_st_times(culsh_st* st, culsh_times* out);
/* Per-query candidate counts of the last search (host pointers, [Q]). */
const unsigned long long* culsh_st_query_cnu(culsh_st* st);
const unsigned long long* culsh_st_query_cu(culsh_st* st);
/* Candidates re-evaluated in FP64 by the last search (distance boundary). */
unsigned long long culsh_st_rechecked(culsh_st* st);
/* Device time (ms, HIP events on the search stream) of the last search's
 * probe kernel (projection + multiprobe + lookup) and candidate kernel
 * (gather + dedup + distances + top-K). */
int culsh_st_kernel_ms(culsh_st* st, float* probe_ms, float* candidates_ms);

/* ---- Batches of independent requests (batching form) -------------------
 * No counterpart in the reference, which searches one (K, T) per call
 * (cuLSH_SearchTables.cu:197-463).  A batch holds requests with their own
 * Q, K and T against one index; each request's results are exactly those of
 * culsh_st_reset(ht, K, T) + culsh_st_search_device on its queries alone.
 * The whole batch runs as one hipGraph launch: a descriptor copy, one
 * expansion kernel, the probe and candidate stages (one persistent fused
 * kernel, or a probe kernel and a persistent candidate kernel; see
 * culsh_schedule.fused) and one kernel for queries whose candidates overflow
 * on-chip memory. */
typedef struct culsh_bt culsh_bt;

typedef struct {
  int Q, K, T;         /* T <= 64 */
  int reserved;        /* 0 */
  long long q_row;     /* first query row in d_queries (row-major [rows][D]) */
  long long out_offset;/* first element of its [K x Q] results in d_ids / d_dists */
} culsh_request;

/* Schedule of the persistent kernel (picked by culsh_bt_autotune). */
typedef struct {
  int probe_queries;   /* projection tile: queries per pass (32, 16 or 8) */
  int probe_tables;    /* projection tile: tables per pass (32 / probe_queries) */
  int queries_per_item;/* probe-group width: queries per work item (a multiple
                          of probe_queries, <= 256) */
  int rows_in_flight;  /* candidate chunk: rows per 16-lane group (2, 4, 8) */
  int wgs_per_cu;      /* persistent workgroups per compute unit (1..3); 0: the
                          split schedule's candidate stage runs one workgroup
                          per query instead */
  int fused;           /* 1: one persistent kernel, a work item = a block of
                          queries (probes over all tables, then candidates);
                          0: a probe kernel over every request, then a
                          persistent candidate kernel */
} culsh_schedule;

/* A plan for batches against `ht`.  The first batch on a dataset builds an
 * FP16 copy of its rows (2 N D bytes, held by the index) for the batches'
 * distance pass; results are unchanged.  No counterpart in the reference. */
int culsh_bt_create(culsh_ht* ht, culsh_bt** out);
void culsh_bt_destroy(culsh_bt* bt);
int culsh_bt_set_schedule(culsh_bt* bt, const culsh_schedule* s);
int culsh_bt_get_schedule(culsh_bt* bt, culsh_schedule* s);
/* Enqueues the batch on `stream` and returns; d_queries, d_ids and d_dists
 * are device pointers that stay valid until culsh_bt_wait returns.  Waits
 * first for this object's previous batch. */
int culsh_bt_search_device(culsh_bt* bt, const culsh_request* reqs, int nreq,
                           const float* d_queries, int* d_ids, float* d_dists, void* stream);
/* Waits for the last batch; CULSH_E_LIMIT if a query exceeded a capacity
 * (its results are then not written). */
int culsh_bt_wait(culsh_bt* bt);
/* Device time (ms) of the last batch, HIP events around the graph launch. */
int culsh_bt_last_ms(culsh_bt* bt, float* ms);
/* Candidates re-evaluated in FP64 by the last batch. */
unsigned long long culsh_bt_rechecked(culsh_bt* bt);
/* Runs the batch once without the graph, the same kernels on `stream` with
 * HIP events between its stages, and waits: stage_ms[0] = descriptor
 * expansion + probe stage, [1] = locality order of the candidate stage,
 * [2] = candidate stage (k_batch_cand; the fused schedule's one kernel),
 * [3] = overflow kernel + header download.  Results are written as by
 * culsh_bt_search_device. */
int culsh_bt_profile(culsh_bt* bt, const culsh_request* reqs, int nreq, const float* d_queries,
                     int* d_ids, float* d_dists, void* stream, float* stage_ms);
/* Coordinate search over the schedule on this batch (reps timed runs per
 * candidate schedule, best kept), starting from a fixed naive schedule
 * (fused, 32x1 probe tile, 32 queries per item, 4 rows in flight, one
 * workgroup per CU) and sweeping until a pass changes nothing; leaves the
 * best schedule set.  Reports the naive and tuned batch times (ms). */
int culsh_bt_autotune(culsh_bt* bt, const culsh_request* reqs, int nreq, const float* d_queries,
                      int* d_ids, float* d_dists, int reps, float* default_ms, float* tuned_ms);

/* ---- Serving (continuous batching over device-resident indexes) ---------
 * No counterpart in the reference, whose search takes one query set per call
 * and sizes its candidate batches from the free device memory
 * (cuLSH_SearchTables.cu:283-311).  A request loop on the calling thread:
 * requests arrive at their arrival times, wait in a per-index queue, and are
 * admitted into the next batch of that index while the batch fits a slot.
 * Slots are sized at creation from the free HBM; a request larger than an
 * empty slot grows the slot when the HBM free at that moment (times
 * hbm_fraction) covers it, and otherwise runs in pieces over successive
 * batches -- no request is rejected for its size.  Every slot holds one
 * culsh_bt; batches of different slots run concurrently on their own
 * streams.  A request's results equal a SearchTables search of it alone. */
typedef struct culsh_srv culsh_srv;

typedef struct {
  double arrival_s;      /* arrival time, seconds after the run starts */
  int index;             /* which of the server's indexes */
  int Q, K, T;           /* T <= 64 */
  long long q_row;       /* first query row in that index's query array */
  long long out_offset;  /* first element of its [K x Q] results */
} culsh_srv_request;

typedef struct {
  int slots_per_index;    /* concurrent batches per index (>= 1) */
  int batching;           /* 1: continuous batching; 0: one request at a time
                             through the search path (the plain baseline) */
  double hbm_fraction;    /* share of free HBM the slots may use (0 < f <= 1) */
  long long max_batch_queries; /* per batch (0 = 4096); HBM may lower it.  The
                                  batch plans' fixed grids scale with it, so a
                                  smaller cap means less per-batch overhead */
  int no_growth;          /* 1: slots keep their initial size; a request
                             larger than a slot always runs in pieces */
} culsh_srv_config;

typedef struct {
  double p50_ms, p99_ms, mean_ms, max_ms; /* request latency: completion on the
                                             GPU timeline minus arrival */
  double queries_per_s;   /* queries / (last completion - first arrival) */
  double requests_per_s;
  long long batches, queries, requests;
  double mean_batch_requests;
  long long slot_bytes;   /* HBM reserved per slot at creation */
  long long deferred;     /* admissions postponed for lack of slot room */
  long long split_requests; /* requests that ran in pieces (larger than a slot) */
  long long grown_slots;  /* slot growths by the free HBM at admission */
  int host_io;            /* 1: latencies include query upload + result download */
} culsh_srv_stats;

int culsh_srv_create(culsh_ht* const* indexes, int nindex, const culsh_srv_config* cfg,
                     culsh_srv** out);
void culsh_srv_destroy(culsh_srv* srv);
/* Serves the requests (sorted by arrival) in real time; d_queries[i],
 * d_ids[i], d_dists[i] are index i's device arrays.  latency_ms (n entries,
 * may be null) receives each request's latency.  Returns after the last
 * completion. */
int culsh_srv_run(culsh_srv* srv, const culsh_srv_request* reqs, int n,
                  const float* const* d_queries, int* const* d_ids, float* const* d_dists,
                  double* latency_ms, culsh_srv_stats* stats);
/* Same, with the queries in host memory and the results wanted there: each
 * batch uploads its requests' query rows (h_queries[i] -> d_queries[i], same
 * row offsets) on the slot's stream before its graph launch and downloads
 * their results (d_ids[i] / d_dists[i] -> h_ids[i] / h_dists[i], same
 * offsets) after it; a request completes when its results are in host
 * memory.  Host arrays should be pinned (hipHostMalloc) for the copies to
 * overlap. */
int culsh_srv_run_host(culsh_srv* srv, const culsh_srv_request* reqs, int n,
                       const float* const* h_queries, int* const* h_ids, float* const* h_dists,
                       const float* const* d_queries, int* const* d_ids, float* const* d_dists,
                       double* latency_ms, culsh_srv_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* CULSH_C_H */
