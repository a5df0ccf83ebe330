// culsh_cli — command-line driver with the reference CLI's arguments
// (cuLSH_main.cu:21-108): a raw float32 [D x N] file, the first Q vectors as
// queries, random projections, index, search, timing and selectivity.
// Built with a plain C++ compiler against include/cuLSH.h and libculsh.so.
//
//   culsh_cli <file> <D> <N> <L> <M> <W> <Q> <K> <T>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cuLSH.h"

int main(int argc, char** argv) {
  if (argc != 10) {
    printf("%d given arguments, while %d are required...\n", argc - 1, 9);
    printf("Arguments: [filename] [D] [N] [L] [M] [W] [Q] [K] [T]\n");
    return 1;
  }
  const int D = atoi(argv[2]), N = atoi(argv[3]), L = atoi(argv[4]), M = atoi(argv[5]);
  const float W = (float)atof(argv[6]);
  const int Q = atoi(argv[7]), K = atoi(argv[8]), T = atoi(argv[9]);
  if (D <= 0 || N <= 0 || Q <= 0 || Q > N) {
    printf("Invalid sizes (need D, N > 0 and 0 < Q <= N)\n");
    return 1;
  }
  FILE* fp = fopen(argv[1], "rb");
  if (!fp) {
    printf("Could not open file with data matrix...\n");
    return 1;
  }
  std::vector<float> dataset((size_t)D * N);
  if (fread(dataset.data(), sizeof(float), dataset.size(), fp) != dataset.size()) {
    printf("Could not load data matrix from file...\n");
    fclose(fp);
    return 1;
  }
  fclose(fp);
  std::vector<float> queries(dataset.begin(), dataset.begin() + (size_t)D * Q);
  printf("data: [%d x %d], queries: [%d x %d], (L, M, W) = (%d, %d, %.2f), (Q, K, T) = (%d, %d, %d)\n",
         D, N, D, Q, L, M, W, Q, K, T);

  cuLSH::HashTables hashtables;
  cuLSH::SearchTables searchtables;
  if (!hashtables.reset(N, D, L, M, W)) {
    printf("Could not reset tables...\n");
    return 1;
  }
  searchtables.reset(&hashtables, K, T);
  auto t0 = std::chrono::steady_clock::now();
  if (!hashtables.index(dataset.data())) {
    printf("Failed to perform indexing...\n");
    return 1;
  }
  auto t1 = std::chrono::steady_clock::now();
  if (!searchtables.search(queries.data(), Q, dataset.data())) {
    printf("Failed to perform searching...\n");
    return 1;
  }
  auto t2 = std::chrono::steady_clock::now();
  cuLSH::SearchTables::timesStruct times;
  searchtables.getTimes(&times);
  const double ti = std::chrono::duration<double>(t1 - t0).count();
  const double ts = std::chrono::duration<double>(t2 - t1).count();
  printf("%f seconds to index\n", ti);
  printf("%f seconds to search\n", ts);
  printf("\tBuckets:    %f\n", times.buckets);
  printf("\tCandidates: %f\n", times.distances);
  printf("Selectivity = %.6f%%\n", searchtables.getSelectivity() * 100.0);
  printf("Duplicate candidates = %.3f%%\n", searchtables.getNonUniquePercentage() * 100.0);
  const int* ids = searchtables.getKnnIds();
  const float* d = searchtables.getKnnDistances();
  printf("query 0 neighbours:");
  for (int k = 0; k < K && k < 10; ++k) printf(" %d(%.4g)", ids[k], d[k]);
  printf("\n");
  return 0;
}
