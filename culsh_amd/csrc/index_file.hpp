// Reader of the index file of HashTables::save / load
// (cuLSH_HashTables.cu:395-483): native endianness, no header magic,
//   int32 N, D, L, M; float W
//   per table: float A_t[D*M], float b_t[M]
//   uint32 B[L]
//   per table: uint32 totalIndices[N], startingIndices[B_t], sizes[B_t],
//              float codes[B_t*M] (column-major [B_t x M])
// Plain host C++ (no HIP): the file is read and checked in full before the
// caller touches the device, and every allocation is bounded by the file's
// size, so a truncated or corrupted file is an error, never an oversized
// allocation or an out-of-range read.
#pragma once
#include <sys/types.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/culsh_c.h"

namespace culsh_file {

constexpr int kCodeLimit = 32766;  // codes are kept as int16 on the device

struct IndexFile {
  int N = 0, D = 0, L = 0, M = 0;
  float W = 0.f;
  std::vector<float> A, b;  // [L][D*M], [L][M]
  std::vector<unsigned> B;  // [L]
  std::vector<std::vector<unsigned>> total, start, size;
  std::vector<std::vector<float>> codes;
};

// Size limits of the index (shared with HashTables::reset); "" when within.
inline std::string limits_message(long long N, long long D, long long L, long long M) {
  char buf[160];
  if (M > 64) {
    snprintf(buf, sizeof(buf), "M = %lld exceeds the supported maximum of 64", M);
    return buf;
  }
  if (L * N >= (1ll << 31)) {
    snprintf(buf, sizeof(buf), "L*N = %lld exceeds 2^31 - 1", L * N);
    return buf;
  }
  if (L * M >= (1ll << 24)) return "L*M exceeds 2^24";
  if (D > (1ll << 20)) return "D exceeds 2^20";
  return "";
}

// Reads and checks the whole file.  Returns CULSH_OK, or an error code with
// the message in `err` (CULSH_E_IO: unreadable, truncated or inconsistent
// file; CULSH_E_ARG: parameters not positive; CULSH_E_LIMIT: beyond this
// implementation's limits; CULSH_E_RANGE: a code that is not a 16-bit
// integer).
inline int read_index_file(FILE* fp, const char* name, IndexFile& f, std::string& err) {
  char buf[320];
  auto bad = [&](int code, const char* fmt, auto... args) {
    snprintf(buf, sizeof(buf), fmt, args...);
    err = buf;
    return code;
  };
  if (fseeko(fp, 0, SEEK_END) != 0) return bad(CULSH_E_IO, "\"%s\": cannot seek", name);
  const long long fsize = (long long)ftello(fp);
  if (fsize < 0 || fseeko(fp, 0, SEEK_SET) != 0)
    return bad(CULSH_E_IO, "\"%s\": cannot read its size", name);
  auto rd = [fp](void* p, size_t sz, size_t n) { return n == 0 || fread(p, sz, n, fp) == n; };
  int hdr[4];
  if (!rd(hdr, 4, 4) || !rd(&f.W, 4, 1)) return bad(CULSH_E_IO, "\"%s\": truncated header", name);
  f.N = hdr[0];
  f.D = hdr[1];
  f.L = hdr[2];
  f.M = hdr[3];
  const long long N = f.N, D = f.D, L = f.L, M = f.M;
  if (!(N > 0 && D > 0 && L > 0 && M > 0 && f.W > 0.0f))
    return bad(CULSH_E_ARG, "%s",
               "Projection parameters must be positive! Have you reset the index structure "
               "properly???");
  const std::string lim = limits_message(N, D, L, M);
  if (!lim.empty()) return bad(CULSH_E_LIMIT, "\"%s\": %s", name, lim.c_str());
  // Header, projections and bucket counts must fit the file before any
  // allocation sized by them.
  const long long head = 20 + 4 * L * (D * M + M) + 4 * L;
  if (head > fsize)
    return bad(CULSH_E_IO, "\"%s\": truncated projections (%lld bytes, at least %lld needed)",
               name, fsize, head);
  f.A.resize((size_t)(L * D * M));
  f.b.resize((size_t)(L * M));
  for (long long t = 0; t < L; ++t)
    if (!rd(f.A.data() + t * D * M, 4, (size_t)(D * M)) || !rd(f.b.data() + t * M, 4, (size_t)M))
      return bad(CULSH_E_IO, "\"%s\": truncated projection of table %lld", name, t);
  f.B.assign((size_t)L, 0u);
  if (!rd(f.B.data(), 4, (size_t)L)) return bad(CULSH_E_IO, "\"%s\": truncated bucket counts", name);
  long long need = head;
  for (long long t = 0; t < L; ++t) {
    const long long Bt = f.B[(size_t)t];
    if (Bt == 0 || Bt > N)
      return bad(CULSH_E_IO, "\"%s\": table %lld has %lld buckets for N = %lld", name, t, Bt, N);
    need += 4 * (N + 2 * Bt + Bt * M);
  }
  if (need > fsize)
    return bad(CULSH_E_IO, "\"%s\": truncated buckets (%lld bytes, %lld needed)", name, fsize,
               need);
  f.total.assign((size_t)L, {});
  f.start.assign((size_t)L, {});
  f.size.assign((size_t)L, {});
  f.codes.assign((size_t)L, {});
  std::vector<char> seen((size_t)N);
  for (long long t = 0; t < L; ++t) {
    const size_t Bt = f.B[(size_t)t];
    auto& tot = f.total[(size_t)t];
    auto& st = f.start[(size_t)t];
    auto& sz = f.size[(size_t)t];
    auto& cd = f.codes[(size_t)t];
    tot.resize((size_t)N);
    st.resize(Bt);
    sz.resize(Bt);
    cd.resize(Bt * (size_t)M);
    if (!rd(tot.data(), 4, (size_t)N) || !rd(st.data(), 4, Bt) || !rd(sz.data(), 4, Bt) ||
        !rd(cd.data(), 4, Bt * (size_t)M))
      return bad(CULSH_E_IO, "\"%s\": truncated buckets of table %lld", name, t);
    // Sizes tile [0, N) in order and every id appears once.
    unsigned long long pos = 0;
    for (size_t k = 0; k < Bt; ++k) {
      if (st[k] != pos || sz[k] == 0 || sz[k] > (unsigned long long)N - pos)
        return bad(CULSH_E_IO, "\"%s\": table %lld bucket %zu has an inconsistent start/size",
                   name, t, k);
      pos += sz[k];
    }
    if (pos != (unsigned long long)N)
      return bad(CULSH_E_IO, "\"%s\": table %lld bucket sizes sum to %llu, not N", name, t, pos);
    std::fill(seen.begin(), seen.end(), 0);
    for (long long i = 0; i < N; ++i) {
      const uint32_t id = tot[(size_t)i];
      if (id >= (uint32_t)N || seen[id])
        return bad(CULSH_E_IO, "\"%s\": table %lld lists id %u twice or out of range", name, t,
                   id);
      seen[id] = 1;
    }
    for (size_t e = 0; e < cd.size(); ++e) {
      const float c = cd[e];
      if (!(c == std::floor(c)) || std::fabs(c) > (float)kCodeLimit)
        return bad(CULSH_E_RANGE, "table %lld bucket %zu: code %g is not a 16-bit integer", t,
                   e % Bt, (double)c);
    }
  }
  return CULSH_OK;
}

}  // namespace culsh_file
