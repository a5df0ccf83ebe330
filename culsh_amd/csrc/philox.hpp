// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) and the draws built
// on it.  Host-only.
#pragma once
#include <cmath>
#include <cstdint>

namespace culsh_rng {

inline void philox_block(uint32_t c[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c[0];
    const uint64_t p1 = (uint64_t)M1 * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = (uint32_t)p1;
    c[2] = n2;
    c[3] = (uint32_t)p0;
    k0 += W0;
    k1 += W1;
  }
}

// Four 32-bit words for (seed, stream, index).
inline void draw4(uint64_t seed, uint32_t stream, uint64_t index, uint32_t out[4]) {
  out[0] = (uint32_t)index;
  out[1] = (uint32_t)(index >> 32);
  out[2] = stream;
  out[3] = 0x5EEDu;
  philox_block(out, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// 53-bit uniform in (0, 1].
inline double u01_open0(uint32_t a, uint32_t b) {
  const uint64_t k = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return (double)(k + 1) * 0x1.0p-53;
}

// 53-bit uniform in [0, 1).
inline double u01_open1(uint32_t a, uint32_t b) {
  const uint64_t k = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return (double)k * 0x1.0p-53;
}

inline double uniform(uint64_t seed, uint32_t stream, uint64_t i) {
  uint32_t w[4];
  draw4(seed, stream, i, w);
  return u01_open1(w[0], w[1]);
}

inline double uniform_open0(uint64_t seed, uint32_t stream, uint64_t i) {
  uint32_t w[4];
  draw4(seed, stream, i, w);
  return u01_open0(w[0], w[1]);
}

// Standard normal number i of a stream (Box-Muller on one Philox block).
inline double normal(uint64_t seed, uint32_t stream, uint64_t i) {
  uint32_t w[4];
  draw4(seed, stream, i, w);
  const double u1 = u01_open0(w[0], w[1]);
  const double u2 = u01_open1(w[2], w[3]);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

}  // namespace culsh_rng
