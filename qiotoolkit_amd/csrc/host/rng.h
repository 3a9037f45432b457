// Host-side counter-based RNG of the device contract (Philox4x32-10, Salmon
// et al. SC'11) for the few draws the host solvers make.
#pragma once
#include <cstdint>

namespace qiotk {

struct Words4 {
  uint32_t w[4];
};

inline Words4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return Words4{{c0, c1, c2, c3}};
}

// (w0 + w1 * 2^32) / 2^64, below 1
inline double canonical2(uint32_t w0, uint32_t w1) {
  double r = ((double)w0 + (double)w1 * 4294967296.0) / 18446744073709551616.0;
  return r >= 1.0 ? 0.99999999999999988898 : r;
}

}  // namespace qiotk
