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

#include <cmath>

namespace qiotk {

// Counter domains used by the host solvers (the device uses 0..3).
enum : uint32_t { DOM_POISSON = 4, DOM_PICK = 5 };

// Poisson variate with mean `lam` from a stream of counter-based uniforms
// u(k), k = 0, 1, ...  Inversion by sequential search for lam < 30; above,
// the PTRS transformed-rejection method of W. Hormann, "The transformed
// rejection method for generating Poisson random variables", Insurance:
// Mathematics and Economics 12 (1993) 39-45.
template <class Uniform>
uint32_t poisson_draw(double lam, Uniform&& u) {
  if (!(lam > 0)) return 0;
  if (lam < 30.0) {
    const double x = u(0);
    double p = std::exp(-lam), F = p;
    uint32_t k = 0;
    while (x > F && k < 10000) {
      ++k;
      p *= lam / (double)k;
      F += p;
      if (p < 1e-300 && (double)k > lam) break;
    }
    return k;
  }
  const double slam = std::sqrt(lam), loglam = std::log(lam);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (uint32_t it = 0;; it += 2) {
    const double U = u(it) - 0.5;
    const double V = u(it + 1);
    const double us = 0.5 - std::fabs(U);
    const double kd = std::floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return (uint32_t)kd;
    if (kd < 0 || (us < 0.013 && V > us)) continue;
    if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <=
        -lam + kd * loglam - std::lgamma(kd + 1.0))
      return (uint32_t)kd;
  }
}

}  // namespace qiotk
