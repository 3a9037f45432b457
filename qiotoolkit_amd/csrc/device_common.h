// Device-side building blocks shared by the sweep kernels: Philox4x32-10,
// the ln-table Metropolis accept, and the counter layout of the RNG contract.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qmc {

// Counter domains: ctr = (c0, c1, c2, domain).
enum : uint32_t { DOM_SWEEP = 0, DOM_REDRAW = 1, DOM_INIT = 2, DOM_EXCHANGE = 3, DOM_POISSON = 4, DOM_PICK = 5 };

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  The 32x32->64 products
// map to v_mad_u64_u32; the key schedule is wave-uniform when the key is.
__device__ __forceinline__ uint4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                               uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    // three-input XORs (v_bitop3_b32, truth table 0x96 = a ^ b ^ c)
    const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
    const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return make_uint4(c0, c1, c2, c3);
}

__device__ __forceinline__ uint32_t pick(const uint4& v, uint32_t w) {
  return w == 0 ? v.x : (w == 1 ? v.y : (w == 2 ? v.z : v.w));
}

// (w0 + w1 * 2^32) / 2^64 clamped below 1: generate_canonical<double,53> over
// two 32-bit draws, as std::uniform_real_distribution<double>(0,1) computes it.
__device__ __forceinline__ double canonical2(uint32_t w0, uint32_t w1) {
  double s = __dadd_rn((double)w0, __dmul_rn((double)w1, 4294967296.0));
  double r = __dmul_rn(s, 5.421010862427522170037264e-20);  // exact: 2^-64
  return r >= 1.0 ? 0.99999999999999988898 : r;
}

// k / 16777214.0 correctly rounded for integer 0 <= k < 2^24 via one product
// and one FMA correction (checked exhaustively against IEEE division for every
// k in range).
__device__ __forceinline__ double div_ffffe(uint32_t k) {
  const double D = 16777214.0;
  const double inv = 5.9604648328104515558750364705e-08;  // RN(1/D)
  const double kd = (double)k;
  const double q0 = __dmul_rn(kd, inv);
  const double r = __fma_rn(-q0, D, kd);
  return __fma_rn(r, inv, q0);
}

// Metropolis acceptance with the reference's 4x256 ln table
// (markov/metropolis.h:64-104, utils/utils.cc:12-32).  `w0` is the first draw
// of the attempt; `redraw` yields rows 1..3 (only when bucket 0 is hit);
// lut(row, i) reads entry i of table row `row`.
template <class Lut, class Redraw>
__device__ __forceinline__ bool metropolis_accept(double cost_diff, double beta, uint32_t w0, Lut&& lut,
                                                  Redraw&& redraw) {
  if (cost_diff <= 0.0) return true;
  uint32_t r = w0;
  uint32_t bucket = r & 255u;
  int row = 0;
  if (bucket == 0u) {  // probability 1/256: zoom into the next table rows
    const uint4 w = redraw();
    r = w.x;
    bucket = r & 255u;
    row = 1;
    if (bucket == 0u) {
      r = w.y;
      bucket = r & 255u;
      row = 2;
      if (bucket == 0u) {
        r = w.z;
        bucket = r & 255u;
        row = 3;
        if (bucket == 0u) return true;
      }
    }
  }
  const double t = __dmul_rn(cost_diff, -beta);
  // Row 0, bucket >= 2: value >= ln(1/256) = -5.5452, so t below that can
  // never accept; bucket 1 interpolates from ln(0) = -inf and yields NaN,
  // which never accepts either.
  if (row == 0 && (bucket == 1u || t < -5.6)) return false;
  const double y0 = lut(row, (int)bucket - 1);
  const double y1 = lut(row, (int)bucket);
  const double distance = div_ffffe(r >> 8);
  const double value = __dadd_rn(y0, __dmul_rn(distance, __dsub_rn(y1, y0)));
  return value < t;
}

// The same decision without branches for a draw `r` of table row 0 or 1
// (the caller resolves the 2^-16 of draws whose row-1 bucket is 0 as well
// with metropolis_accept's row walk).  Bucket 1 reads y0 = ln(0) = -inf and
// the interpolation yields NaN, which never accepts; t below row 0's range
// (ln(1/256)) cannot accept either, so no early exits are needed.
template <class Lut>
__device__ __forceinline__ double row_value(uint32_t r, int row, Lut&& lut) {
  const uint32_t bucket = r & 255u;
  const int i1 = (int)bucket, i0 = i1 > 0 ? i1 - 1 : 0;
  const double y0 = lut(row, i0);
  const double y1 = lut(row, i1);
  const double distance = div_ffffe(r >> 8);
  return __dadd_rn(y0, __dmul_rn(distance, __dsub_rn(y1, y0)));
}
// The threshold side of the decision, `value` = row_value(r, row, lut),
// depends on the draw only: it can be formed before the cost difference.
__device__ __forceinline__ bool accept_value(double cost_diff, double beta, double value) {
  return cost_diff <= 0.0 || value < __dmul_rn(cost_diff, -beta);
}
template <class Lut>
__device__ __forceinline__ bool accept_row(double cost_diff, double beta, uint32_t r, int row, Lut&& lut) {
  return accept_value(cost_diff, beta, row_value(r, row, lut));
}

}  // namespace qmc
