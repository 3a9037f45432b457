// MI355X (gfx950) kernels of the Metropolis sweep engine.
//
// Data layout (per problem, in "position space"): variables are renumbered so
// that every colour class is a contiguous range of positions; inside a class
// the positions follow descending degree, then variable id.  Position N is a
// ghost that is always 0, so one-variable terms need no special case.
//
// Per walker group ("unit"): up to 32 replicas of one problem share a
// work-group.  Spins are packed bits, byte g of position q holding replicas
// 8g..8g+7 of the unit: spins[q*G + g].  Thread (g, tp) owns byte lane g and
// positions rho = k*lanes + tp of every class, so a wavefront reads
// consecutive bytes (conflict-free LDS) and one neighbour byte serves eight
// replicas.
//
// Sweep order: classes in order; inside a class, position-lanes tp in order,
// each lane's positions k = 0,1,.. in order.  Members of one class share no
// term, so the parallel update equals that sequential order exactly.  The
// reference's per-flip best tracking (markov/walker.h:277-280) is kept exact:
// each thread accumulates the prefix sum S and minimum prefix m of its
// accepted dE (and the round k of its first minimum); DPP wavefront scans and
// a cross-wave min-reduction give the minimum over the class's sequence and
// where it is attained, and the lowest state is rebuilt from the pre-class
// state and the flips recorded in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "qmc_internal.h"

namespace qmc {

constexpr int kMaxRep = 32;          // replicas per unit
constexpr int kMaxWavesPerGroup = 8; // lanes <= 512
constexpr int kBlock = 512;

template <typename T>
__device__ __forceinline__ T min_ct(T a, T b) { return b < a ? b : a; }

// Accumulator of a class's accepted dE: exact int32 for integer couplings,
// double for fp32 couplings (the reference accumulates its cost in double).
template <typename CT> struct AccT { using T = int32_t; };
template <> struct AccT<float> { using T = double; };

struct Smem {
  double lut[1024];         // the 4 x 256 ln table
  double rcost[kMaxRep];
  double rbest[kMaxRep];
  double rbeta[kMaxRep];
  double lbeta[kMaxRep];
  int32_t label[kMaxRep];
  int32_t inv[kMaxRep];     // ladder slot -> beta index
  double lcost[kMaxRep];    // ladder energies at the exchange
  int32_t abort_flag;
  int32_t any_new_best;
  uint32_t racc[kMaxRep];
  int32_t nb_c[kMaxRep];    // class of the sweep's new lowest cost, -1 none
  int32_t nb_tp[kMaxRep];   // its lane in the class sequence
  int32_t kstar[kMaxRep];   // its round (-1: before the lane's first flip)
};

// Per-wave partials of the class scans, in LDS right after Smem: entry
// ((p * GS + g) * 8 + b) * 8 + w for class parity p, byte group g, replica b
// of the group and wave w of the group (up to 8) of four arrays -- the sums
// and the minimum prefixes (int32 or fp64 in 8-byte slots), the lane and the
// round of each minimum.  GS, the byte groups laid out, is 1 below 8
// replicas per byte lane and else at most threads / 64 (and 4), so small
// work-groups keep little LDS for them.  A class's combine reads its
// parity's slots while the next class fills the other ones.
__host__ __device__ constexpr int slot_groups(int nb, int threads) {
  return nb < 8 ? 1 : (threads / 64 < 4 ? (threads / 64 < 1 ? 1 : threads / 64) : 4);
}
__host__ __device__ constexpr size_t slot_bytes(int groups) {
  return (size_t)groups * 2 * 8 * kMaxWavesPerGroup * 24;
}

template <typename AT> struct Slots {
  using lds_at = __attribute__((address_space(3))) AT;
  using lds_i32 = __attribute__((address_space(3))) int32_t;
  lds_at* sum;
  lds_at* mn;
  lds_i32* lane;
  lds_i32* k;
  int GS;
  __device__ Slots(unsigned char* base, int GS_) : GS(GS_) {
    const size_t n = (size_t)GS_ * 2 * 8 * kMaxWavesPerGroup;
    sum = (lds_at*)(base);
    mn = (lds_at*)(base + 8 * n);
    lane = (lds_i32*)(base + 16 * n);
    k = (lds_i32*)(base + 20 * n);
  }
  __device__ int at(int p, int g, int b, int w) const { return ((p * GS + g) * 8 + b) * kMaxWavesPerGroup + w; }
};

// ---------------------------------------------------------------------------
// Wavefront scans with DPP (gfx9 row_shr / row_bcast / wave_shr): lanes
// without a source take `ident`.
enum : int { DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114, DPP_ROW_SHR8 = 0x118,
             DPP_WAVE_SHR1 = 0x138, DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143 };

template <int CTRL, int ROWS>
__device__ __forceinline__ int32_t dpp(int32_t x, int32_t ident) {
  return __builtin_amdgcn_update_dpp(ident, x, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp(double x, double ident) {
  const long long xi = __double_as_longlong(x), oi = __double_as_longlong(ident);
  const int lo = __builtin_amdgcn_update_dpp((int)oi, (int)xi, CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(oi >> 32), (int)(xi >> 32), CTRL, ROWS, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Lanes without a source read 0 (bound_ctrl), every row written: no
// identity has to be placed in the destination first.
template <int CTRL>
__device__ __forceinline__ int32_t dpp_z(int32_t x) {
  return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xf, 0xf, true);
}
template <int CTRL>
__device__ __forceinline__ double dpp_z(double x) {
  const long long xi = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)xi, CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(xi >> 32), CTRL, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// Lanes without a source (and rows outside ROWS) keep `old`.
template <int CTRL, int ROWS>
__device__ __forceinline__ int32_t dpp_keep(int32_t old, int32_t x) {
  return __builtin_amdgcn_update_dpp(old, x, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_keep(double old, double x) {
  const long long xi = __double_as_longlong(x), oi = __double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)oi, (int)xi, CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(oi >> 32), (int)(xi >> 32), CTRL, ROWS, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <typename AT> __device__ __forceinline__ AT max_of();
template <> __device__ __forceinline__ double max_of<double>() { return __builtin_huge_val(); }
template <> __device__ __forceinline__ int32_t max_of<int32_t>() { return 0x7fffffff; }

// Inclusive prefix sums and prefix minima over the wave of the NB values of
// a lane's replicas, step by step, so the replicas' dependency chains
// interleave.  Sums: the row shifts read 0 where they have no source; the
// row broadcasts leave half the rows unwritten, which must read 0.
template <int CTRL, typename AT, int NB>
__device__ __forceinline__ void add_step_z(AT (&x)[8]) {
#pragma unroll
  for (int b = 0; b < NB; ++b) x[b] += dpp_z<CTRL>(x[b]);
}
template <int CTRL, int ROWS, typename AT, int NB>
__device__ __forceinline__ void add_step(AT (&x)[8]) {
#pragma unroll
  for (int b = 0; b < NB; ++b) x[b] += dpp<CTRL, ROWS>(x[b], AT(0));
}
template <typename AT, int NB>
__device__ __forceinline__ void scan_add_n(AT (&x)[8]) {
  add_step_z<DPP_ROW_SHR1, AT, NB>(x);
  add_step_z<DPP_ROW_SHR2, AT, NB>(x);
  add_step_z<DPP_ROW_SHR4, AT, NB>(x);
  add_step_z<DPP_ROW_SHR8, AT, NB>(x);
  add_step<DPP_ROW_BCAST15, 0xa, AT, NB>(x);
  add_step<DPP_ROW_BCAST31, 0xc, AT, NB>(x);
}
// Minima: a lane without a source (or in an unwritten row) keeps the
// previous step's value in t, which is a minimum over lanes at or before
// it in its own prefix, so it is never below the lane's running minimum
// and min() leaves that unchanged; t starts as x itself.
template <int CTRL, int ROWS, typename AT, int NB>
__device__ __forceinline__ void min_step(AT (&x)[8], AT (&t)[8]) {
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    t[b] = dpp_keep<CTRL, ROWS>(t[b], x[b]);
    x[b] = min_ct(x[b], t[b]);
  }
}
template <typename AT, int NB>
__device__ __forceinline__ void scan_min_n(AT (&x)[8]) {
  AT t[8];
#pragma unroll
  for (int b = 0; b < NB; ++b) t[b] = x[b];
  min_step<DPP_ROW_SHR1, 0xf, AT, NB>(x, t);
  min_step<DPP_ROW_SHR2, 0xf, AT, NB>(x, t);
  min_step<DPP_ROW_SHR4, 0xf, AT, NB>(x, t);
  min_step<DPP_ROW_SHR8, 0xf, AT, NB>(x, t);
  min_step<DPP_ROW_BCAST15, 0xa, AT, NB>(x, t);
  min_step<DPP_ROW_BCAST31, 0xc, AT, NB>(x, t);
}

__device__ __forceinline__ int32_t lane_value(int32_t x, int lane) { return __builtin_amdgcn_readlane(x, lane); }
__device__ __forceinline__ double lane_value(double x, int lane) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)v, lane), hi = __builtin_amdgcn_readlane((int)(v >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Global (addrspace 1) and constant (addrspace 4, read-only for the whole
// launch: scalar loads when the address is uniform) views of device memory.
// Plain pointers loaded from the unit descriptor are generic, and generic
// loads are flat loads that also wait on LDS traffic.
template <class T> using gptr = const __attribute__((address_space(1))) T*;
template <class T> using cptr = const __attribute__((address_space(4))) T*;
template <class T> __device__ __forceinline__ gptr<T> glob(const T* p) { return (gptr<T>)p; }
template <class T> __device__ __forceinline__ cptr<T> cnst(const T* p) { return (cptr<T>)p; }
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
// A wave-uniform value held in a vector register: the Philox round keys
// derived from it are then vector values too, instead of twenty scalar
// registers live across the whole sweep.
__device__ __forceinline__ uint32_t in_vgpr(uint32_t x) {
  uint32_t v;
  asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(x));
  return v;
}

// LDS bytes: index arithmetic stays 32-bit (ds_read/ds_write take a 32-bit
// address), where generic pointers into LDS make every gather a 64-bit
// multiply-add.
using lds_u8 = __attribute__((address_space(3))) uint8_t;
using lds_u32 = __attribute__((address_space(3))) uint32_t;
// Byte index of position q in byte lane g of a unit's LDS spin array: q * G +
// g with a full-rate 24-bit multiply-add (q < 2^24: a spin array in LDS has
// fewer than 160 Ki bytes); units of at most 4 replicas (NB < 8) have one
// byte lane, so the index is q itself.
template <int NB>
__device__ __forceinline__ uint32_t lds_at(uint32_t q, int G, int g) {
  if constexpr (NB < 8) return q;
  return __umul24(q, (uint32_t)G) + (uint32_t)g;
}

template <typename CT> __device__ __forceinline__ CT coef_of(uint32_t bits);
template <> __device__ __forceinline__ float coef_of<float>(uint32_t bits) { return __uint_as_float(bits); }
template <> __device__ __forceinline__ int32_t coef_of<int32_t>(uint32_t bits) { return (int32_t)bits; }

// The first (up to) 8 sliced-ELL entries of one position of a lane and the
// position's variable id, loaded ahead of use.  PAIR 1: 8-byte entries (low
// word: neighbour position, high word: coefficient bits); PAIR 2: 4-byte
// entries of integer problems whose coefficients fit int8 (neighbour
// position in the low 24 bits, the coefficient as int8 in the high 8).
template <int PAIR> struct EllRow {
  uint64_t ent[8];
  uint32_t var;
};
template <> struct EllRow<2> {
  uint32_t ent[8];
  uint32_t var;
};

// The draws of one lane's replicas at one position: the interpolated
// ln-table value of each replica's draw (row 0, or row 1 after a row-0
// bucket 0), and the replicas whose row-1 bucket is 0 as well (the rest of
// the row walk runs only for those, on an uphill move).
struct Draw {
  double val[8];
  uint32_t deep;
};

// A pointer every lane of the wave holds, moved to scalar registers.
template <class T> __device__ __forceinline__ T* uni_ptr(T* p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return reinterpret_cast<T*>(((uint64_t)hi << 32) | lo);
}

// The unit's graph tables, read once from its descriptor into scalar
// registers: re-reading the descriptor inside the round loop would be a
// vector load whose wait (vmcnt counts in order) drains the graph row
// prefetched for the next round.
struct Rows {
  cptr<int32_t> slice;   // {entry offset, width, first position, lane split} per slice
  gptr<uint64_t> ent;    // pairwise entries {position, coefficient bits}
  gptr<uint64_t> ent4;   // general 16-byte entries as two 64-bit words (or null)
  gptr<uint32_t> ent32;  // 4-byte pairwise entries {position | int8 coefficient << 24}
};
__device__ __forceinline__ Rows rows_of(const UnitDev& U) {
  Rows r;
  r.slice = cnst(uni_ptr(reinterpret_cast<const int32_t*>(U.ell_slice)));
  r.ent = glob(uni_ptr(reinterpret_cast<const uint64_t*>(U.ell_ent)));
  r.ent4 = glob(uni_ptr(reinterpret_cast<const uint64_t*>(U.ell4_ent)));
  r.ent32 = glob(uni_ptr(U.ell_ent32));
  return r;
}

__device__ __forceinline__ int4 load_slice(const Rows& R, int s) {
  return make_int4(R.slice[4 * s], R.slice[4 * s + 1], R.slice[4 * s + 2], R.slice[4 * s + 3]);
}

// slice = {offset, width} (uniform); lane tp of a slice with `lanes` = 1 << lg.
// Always E loads (the row width of the kernel: 4, 6 or 8), without
// branches, so the compiler can keep them in flight across the processing of
// the previous row.  Every slice stores at least 8 entries (zero
// coefficients on the ghost position past its width), so the first E are
// always the slice's own and need no mask.
template <int PAIR, int E>
__device__ __forceinline__ void ell_fetch(const Rows& R, int4 slice, int tp, int lg, EllRow<PAIR>& e) {
  const int step = 1 << lg;
  if constexpr (PAIR == 2) {
    gptr<uint32_t> p = R.ent32 + slice.x + tp;
#pragma unroll
    for (int j = 0; j < E; ++j) {
      e.ent[j] = *p;
      p += step;
    }
  } else {
    gptr<uint64_t> p = R.ent + slice.x + tp;
#pragma unroll
    for (int j = 0; j < E; ++j) {
      e.ent[j] = *p;
      p += step;
    }
  }
}

template <int KIND, typename CT, int NB>
__device__ __forceinline__ void accumulate(uint32_t wv, CT cc, CT (&h)[8]) {
  if constexpr (KIND == QMC_ISING && std::is_same<CT, float>::value) {
    // -c is c with its sign bit flipped: move replica b's bit into bit 31
    // and XOR it into the coefficient (the same value as `bit ? -c : c`)
    const uint32_t cb = __float_as_uint(cc);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const uint32_t sh = wv << (31 - b);
      // (sh & 0x80000000) ^ cb: v_bitop3_b32 truth table 0x6a
      const uint32_t t = b == 0 ? (sh ^ cb) : __builtin_amdgcn_bitop3_b32(sh, 0x80000000u, cb, 0x6a);
      h[b] += __uint_as_float(t);
    }
    return;
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const bool bit = (wv >> b) & 1u;
    if (KIND == QMC_ISING) h[b] += bit ? -cc : cc;
    else h[b] += bit ? CT(0) : cc;
  }
}

// Pairwise rows: the spin bytes of the first E entries' neighbours (LDS
// reads issued here, consumed by pair_field).
template <int PAIR, int NB, int E>
__device__ __forceinline__ void pair_gather(const lds_u8* sp, const EllRow<PAIR>& e, int G, int g,
                                            uint32_t (&wv)[8]) {
#pragma unroll
  for (int j = 0; j < E; ++j)
    wv[j] = sp[lds_at<NB>(PAIR == 2 ? (uint32_t)(e.ent[j] & 0xFFFFFFu) : (uint32_t)e.ent[j], G, g)];
}

// The local field of a pairwise row from its gathered neighbour bytes (the
// first E entries; entries from E to the width are loaded here).  Padding
// entries have zero coefficients: they add nothing (at most the sign of a
// zero field, which no decision depends on).
template <int KIND, int PAIR, typename CT, int NB, int E>
__device__ __forceinline__ void pair_field(const Rows& R, const lds_u8* sp, const EllRow<PAIR>& e,
                                           const uint32_t (&wv)[8], int4 slice, int tp, int lg, int G, int g,
                                           CT (&h)[8]) {
#pragma unroll
  for (int b = 0; b < NB; ++b) h[b] = CT(0);
  if constexpr (PAIR == 2) {
    // int8 entries: position in the low 24 bits, the coefficient in the high 8
    static_assert(std::is_same<CT, int32_t>::value, "int8 entries carry integer coefficients");
#pragma unroll
    for (int j = 0; j < E; ++j) accumulate<KIND, CT, NB>(wv[j], (int32_t)e.ent[j] >> 24, h);
    if (slice.y > E) {
      gptr<uint32_t> p = R.ent32 + slice.x + tp;
      for (int j = E; j < slice.y; ++j) {
        const uint32_t t = p[j << lg];
        accumulate<KIND, CT, NB>(sp[lds_at<NB>(t & 0xFFFFFFu, G, g)], (int32_t)t >> 24, h);
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < E; ++j) accumulate<KIND, CT, NB>(wv[j], coef_of<CT>((uint32_t)(e.ent[j] >> 32)), h);
    if (slice.y > E) {
      gptr<uint64_t> p = R.ent + slice.x + tp;
      for (int j = E; j < slice.y; ++j) {
        const uint64_t t = p[j << lg];
        accumulate<KIND, CT, NB>(sp[lds_at<NB>((uint32_t)t, G, g)], coef_of<CT>((uint32_t)(t >> 32)), h);
      }
    }
  }
}

// Local field of the replicas of byte lane g at position q:
//   Ising: h_b = sum_e c_e * (-1)^(parity of the other members, replica b)
//   PUBO : h_b = sum_e c_e * [all other members are 1 in replica b]
// Sums run over the variable's incident terms in term order, as the
// reference's calculate_cost_difference does.  Pairwise problems read the
// prefetched ELL row (entries beyond the first 8 are loaded here).
template <int KIND, int PAIR, typename CT, int NB, int E>
__device__ __forceinline__ void local_field(const UnitDev& U, const Rows& R, const lds_u8* sp, int q,
                                            const EllRow<PAIR>& e, int4 slice, int tp, int lg, int G, int g,
                                            CT (&h)[8], int sub = 0, int lgs = 0) {
  if constexpr (PAIR != 0) {
    uint32_t wv[8];
    pair_gather<PAIR, NB, E>(sp, e, G, g, wv);
    pair_field<KIND, PAIR, CT, NB, E>(R, sp, e, wv, slice, tp, lg, G, g, h);
    return;
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) h[b] = CT(0);
  if (R.ent4) {
    // sliced entries {coefficient, up to 3 other members} in term order;
    // the ghost position pads missing members (its spin is always 0).  With
    // 2^lgs lanes per position (tp = the position's lane), this lane takes
    // entries sub, sub + 2^lgs, ...
    gptr<uint64_t> p = R.ent4 + 2 * (slice.x + tp);
    const int step = 2 << lg;
    auto entry = [&](int j) {
      const uint64_t lo = p[j * step], hi = p[j * step + 1];
      const uint32_t w0 = sp[lds_at<NB>((uint32_t)(lo >> 32), G, g)];
      const uint32_t w1 = sp[lds_at<NB>((uint32_t)hi, G, g)];
      const uint32_t w2 = sp[lds_at<NB>((uint32_t)(hi >> 32), G, g)];
      const uint32_t w = (KIND == QMC_ISING) ? (w0 ^ w1 ^ w2) : (w0 | w1 | w2);
      accumulate<KIND, CT, NB>(w, coef_of<CT>((uint32_t)lo), h);
    };
    if (lgs == 0) {
#pragma unroll 8  // eight entries' loads in flight (deeper unrolling runs out of registers)
      for (int j = 0; j < slice.y; ++j) entry(j);
    } else {
#pragma unroll 4
      for (int j = sub; j < slice.y; j += 1 << lgs) entry(j);
    }
    return;
  }
  gptr<CT> coef = glob(static_cast<const CT*>(U.coef));
  gptr<int32_t> inc = glob(U.inc_off);
  gptr<int32_t> ooff = glob(U.oth_off);
  gptr<int32_t> opos = glob(U.oth_pos);
  const int e0 = inc[q], e1 = inc[q + 1];
  for (int e = e0; e < e1; ++e) {
    const CT cc = coef[e];
    uint32_t w = 0;
    const int o0 = ooff[e], o1 = ooff[e + 1];
    for (int o = o0; o < o1; ++o) {
      const uint32_t x = sp[lds_at<NB>(opos[o], G, g)];
      w = (KIND == QMC_ISING) ? (w ^ x) : (w | x);
    }
    accumulate<KIND, CT, NB>(w, cc, h);
  }
}

template <int KIND, typename CT>
__device__ __forceinline__ CT delta_from_field(CT h, uint32_t own_bit) {
  if (KIND == QMC_ISING) return own_bit ? CT(2) * h : CT(-2) * h;
  return own_bit ? h : -h;
}

// Energy hand-off between the work-groups of one PT ladder (cooperative
// launches only, so every member is resident).  Step s publishes each slot's
// energy into buffer s % 3 of the ladder; a slot holds NaN until its step's
// value arrives, so readers poll the values themselves and no separate
// arrival counter is needed.  After reading step s, a work-group resets its
// own slots of buffer (s + 2) % 3 -- every work-group has finished reading
// that buffer (step s - 1) by then, since it has published step s -- and the
// reset is complete before the slot is published again one step later.
// Energies travel as agent-scope atomics (coherent across the XCDs), and a
// writer waits for its stores to complete before anything that depends on
// them; agent-scope fences would also write back and invalidate the whole L2
// of the XCD, which every later read of the coupling graph would pay for.
// Bounded: after ~2 s of polling the error flag is raised and all members
// abandon the run instead of hanging.
__device__ bool await_energy(const UnitDev& U, const double* slot, double& out) {
  using g64 = __attribute__((address_space(1))) unsigned long long;
  using g32 = __attribute__((address_space(1))) int32_t;
  g64* sl = (g64*)slot;
  g32* err = (g32*)U.err;
  const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
  while (true) {
    const double v = __longlong_as_double((long long)__hip_atomic_load(sl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (v == v) {
      out = v;
      return true;
    }
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    if (__builtin_amdgcn_s_memrealtime() - t0 > 200000000ull) {  // 100 MHz clock: 2 s
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// Work-group barrier that orders LDS traffic only: vector memory loads in
// flight (the next graph rows) stay in flight across it, where
// __syncthreads() would wait for them.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Orders LDS accesses among the lanes of one wavefront (its LDS
// instructions execute in order): no work-group barrier.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// s_waitcnt vmcnt(0): every vector memory access of this lane issued so far,
// stores included, has completed.
__device__ __forceinline__ void wait_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }

__device__ __forceinline__ void store_agent(double* p, double v) {
  using g64 = __attribute__((address_space(1))) unsigned long long;
  __hip_atomic_store((g64*)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_agent(const double* p) {
  using g64 = __attribute__((address_space(1))) unsigned long long;
  return __longlong_as_double((long long)__hip_atomic_load((g64*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// LADDER: the launch may hold parallel tempering ladders (exchanges in the
// kernel); without it the exchange code, and the registers it needs, are
// compiled out.  E: entries of a pairwise row fetched ahead (4, 6 or 8; the
// smallest that covers every slice's width, else 8 and the rest per row).
template <int KIND, int PAIR, typename CT, int NB, bool LADDER, int E>
__global__ void __launch_bounds__(kBlock)
k_sweep(const UnitDev* units, SweepArgs a) {
  using AT = typename AccT<CT>::T;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem& S = *reinterpret_cast<Smem*>(smem_raw);
  // The descriptor stays in memory (scalar loads at each use): fields needed
  // only at the start and end are not held in registers across the sweeps.
  const UnitDev& U = units[uni(a.unit_of_block ? a.unit_of_block[blockIdx.x] : (int)blockIdx.x)];
  // at most 4 replicas per unit (NB < 8) means a single byte group
  const int N = U.N, G = NB < 8 ? 1 : U.G, TP = U.lanes, lg = U.lanes_log2;
  const int stride = U.stride;
  cptr<int32_t> class_off = cnst(U.class_off);
  cptr<int32_t> slice_base = cnst(U.ell_slice_base);
  // general problems read sliced entries when their terms are at most 4-local
  const Rows rows = rows_of(U);
  const bool gen_ell = !PAIR && rows.ent4 != nullptr;
  const bool slices = PAIR || gen_ell;
  gptr<int32_t> var_of_pos = glob(uni_ptr(U.var_of_pos));
  const int tid = threadIdx.x;
  const int nthreads = blockDim.x;
  const int GS = slot_groups(NB, nthreads);
  const Slots<AT> W_(smem_raw + sizeof(Smem), GS);
  unsigned char* const state_lds = smem_raw + sizeof(Smem) + slot_bytes(GS);
  lds_u8* sp = (lds_u8*)state_lds;  // generic -> LDS address space
  lds_u8* fl = sp + stride;
  // lowest states: in LDS for the launch when they fit, else in HBM
  uint8_t* bst = a.best_lds ? reinterpret_cast<uint8_t*>(state_lds + 2 * stride) : U.best;
  const bool active = tid < TP * G;
  const int g = (NB < 8 || !active) ? 0 : tid >> lg;
  const int tp = tid - (g << lg);
  const int lane = tid & 63;
  const int wave_in_group = tp >> 6;
  const int nrep = unit_nrep(U);
  if (nrep == 0) return;  // idle unit of a device-sized population (uniform)
  const int nb_g = active ? min(NB, nrep - 8 * g) : 0;
  const uint32_t rbase = U.r0 + 8u * (uint32_t)g;
  const uint32_t key0 = in_vgpr(U.seed), key1 = in_vgpr(U.stream);

  {
    gptr<double> lut_g = glob(a.lut);
    for (int i = tid; i < 1024; i += nthreads) S.lut[i] = lut_g[i];
  }
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(U.spins);
    lds_u32* dst = reinterpret_cast<lds_u32*>(sp);
    for (int i = tid; i < stride / 4; i += nthreads) dst[i] = src[i];
    if (a.best_lds) {
      const uint32_t* bsrc = reinterpret_cast<const uint32_t*>(U.best);
      uint32_t* bdst = reinterpret_cast<uint32_t*>(bst);
      for (int i = tid; i < stride / 4; i += nthreads) bdst[i] = bsrc[i];
    }
  }
  if (tid < nrep) {
    S.rcost[tid] = U.cost[tid];
    S.rbest[tid] = U.best_cost[tid];
    S.rbeta[tid] = U.beta[tid];
    S.racc[tid] = 0;
    S.nb_c[tid] = -1;
  }
  if (tid == 0) {
    S.abort_flag = 0;
    S.any_new_best = 0;
  }
  if (LADDER && U.ladder > 0 && tid < U.ladder) {
    S.label[tid] = U.label[tid];
    S.lbeta[tid] = U.ladder_beta[tid];
  }
  __syncthreads();
  if (LADDER && U.ladder > 0 && tid < U.ladder) S.inv[S.label[tid]] = tid;
  __syncthreads();

  auto lut = [&](int row, int i) { return S.lut[row * 256 + i]; };

  uint32_t acc[8];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = 0;

  const int64_t sweep_begin = U.progress[0];
  const int n_run = (int)min((int64_t)a.n_sweeps, U.progress[1]);
  // Every global access inside the sweep loop goes through a global (not
  // generic) pointer: one flat access in the loop makes the compiler's wait
  // counts conservative (s_waitcnt vmcnt(0)), which waits for the graph row
  // just prefetched for the next round.
  gptr<double> sched = glob(U.sched);
  const int64_t sched_base = U.sched_base;
  const int n_classes = U.n_classes;
  // beta of the next sweep, loaded one sweep ahead
  double beta_next = (sched && tid < nrep && n_run > 0) ? sched[sweep_begin - sched_base] : 0.0;

  // The graph rows stream through two buffers across class and sweep
  // boundaries: round j of class c lives in `ra` when the class starts, and
  // rounds past a class's end are the next class's first rounds (the last
  // class is followed by class 0 of the next sweep), so the first row of a
  // class is already in flight while the previous class finishes.
  EllRow<PAIR> ra, rb;
  int4 sla = make_int4(0, 0, 0, 0), slb = make_int4(0, 0, 0, 0);
  // Small classes of general problems with integer couplings share each
  // position among 2^lgs lanes (at most 8, while positions x lanes per
  // position fit the lanes): each sums every 2^lgs-th incident term and a
  // butterfly adds the partial fields (integer sums: the order does not
  // matter), so a round waits for a fraction of the position's loads.
  constexpr bool kSplit = !PAIR && std::is_same<CT, int32_t>::value;
  auto split_of = [&](int ncls) {
    int l = 0;
    if (kSplit && gen_ell)
      while (l < 3 && (ncls << (l + 1)) <= TP) ++l;
    return l;
  };
  // Slice tables run through the classes in order and end with copies of the
  // first two slices, so round k past a class's end is slice
  // slice_base[c] + k whichever class follows; each slice carries its first
  // position and lane split, so the variable of this lane's position needs
  // no class lookup.
  auto fetch_slice = [&](int sidx, int4& sl, EllRow<PAIR>& e) {
    sl = load_slice(rows, sidx);
    if (PAIR) ell_fetch<PAIR, E>(rows, sl, tp, lg, e);
    e.var = (uint32_t)var_of_pos[sl.z + (tp >> sl.w)];  // padded past position N - 1
  };
  // problems without slices (terms of more than 4 variables): position of
  // round k of the class starting at cc0
  auto fetch_csr = [&](int cc0, int k, EllRow<PAIR>& e) { e.var = (uint32_t)var_of_pos[cc0 + (k << lg) + tp]; };
  if (active) {
    if (slices) fetch_slice(slice_base[0], sla, ra);
    else fetch_csr(class_off[0], 0, ra);
  }

  // The draws of the replicas of this lane at the position of `var` in
  // sweep `sw`, and their interpolated ln-table values: independent of the
  // field and of beta, so they are formed while the neighbour bytes are
  // read (or during a ladder exchange, for the next sweep's first class).
  auto draw = [&](uint32_t var, uint32_t sw, Draw& d) {
    uint4 quad = make_uint4(0, 0, 0, 0);
    d.deep = 0;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const uint32_t r = rbase + (uint32_t)b;
      // one draw serves a replica pair: words 0, 1 = their row-0
      // draws, words 2, 3 = their row-1 draws (used after bucket 0)
      if (b == 0 || (r & 1u) == 0u) quad = philox4x32_10(var, sw, r >> 1, DOM_SWEEP, key0, key1);
      const uint32_t w0 = (r & 1u) ? quad.y : quad.x;
      const uint32_t w1 = (r & 1u) ? quad.w : quad.z;
      const bool z0 = (w0 & 255u) == 0u;
      d.val[b] = row_value(z0 ? w1 : w0, z0 ? 1 : 0, lut);
      d.deep |= (z0 && (w1 & 255u) == 0u) ? 1u << b : 0u;
    }
  };

  // Split ladders: while the ladder's energies travel between work-groups
  // after a sweep, the local fields of the next sweep's colour class 0 are
  // computed into LDS (they depend only on the spins of the other classes,
  // final at the end of a sweep, and not on beta), and with room for them
  // the draws of that class too; the class then only decides.  Same fields,
  // same draws, same decisions: the results do not change.
  constexpr bool kEarly = PAIR && LADDER;
  using lds_ct = __attribute__((address_space(3))) CT;
  using lds_f64 = __attribute__((address_space(3))) double;
  // [round][replica][thread] fields, [round][thread] variables, then (when
  // a.pre_draws) [round][replica][thread] draw values, [round][thread] masks
  lds_ct* hpre = (lds_ct*)(smem_raw + a.pre_off);
  lds_u32* vpre = (lds_u32*)(smem_raw + a.pre_off + (size_t)a.pre_rounds * NB * nthreads * 4);
  lds_f64* dpre = (lds_f64*)(smem_raw + a.pre_off + (size_t)a.pre_rounds * (NB + 1) * nthreads * 4);
  lds_u32* mpre = (lds_u32*)(smem_raw + a.pre_off + (size_t)a.pre_rounds * ((NB + 1) * 4 + NB * 8) * nthreads);
  // (8 replicas per byte lane: fields only, the draws' registers would spill)
  const bool pre_draws = NB <= 4 && a.pre_draws != 0;
  const int nc0 = class_off[1] - class_off[0];
  const int K0 = (nc0 + TP - 1) >> lg;
  const bool early_ok = kEarly && U.ladder > 0 && U.upl > 1 && n_classes >= 2 && K0 <= a.pre_rounds;
  bool early_ready = false;
  // round k of class 0, or round k - K0 of class 1 (pairwise problems only)
  auto fetch_in = [&](int k, int4& sl, EllRow<PAIR>& e) { fetch_slice(slice_base[0] + k, sl, e); };
  auto field_early = [&](int k, const int4& sl, const EllRow<PAIR>& e, uint32_t sw) {
    const int rho = (k << lg) + tp;
    const int q = rho < nc0 ? class_off[0] + rho : N;
    CT h[8];
    if constexpr (kEarly && NB <= 4) {
      if (pre_draws) {
        uint32_t wv[8];
        pair_gather<PAIR, NB, E>(sp, e, G, g, wv);
        Draw d;
        draw(e.var, sw, d);
        pair_field<KIND, PAIR, CT, NB, E>(rows, sp, e, wv, sl, tp, lg, G, g, h);
#pragma unroll
        for (int b = 0; b < NB; ++b) dpre[(k * NB + b) * nthreads + tid] = d.val[b];
        mpre[k * nthreads + tid] = d.deep;
      } else {
        local_field<KIND, PAIR, CT, NB, E>(U, rows, sp, q, e, sl, tp, lg, G, g, h);
      }
    } else {
      local_field<KIND, PAIR, CT, NB, E>(U, rows, sp, q, e, sl, tp, lg, G, g, h);
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) hpre[(k * NB + b) * nthreads + tid] = h[b];
    vpre[k * nthreads + tid] = e.var;
  };

  // ladder fields of the descriptor, read once (uniform) before the sweeps
  const int x_ladder = LADDER ? uni(U.ladder) : 0;
  const int x_spr = LADDER ? uni(U.spr) : 1;
  const int x_upl = LADDER ? uni(U.upl) : 1;
  const int x_slot0 = LADDER ? uni(U.slot0) : 0;
  double* const x_cbuf = LADDER ? uni_ptr(U.cbuf) : nullptr;
  const uint32_t x_seed = uni(U.seed), x_stream = uni(U.stream);
  // parallel tempering: exchange step of the current sweep, sweeps run into
  // it and its energy buffer (step mod 3), advanced per sweep
  int64_t xstep = 0;
  int xphase = 0, xbuf = 0;
  if (LADDER && x_ladder > 0) {
    xstep = sweep_begin / x_spr;
    xphase = (int)(sweep_begin - xstep * x_spr);
    xbuf = (int)(xstep % 3);
  }

  for (int s = 0; s < n_run; ++s) {
    const int64_t sweep64 = sweep_begin + s;
    const uint32_t sweep = (uint32_t)sweep64;
    // Split ladders: the energy-slot resets of the last exchange complete
    // here, before the next publish; the only other access in flight is the
    // row fetched before that exchange, long arrived, so the wait is short.
    if (LADDER && x_upl > 1 && tid < nrep) wait_stores();
    if (sched) {
      if (tid < nrep) {
        S.rbeta[tid] = beta_next;
        if (s + 1 < n_run) beta_next = sched[sweep64 + 1 - sched_base];
      }
      lds_barrier();
    }
    // (otherwise the betas were last written before the barrier that ends
    // the previous sweep, or before the launch's set-up barrier)
    double beta[8];
#pragma unroll
    for (int b = 0; b < NB; ++b) beta[b] = (b < nb_g) ? S.rbeta[8 * g + b] : 0.0;

    for (int c = 0; c < n_classes; ++c) {
      const int c0 = class_off[c];
      const int nc = class_off[c + 1] - c0;
      const int K = (nc + TP - 1) >> lg;
      const int par = c & 1;
      const int sbase = slices ? slice_base[c] : 0;
      const int cn = c + 1 < n_classes ? c + 1 : 0;  // the class after this one
      const int c0n = class_off[cn];
      const int lgs = split_of(nc);
      // per replica: sum of the accepted dE of this lane's positions, the
      // minimum prefix of that sum and the round where it is first reached
      AT Ssum[8], Smin[8];
      int kmin[8];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        Ssum[b] = AT(0);
        Smin[b] = AT(0);
        kmin[b] = -1;
      }
      if (active) {
        // Position k of this lane is rho = k * lanes + tp.  The rounds are
        // uniform; the graph row of round k + 1 is fetched into the other
        // buffer while round k is processed.
        // round k of this class, or round k - K of the next (no branch: the
        // loads stay in flight across the processing of the other buffer)
        auto fetch = [&](int k, int4& sl, EllRow<PAIR>& e) {
          if (slices) {
            fetch_slice(sbase + k, sl, e);
          } else {
            const bool here = k < K;
            fetch_csr(here ? c0 : c0n, here ? k : k - K, e);
          }
        };
        // the decision for the replicas of this lane at position q (round
        // k) from their local fields h and draws d
        auto decide = [&](int k, int q, bool valid, const CT (&h)[8], const Draw& d, uint32_t var) {
          const uint32_t own = sp[lds_at<NB>(q, G, g)];
          CT dE[8];
          uint32_t flips = 0, deep = 0;
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            dE[b] = delta_from_field<KIND, CT>(h[b], (own >> b) & 1u);
            const bool live = valid && b < nb_g;
            const bool ok = accept_value((double)dE[b], beta[b], d.val[b]);
            const bool dp = ((d.deep >> b) & 1u) && dE[b] > CT(0);
            flips |= (live && ok && !dp) ? 1u << b : 0u;
            deep |= (live && dp) ? 1u << b : 0u;
          }
          // bucket 0 in rows 0 and 1 of an uphill draw (2^-16): the rest of
          // the reference's row walk, one replica at a time, with that
          // replica's redraw counter for rows 2 and 3
          while (deep) {
            const int b = __builtin_ctz(deep);
            deep &= deep - 1;
            const uint32_t r = rbase + (uint32_t)b;
            CT d = dE[0];
            double bb = beta[0];
#pragma unroll
            for (int j = 1; j < NB; ++j) {
              d = j == b ? dE[j] : d;
              bb = j == b ? beta[j] : bb;
            }
            const uint4 w = philox4x32_10(var, sweep, r, DOM_REDRAW, key0, key1);
            if (metropolis_accept((double)d, bb, 0u, lut, [&]() { return make_uint4(0u, w.x, w.y, 0u); }))
              flips |= 1u << b;
          }
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            const bool ok = (flips >> b) & 1u;
            Ssum[b] += ok ? (AT)dE[b] : AT(0);
            const bool lower = Ssum[b] < Smin[b];  // first time at the new minimum
            Smin[b] = lower ? Ssum[b] : Smin[b];
            kmin[b] = lower ? k : kmin[b];
            acc[b] += ok ? 1u : 0u;
          }
          sp[lds_at<NB>(q, G, g)] = (uint8_t)(own ^ flips);
          fl[lds_at<NB>(q, G, g)] = (uint8_t)flips;
        };
        auto process = [&](int k, const int4& sl, const EllRow<PAIR>& e) {
          const int sub = tp & ((1 << lgs) - 1);
          const int rho = (k << lg) + (tp >> lgs);
          // The term lists of general problems end at N - 1; pairwise rows
          // past the class end are ghost padding, so those lanes run the
          // same branch-free path on the ghost position (always 0, its flips
          // forced to 0) and the row's loads are consumed on every path.
          if (!PAIR && rho >= nc) return;
          const bool valid = rho < nc;
          const int q = valid ? c0 + rho : N;
          CT h[8];
          Draw d;
          if constexpr (PAIR != 0) {
            uint32_t wv[8];
            pair_gather<PAIR, NB, E>(sp, e, G, g, wv);
            draw(e.var, sweep, d);
            pair_field<KIND, PAIR, CT, NB, E>(rows, sp, e, wv, sl, tp, lg, G, g, h);
          } else {
            local_field<KIND, PAIR, CT, NB, E>(U, rows, sp, q, e, sl, tp >> lgs, lg, G, g, h, sub, lgs);
            if (kSplit && lgs) {
              // every lane of a position is active here (they share rho)
#pragma unroll
              for (int b = 0; b < NB; ++b)
                for (int o = 1; o < (1 << lgs); o <<= 1) h[b] += __shfl_xor(h[b], o);
              if (sub) return;
            }
            draw(e.var, sweep, d);
          }
          decide(k, q, valid, h, d, e.var);
        };
        if (kEarly && c == 0 && early_ready) {
          // fields (and draws) computed during the last exchange; the row
          // pipeline already holds class 1's first row
          for (int k = 0; k < K; ++k) {
            CT h[8];
#pragma unroll
            for (int b = 0; b < NB; ++b) h[b] = hpre[(k * NB + b) * nthreads + tid];
            const int rho = (k << lg) + tp;
            const uint32_t var = vpre[k * nthreads + tid];
            Draw d;
            if (pre_draws) {
#pragma unroll
              for (int b = 0; b < NB; ++b) d.val[b] = dpre[(k * NB + b) * nthreads + tid];
              d.deep = mpre[k * nthreads + tid];
            } else {
              draw(var, sweep, d);
            }
            decide(k, rho < nc ? c0 + rho : N, rho < nc, h, d, var);
          }
        } else {
          // Fetches past the last round read sentinel slices and padding
          // (the tables are padded for it), so no fetch sits behind a branch
          // and the loads stay in flight while the other buffer is processed.
          // The loop has one exit: an odd K's last pair fetches the next
          // class's first row (round K) into `ra` and skips the second half.
          for (int k = 0; k < K; k += 2) {
            fetch(k + 1, slb, rb);
            process(k, sla, ra);
            // the last pair of an odd K: `ra` takes the next class's first
            // row (round K) instead of round K + 1
            fetch(min(k + 2, K), sla, ra);
            if (k + 1 < K) process(k + 1, slb, rb);
          }
          if (K == 0) fetch(0, sla, ra);
        }

        // ---- per replica: minimum of the cost along the class sequence ----
        // The sequence is lane-major, so lane t's prefixes follow the sums of
        // all earlier lanes: per wave, an exclusive scan of the sums plus each
        // lane's minimum prefix, minimised (first lane on ties); the waves
        // are combined in order below.
        // (All NB replicas of the lane are scanned together, absent ones
        // with zero sums, so the chains interleave; only present ones store.)
        AT incl[8], mv[8], mn[8];
#pragma unroll
        for (int b = 0; b < NB; ++b) incl[b] = Ssum[b];
        scan_add_n<AT, NB>(incl);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const AT excl = dpp_z<DPP_WAVE_SHR1>(incl[b]);
          mv[b] = excl + Smin[b];
          mn[b] = mv[b];
        }
        scan_min_n<AT, NB>(mn);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const AT m = lane_value(mn[b], 63);
          const unsigned long long hit = __builtin_amdgcn_ballot_w64(mv[b] == m);
          const int wl = (int)__builtin_ctzll(hit);
          const int wk = __builtin_amdgcn_readlane(kmin[b], wl);
          const AT tot = lane_value(incl[b], 63);
          if (lane == 0 && b < nb_g) {
            const int at = W_.at(par, g, b, wave_in_group);
            W_.sum[at] = tot;
            W_.mn[at] = m;
            W_.lane[at] = wl;
            W_.k[at] = wk;
          }
        }
      }
      lds_barrier();
      if constexpr (std::is_same<AT, int32_t>::value) {
        // Integer sums (exact in any order): wave 0 combines every
        // replica's wave partials at once, lane = replica x W + wave --
        // segmented prefix sums of the totals, then a segmented prefix
        // minimum of (earlier totals + the wave's minimum prefix), ties to
        // the earlier wave: the wave-order fold, in log2(W) steps.
        if (tid < 64) {
          const int W = TP >> 6;  // waves per byte group: 1, 2, 4 or 8
          const int lw = W >= 8 ? 3 : (W >= 4 ? 2 : (W >= 2 ? 1 : 0));
          const int rr = lane >> lw, w = lane & (W - 1);
          const bool live = rr < nrep;
          const int gg = rr >> 3, bb = rr & 7;
          int32_t sm = 0, m = 0, wl = 0, wk = -1;
          if (live) {
            const int at = W_.at(par, gg, bb, w);
            sm = W_.sum[at];
            m = W_.mn[at];
            wl = W_.lane[at];
            wk = W_.k[at];
          }
          int32_t incl = sm;
          const int32_t o1 = dpp_z<DPP_ROW_SHR1>(incl);
          incl = (W > 1 && w >= 1) ? incl + o1 : incl;
          const int32_t o2 = dpp_z<DPP_ROW_SHR2>(incl);
          incl = (W > 2 && w >= 2) ? incl + o2 : incl;
          const int32_t o4 = dpp_z<DPP_ROW_SHR4>(incl);
          incl = (W > 4 && w >= 4) ? incl + o4 : incl;
          int32_t cnd = incl - sm + m;
          int32_t wi = w;
          auto min_step_seg = [&](auto shift_of, int sh) {
            const int32_t oc = shift_of(cnd), owi = shift_of(wi), owl = shift_of(wl), owk = shift_of(wk);
            // the earlier part wins unless this lane's candidate is lower
            const bool take = sh < W && w >= sh && !(cnd < oc);
            cnd = take ? oc : cnd;
            wi = take ? owi : wi;
            wl = take ? owl : wl;
            wk = take ? owk : wk;
          };
          min_step_seg([](int32_t x) { return dpp_z<DPP_ROW_SHR1>(x); }, 1);
          min_step_seg([](int32_t x) { return dpp_z<DPP_ROW_SHR2>(x); }, 2);
          min_step_seg([](int32_t x) { return dpp_z<DPP_ROW_SHR4>(x); }, 4);
          if (live && w == W - 1) {
            const double before = S.rcost[rr], best_before = S.rbest[rr];
            S.rcost[rr] = before + (double)incl;
            const double cand = before + (double)cnd;
            if (cand < best_before) {
              // the state there is rebuilt once, at the end of the sweep
              S.rbest[rr] = cand;
              S.nb_c[rr] = c;
              // the lane's position (lanes share positions in split classes)
              S.nb_tp[rr] = (wi * 64 + wl) >> lgs;
              S.kstar[rr] = wk;
              S.any_new_best = 1;
            }
          }
        }
      } else {
        if (tid < nrep) {
          const int gg = tid >> 3, bb = tid & 7;
          const int W = TP >> 6;
          // every wave's slots read at once (one LDS round trip), then
          // combined in wave order; slots of absent waves are read, not used
          AT mnv[kMaxWavesPerGroup], smv[kMaxWavesPerGroup];
          int wlv[kMaxWavesPerGroup], wkv[kMaxWavesPerGroup];
          const int at0 = W_.at(par, gg, bb, 0);
#pragma unroll
          for (int w = 0; w < kMaxWavesPerGroup; ++w) {
            mnv[w] = W_.mn[at0 + w];
            smv[w] = W_.sum[at0 + w];
            wlv[w] = W_.lane[at0 + w];
            wkv[w] = W_.k[at0 + w];
          }
          const double before = S.rcost[tid], best_before = S.rbest[tid];
          AT run = AT(0), bestc = AT(0);
          int bt = -1, bk = -1;
#pragma unroll
          for (int w = 0; w < kMaxWavesPerGroup; ++w) {
            if (w < W) {
              const AT cnd = run + mnv[w];
              if (bt < 0 || cnd < bestc) {
                bestc = cnd;
                // the lane's position (lanes share positions in split classes)
                bt = (w * 64 + wlv[w]) >> lgs;
                bk = wkv[w];
              }
              run += smv[w];
            }
          }
          S.rcost[tid] = before + (double)run;
          const double cand = before + (double)bestc;
          if (cand < best_before) {
            // the state there is rebuilt once, at the end of the sweep
            S.rbest[tid] = cand;
            S.nb_c[tid] = c;
            S.nb_tp[tid] = bt;
            S.kstar[tid] = bk;
            S.any_new_best = 1;
          }
        }
      }
      // no barrier: the next class writes the other parity's slots, and the
      // one after it passes the next class's barrier first
      if (kEarly && c == 0) early_ready = false;
    }  // classes
    lds_barrier();

    // ---- lowest states reached during this sweep ----
    // Every position is visited once per sweep and fl holds its flips of
    // this sweep, so the state at a replica's new minimum (class c*, lane
    // t*, round k*) has the end-of-sweep bits of the positions visited up to
    // that point and the start-of-sweep bits (end ^ flips) of the rest.
    if (S.any_new_best) {
      if (active) {
        uint32_t newmask = 0;
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if (b < nb_g && S.nb_c[8 * g + b] >= 0) newmask |= 1u << b;
        if (newmask) {
          // class by class (uniform bounds), each lane over its positions
          auto rebuild = [&](auto* out) {
            int cs[8], ts[8], ks[8];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
              cs[b] = S.nb_c[8 * g + b];
              ts[b] = S.nb_tp[8 * g + b];
              ks[b] = S.kstar[8 * g + b];
            }
            for (int c = 0; c < n_classes; ++c) {
              const int cb0 = class_off[c], cb1 = class_off[c + 1];
              for (int q = cb0 + tp; q < cb1; q += TP) {
                const int rho = q - cb0;
                const int tq = rho & (TP - 1), kq = rho >> lg;
                const uint32_t at = lds_at<NB>((uint32_t)q, G, g);
                const uint32_t cur = sp[at];
                const uint32_t f = fl[at];
                uint32_t val = cur;
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                  if ((newmask >> b) & 1u) {
                    const bool after = c > cs[b] || (c == cs[b] && (tq > ts[b] || (tq == ts[b] && kq > ks[b])));
                    if (after) val ^= f & (1u << b);
                  }
                }
                const uint32_t old = out[at];
                out[at] = (uint8_t)((old & ~newmask) | (val & newmask));
              }
            }
          };
          // explicit address spaces: LDS or global stores, never flat ones
          if (a.best_lds) {
            rebuild((lds_u8*)bst);
          } else {
            rebuild((__attribute__((address_space(1))) uint8_t*)bst);
          }
        }
      }
      lds_barrier();
      if (tid < nrep) S.nb_c[tid] = -1;
      if (tid == 0) S.any_new_best = 0;
      lds_barrier();
    }

    // ---- parallel tempering exchange (parallel_tempering.h:126-189) ----
    // Every work-group of the ladder computes the same decisions from the
    // same energies, so labels stay identical without another barrier.
    if (LADDER && x_ladder > 0 && xphase == x_spr - 1) {
      const uint32_t step = (uint32_t)xstep;
      const int L = x_ladder;
      // this thread's pair and its uniform (drawn after the energy is
      // published, while the ladder's energies travel)
      const int i = (int)(step & 1u) + 2 * tid;
      double u = 0.0;
      auto exchange_uniform = [&]() {
        if (i < L - 1) {
          const uint4 w = philox4x32_10((uint32_t)i, step, 0u, DOM_EXCHANGE, x_seed, x_stream);
          u = canonical2(w.x, w.y);
        }
      };
      if (x_upl > 1) {
        double* cb = x_cbuf + xbuf * L;
        if (tid < nrep) {
          // this slot's reset (one step ago) completed at the start of the
          // sweep
          store_agent(cb + x_slot0 + tid, S.rcost[tid]);
        }
        // while the energies travel: the fields of class 0 of the next sweep
        // (its row pipeline moves on to class 1's first row)
        if (early_ok && s + 1 < n_run) {
          if (active) {
            const uint32_t next = (uint32_t)(sweep64 + 1);
            for (int k = 0; k < K0; k += 2) {
              fetch_in(k + 1, slb, rb);
              field_early(k, sla, ra, next);
              fetch_in(min(k + 2, K0), sla, ra);
              if (k + 1 < K0) field_early(k + 1, slb, rb, next);
            }
            if (K0 == 0) fetch_in(0, sla, ra);
          }
          early_ready = true;
        }
        exchange_uniform();
        if (tid < L) {
          double e = 0.0;
          if (!await_energy(U, cb + tid, e)) S.abort_flag = 1;
          S.lcost[tid] = e;
        }
        if (tid < nrep) {
          const double nan = __longlong_as_double(0x7ff8000000000000ll);
          store_agent(x_cbuf + (xbuf == 0 ? 2 : xbuf - 1) * L + x_slot0 + tid, nan);
        }
      } else {
        exchange_uniform();
        if (tid < L) S.lcost[tid] = S.rcost[tid];
      }
      // Energies, decisions and new betas of a ladder (at most 32 replicas)
      // all live in wavefront 0: ordered within it, one work-group barrier
      // at the end of the sweep publishes the betas (and an abort).
      wave_lds_sync();
      if (i < L - 1) {
        const int sa = S.label[i], sb = S.label[i + 1];
        const double bi = S.lbeta[i], bj = S.lbeta[i + 1];
        const double arg = __dmul_rn(__dsub_rn(S.lcost[sa], S.lcost[sb]), __dsub_rn(bi, bj));
        if (u < exp(arg)) {
          S.label[i] = sb;
          S.label[i + 1] = sa;
          S.inv[sa] = i + 1;
          S.inv[sb] = i;
        }
      }
      wave_lds_sync();
      if (tid < nrep) S.rbeta[tid] = S.lbeta[S.inv[x_slot0 + tid]];
    }
    if (LADDER && x_ladder > 0 && ++xphase == x_spr) {
      xphase = 0;
      ++xstep;
      xbuf = xbuf == 2 ? 0 : xbuf + 1;
    }
    lds_barrier();
    if (LADDER && x_upl > 1 && S.abort_flag) break;
  }  // sweeps

  // ---- write back ----
  if (active) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
      if (b < nb_g && acc[b]) atomicAdd(&S.racc[8 * g + b], acc[b]);
  }
  {
    const lds_u32* src = reinterpret_cast<const lds_u32*>(sp);
    uint32_t* dst = reinterpret_cast<uint32_t*>(U.spins);
    for (int i = tid; i < stride / 4; i += nthreads) dst[i] = src[i];
    if (a.best_lds) {
      const uint32_t* bsrc = reinterpret_cast<const uint32_t*>(bst);
      uint32_t* bdst = reinterpret_cast<uint32_t*>(U.best);
      for (int i = tid; i < stride / 4; i += nthreads) bdst[i] = bsrc[i];
    }
  }
  __syncthreads();
  if (tid < nrep) {
    U.cost[tid] = S.rcost[tid];
    U.best_cost[tid] = S.rbest[tid];
    U.beta[tid] = S.rbeta[tid];
    U.accepted[tid] += S.racc[tid];
  }
  if (LADDER && U.ladder > 0 && U.slot0 == 0 && tid < U.ladder) U.label[tid] = S.label[tid];
  if (tid == 0) {
    U.progress[0] = sweep_begin + n_run;
    U.progress[1] -= n_run;
  }
}

// ---------------------------------------------------------------------------
// Reference-order sweep: one lane per replica, variables i = 0..N-1 in index
// order (markov/walker.h:144-153).  Graph reads are wave-uniform.  The
// lowest state uses the ordered-sweep merge: the position of the sweep's
// running minimum selects end-of-sweep bits up to it and start-of-sweep bits
// after it (each variable is touched once per sweep).
template <int KIND, bool PAIR, typename CT>
__global__ void __launch_bounds__(64)
k_sweep_ordered(const UnitDev* __restrict__ units, SweepArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* lut = reinterpret_cast<double*>(smem_raw);
  uint8_t* sp = smem_raw + 8192;            // unit layout, stride bytes
  const UnitDev U = units[blockIdx.x];
  const int G = U.G, N = U.N, stride = U.stride;
  uint8_t* start = sp + stride;             // start-of-sweep copy
  const int lane = threadIdx.x;
  const int nrep = unit_nrep(U);
  if (nrep == 0) return;  // idle unit of a device-sized population (uniform)
  for (int i = lane; i < 1024; i += 64) lut[i] = a.lut[i];
  for (int i = lane; i < stride / 4; i += 64)
    reinterpret_cast<uint32_t*>(sp)[i] = reinterpret_cast<const uint32_t*>(U.spins)[i];
  __syncthreads();
  const bool act = lane < nrep;
  const int graw = lane >> 3;
  const int g = min(graw, G - 1), b = lane & 7;
  const bool byte_writer = (b == 0) && (graw < G) && (8 * graw < nrep);
  const uint32_t r = U.r0 + (uint32_t)lane;
  double cost = act ? U.cost[lane] : 0.0;
  double best = act ? U.best_cost[lane] : 0.0;
  double beta = act ? U.beta[lane] : 0.0;
  unsigned long long accn = 0;
  const CT* coef = static_cast<const CT*>(U.coef);
  const int64_t sweep_begin = U.progress[0];
  const int n_run = (int)min((int64_t)a.n_sweeps, U.progress[1]);
  for (int s = 0; s < n_run; ++s) {
    const int64_t sweep64 = sweep_begin + s;
    const uint32_t sweep = (uint32_t)sweep64;
    if (U.sched) beta = U.sched[sweep64 - U.sched_base];
    for (int i = lane; i < stride / 4; i += 64)
      reinterpret_cast<uint32_t*>(start)[i] = reinterpret_cast<uint32_t*>(sp)[i];
    __syncthreads();
    double run = cost;
    double mincost = best;
    int kmin = -1;
    for (int i = 0; i < N; ++i) {
      const int q = U.pos_of_var[i];
      const int e0 = U.inc_off[q], e1 = U.inc_off[q + 1];
      CT h = CT(0);
      for (int e = e0; e < e1; ++e) {
        const CT cc = coef[e];
        uint32_t bit;
        if (PAIR) {
          bit = (sp[U.nbr[e] * G + g] >> b) & 1u;
        } else {
          bit = 0;
          for (int o = U.oth_off[e]; o < U.oth_off[e + 1]; ++o) {
            const uint32_t x = (sp[U.oth_pos[o] * G + g] >> b) & 1u;
            bit = (KIND == QMC_ISING) ? (bit ^ x) : (bit | x);
          }
        }
        if (KIND == QMC_ISING) h += bit ? -cc : cc;
        else h += bit ? CT(0) : cc;
      }
      const uint32_t own = (sp[q * G + g] >> b) & 1u;
      const CT dE = delta_from_field<KIND, CT>(h, own);
      bool ok = false;
      if (act) {
        const uint4 quad = philox4x32_10((uint32_t)i, sweep, r >> 1, DOM_SWEEP, U.seed, U.stream);
        const uint32_t w0 = (r & 1u) ? quad.y : quad.x, w1 = (r & 1u) ? quad.w : quad.z;
        ok = metropolis_accept((double)dE, beta, w0, [&](int row, int i) { return lut[row * 256 + i]; }, [&]() {
          const uint4 w = philox4x32_10((uint32_t)i, sweep, r, DOM_REDRAW, U.seed, U.stream);
          return make_uint4(w1, w.x, w.y, 0u);
        });
      }
      // one writer per byte: lanes 8g..8g+7 combine their flips
      const unsigned long long bal = __ballot(ok);
      if (ok) {
        run += (double)dE;
        accn++;
        if (run < mincost) {
          mincost = run;
          kmin = i;
        }
      }
      if (byte_writer) {
        const uint32_t fm = (uint32_t)((bal >> (8 * g)) & 0xffu);
        sp[q * G + g] ^= (uint8_t)fm;
      }
      __syncthreads();
    }
    cost = run;
    // lowest state: bits of variables <= kmin from the end state, others from the start
    const unsigned long long need = __ballot(act && kmin >= 0);
    if (need) {
      for (int q = 0; q < N; ++q) {
        // every lane handles its own bit; bytes assembled by lane 8g
        const int i = U.var_of_pos[q];
        const uint32_t cb = (sp[q * G + g] >> b) & 1u;
        const uint32_t sb = (start[q * G + g] >> b) & 1u;
        const bool mine = act && kmin >= 0;
        const uint32_t v = (i <= kmin) ? cb : sb;
        const unsigned long long vb = __ballot(mine && v);
        const unsigned long long mb = __ballot(mine);
        if (byte_writer) {
          const uint32_t m8 = (uint32_t)((mb >> (8 * g)) & 0xffu);
          const uint32_t v8 = (uint32_t)((vb >> (8 * g)) & 0xffu);
          if (m8) {
            const uint32_t old = U.best[q * G + g];
            U.best[q * G + g] = (uint8_t)((old & ~m8) | (v8 & m8));
          }
        }
      }
      if (act && kmin >= 0) best = mincost;
    }
    __syncthreads();
  }
  for (int i = lane; i < stride / 4; i += 64)
    reinterpret_cast<uint32_t*>(U.spins)[i] = reinterpret_cast<uint32_t*>(sp)[i];
  if (act) {
    U.cost[lane] = cost;
    U.best_cost[lane] = best;
    U.beta[lane] = beta;
    U.accepted[lane] += accn;
  }
  if (lane == 0) {
    U.progress[0] = sweep_begin + n_run;
    U.progress[1] -= n_run;
  }
}

// ---------------------------------------------------------------------------
// PT exchange across units (ladders spanning several work-groups).
// One thread per ladder pair.  costs/betas/labels are per global replica.
// The step is derived from the ladder's sweep progress (next sweep / spr - 1)
// so graph replays need no host-side argument updates; last_step[ladder]
// keeps an exchange from running twice for one step.
__global__ void k_exchange(double* __restrict__ cost, double* __restrict__ beta,
                           int32_t* __restrict__ label, const double* __restrict__ lbeta,
                           int n_ladders, int L, const int64_t* __restrict__ progress,
                           int units_per_ladder, int spr, int64_t* __restrict__ last_step,
                           uint32_t seed0, uint32_t stream, PtObs obs) {
  const int ld = blockIdx.x;
  if (ld >= n_ladders) return;
  const int64_t next = progress[2 * (int64_t)ld * units_per_ladder];
  const int64_t step64 = next / spr - 1;
  __shared__ int go;
  if (threadIdx.x == 0) go = (next % spr == 0) && step64 > last_step[ld];
  __syncthreads();
  if (!go) return;
  const uint32_t step = (uint32_t)step64;
  const int p = threadIdx.x;
  const int i = (int)(step & 1u) + 2 * p;
  // observations of the step, before its swaps
  const bool ob = obs.log != nullptr && step64 < obs.log_steps;
  double* rec = ob ? obs.log + ((int64_t)ld * obs.log_steps + step64) * L * 4 : nullptr;
  int32_t* dir = ob ? obs.dir + ld * L : nullptr;
  if (ob && p < L) {
    const int slot = ld * L + label[ld * L + p];
    const unsigned long long a = obs.accepted[slot];
    rec[4 * p + 0] = cost[slot];
    rec[4 * p + 1] = (double)(a - obs.acc_prev[slot]);
    obs.acc_prev[slot] = a;
    rec[4 * p + 2] = (double)dir[p];
    rec[4 * p + 3] = -1.0;
  }
  __syncthreads();
  if (i < L - 1) {
    int32_t* lab = label + ld * L;
    const int sa = ld * L + lab[i], sb = ld * L + lab[i + 1];
    const uint32_t seed = seed0 + (uint32_t)ld;
    const uint4 w = philox4x32_10((uint32_t)i, step, 0u, DOM_EXCHANGE, seed, stream);
    const double u = canonical2(w.x, w.y);
    const double bi = lbeta[i], bj = lbeta[i + 1];
    const double arg = __dmul_rn(__dsub_rn(cost[sa], cost[sb]), __dsub_rn(bi, bj));
    const bool swap = u < exp(arg);
    if (swap) {
      lab[i] = sb - ld * L;
      lab[i + 1] = sa - ld * L;
      beta[sa] = bj;
      beta[sb] = bi;
    }
    if (ob) {
      rec[4 * i + 3] = swap ? 1.0 : 0.0;
      if (swap) {  // the direction labels travel with the states
        const int32_t t = dir[i];
        dir[i] = dir[i + 1];
        dir[i + 1] = t;
      }
    }
  }
  __syncthreads();
  if (ob && threadIdx.x == 0) {
    dir[0] = 0;
    dir[L - 1] = 1;
  }
  if (threadIdx.x == 0) last_step[ld] = step64;
}

// ---------------------------------------------------------------------------
// Initial random states: flip variable i of replica r if bit i%32 of word
// i/32 is set (model/ising.h:23-41), words from the INIT counter domain.
__global__ void k_init_states(const UnitDev* __restrict__ units, uint32_t epoch) {
  const UnitDev U = units[blockIdx.y];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > U.N) return;
  for (int g = 0; g < U.G; ++g) {
    uint32_t byte = 0;
    if (q < U.N) {
      const uint32_t i = (uint32_t)U.var_of_pos[q];
      const uint32_t wi = i >> 5;
      for (int b = 0; b < 8; ++b) {
        const int rr = 8 * g + b;
        if (rr < unit_nrep(U)) {
          const uint32_t r = U.r0 + (uint32_t)rr;
          const uint4 w = philox4x32_10(wi >> 2, r, epoch, DOM_INIT, U.seed, U.stream);
          const uint32_t word = pick(w, wi & 3u);
          byte |= ((word >> (i & 31u)) & 1u) << b;
        }
      }
    }
    U.spins[q * U.G + g] = (uint8_t)byte;
    U.best[q * U.G + g] = (uint8_t)byte;
  }
}

// One work-group of 256 threads per set-up entry: counters, progress,
// initial states (as k_init_states, staged in LDS) and energies.  Eight
// replicas of a byte lane are summed together: thread t adds terms t,
// t + 256, ... in order and the 256 partial sums meet in the fixed tree of
// state_energy_256, so each replica's energy equals launch_energy's bit for
// bit.
__global__ void __launch_bounds__(256) k_setup_units(const SetupDesc* __restrict__ descs) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* part = reinterpret_cast<double*>(smem_raw);  // [8][256]
  uint8_t* sp = smem_raw + 8 * 256 * sizeof(double);
  const SetupDesc& D = descs[blockIdx.x];
  const UnitDev& U = *D.unit;
  const int tid = threadIdx.x;
  const int N = U.N, G = U.G;
  const int nrep = unit_nrep(U);
  const uint32_t epoch = D.epoch;
  if (D.clear) {
    for (int r = tid; r < (int)U.nrep; r += 256) {
      U.accepted[r] = 0ull;
      if (!D.keep_ladder) {
        U.beta[r] = 0.0;
        U.label[r] = 0;
      }
    }
    if (U.rbase == 0 && tid == 0) *U.err = 0;
    if (!D.set_progress && tid < 2) U.progress[tid] = 0;
  }
  if (D.set_progress && tid == 0) {
    U.progress[0] = D.sweep0;
    U.progress[1] = D.n_sweeps;
  }
  for (int q = tid; q <= N; q += 256) {
    for (int g = 0; g < G; ++g) {
      uint32_t byte = 0;
      if (q < N) {
        const uint32_t i = (uint32_t)U.var_of_pos[q];
        const uint32_t wi = i >> 5;
        for (int b = 0; b < 8; ++b) {
          const int rr = 8 * g + b;
          if (rr < nrep) {
            const uint4 w = philox4x32_10(wi >> 2, U.r0 + (uint32_t)rr, epoch, DOM_INIT, U.seed, U.stream);
            byte |= ((pick(w, wi & 3u) >> (i & 31u)) & 1u) << b;
          }
        }
      }
      U.spins[q * G + g] = (uint8_t)byte;
      U.best[q * G + g] = (uint8_t)byte;
      sp[q * G + g] = (uint8_t)byte;
    }
  }
  __syncthreads();
  const int64_t* __restrict__ term_off = D.term_off;
  const int32_t* __restrict__ term_var = D.term_var;
  const double* __restrict__ term_c = D.term_c;
  const int64_t M = D.M;
  const int kind = D.kind;
  for (int g = 0; g * 8 < nrep; ++g) {
    double acc[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[b] = 0.0;
    for (int64_t j = tid; j < M; j += 256) {
      uint32_t x = 0;
      for (int64_t k = term_off[j]; k < term_off[j + 1]; ++k) {
        const uint32_t byte = sp[U.pos_of_var[term_var[k]] * G + g];
        x = (kind == QMC_ISING) ? (x ^ byte) : (x | byte);
      }
      const double c = term_c[j];
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const uint32_t bit = (x >> b) & 1u;
        if (kind == QMC_ISING) acc[b] += bit ? -c : c;
        else if (!bit) acc[b] += c;
      }
    }
#pragma unroll
    for (int b = 0; b < 8; ++b) part[b * 256 + tid] = acc[b];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (tid < w) {
#pragma unroll
        for (int b = 0; b < 8; ++b) part[b * 256 + tid] += part[b * 256 + tid + w];
      }
      __syncthreads();
    }
    if (tid < 8 && 8 * g + tid < nrep) {
      const double e = part[tid * 256];
      U.cost[8 * g + tid] = e;
      U.best_cost[8 * g + tid] = e;
    }
    __syncthreads();
  }
}

// plain states [r][var] <-> unit bytes
__global__ void k_unpack(const UnitDev* __restrict__ units, int B, int r0, int R, int N, int which,
                         uint8_t* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)R * N) return;
  const int rr = (int)(t / N), v = (int)(t - (int64_t)rr * N);
  const int r = r0 + rr;
  const UnitDev& U = units[r / B];
  const int b = r % B;
  const int q = U.pos_of_var[v];
  const uint8_t* src = which ? U.best : U.spins;
  out[t] = (src[q * U.G + (b >> 3)] >> (b & 7)) & 1u;
}

__global__ void k_pack(const UnitDev* __restrict__ units, int B, int R, const uint8_t* __restrict__ in,
                       int which) {
  const UnitDev U = units[blockIdx.y];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > U.N) return;
  const int u = blockIdx.y;
  for (int g = 0; g < U.G; ++g) {
    uint32_t byte = 0;
    if (q < U.N) {
      const int v = U.var_of_pos[q];
      for (int b = 0; b < 8; ++b) {
        const int rr = 8 * g + b;
        const int r = u * B + rr;
        if (rr < unit_nrep(U) && r < R) byte |= ((uint32_t)(in[(int64_t)r * U.N + v] & 1u)) << b;
      }
    }
    if (which == 0 || which == 2) U.spins[q * U.G + g] = (uint8_t)byte;
    if (which == 1 || which == 2) U.best[q * U.G + g] = (uint8_t)byte;
  }
}

// Population resampling: replica r of the new population is a copy of
// replica src[r] of the old one (plain layout in, unit layout out).
__global__ void k_pack_gather(const UnitDev* __restrict__ units, int B, int R,
                              const int32_t* __restrict__ src, const uint8_t* __restrict__ cur,
                              const uint8_t* __restrict__ low) {
  const UnitDev U = units[blockIdx.y];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > U.N) return;
  const int u = blockIdx.y;
  for (int g = 0; g < U.G; ++g) {
    uint32_t bc = 0, bl = 0;
    if (q < U.N) {
      const int v = U.var_of_pos[q];
      for (int b = 0; b < 8; ++b) {
        const int rr = 8 * g + b;
        const int r = u * B + rr;
        if (rr < unit_nrep(U) && r < R) {
          const int64_t s = (int64_t)src[r] * U.N + v;
          bc |= ((uint32_t)(cur[s] & 1u)) << b;
          bl |= ((uint32_t)(low[s] & 1u)) << b;
        }
      }
    }
    U.spins[q * U.G + g] = (uint8_t)bc;
    U.best[q * U.G + g] = (uint8_t)bl;
  }
}

__global__ void k_gather_replicas(int n, const int32_t* __restrict__ src, const double* __restrict__ c0,
                                  const double* __restrict__ c1, const double* __restrict__ c2,
                                  const unsigned long long* __restrict__ a0, double* __restrict__ o0,
                                  double* __restrict__ o1, double* __restrict__ o2,
                                  unsigned long long* __restrict__ oa) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int s = src[r];
  o0[r] = c0[s];
  o1[r] = c1[s];
  o2[r] = c2[s];
  oa[r] = a0[s];
}

// FP64 energy of plain states [s][var]: term CSR in variable ids.
// FP64 energy of one plain state by a 256-thread work-group: strided partial
// sums over the terms, then a fixed tree (the same order for every caller, so
// single and batched read-backs agree bit for bit).
__device__ double state_energy_256(const int64_t* __restrict__ term_off, const int32_t* __restrict__ term_var,
                                   const double* __restrict__ term_c, int64_t M, int kind,
                                   const uint8_t* st, double* part) {
  double acc = 0.0;
  for (int64_t j = threadIdx.x; j < M; j += 256) {
    uint32_t x = 0;
    for (int64_t k = term_off[j]; k < term_off[j + 1]; ++k) {
      const uint32_t bit = st[term_var[k]] & 1u;
      x = (kind == QMC_ISING) ? (x ^ bit) : (x | bit);
    }
    const double c = term_c[j];
    if (kind == QMC_ISING) acc += x ? -c : c;
    else if (!x) acc += c;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  return part[0];
}

__global__ void __launch_bounds__(256)
k_energy(const int64_t* __restrict__ term_off, const int32_t* __restrict__ term_var,
         const double* __restrict__ term_c, int64_t M, int kind, int N,
         const uint8_t* __restrict__ states, double* __restrict__ out, double* __restrict__ out2) {
  __shared__ double part[256];
  const uint8_t* st = states + (int64_t)blockIdx.x * N;
  const double e = state_energy_256(term_off, term_var, term_c, M, kind, st, part);
  if (threadIdx.x == 0) {
    out[blockIdx.x] = e;
    if (out2) out2[blockIdx.x] = e;
  }
}

// Work-group row y zeroes span y: 16-byte stores over the aligned body,
// byte stores over the tail.
__global__ void __launch_bounds__(256) k_zero_spans(ZeroSpans z) {
  uint8_t* p = (uint8_t*)z.p[blockIdx.y];
  const size_t n = z.n[blockIdx.y];
  const size_t n16 = n >> 4;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (size_t i = t; i < n16; i += step) ((uint4*)p)[i] = make_uint4(0u, 0u, 0u, 0u);
  for (size_t i = (n16 << 4) + t; i < n; i += step) p[i] = 0;
}

// Batched read-back: work-group = one global replica.  Unpacks its lowest
// state into the plain output (variable order), then sums its energy from
// that copy (written by this work-group, visible after the barrier).
__global__ void __launch_bounds__(256)
k_batch_read(const ReadDesc* __restrict__ desc, const int32_t* __restrict__ rep_member,
             uint8_t* __restrict__ states, double* __restrict__ energy,
             unsigned long long* __restrict__ accepted) {
  __shared__ double part[256];
  const int64_t gr = blockIdx.x;
  const ReadDesc D = desc[rep_member[gr]];
  const int r = (int)(gr - D.roff);
  const UnitDev& U = D.units[r / D.B];
  const int b = r % D.B;
  const int G = U.G;
  uint8_t* st = states + D.soff + (int64_t)r * D.N;
  for (int v = threadIdx.x; v < D.N; v += 256) {
    const int q = U.pos_of_var[v];
    st[v] = (U.best[q * G + (b >> 3)] >> (b & 7)) & 1u;
  }
  __syncthreads();
  const double e = state_energy_256(D.term_off, D.term_var, D.term_c, D.M, D.kind, st, part);
  if (threadIdx.x == 0) {
    energy[gr] = e;
    accepted[gr] = U.accepted[b];
  }
}

// dE of every variable for plain states, with the sweep kernels' arithmetic.
template <int KIND, typename CT>
__global__ void k_delta_all(const ProblemDev P, int n_states, const uint8_t* __restrict__ states,
                            double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n_states * P.N) return;
  const int s = (int)(t / P.N), i = (int)(t - (int64_t)s * P.N);
  const uint8_t* st = states + (int64_t)s * P.N;
  const int q = P.pos_of_var[i];
  const CT* coef = static_cast<const CT*>(P.coef);
  CT h = CT(0);
  for (int e = P.inc_off[q]; e < P.inc_off[q + 1]; ++e) {
    uint32_t bit = 0;
    for (int o = P.oth_off[e]; o < P.oth_off[e + 1]; ++o) {
      const uint32_t x = st[P.var_of_pos[P.oth_pos[o]]] & 1u;
      bit = (KIND == QMC_ISING) ? (bit ^ x) : (bit | x);
    }
    const CT cc = coef[e];
    if (KIND == QMC_ISING) h += bit ? -cc : cc;
    else h += bit ? CT(0) : cc;
  }
  out[t] = (double)delta_from_field<KIND, CT>(h, st[i] & 1u);
}

__global__ void k_philox(int n, const uint32_t* __restrict__ ctr, uint32_t k0, uint32_t k1,
                         uint32_t* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint4 w = philox4x32_10(ctr[4 * t], ctr[4 * t + 1], ctr[4 * t + 2], ctr[4 * t + 3], k0, k1);
  out[4 * t] = w.x;
  out[4 * t + 1] = w.y;
  out[4 * t + 2] = w.z;
  out[4 * t + 3] = w.w;
}

// ---------------------------------------------------------------------------
// Population annealing on the device (solver/population_annealing.h:262-616).
// The population size lives in PaDev::R; units past it are idle.  One step
// runs after the step's sweeps as a chain of launches on one stream, each
// reading the step's decisions from PaDev: statistics, weights, Poisson
// offspring and their scan in one 1024-thread work-group (k_pa_step), then
// the copies of the resampled citizens and, on a restart, fresh ones.

// Poisson variate with mean `lam` from the counter uniforms
// u(k) = canonical2(Philox(i, step, k, DOM_POISSON) words 0, 1) under the key
// (seed, 0): inversion by sequential search below 30, else the PTRS
// transformed rejection of W. Hormann, "The transformed rejection method for
// generating Poisson random variables", Insurance: Mathematics and Economics
// 12 (1993) 39-45.
__device__ uint32_t pa_poisson(double lam, uint32_t i, uint32_t step, uint32_t seed) {
  auto u = [&](uint32_t k) {
    const uint4 w = philox4x32_10(i, step, k, DOM_POISSON, seed, 0u);
    return canonical2(w.x, w.y);
  };
  if (!(lam > 0.0)) return 0;
  if (lam < 30.0) {
    const double x = u(0);
    double p = exp(-lam), F = p;
    uint32_t k = 0;
    while (x > F && k < 10000) {
      ++k;
      p *= lam / (double)k;
      F += p;
      if (p < 1e-300 && (double)k > lam) break;
    }
    return k;
  }
  const double slam = sqrt(lam), loglam = log(lam);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (uint32_t it = 0; it < (1u << 20); it += 2) {
    const double U = u(it) - 0.5;
    const double V = u(it + 1);
    const double us = 0.5 - fabs(U);
    const double kd = floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return (uint32_t)kd;
    if (kd < 0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + kd * loglam - lgamma(kd + 1.0))
      return (uint32_t)kd;
  }
  return (uint32_t)floor(lam);  // not reached in practice (acceptance > 0.9 per pair)
}

// Work-group reductions of a 1024-thread block: wavefront shuffles, then the
// wave results combined in a fixed order (deterministic).
template <class T, class Op>
__device__ T pa_reduce(T v, Op op, T* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wv] = v;
  __syncthreads();
  T r = lds[0];
  for (int k = 1; k < nw; ++k) r = op(r, lds[k]);
  return r;
}

// Exclusive prefix sum over the work-group (wavefront shuffle scan, then the
// wave totals); *total = the sum over the work-group.
__device__ int pa_exclusive_scan(int v, int* lds, int* total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  __syncthreads();
  if (lane == 63) lds[wv] = x;
  __syncthreads();
  int before = 0, all = 0;
  for (int k = 0; k < nw; ++k) {
    before += k < wv ? lds[k] : 0;
    all += lds[k];
  }
  *total = all;
  return before + x - v;
}

struct MinOp {
  template <class T> __device__ T operator()(T a, T b) const { return b < a ? b : a; }
};
struct AddOp {
  template <class T> __device__ T operator()(T a, T b) const { return a + b; }
};
struct MaxOp {
  template <class T> __device__ T operator()(T a, T b) const { return b > a ? b : a; }
};

// Sum over the 64 lanes of a wavefront (butterfly: every lane ends with the
// same bits, as x + y == y + x).
__device__ __forceinline__ double pa_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// CullingEstimator (population_annealing.h:175-212) at dbeta, by one
// wavefront: weights w_i = exp(-dbeta (E_i - E_min) scale), Q = sum w,
// result = sum_i max(0, 1 - w_i target / Q) / R - epsilon0.
__device__ double pa_culling_wave(const PaArgs& A, int R, int target, double mn, double db) {
  const int lane = threadIdx.x & 63;
  double q = 0.0;
  for (int i = lane; i < R; i += 64) q += exp(-db * (A.cost[i] - mn) * A.scale);
  q = pa_wave_sum(q);
  const double rrq = (double)target / q;
  double s = 0.0;
  for (int i = lane; i < R; i += 64) {
    const double v = 1.0 - exp(-db * (A.cost[i] - mn) * A.scale) * rrq;
    if (v > 0.0) s += v;
  }
  s = pa_wave_sum(s);
  return s / (double)R - A.strategy_const;
}

// find_delta_beta (population_annealing.h:218-260): the first e in
// [-20, 20) where the estimate at 2^e changes sign against the estimate at
// 0 brackets the root (all 40 estimates at once, one per wavefront and
// round), then bisection by halving steps down to 2^-20 from the lower end.
// Every thread returns the same value.
__device__ double pa_find_delta_beta(const PaArgs& A, int R, int target, double mn, double* ce, double* bcast) {
  const int wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int j = wv; j <= 40; j += nw) {
    const double v = pa_culling_wave(A, R, target, mn, j < 40 ? ldexp(1.0, j - 20) : 0.0);
    if ((threadIdx.x & 63) == 0) ce[j] = v;
  }
  __syncthreads();
  if (wv == 0) {
    double lower = 0.0, lower_value = ce[40];
    int e;
    for (e = -20; e < 20; ++e) {
      const double uv = ce[e + 20];
      if ((lower_value < 0) == (uv < 0)) {
        lower = ldexp(1.0, e);
        lower_value = uv;
      } else {
        break;
      }
    }
    for (e -= 2; e >= -20; --e) {
      const double upper = lower + ldexp(1.0, e);
      const double uv = pa_culling_wave(A, R, target, mn, upper);
      if ((lower_value < 0) == (uv < 0)) {
        lower = upper;
        lower_value = uv;
      }
    }
    if ((threadIdx.x & 63) == 0) *bcast = lower;
  }
  __syncthreads();
  return *bcast;
}

// Prefix sum of a selection tree (RandomSelector::get_cumulative_weight,
// random_selector.cc:22-35): weight of the positions below `position`.
__device__ __forceinline__ double pa_tree_cum(const double* tree, int P, int position) {
  int p = 0;
  double w = 0.0;
  for (int k = P; k > 0; k >>= 1)
    if (position & k) {
      p |= k;
      w += tree[p];
    }
  return w;
}

// RandomSelector::add_weight (random_selector.cc:37-50).
__device__ __forceinline__ void pa_tree_add(double* tree, int P, int position, double w) {
  for (int k = 1; k <= P; k <<= 1) {
    if (position & k) position ^= k;
    else tree[position + k] += w;
  }
}

// Residual resampling (population_annealing.h:540-578): `fills` citizens
// drawn without replacement with probability proportional to their residual
// weights res[0..R), each adding one offspring to count[].  The selection
// tree is utils::RandomSelector's (random_selector.cc): node q (1..P, P the
// power of two >= R) holds the residuals of positions [q - lowbit(q), q)
// summed in position order, exactly as R sequential insertions leave it;
// pick j's uniform is canonical2 of Philox(j, step, 1, DOM_PICK) under
// (seed, 0).  The tree is built by all threads, the picks -- a sequential
// process -- by thread 0.
__device__ void pa_residual_fill(const PaArgs& A, int R, int fills, uint32_t step, double* tree_lds,
                                 int32_t* ids_lds) {
  int P = 1;
  while (P < R) P <<= 1;
  double* tree = P <= kPaTreeLds ? tree_lds : A.tree;
  int32_t* ids = P <= kPaTreeLds ? ids_lds : A.ids;
  const double* res = A.weight;
  for (int q = threadIdx.x; q <= P; q += blockDim.x) {
    double s = 0.0;
    if (q > 0) {
      const int lo = q - (q & -q), hi = q < R ? q : R;
      for (int j = lo; j < hi; ++j) s += res[j];
    }
    tree[q] = s;
  }
  for (int i = threadIdx.x; i < R; i += blockDim.x) ids[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    double total = 0.0;
    for (int j = 0; j < R; ++j) total += res[j];
    int n = R;
    for (int j = 0; j < fills && n > 0; ++j) {
      const uint4 w = philox4x32_10((uint32_t)j, step, 1u, DOM_PICK, A.seed, 0u);
      const double u = canonical2(w.x, w.y);
      // select_position (random_selector.cc:67-94); a descent past the last
      // node (possible only when rounding leaves the root below u * total)
      // stops there
      int pos = 0;
      if (n > 1) {
        double x = u * total;
        for (int k = P; k > 0; k >>= 1)
          if (pos + k <= P && x >= tree[pos + k]) {
            pos += k;
            x -= tree[pos];
          }
        if (pos >= n) pos = n - 1;
      }
      const int id = ids[pos];
      // remove_position (random_selector.cc:96-107)
      ids[pos] = ids[n - 1];
      n -= 1;
      const double before = pa_tree_cum(tree, P, pos + 1) - pa_tree_cum(tree, P, pos);
      const double after = pa_tree_cum(tree, P, n + 1) - pa_tree_cum(tree, P, n);
      total -= before;
      pa_tree_add(tree, P, pos, after - before);
      pa_tree_add(tree, P, n, -after);
      A.count[id] += 1;
    }
  }
  __syncthreads();
}

// The step's decisions (population_annealing.h:262-488 with resample
// :490-616): lowest-cost update, cost statistics, the stepping size (schedule
// difference, friction tensor, energy variance or constant culling), weights
// exp(-dbeta (E - E_min) scale), offspring counts -- Poisson with mean
// target * w / sum(w), or floor(target * w / sum(w)) plus a residual fill to
// the population size -- their scan into the copy map src, the family census
// rho = sum(family size^2) / R, and the restart rule R / rho < alpha or zero
// cost variance.  One work-group of 1024 threads.
__global__ void __launch_bounds__(1024) k_pa_step(PaArgs A) {
  __shared__ double rd[16];
  __shared__ int ri[16];
  __shared__ long long rl[16];
  __shared__ double ce[41];
  __shared__ double s_tree[kPaTreeLds + 1];
  __shared__ int32_t s_ids[kPaTreeLds];
  PaDev* P = A.pa;
  const int tid = threadIdx.x, bd = blockDim.x;
  const int R = P->R;
  const int target = P->target;
  const int64_t step = P->step;
  const double inf = __longlong_as_double(0x7ff0000000000000ll);

  double mn = inf, sum = 0.0, bmin_t = inf;
  int barg_t = 0x7fffffff;
  for (int i = tid; i < R; i += bd) {
    const double c = A.cost[i];
    mn = c < mn ? c : mn;
    sum += c;
    const double b = A.best_cost[i];
    if (b < bmin_t) {
      bmin_t = b;
      barg_t = i;
    }
  }
  mn = pa_reduce(mn, MinOp(), rd);
  sum = pa_reduce(sum, AddOp(), rd);
  const double bmin = pa_reduce(bmin_t, MinOp(), rd);
  const int barg = pa_reduce(bmin_t == bmin ? barg_t : 0x7fffffff, MinOp(), ri);
  const double avg = sum / (double)R;
  double vs = 0.0;
  for (int i = tid; i < R; i += bd) {
    const double d = A.cost[i] - avg;
    vs += d * d;
  }
  const double var = pa_reduce(vs, AddOp(), rd) / (double)R;

  // sigma = sqrt(var_before var_after), gamma = cov(before, after), and
  // tau = sweeps / (ln sigma - ln |gamma|), zeta = sigma tau where defined
  // (population_annealing.h:347-373), from the costs before the sweeps
  const double nan = __longlong_as_double(0x7ff8000000000000ll);
  double sigma = nan, tau = nan, zeta = nan;
  if (A.cost0) {
    double sb = 0.0;
    for (int i = tid; i < R; i += bd) sb += A.cost0[i];
    const double avg_b = pa_reduce(sb, AddOp(), rd) / (double)R;
    double vb = 0.0, g = 0.0;
    for (int i = tid; i < R; i += bd) {
      const double d0 = A.cost0[i] - avg_b, d1 = A.cost[i] - avg;
      vb += d0 * d0;
      g += d0 * d1;
    }
    const double var_b = pa_reduce(vb, AddOp(), rd) / (double)R;
    const double gamma = pa_reduce(g, AddOp(), rd) / (double)R;
    sigma = sqrt(var_b * var);
    if (sigma > 0.0 && gamma != 0.0) {
      const double delta = log(sigma) - log(fabs(gamma));
      if (fabs(delta) > 1e-11) {
        tau = (double)A.sweeps_per_step / delta;
        zeta = sigma * tau;
      }
    }
  }
  double db = 0.0;
  if (A.strategy == QMC_PA_SCHEDULE) {
    const int64_t k = step - P->base_step;
    const int64_t last = A.n_table - 1;
    db = A.beta_table[k + 1 < last ? k + 1 : last] - A.beta_table[k < last ? k : last];
  } else if (A.strategy == QMC_PA_CONSTANT_CULLING) {
    db = pa_find_delta_beta(A, R, target, mn, ce, &rd[0]);
  } else if (A.strategy == QMC_PA_FRICTION_TENSOR) {
    if (zeta > 0.0) db = A.strategy_const / sqrt(zeta);
  } else if (sigma > 0.0) {
    db = A.strategy_const / sqrt(sigma);
  }
  double rho = 1.0;
  int Rn = R;
  bool moved = false;
  if (db > 0.0) {
    double ws = 0.0;
    for (int i = tid; i < R; i += bd) {
      const double w = exp(-db * (A.cost[i] - mn) * A.scale);
      A.weight[i] = w;
      ws += w;
    }
    const double total = pa_reduce(ws, AddOp(), rd);
    if (A.constant_population) {
      int fl = 0;
      for (int i = tid; i < R; i += bd) {
        const double rw = (double)target * A.weight[i] / total;
        const int c = (int)floor(rw);
        A.count[i] = c;
        A.weight[i] = rw - (double)c;
        fl += c;
      }
      const int floors = pa_reduce(fl, AddOp(), ri);
      if (floors < R) pa_residual_fill(A, R, R - floors, (uint32_t)step, s_tree, s_ids);
    } else {
      const double rrq = (double)target / total;
      for (int i = tid; i < R; i += bd) {
        const double mean = rrq * A.weight[i];
        A.count[i] = mean > 0.0 ? (int)pa_poisson(mean, (uint32_t)i, (uint32_t)step, A.seed) : 0;
      }
    }
    // offspring of citizen i fill src[off_i .. off_i + count_i)
    int carry = 0;
    for (int base = 0; base < R; base += bd) {
      const int i = base + tid;
      const int c = i < R ? A.count[i] : 0;
      int tot = 0;
      const int off = carry + pa_exclusive_scan(c, ri, &tot);
      for (int j = 0; j < c; ++j) {
        const int d = off + j;
        if (d < A.cap) {
          A.src[d] = i;
          A.family_next[d] = A.family[i];
        }
      }
      carry += tot;
    }
    int current = carry;
    if (current == 0) {  // keep the population alive
      if (tid == 0) {
        const uint4 w = philox4x32_10(0u, (uint32_t)step, 0u, DOM_PICK, A.seed, 0u);
        const int i = (int)floor(canonical2(w.x, w.y) * (double)R);
        A.src[0] = i;
        A.family_next[0] = A.family[i];
      }
      current = 1;
    }
    if (current > A.cap) {
      if (tid == 0) P->err = 1;
      current = A.cap;
    }
    for (int f = tid; f < target; f += bd) A.hist[f] = 0;
    __syncthreads();
    for (int d = tid; d < current; d += bd) {
      const int f = A.family_next[d];
      if ((unsigned)f < (unsigned)target) atomicAdd(&A.hist[f], 1);
    }
    __syncthreads();
    long long sq = 0;
    int nfam = 0, fmax = -1;
    for (int f = tid; f < target; f += bd) {
      const long long h = A.hist[f];
      sq += h * h;
      nfam += h > 0 ? 1 : 0;
      fmax = h > 0 ? f : fmax;
    }
    sq = pa_reduce(sq, AddOp(), rl);
    rho = (double)sq / (double)R;
    if (A.log) {
      nfam = pa_reduce(nfam, AddOp(), ri);
      fmax = pa_reduce(fmax, MaxOp(), ri);
      if (tid == 0 && step < A.log_steps) {
        double* rec = A.log + step * kPaLogFields;
        rec[7] = (double)nfam;
        rec[9] = fmax >= 0 ? (double)A.hist[fmax] : 0.0;
      }
    }
    Rn = current;
    moved = true;
  }

  const bool restart = (double)R / rho < A.alpha || var == 0.0;
  int Rf = Rn;
  // a restart with exp_repopulation doubles the target first
  // (population_annealing.h:463-470), within the capacity
  int tnew = target;
  if (restart && A.exp_repopulation) {
    tnew = 2 * target;
    if (tnew > A.cap) {
      if (tid == 0) P->err = 1;
      tnew = A.cap;
    }
  }
  if (restart) {
    if (!moved)
      for (int i = tid; i < Rn; i += bd) {
        A.src[i] = i;
        A.family_next[i] = A.family[i];
      }
    __syncthreads();
    if (Rn < tnew) {  // top up with copies of the last citizen
      const int lastsrc = A.src[Rn - 1];
      for (int i = Rn + tid; i < tnew; i += bd) A.src[i] = lastsrc;
      Rf = tnew;
    }
    for (int i = tid; i < Rf && i < tnew; i += bd) A.family_next[i] = i;
    moved = true;
  }
  __syncthreads();
  if (moved)
    for (int i = tid; i < Rf; i += bd) A.family[i] = A.family_next[i];
  __syncthreads();
  if (tid == 0) {
    const bool imp = !P->have_best || bmin < P->best;
    P->best_src = imp ? barg : -1;
    if (imp) {
      P->best = bmin;
      P->have_best = 1;
    }
    P->attempts += (int64_t)R * A.sweeps_per_step * A.n_vars;
    P->evals += (int64_t)R * A.sweeps_per_step * A.n_vars + P->pending_inits;
    P->pending_inits = restart ? Rf : 0;
    P->R_prev = R;
    P->R_resampled = Rn;
    P->R = Rf;
    P->moved = moved ? 1 : 0;
    P->restart = restart ? 1 : 0;
    P->rho = rho;
    P->delta_beta = db;
    P->target = tnew;
    if (A.log && step < A.log_steps) {
      double* rec = A.log + step * kPaLogFields;
      rec[0] = (double)P->epoch;
      rec[1] = avg;
      rec[2] = (double)R;
      rec[3] = sigma;
      rec[4] = tau;
      rec[5] = zeta;
      rec[6] = db;
      if (!(db > 0.0)) rec[7] = rec[9] = 0.0;
      rec[8] = (double)Rn;
      rec[10] = rho;
      rec[11] = restart ? 1.0 : 0.0;
      rec[12] = db > 0.0 ? 1.0 : 0.0;
    }
    P->step = step + 1;
    if (restart) {
      P->cur_beta = A.beta_start;
      P->base_step = step;
      P->epoch += 1;
      P->restarts += 1;
    } else {
      P->cur_beta += db;
    }
  }
}

// Lowest state of a new global lowest (replica best_src of the population
// the step started from) into best_state, variable order.
__global__ void k_pa_copy_best(PaArgs A) {
  const int r = A.pa->best_src;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < 0 || v >= A.n_vars) return;
  const UnitDev& U = A.units[r / A.B];
  const int b = r % A.B;
  const int q = U.pos_of_var[v];
  A.best_state[v] = (U.best[q * U.G + (b >> 3)] >> (b & 7)) & 1u;
}

// Plain copies of the current (and lowest) states of replicas 0..n-1, n =
// R_prev (resampling) or R (a restart's fresh states).
__global__ void k_pa_unpack(PaArgs A, int fresh) {
  const PaDev* P = A.pa;
  if (fresh ? !P->restart : !P->moved) return;
  const int R = fresh ? P->R : P->R_prev;
  const int64_t n = (int64_t)R * A.n_vars;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(t / A.n_vars), v = (int)(t - (int64_t)r * A.n_vars);
    const UnitDev& U = A.units[r / A.B];
    const int b = r % A.B;
    const int at = U.pos_of_var[v] * U.G + (b >> 3);
    A.plane_cur[t] = (U.spins[at] >> (b & 7)) & 1u;
    if (!fresh) A.plane_low[t] = (U.best[at] >> (b & 7)) & 1u;
  }
}

// Per-replica values of the copies, read before any is overwritten.
__global__ void k_pa_gather(PaArgs A) {
  const PaDev* P = A.pa;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P->R) return;
  if (!P->moved) {
    A.beta[r] = P->cur_beta;
    return;
  }
  const int s = A.src[r];
  A.tcost[r] = A.cost[s];
  A.tbest[r] = A.best_cost[s];
  A.tbeta[r] = A.beta[s];
  A.tacc[r] = A.accepted[s];
}

__global__ void k_pa_scatter(PaArgs A) {
  const PaDev* P = A.pa;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (!P->moved || r >= P->R) return;
  A.cost[r] = A.tcost[r];
  A.best_cost[r] = A.tbest[r];
  A.accepted[r] = A.tacc[r];
  // after a restart the citizens keep their beta until the next step sets
  // it; the top-up starts at beta_start
  A.beta[r] = P->restart ? (r < P->R_resampled ? A.tbeta[r] : A.beta_start) : P->cur_beta;
}

// Unit bytes of the new population from the plain copies (k_pack_gather
// with the copy map and population size on the device).
__global__ void k_pa_pack(PaArgs A) {
  const PaDev* P = A.pa;
  const UnitDev& U = A.units[blockIdx.y];
  if (!P->moved || unit_nrep(U) == 0) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > U.N) return;
  const int u = blockIdx.y, nrep = unit_nrep(U);
  for (int g = 0; g < U.G; ++g) {
    uint32_t bc = 0, bl = 0;
    if (q < U.N) {
      const int v = U.var_of_pos[q];
      for (int b = 0; b < 8; ++b) {
        const int rr = 8 * g + b;
        if (rr < nrep) {
          const int64_t s = (int64_t)A.src[u * A.B + rr] * U.N + v;
          bc |= ((uint32_t)(A.plane_cur[s] & 1u)) << b;
          bl |= ((uint32_t)(A.plane_low[s] & 1u)) << b;
        }
      }
    }
    U.spins[q * U.G + g] = (uint8_t)bc;
    U.best[q * U.G + g] = (uint8_t)bl;
  }
}

// A restart's fresh random citizens (model/ising.h:23-41 draws, epoch from
// the device state), as k_init_states.
__global__ void k_pa_init(PaArgs A) {
  const PaDev* P = A.pa;
  const UnitDev& U = A.units[blockIdx.y];
  if (!P->restart || unit_nrep(U) == 0) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > U.N) return;
  const uint32_t epoch = P->epoch;
  const int nrep = unit_nrep(U);
  for (int g = 0; g < U.G; ++g) {
    uint32_t byte = 0;
    if (q < U.N) {
      const uint32_t i = (uint32_t)U.var_of_pos[q];
      const uint32_t wi = i >> 5;
      for (int b = 0; b < 8; ++b) {
        const int rr = 8 * g + b;
        if (rr < nrep) {
          const uint4 w = philox4x32_10(wi >> 2, U.r0 + (uint32_t)rr, epoch, DOM_INIT, U.seed, U.stream);
          byte |= ((pick(w, wi & 3u) >> (i & 31u)) & 1u) << b;
        }
      }
    }
    U.spins[q * U.G + g] = (uint8_t)byte;
    U.best[q * U.G + g] = (uint8_t)byte;
  }
}

// Energies of a restart's fresh citizens: cost = lowest cost = E(state);
// with cost0, every citizen's cost is kept as the next step's cost before
// its sweeps.
__global__ void __launch_bounds__(256)
k_pa_energy(PaArgs A, const int64_t* __restrict__ term_off, const int32_t* __restrict__ term_var,
            const double* __restrict__ term_c, int64_t M, int kind) {
  __shared__ double part[256];
  const PaDev* P = A.pa;
  if ((int)blockIdx.x >= P->R) return;
  if (!P->restart) {
    // the next step's costs before its sweeps
    if (A.cost0 && threadIdx.x == 0) A.cost0[blockIdx.x] = A.cost[blockIdx.x];
    return;
  }
  const double e = state_energy_256(term_off, term_var, term_c, M, kind,
                                    A.plane_cur + (int64_t)blockIdx.x * A.n_vars, part);
  if (threadIdx.x == 0) {
    A.cost[blockIdx.x] = e;
    A.best_cost[blockIdx.x] = e;
    if (A.cost0) A.cost0[blockIdx.x] = e;
  }
}

// LDS read bandwidth probe: consecutive lanes read consecutive 16-byte words
// (no bank conflicts); the accumulated words keep the loads alive.
__global__ void __launch_bounds__(1024) k_lds_probe(int iters, uint32_t* __restrict__ out) {
  __shared__ uint4 buf[4096];
  for (int i = threadIdx.x; i < 4096; i += blockDim.x)
    buf[i] = make_uint4((uint32_t)i, (uint32_t)i * 3u, (uint32_t)i * 5u, (uint32_t)i * 7u);
  __syncthreads();
  uint4 acc = make_uint4(0u, 0u, 0u, 0u);
  int idx = threadIdx.x;
#pragma unroll 8
  for (int it = 0; it < iters; ++it) {
    const uint4 v = buf[idx];
    acc.x ^= v.x;
    acc.y += v.y;
    acc.z ^= v.z;
    acc.w += v.w;
    idx = (idx + 64) & 4095;
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = acc.x ^ acc.y ^ acc.z ^ acc.w;
}

// Staged uploads: record b copies rec[b].bytes bytes from the device copy of
// the staging ring to rec[b].dst, in 4-byte words when everything is aligned.
__global__ void __launch_bounds__(256) k_scatter(const uint8_t* __restrict__ stage,
                                                 const ScatterRec* __restrict__ rec) {
  const ScatterRec r = rec[blockIdx.x];
  const uint8_t* src = stage + r.src;
  uint8_t* dst = reinterpret_cast<uint8_t*>(r.dst);
  if (((r.src | r.dst | r.bytes) & 3u) == 0) {
    const uint32_t* s4 = reinterpret_cast<const uint32_t*>(src);
    uint32_t* d4 = reinterpret_cast<uint32_t*>(dst);
    for (uint64_t i = threadIdx.x; i < r.bytes / 4; i += blockDim.x) d4[i] = s4[i];
  } else {
    for (uint64_t i = threadIdx.x; i < r.bytes; i += blockDim.x) dst[i] = src[i];
  }
}

// ---------------------------------------------------------------------------
// launch helpers (host side of this translation unit)
size_t sweep_smem_bytes(int max_stride, int ordered, int nb, int threads, int* best_lds, int allow_best_lds) {
  if (best_lds) *best_lds = 0;
  if (ordered) return 8192 + 2 * (size_t)max_stride;
  const size_t fixed = sizeof(Smem) + slot_bytes(slot_groups(nb, threads));
  const size_t with_best = fixed + 3 * (size_t)max_stride;
  if (allow_best_lds && with_best <= 160 * 1024) {
    if (best_lds) *best_lds = 1;
    return with_best;
  }
  return fixed + 2 * (size_t)max_stride;
}

// The dynamic-LDS limit is raised once per kernel to the full 160 KiB, so
// launches (also inside graph capture) need no attribute calls.
template <class F>
static hipError_t raise_lds_limit(F fn) {
  static hipError_t once = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               160 * 1024);
  return once;
}

template <int KIND, int PAIR, typename CT, int NB, int E>
static hipError_t launch_sweep_t(const UnitDev* units, int n_units, int block, size_t smem,
                                 const SweepArgs& a, int ordered, int cooperative, int ladder, hipStream_t st) {
  if (ordered) {
    auto fn = k_sweep_ordered<KIND, (PAIR != 0), CT>;
    hipError_t e = raise_lds_limit(fn);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3(n_units), dim3(64), smem, st, units, a);
    return hipGetLastError();
  }
  auto fn = ladder ? k_sweep<KIND, PAIR, CT, NB, true, E> : k_sweep<KIND, PAIR, CT, NB, false, E>;
  hipError_t e = ladder ? raise_lds_limit(k_sweep<KIND, PAIR, CT, NB, true, E>)
                        : raise_lds_limit(k_sweep<KIND, PAIR, CT, NB, false, E>);
  if (e != hipSuccess) return e;
  if (cooperative) {
    SweepArgs args = a;
    const UnitDev* u = units;
    void* kargs[] = {(void*)&u, (void*)&args};
    return hipLaunchCooperativeKernel((const void*)fn, dim3(n_units), dim3(block), kargs, (unsigned)smem, st);
  }
  hipLaunchKernelGGL(fn, dim3(n_units), dim3(block), smem, st, units, a);
  return hipGetLastError();
}

// Rows narrower than 8 entries get kernels of their own for up to 2
// replicas per byte lane (the lattice shapes); the rest read 8.
template <int KIND, int PAIR, typename CT, int NB>
static hipError_t launch_e(int row_e, const UnitDev* units, int n_units, int block, size_t smem,
                           const SweepArgs& a, int ordered, int coop, int ladder, hipStream_t st) {
  if constexpr (PAIR != 0 && NB <= 2) {
    if (row_e == 4)
      return launch_sweep_t<KIND, PAIR, CT, NB, 4>(units, n_units, block, smem, a, ordered, coop, ladder, st);
    if (row_e == 6)
      return launch_sweep_t<KIND, PAIR, CT, NB, 6>(units, n_units, block, smem, a, ordered, coop, ladder, st);
  }
  return launch_sweep_t<KIND, PAIR, CT, NB, 8>(units, n_units, block, smem, a, ordered, coop, ladder, st);
}

template <int KIND, int PAIR, typename CT>
static hipError_t launch_nb(int nb, int row_e, const UnitDev* units, int n_units, int block, size_t smem,
                            const SweepArgs& a, int ordered, int coop, int ladder, hipStream_t st) {
  if (nb <= 1) return launch_e<KIND, PAIR, CT, 1>(row_e, units, n_units, block, smem, a, ordered, coop, ladder, st);
  if (nb <= 2) return launch_e<KIND, PAIR, CT, 2>(row_e, units, n_units, block, smem, a, ordered, coop, ladder, st);
  if (nb <= 4) return launch_e<KIND, PAIR, CT, 4>(row_e, units, n_units, block, smem, a, ordered, coop, ladder, st);
  return launch_e<KIND, PAIR, CT, 8>(row_e, units, n_units, block, smem, a, ordered, coop, ladder, st);
}

hipError_t launch_sweep(int kind, int pairwise, int coupling, int nb, int row_e, const UnitDev* units, int n_units,
                        int block, size_t smem, const SweepArgs& a, int ordered, int coop, int ladder,
                        hipStream_t st) {
  const bool i32 = coupling == QMC_COUPLING_I32;
#define QMC_LAUNCH(K, P)                                                                                       \
  return i32 ? launch_nb<K, P, int32_t>(nb, row_e, units, n_units, block, smem, a, ordered, coop, ladder, st) \
             : launch_nb<K, P, float>(nb, row_e, units, n_units, block, smem, a, ordered, coop, ladder, st)
  // int8 entries only for integer coefficients
  if (pairwise == 2 && !i32) return hipErrorInvalidValue;
  if (kind == QMC_ISING) {
    if (pairwise == 2)
      return launch_nb<QMC_ISING, 2, int32_t>(nb, row_e, units, n_units, block, smem, a, ordered, coop, ladder, st);
    if (pairwise) QMC_LAUNCH(QMC_ISING, 1);
    QMC_LAUNCH(QMC_ISING, 0);
  }
  if (pairwise == 2)
    return launch_nb<QMC_PUBO, 2, int32_t>(nb, row_e, units, n_units, block, smem, a, ordered, coop, ladder, st);
  if (pairwise) QMC_LAUNCH(QMC_PUBO, 1);
  QMC_LAUNCH(QMC_PUBO, 0);
#undef QMC_LAUNCH
}

template <int KIND, int PAIR, typename CT, int NB, int E>
static int coresident_t(int block, size_t smem) {
  int per_cu = 0;
  // cooperative launches carry ladders
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_sweep<KIND, PAIR, CT, NB, true, E>, block,
                                                   smem) != hipSuccess)
    return 0;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return per_cu * cus;
}

template <int KIND, int PAIR, typename CT, int NB>
static int coresident_e(int row_e, int block, size_t smem) {
  if constexpr (PAIR != 0 && NB <= 2) {
    if (row_e == 4) return coresident_t<KIND, PAIR, CT, NB, 4>(block, smem);
    if (row_e == 6) return coresident_t<KIND, PAIR, CT, NB, 6>(block, smem);
  }
  return coresident_t<KIND, PAIR, CT, NB, 8>(block, smem);
}

template <int KIND, int PAIR, typename CT>
static int coresident_nb(int nb, int row_e, int block, size_t smem) {
  if (nb <= 1) return coresident_e<KIND, PAIR, CT, 1>(row_e, block, smem);
  if (nb <= 2) return coresident_e<KIND, PAIR, CT, 2>(row_e, block, smem);
  if (nb <= 4) return coresident_e<KIND, PAIR, CT, 4>(row_e, block, smem);
  return coresident_e<KIND, PAIR, CT, 8>(row_e, block, smem);
}

int max_coresident_blocks(int kind, int pairwise, int coupling, int nb, int row_e, int block, size_t smem) {
  const bool i32 = coupling == QMC_COUPLING_I32;
#define QMC_CORES(K, P)                                                \
  return i32 ? coresident_nb<K, P, int32_t>(nb, row_e, block, smem) \
             : coresident_nb<K, P, float>(nb, row_e, block, smem)
  if (pairwise == 2 && !i32) return 0;
  if (kind == QMC_ISING) {
    if (pairwise == 2) return coresident_nb<QMC_ISING, 2, int32_t>(nb, row_e, block, smem);
    if (pairwise) QMC_CORES(QMC_ISING, 1);
    QMC_CORES(QMC_ISING, 0);
  }
  if (pairwise == 2) return coresident_nb<QMC_PUBO, 2, int32_t>(nb, row_e, block, smem);
  if (pairwise) QMC_CORES(QMC_PUBO, 1);
  QMC_CORES(QMC_PUBO, 0);
#undef QMC_CORES
}

hipError_t launch_exchange(double* cost, double* beta, int32_t* label, const double* lbeta,
                           int n_ladders, int L, const int64_t* progress, int units_per_ladder,
                           int spr, int64_t* last_step, uint32_t seed, uint32_t stream,
                           const PtObs& obs, hipStream_t st) {
  if (n_ladders <= 0 || L < 2) return hipSuccess;
  if (obs.log && L > 32) return hipErrorInvalidValue;  // one thread per ladder position
  hipLaunchKernelGGL(k_exchange, dim3(n_ladders), dim3(32), 0, st, cost, beta, label, lbeta,
                     n_ladders, L, progress, units_per_ladder, spr, last_step, seed, stream, obs);
  return hipGetLastError();
}

hipError_t launch_init(const UnitDev* units, int n_units, int N, uint32_t epoch, hipStream_t st) {
  hipLaunchKernelGGL(k_init_states, dim3((N + 1 + 127) / 128, n_units), dim3(128), 0, st, units, epoch);
  return hipGetLastError();
}

bool setup_fits(int max_stride) { return 8 * 256 * sizeof(double) + (size_t)max_stride <= 160 * 1024; }

hipError_t launch_setup(const SetupDesc* descs, int n, int max_stride, hipStream_t st) {
  if (!setup_fits(max_stride)) return hipErrorInvalidValue;
  if (n <= 0) return hipSuccess;
  static hipError_t once = hipFuncSetAttribute((const void*)k_setup_units,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (once != hipSuccess) return once;
  const size_t smem = 8 * 256 * sizeof(double) + (size_t)max_stride;
  hipLaunchKernelGGL(k_setup_units, dim3(n), dim3(256), smem, st, descs);
  return hipGetLastError();
}

hipError_t launch_unpack(const UnitDev* units, int B, int r0, int R, int N, int which, uint8_t* out,
                         hipStream_t st) {
  const int64_t n = (int64_t)R * N;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_unpack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, units, B, r0, R, N, which, out);
  return hipGetLastError();
}

hipError_t launch_pack(const UnitDev* units, int n_units, int B, int R, int N, const uint8_t* in,
                       int which, hipStream_t st) {
  hipLaunchKernelGGL(k_pack, dim3((N + 1 + 127) / 128, n_units), dim3(128), 0, st, units, B, R, in, which);
  return hipGetLastError();
}

hipError_t launch_pack_gather(const UnitDev* units, int n_units, int B, int R, int N,
                              const int32_t* src, const uint8_t* cur, const uint8_t* low, hipStream_t st) {
  hipLaunchKernelGGL(k_pack_gather, dim3((N + 1 + 127) / 128, n_units), dim3(128), 0, st, units, B, R,
                     src, cur, low);
  return hipGetLastError();
}

hipError_t launch_gather_replicas(int n, const int32_t* src, const double* c0, const double* c1,
                                  const double* c2, const unsigned long long* a0, double* o0,
                                  double* o1, double* o2, unsigned long long* oa, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_replicas, dim3((n + 255) / 256), dim3(256), 0, st, n, src, c0, c1, c2,
                     a0, o0, o1, o2, oa);
  return hipGetLastError();
}

hipError_t launch_energy(const ProblemDev& P, const uint8_t* states, int n, double* out, hipStream_t st,
                         double* out2) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_energy, dim3(n), dim3(256), 0, st, P.term_off, P.term_var, P.term_c, P.M,
                     P.kind, P.N, states, out, out2);
  return hipGetLastError();
}

hipError_t launch_pa_step(const PaArgs& a, const ProblemDev& P, hipStream_t st) {
  const int N = a.n_vars;
  hipLaunchKernelGGL(k_pa_step, dim3(1), dim3(1024), 0, st, a);
  hipLaunchKernelGGL(k_pa_copy_best, dim3((N + 255) / 256), dim3(256), 0, st, a);
  const int64_t planes = (int64_t)a.cap * N;
  const unsigned ublocks = (unsigned)std::min<int64_t>((planes + 255) / 256, 8192);
  hipLaunchKernelGGL(k_pa_unpack, dim3(ublocks), dim3(256), 0, st, a, 0);
  hipLaunchKernelGGL(k_pa_gather, dim3((a.cap + 255) / 256), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_pa_scatter, dim3((a.cap + 255) / 256), dim3(256), 0, st, a);
  const dim3 ugrid((N + 1 + 127) / 128, a.n_units);
  hipLaunchKernelGGL(k_pa_pack, ugrid, dim3(128), 0, st, a);
  hipLaunchKernelGGL(k_pa_init, ugrid, dim3(128), 0, st, a);
  hipLaunchKernelGGL(k_pa_unpack, dim3(ublocks), dim3(256), 0, st, a, 1);
  hipLaunchKernelGGL(k_pa_energy, dim3(a.cap), dim3(256), 0, st, a, P.term_off, P.term_var, P.term_c, P.M,
                     P.kind);
  return hipGetLastError();
}

hipError_t launch_zero(const ZeroSpans& z, int count, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (count > kZeroSpans) return hipErrorInvalidValue;
  size_t most = 0;
  for (int i = 0; i < count; ++i) {
    if ((uintptr_t)z.p[i] & 15u) return hipErrorInvalidValue;
    most = std::max(most, z.n[i]);
  }
  if (most == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>((most / 16 + 255) / 256 + 1, 1024);
  hipLaunchKernelGGL(k_zero_spans, dim3((unsigned)blocks, count), dim3(256), 0, st, z);
  return hipGetLastError();
}

hipError_t launch_batch_read(const ReadDesc* desc, const int32_t* rep_member, int64_t n_replicas,
                             uint8_t* states, double* energy, unsigned long long* accepted,
                             hipStream_t st) {
  if (n_replicas <= 0) return hipSuccess;
  if (n_replicas > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_batch_read, dim3((unsigned)n_replicas), dim3(256), 0, st, desc, rep_member, states,
                     energy, accepted);
  return hipGetLastError();
}

hipError_t launch_delta_all(const ProblemDev& P, int n_states, const uint8_t* states, double* out,
                            hipStream_t st) {
  const int64_t n = (int64_t)n_states * P.N;
  if (n == 0) return hipSuccess;
  dim3 grid((unsigned)((n + 255) / 256));
  if (P.kind == QMC_ISING) {
    if (P.coupling == QMC_COUPLING_I32) hipLaunchKernelGGL((k_delta_all<QMC_ISING, int32_t>), grid, dim3(256), 0, st, P, n_states, states, out);
    else hipLaunchKernelGGL((k_delta_all<QMC_ISING, float>), grid, dim3(256), 0, st, P, n_states, states, out);
  } else {
    if (P.coupling == QMC_COUPLING_I32) hipLaunchKernelGGL((k_delta_all<QMC_PUBO, int32_t>), grid, dim3(256), 0, st, P, n_states, states, out);
    else hipLaunchKernelGGL((k_delta_all<QMC_PUBO, float>), grid, dim3(256), 0, st, P, n_states, states, out);
  }
  return hipGetLastError();
}

double launch_lds_probe(int blocks, int iters, uint32_t* out, hipStream_t st) {
  hipLaunchKernelGGL(k_lds_probe, dim3(blocks), dim3(1024), 0, st, iters, out);
  return hipGetLastError() == hipSuccess ? (double)blocks * 1024.0 * iters * 16.0 : -1.0;
}

hipError_t launch_scatter(const uint8_t* stage, const ScatterRec* rec, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)n), dim3(256), 0, st, stage, rec);
  return hipGetLastError();
}

hipError_t launch_philox(int n, const uint32_t* ctr, uint32_t k0, uint32_t k1, uint32_t* out,
                         hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_philox, dim3((n + 255) / 256), dim3(256), 0, st, n, ctr, k0, k1, out);
  return hipGetLastError();
}

}  // namespace qmc
