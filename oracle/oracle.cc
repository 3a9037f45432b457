// CPU oracle for the QIOToolkit annealing hot path -- TEST INFRASTRUCTURE.
//
// A plain C++ restatement of the reference's algorithm, used only by tests/,
// the smoke() entry point and bench.py's CPU-baseline leg, as the checker.
// The product library never links it.  Every function cites the reference
// file:line it restates (paths relative to the reference's cpp/ directory).
//
// Parity pins (see tests/test_oracle_golden.py):
//   pcg32/uniform/fork ....... utils/test/random_generator_test.cc:38-78 golden vectors
//   Philox4x32-10 ............ Random123 known-answer vectors (Salmon et al., SC'11)
//   ln-LUT accept ............ closed form + exhaustive bucket/row cells
//   cost / delta ............. brute force over all 2^N states, cost+dE==cost',
//                              model/test known answers (ring-10, constants)
//   graph normalisation ...... graph/test and model/test id-remap answers
//   schedules ................ schedule/test known values
//   SA / PT / PA ............. data/*.json fixture optima, planted optima in
//                              data/input_e2e_pipeline/*/energies.json
//   std::poisson_distribution  parity unpinned (libstdc++ algorithm, used as-is)
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// pcg32 = pcg_engines::setseq_xsh_rr_64_32 as used by utils::PCG
// (utils/random_generator.h:79-96, random_generator.cc:47-52).  Written from
// O'Neill, "PCG: A Family of Simple Fast Space-Efficient Statistically Good
// Algorithms for Random Number Generation" (2014): 64-bit LCG, output
// XSH-RR of the *previous* state; seed(s) sets state = bump(s + inc).
// One cache line each: the replicas' generators and walkers sit side by side
// in vectors and are advanced by different OpenMP threads.
struct alignas(64) Pcg32 {
  using result_type = uint32_t;
  static constexpr uint64_t kMult = 6364136223846793005ULL;
  static constexpr uint64_t kInc = 1442695040888963407ULL;
  uint64_t state;
  Pcg32() { seed64(0xcafef00dd15ea5e5ULL); }
  explicit Pcg32(uint32_t s) { seed64(s); }
  void seed64(uint64_t s) { state = (s + kInc) * kMult + kInc; }
  static constexpr result_type min() { return 0; }
  static constexpr result_type max() { return 0xffffffffu; }
  result_type operator()() {
    uint64_t old = state;
    state = old * kMult + kInc;
    uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
    uint32_t rot = (uint32_t)(old >> 59);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
  // utils::PCG::uniform: std::uniform_real_distribution<double>(0,1)
  // (random_generator.cc:11-12,49) -- generate_canonical with two draws.
  double uniform() {
    std::uniform_real_distribution<double> d(0.0, 1.0);
    return d(*this);
  }
  // utils::PCG::poisson (random_generator.cc:54-58): libstdc++'s
  // std::poisson_distribution<int>, constructed per call.
  uint32_t poisson(double mean) {
    std::poisson_distribution<> p(mean);
    return (uint32_t)p(*this);
  }
  // utils::PCG::fork (random_generator.cc:37-45)
  Pcg32 fork() { return Pcg32((*this)()); }
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): the counter-based generator of the GPU contract.
void philox4x32_10(const uint32_t in_ctr[4], const uint32_t in_key[2], uint32_t out[4]) {
  uint32_t c0 = in_ctr[0], c1 = in_ctr[1], c2 = in_ctr[2], c3 = in_ctr[3];
  uint32_t k0 = in_key[0], k1 = in_key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Counter domains of the GPU RNG contract.
enum : uint32_t { DOM_SWEEP = 0, DOM_REDRAW = 1, DOM_INIT = 2, DOM_EXCHANGE = 3, DOM_POISSON = 4, DOM_PICK = 5 };

inline uint32_t philox_word(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2,
                            uint32_t c3, int w) {
  uint32_t ctr[4] = {c0, c1, c2, c3}, key[2] = {k0, k1}, o[4];
  philox4x32_10(ctr, key, o);
  return o[w];
}

// generate_canonical<double, 53> over two 32-bit words (libstdc++ random.tcc):
// (w0 + w1 * 2^32) / 2^64, clamped below 1.
inline double canonical2(uint32_t w0, uint32_t w1) {
  double s = (double)w0 + (double)w1 * 4294967296.0;
  double r = s / 18446744073709551616.0;
  if (r >= 1.0) r = std::nextafter(1.0, 0.0);
  return r;
}

// Poisson variate with mean lam in counter mode: the uniforms are
// u(k) = canonical2(Philox(i, step, k, DOM_POISSON) words 0, 1) under the key
// (seed, 0).  Inversion by sequential search for lam < 30; above, the PTRS
// transformed rejection of W. Hormann, Insurance: Mathematics and Economics
// 12 (1993) 39-45.  (The reference draws std::poisson_distribution from its
// master pcg32, random_generator.cc:54-58.)
uint32_t poisson_counter(double lam, uint32_t i, uint32_t step, uint32_t seed) {
  auto u = [&](uint32_t k) {
    uint32_t ctr[4] = {i, step, k, DOM_POISSON}, key[2] = {seed, 0u}, w[4];
    philox4x32_10(ctr, key, w);
    return canonical2(w[0], w[1]);
  };
  if (!(lam > 0.0)) return 0;
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
  for (uint32_t it = 0; it < (1u << 20); it += 2) {
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
  return (uint32_t)std::floor(lam);
}

// ---------------------------------------------------------------------------
// utils::uniform_ln (utils/utils.cc:12-32): 4 rows x 256 buckets of
// ln(i / 256^(row+1)).
struct UniformLn {
  double t[4][256];
  UniformLn() {
    for (int r = 0; r < 4; ++r)
      for (int i = 0; i < 256; ++i) t[r][i] = std::log((double)i / std::pow(256.0, (double)(r + 1)));
  }
};
const UniformLn& uniform_ln() {
  static UniformLn u;
  return u;
}

// markov::Metropolis::accept (markov/metropolis.h:64-104).  `draw(row)`
// returns the row-th 32-bit word (row 0 is the first draw).
template <class Draw>
inline bool metropolis_accept(double cost_diff, double beta, Draw&& draw) {
  if (cost_diff <= 0) return true;
  uint32_t bucket = 0, r = 0;
  int row = -1;
  while (bucket == 0) {
    if (row < 3) {
      r = draw(row + 1);
      bucket = r & 255u;
      row++;
    } else {
      return true;
    }
  }
  const auto& L = uniform_ln();
  double y0 = L.t[row][bucket - 1];
  double y1 = L.t[row][bucket];
  double distance = (double)(r >> 8) / (double)(0xFFFFFEu);
  double value = y0 + distance * (y1 - y0);
  return value < (cost_diff * -beta);
}

// ---------------------------------------------------------------------------
// graph::normalize_edges (graph/graph.h:74-177) + populate_node_edge_ids
// (:517-534): internal ids by first appearance, ids sorted within a term,
// per-variable incident terms in term order, empty terms -> const cost,
// duplicate ids an error (Ising) or merged (PUBO, pubo.h:82).
struct Problem {
  int kind = ORC_ISING;
  int64_t n = 0;
  std::vector<std::vector<int32_t>> edges;   // internal ids, sorted
  std::vector<double> c;
  std::vector<float> cf;                     // fp32 coefficients (F32 arithmetic)
  std::vector<std::vector<int32_t>> node_edges;
  std::vector<int32_t> names;                // internal -> input id
  double const_cost = 0;
  double max_coupling = 0;
  int max_locality = 0;
  double scale_factor() const {              // graph_properties.h:24-40
    return (max_coupling <= 1) ? 1.0 : 1.0 / max_coupling;
  }
};

// State: spins as the reference's bools (1 = true); tstate = Ising term parity
// (IsingTermCachedState::terms) or PUBO zero counters (BinaryWithCounter::zeros).
struct State {
  std::vector<uint8_t> spins;
  std::vector<int32_t> tstate;
};

State empty_state(const Problem& p) {
  State s;
  s.spins.assign(p.n, 0);
  s.tstate.assign(p.edges.size(), 0);
  return s;
}

// IsingTermCached::apply_transition (model/ising.cc:36-44);
// PuboWithCounter::apply_transition (model/pubo.h:369-392).
inline void apply(const Problem& p, State& s, int64_t i) {
  if (p.kind == ORC_ISING) {
    s.spins[i] ^= 1;
    for (int32_t j : p.node_edges[i]) s.tstate[j] ^= 1;
  } else {
    if (s.spins[i]) {
      for (int32_t j : p.node_edges[i]) s.tstate[j]--;
      s.spins[i] = 0;
    } else {
      for (int32_t j : p.node_edges[i]) s.tstate[j]++;
      s.spins[i] = 1;
    }
  }
}

State state_from_spins(const Problem& p, const uint8_t* spins) {
  State s = empty_state(p);
  for (int64_t i = 0; i < p.n; ++i)
    if (spins[i]) apply(p, s, i);
  return s;
}

// AbstractIsing::calculate_cost (ising.h:94-102) with the cached term
// (ising.cc:46-55); PuboWithCounter::calculate_cost (pubo.h:303-333).
double cost(const Problem& p, const State& s) {
  double total = 0.0;
  for (size_t j = 0; j < p.edges.size(); ++j) {
    if (p.kind == ORC_ISING) {
      total += s.tstate[j] ? -p.c[j] : p.c[j];
    } else if (s.tstate[j] == 0) {
      total += p.c[j];
    }
  }
  return total;
}

// AbstractIsing::calculate_cost_difference (ising.h:108-117) and
// PuboWithCounter::calculate_cost_difference (pubo.h:335-367).  F32 mode sums
// the fp32 coefficients in the same (CSR) order in float, as the GPU does.
inline double delta(const Problem& p, const State& s, int64_t i, int arith) {
  if (arith == ORC_ARITH_F64) {
    double diff = 0.0;
    if (p.kind == ORC_ISING) {
      for (int32_t j : p.node_edges[i]) diff += s.tstate[j] ? -p.c[j] : p.c[j];
      return 2.0 * (-diff);
    }
    if (s.spins[i]) {
      for (int32_t j : p.node_edges[i])
        if (s.tstate[j] == 1) diff += p.c[j];
    } else {
      for (int32_t j : p.node_edges[i])
        if (s.tstate[j] == 0) diff -= p.c[j];
    }
    return diff;
  }
  float diff = 0.0f;
  if (p.kind == ORC_ISING) {
    for (int32_t j : p.node_edges[i]) diff += s.tstate[j] ? -p.cf[j] : p.cf[j];
    return (double)(2.0f * (-diff));
  }
  if (s.spins[i]) {
    for (int32_t j : p.node_edges[i])
      if (s.tstate[j] == 1) diff += p.cf[j];
  } else {
    for (int32_t j : p.node_edges[i])
      if (s.tstate[j] == 0) diff -= p.cf[j];
  }
  return (double)diff;
}

// ---------------------------------------------------------------------------
// Walker / Metropolis (markov/walker.h:35-315, metropolis.h:18-116).
struct alignas(64) Walker {
  State state;
  double cost_ = 0;
  double lowest_cost = 0;
  std::vector<uint8_t> lowest_spins;
  double beta = 1.0;
  Pcg32* rng = nullptr;      // reference mode
  int32_t replica = 0;       // counter mode
  uint64_t accepted = 0, evaluations = 0;

  // Metropolis::set_beta / set_temperature (metropolis.h:35-61)
  void set_beta(double b) { beta = (b <= 0) ? 0.0 : b; }
  void set_temperature(double t) {
    beta = (t <= 0) ? std::numeric_limits<double>::infinity() : 1.0 / t;
  }
  void save_lowest() {                       // walker.h:269-273
    lowest_cost = cost_;
    lowest_spins = state.spins;
  }
  void check_lowest() {                      // walker.h:277-280
    if (cost_ < lowest_cost) save_lowest();
  }
};

struct RunCtx {
  const Problem* p;
  const orc_opts* o;
  std::vector<int32_t> order;
};

// Walker::attempt_transition (walker.h:209-229) for one variable.
inline void attempt(const RunCtx& ctx, Walker& w, int64_t i, uint32_t sweep) {
  const Problem& p = *ctx.p;
  double d = delta(p, w.state, i, ctx.o->arith);
  w.evaluations++;
  bool acc;
  if (ctx.o->rng_mode == ORC_RNG_REFERENCE) {
    acc = metropolis_accept(d, w.beta, [&](int) { return (*w.rng)(); });
  } else {
    const uint32_t k0 = ctx.o->seed, k1 = ctx.o->stream, r = (uint32_t)w.replica;
    acc = metropolis_accept(d, w.beta, [&](int row) -> uint32_t {
      // replica pair r >> 1: words r & 1 = row-0 draws, 2 + (r & 1) = row-1
      // draws; rows 2 and 3 from the replica's own redraw counter
      if (row == 0) return philox_word(k0, k1, (uint32_t)i, sweep, r >> 1, DOM_SWEEP, r & 1);
      if (row == 1) return philox_word(k0, k1, (uint32_t)i, sweep, r >> 1, DOM_SWEEP, 2 + (r & 1));
      return philox_word(k0, k1, (uint32_t)i, sweep, r, DOM_REDRAW, row - 2);
    });
  }
  if (acc) {
    apply(p, w.state, i);
    w.accepted++;
    w.cost_ += d;
    w.check_lowest();
  }
}

// Walker::make_sweep (walker.h:144-153): ordered, or in the given order.
inline void make_sweep(const RunCtx& ctx, Walker& w, uint32_t sweep) {
  if (ctx.order.empty()) {
    for (int64_t i = 0; i < ctx.p->n; ++i) attempt(ctx, w, i, sweep);
  } else {
    for (int32_t i : ctx.order) attempt(ctx, w, i, sweep);
  }
}

// get_random_spin_state / get_random_binary_state (ising.h:23-41,
// pubo.h:16-34): one 32-bit word per 32 variables, flip i if bit i%32 set.
State random_state(const RunCtx& ctx, Pcg32* rng, int32_t replica, uint32_t epoch) {
  const Problem& p = *ctx.p;
  State s = empty_state(p);
  uint32_t v = 0;
  for (int64_t i = 0; i < p.n; ++i) {
    if (i % 32 == 0) {
      if (ctx.o->rng_mode == ORC_RNG_REFERENCE) {
        v = (*rng)();
      } else {
        uint32_t wi = (uint32_t)(i / 32);
        v = philox_word(ctx.o->seed, ctx.o->stream, wi >> 2, (uint32_t)replica, epoch, DOM_INIT,
                        (int)(wi & 3));
      }
    }
    if ((v >> (i % 32)) & 1u) apply(p, s, i);
  }
  return s;
}

// Walker::init (walker.h:50-58)
void walker_init(const RunCtx& ctx, Walker& w, uint32_t epoch) {
  w.state = random_state(ctx, w.rng, w.replica, epoch);
  w.cost_ = cost(*ctx.p, w.state);
  w.save_lowest();
}

// ---------------------------------------------------------------------------
// schedule::Schedule::get_value (schedule/schedule.cc:17-25) for the ranged
// generators Linear (linear.cc:15-18), Geometric (geometric.cc:9-12) and
// Constant.
struct Sched {
  int type = 0;  // 0 linear, 1 geometric, 2 constant
  double initial = 0, final_ = 0;
  double start = 0, stop = 1;
  double progress_value(double x) const {
    if (type == 0) return initial + x * (final_ - initial);
    if (type == 1) return initial * std::pow(final_ / initial, x);
    return initial;
  }
  double value(double input) const {
    if (start == stop) return progress_value(0);
    double progress = (input - start) / (stop - start);
    return progress_value(std::min(1., std::max(0., progress)));
  }
};

// SteppingSolver::run (solver/stepping_solver.h:112-268): steps are taken in
// ticks of `steps_per_tick` (default 10) and the step limit is checked only
// after each tick, so fewer than steps_per_tick steps cannot be run.
int64_t steps_to_run(int64_t step_limit, int64_t steps_per_tick) {
  if (steps_per_tick <= 0) steps_per_tick = 10;
  return std::max(step_limit, steps_per_tick);
}

void set_threads(const orc_opts* o) {
#ifdef _OPENMP
  if (o->threads > 0) omp_set_num_threads(o->threads);
#else
  (void)o;
#endif
}

RunCtx make_ctx(const Problem* p, const orc_opts* o) {
  RunCtx ctx{p, o, {}};
  if (o->n_order > 0) ctx.order.assign(o->order, o->order + o->n_order);
  return ctx;
}

void copy_spins(const std::vector<uint8_t>& v, uint8_t* out) {
  if (out) std::memcpy(out, v.data(), v.size());
}

// ---------------------------------------------------------------------------
// utils::RandomSelector (utils/random_selector.h:9-47, random_selector.cc):
// ids drawn with probability proportional to their weights, without
// replacement.  Binary indexed tree over insertion positions: node q >= 1
// holds the weight of positions [q - lowbit(q), q); the root is node P, the
// power of two the tree has grown to (nodes 0..P).  Growing copies the old
// root into the new one (random_selector.cc:52-65).
struct Selector {
  std::vector<size_t> ids;
  std::vector<double> node{0.0, 0.0};
  double total = 0.0;

  size_t root() const { return node.size() - 1; }
  // weight of positions below `position` (get_cumulative_weight, :22-35)
  double below(size_t position) const {
    size_t at = 0;
    double w = 0.0;
    for (size_t k = root(); k > 0; k >>= 1)
      if (position & k) {
        at |= k;
        w += node[at];
      }
    return w;
  }
  // add_weight (:37-50)
  void add(size_t position, double w) {
    for (size_t k = 1; k < node.size(); k <<= 1) {
      if (position & k) position ^= k;
      else node[position + k] += w;
    }
  }
  void insert(size_t id, double w) {
    const size_t position = ids.size();
    ids.push_back(id);
    total += w;
    if (position + 1 >= node.size()) {
      const double all = node.back();
      node.resize(2 * node.size() - 1);
      node.back() = all;
    }
    add(position, w);
  }
  // select_position (:67-94); a descent past the root (only when rounding
  // leaves the root below u * total) stops at the root, where the reference
  // would read past its array
  size_t select_position(double u) const {
    if (ids.size() == 1) return 0;
    double x = u * total;
    size_t position = 0;
    for (size_t k = root(); k > 0; k >>= 1)
      if (position + k <= root() && x >= node[position + k]) {
        position += k;
        x -= node[position];
      }
    return position >= ids.size() ? ids.size() - 1 : position;
  }
  // select_and_remove with remove_position (:96-120)
  size_t select_and_remove(double u) {
    const size_t position = select_position(u);
    const size_t id = ids[position];
    ids[position] = ids.back();
    ids.pop_back();
    const size_t last = ids.size();
    const double before = below(position + 1) - below(position);
    const double after = below(last + 1) - below(last);
    total -= before;
    add(position, after - before);
    add(last, -after);
    return id;
  }
};

// CullingEstimator (solver/population_annealing.h:175-212) over costs
// already shifted to >= 0: expected culled fraction at dbeta minus epsilon0.
double culling_estimate(const std::vector<double>& shifted, double epsilon0, size_t target, double scale,
                        double db) {
  const size_t R = shifted.size();
  std::vector<double> w(R);
  double q = 0.0;
  for (size_t i = 0; i < R; ++i) {
    w[i] = std::exp(-db * shifted[i] * scale);
    q += w[i];
  }
  const double rrq = (double)target / q;
  double culled = 0.0;
  for (size_t i = 0; i < R; ++i) {
    const double v = 1.0 - w[i] * rrq;
    if (v > 0) culled += v;
  }
  return culled / (double)R - epsilon0;
}

// find_delta_beta (population_annealing.h:218-260): widen [lower, 2^e) from
// e = -20 until the estimate changes sign (at most e = 19), then refine the
// lower end by halving steps down to 2^-20.
double find_delta_beta(const std::vector<double>& shifted, double epsilon0, size_t target, double scale) {
  double lower = 0.0, lower_value = culling_estimate(shifted, epsilon0, target, scale, 0.0);
  int e;
  for (e = -20; e < 20; ++e) {
    const double upper = std::ldexp(1.0, e);
    const double v = culling_estimate(shifted, epsilon0, target, scale, upper);
    if ((lower_value < 0) != (v < 0)) break;
    lower = upper;
    lower_value = v;
  }
  for (e -= 2; e >= -20; --e) {
    const double upper = lower + std::ldexp(1.0, e);
    const double v = culling_estimate(shifted, epsilon0, target, scale, upper);
    if ((lower_value < 0) == (v < 0)) {
      lower = upper;
      lower_value = v;
    }
  }
  return lower;
}

// The adaptive stepping sizes (population_annealing.h:347-373, 420-457)
// from the costs before and after a step's sweeps.  strategy 1 friction
// tensor, 2 energy variance, 3 constant culling; c is the strategy's
// constant.
double adaptive_delta_beta(int strategy, double c, const double* before, const double* after, size_t R,
                           size_t target, double scale, int sweeps) {
  const double Rd = (double)R;
  double avg_b = 0, avg_a = 0, mn = std::numeric_limits<double>::max();
  for (size_t i = 0; i < R; ++i) {
    avg_b += before[i];
    avg_a += after[i];
    mn = after[i] < mn ? after[i] : mn;
  }
  avg_b /= Rd;
  avg_a /= Rd;
  if (strategy == 3) {
    std::vector<double> shifted(after, after + R);
    for (double& x : shifted) x -= mn;
    return find_delta_beta(shifted, c, target, scale);
  }
  double var_b = 0, var_a = 0, gamma = 0;
  for (size_t i = 0; i < R; ++i) {
    const double db = before[i] - avg_b, da = after[i] - avg_a;
    var_b += db * db;
    var_a += da * da;
    gamma += db * da;
  }
  var_b /= Rd;
  var_a /= Rd;
  gamma /= Rd;
  const double sigma = std::sqrt(var_b * var_a);
  if (strategy == 1) {
    double zeta = 0;
    if (sigma > 0 && gamma != 0) {
      const double delta = std::log(sigma) - std::log(std::abs(gamma));
      if (std::abs(delta) > 1e-11) zeta = sigma * ((double)sweeps / delta);
    }
    return zeta > 0 ? c / std::sqrt(zeta) : 0.0;
  }
  return sigma > 0 ? c / std::sqrt(sigma) : 0.0;
}

// Offspring counts of the constant-population branch (population_annealing.h
// :540-578): floor(target w_i / sum w), then picks without replacement in
// proportion to the remainders until the counts sum to R; `uniform(j)` is
// pick j's uniform.
template <class U>
void residual_counts(const std::vector<double>& weight, double total, size_t target, std::vector<size_t>& count,
                     U uniform) {
  const size_t R = weight.size();
  count.assign(R, 0);
  std::vector<double> res(R);
  size_t n = 0;
  for (size_t i = 0; i < R; ++i) {
    const double rw = (double)target * weight[i] / total;
    count[i] = (size_t)std::floor(rw);
    res[i] = rw - (double)count[i];
    n += count[i];
  }
  Selector sel;
  for (size_t i = 0; i < R; ++i) sel.insert(i, res[i]);
  for (size_t j = 0; n < R; ++j, ++n) count[sel.select_and_remove(uniform(j))] += 1;
}

}  // namespace

// ===========================================================================
extern "C" {

void orc_pcg_uint32(int has_seed, uint32_t seed, int n, uint32_t* out) {
  Pcg32 r = has_seed ? Pcg32(seed) : Pcg32();
  for (int i = 0; i < n; ++i) out[i] = r();
}

void orc_pcg_uniform(int has_seed, uint32_t seed, int n, double* out) {
  Pcg32 r = has_seed ? Pcg32(seed) : Pcg32();
  for (int i = 0; i < n; ++i) out[i] = r.uniform();
}

void orc_pcg_fork_uniform(uint32_t seed, int n_forks, int n, double* out) {
  Pcg32 master(seed);
  std::vector<Pcg32> forks;
  for (int k = 0; k < n_forks; ++k) forks.push_back(master.fork());
  for (int t = 0; t < n; ++t) out[t] = master.uniform();
  for (int k = 0; k < n_forks; ++k)
    for (int t = 0; t < n; ++t) out[(k + 1) * n + t] = forks[k].uniform();
}

void orc_pcg_poisson(uint32_t seed, double mean, int n, uint32_t* out) {
  Pcg32 r(seed);
  for (int i = 0; i < n; ++i) out[i] = r.poisson(mean);
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox4x32_10(ctr, key, out);
}

void orc_uniform_ln(double* out) {
  const auto& L = uniform_ln();
  for (int r = 0; r < 4; ++r)
    for (int i = 0; i < 256; ++i) out[r * 256 + i] = L.t[r][i];
}

int orc_accept(double cost_diff, double beta, const uint32_t* draws, int* n_used) {
  int used = 0;
  bool a = metropolis_accept(cost_diff, beta, [&](int row) {
    used = row + 1;
    return draws[row];
  });
  if (n_used) *n_used = used;
  return a ? 1 : 0;
}

double orc_schedule_value(int type, double initial, double final_, double start, double stop,
                          double x) {
  Sched s;
  s.type = type;
  s.initial = initial;
  s.final_ = final_;
  s.start = start < 0 ? 0 : start;
  s.stop = stop < 0 ? 1 : stop;
  return s.value(x);
}

// Schedule::get_discretized_values (schedule.cc:38-89) with an explicit count.
int orc_schedule_discretized(int type, double initial, double final_, int count, double* out) {
  if (count < 2) return -1;
  Sched s;
  s.type = type;
  s.initial = initial;
  s.final_ = final_;
  for (int i = 0; i < count; ++i) out[i] = s.progress_value(double(i) / (count - 1));
  return 0;
}

void* orc_problem_new(int kind, int64_t n_terms, const double* c, const int64_t* offsets,
                      const int32_t* ids, char* err, int errlen) {
  auto fail = [&](const std::string& m) -> void* {
    if (err && errlen > 0) {
      std::strncpy(err, m.c_str(), errlen - 1);
      err[errlen - 1] = 0;
    }
    return nullptr;
  };
  if (n_terms == 0) return fail("parameter `terms`: size must be greater than 0, found 0");
  auto* p = new Problem();
  p->kind = kind;
  std::map<int32_t, int32_t> name_to_id;
  for (int64_t e = 0; e < n_terms; ++e) {
    std::vector<int32_t> members;
    std::set<int32_t> seen;
    for (int64_t k = offsets[e]; k < offsets[e + 1]; ++k) {
      int32_t name = ids[k];
      auto it = name_to_id.find(name);
      int32_t id;
      if (it == name_to_id.end()) {
        id = (int32_t)p->names.size();
        name_to_id[name] = id;
        p->names.push_back(name);
        p->node_edges.emplace_back();
      } else {
        id = it->second;
      }
      if (seen.insert(id).second) {
        members.push_back(id);
        p->node_edges[id].push_back((int32_t)p->edges.size());
      } else if (kind == ORC_ISING) {
        delete p;
        return fail("Duplicated ids detected in term!");
      }
    }
    std::sort(members.begin(), members.end());
    if (!members.empty()) {
      p->max_locality = std::max<int>(p->max_locality, (int)members.size());
      p->max_coupling = std::max(p->max_coupling, std::fabs(c[e]));
      p->edges.push_back(members);
      p->c.push_back(c[e]);
      p->cf.push_back((float)c[e]);
    } else {
      p->const_cost += c[e];
    }
  }
  p->n = (int64_t)p->names.size();
  return p;
}

void orc_problem_free(void* p) { delete static_cast<Problem*>(p); }

void orc_problem_dims(void* pp, int64_t* n_vars, int64_t* n_terms, int64_t* n_incidences,
                      double* const_cost, int* max_locality) {
  auto* p = static_cast<Problem*>(pp);
  int64_t inc = 0;
  for (auto& e : p->edges) inc += (int64_t)e.size();
  if (n_vars) *n_vars = p->n;
  if (n_terms) *n_terms = (int64_t)p->edges.size();
  if (n_incidences) *n_incidences = inc;
  if (const_cost) *const_cost = p->const_cost;
  if (max_locality) *max_locality = p->max_locality;
}

void orc_problem_names(void* pp, int32_t* names) {
  auto* p = static_cast<Problem*>(pp);
  std::copy(p->names.begin(), p->names.end(), names);
}

void orc_problem_edges(void* pp, int64_t* offsets, int32_t* ids, double* c) {
  auto* p = static_cast<Problem*>(pp);
  int64_t k = 0;
  for (size_t e = 0; e < p->edges.size(); ++e) {
    offsets[e] = k;
    for (int32_t v : p->edges[e]) ids[k++] = v;
    c[e] = p->c[e];
  }
  offsets[p->edges.size()] = k;
}

double orc_cost(void* pp, const uint8_t* spins) {
  auto* p = static_cast<Problem*>(pp);
  State s = state_from_spins(*p, spins);
  return cost(*p, s);
}

double orc_delta(void* pp, const uint8_t* spins, int64_t i, int arith) {
  auto* p = static_cast<Problem*>(pp);
  State s = state_from_spins(*p, spins);
  return delta(*p, s, i, arith);
}

int orc_check_classes(void* pp, const int32_t* order, int n_classes, const int32_t* class_offsets) {
  auto* p = static_cast<Problem*>(pp);
  std::vector<int> seen(p->n, 0);
  for (int k = 0; k < n_classes; ++k) {
    std::vector<int32_t> term_owner(p->edges.size(), -1);
    for (int32_t q = class_offsets[k]; q < class_offsets[k + 1]; ++q) {
      int32_t i = order[q];
      if (i < 0 || i >= p->n || seen[i]) return 0;
      seen[i] = 1;
      for (int32_t j : p->node_edges[i]) {
        if (term_owner[j] >= 0) return 0;
        term_owner[j] = i;
      }
    }
  }
  for (int64_t i = 0; i < p->n; ++i)
    if (!seen[i]) return 0;
  return 1;
}

void orc_initial_state(void* pp, const orc_opts* o, int32_t replica, uint32_t epoch,
                       uint8_t* spins) {
  auto* p = static_cast<Problem*>(pp);
  RunCtx ctx = make_ctx(p, o);
  State s;
  if (o->rng_mode == ORC_RNG_REFERENCE) {
    // SimulatedAnnealing::init: restart r uses fork number r of PCG(seed).
    Pcg32 master(o->seed);
    Pcg32 f;
    for (int32_t k = 0; k <= replica; ++k) f = master.fork();
    s = random_state(ctx, &f, replica, epoch);
  } else {
    s = random_state(ctx, nullptr, replica, epoch);
  }
  std::memcpy(spins, s.spins.data(), (size_t)p->n);
}

int orc_run_sweeps(void* pp, const orc_opts* o, int R, const int32_t* replica_ids,
                   int64_t sweep0, int n_sweeps, const double* betas, uint8_t* states,
                   double* costs, double* lowest, uint8_t* lowest_states, uint64_t* accepted) {
  auto* p = static_cast<Problem*>(pp);
  if (o->rng_mode != ORC_RNG_COUNTER) return -1;
  RunCtx ctx = make_ctx(p, o);
  set_threads(o);
  const int64_t N = p->n;
#pragma omp parallel for schedule(dynamic, 1)
  for (int r = 0; r < R; ++r) {
    Walker w;
    w.state = state_from_spins(*p, states + (size_t)r * N);
    w.cost_ = costs[r];
    w.lowest_cost = lowest[r];
    w.lowest_spins.assign(lowest_states + (size_t)r * N, lowest_states + (size_t)(r + 1) * N);
    w.replica = replica_ids[r];
    for (int k = 0; k < n_sweeps; ++k) {
      w.set_beta(betas[(size_t)k * R + r]);
      make_sweep(ctx, w, (uint32_t)(sweep0 + k));
    }
    std::memcpy(states + (size_t)r * N, w.state.spins.data(), N);
    costs[r] = w.cost_;
    lowest[r] = w.lowest_cost;
    std::memcpy(lowest_states + (size_t)r * N, w.lowest_spins.data(), N);
    if (accepted) accepted[r] = w.accepted;
  }
  return 0;
}

// SimulatedAnnealing init/make_step/configure (simulated_annealing.h:45-209).
int orc_sa(void* pp, const orc_opts* o, int sched_type, double initial, double final_,
           int use_inverse_temperature, int64_t step_limit, int restarts, double* lowest,
           uint8_t* lowest_states, double* final_costs, uint8_t* final_states,
           int64_t* steps_run) {
  auto* p = static_cast<Problem*>(pp);
  if (step_limit <= 1 || restarts <= 0) return -1;
  RunCtx ctx = make_ctx(p, o);
  set_threads(o);
  Sched sch;
  sch.type = sched_type;
  sch.initial = initial;
  sch.final_ = final_;
  sch.start = 0;
  sch.stop = (double)std::max<int64_t>(1, step_limit - 1);  // stretched (:175-202)
  const int64_t steps = steps_to_run(step_limit, o->steps_per_tick);
  Pcg32 master(o->seed);
  std::vector<Pcg32> rngs;
  for (int r = 0; r < restarts; ++r) rngs.push_back(master.fork());
  std::vector<Walker> ws(restarts);
#pragma omp parallel for schedule(dynamic, 1)
  for (int r = 0; r < restarts; ++r) {
    Walker& w = ws[r];
    w.rng = &rngs[r];
    w.replica = r;
    walker_init(ctx, w, 0);
    for (int64_t s = 0; s < steps; ++s) {
      double v = sch.value((double)s);
      if (use_inverse_temperature) w.set_beta(v); else w.set_temperature(v);
      make_sweep(ctx, w, (uint32_t)s);
    }
  }
  const int64_t N = p->n;
  for (int r = 0; r < restarts; ++r) {
    if (lowest) lowest[r] = ws[r].lowest_cost;
    if (lowest_states) copy_spins(ws[r].lowest_spins, lowest_states + (size_t)r * N);
    if (final_costs) final_costs[r] = ws[r].cost_;
    if (final_states) copy_spins(ws[r].state.spins, final_states + (size_t)r * N);
  }
  if (steps_run) *steps_run = steps;
  return 0;
}

// ParallelTempering init/make_step (parallel_tempering.h:68-189).
int orc_pt(void* pp, const orc_opts* o, const double* values, int n_values, int use_inverse,
           int64_t step_limit, int sweeps_per_replica, double* slot_beta, double* lowest,
           uint8_t* lowest_states, double* final_costs, uint8_t* final_states,
           double* swap_rate, int64_t* steps_run, double* acc_rate, double* avg_cost, double* avg_dir) {
  auto* p = static_cast<Problem*>(pp);
  if (n_values <= 0 || sweeps_per_replica <= 0) return -1;
  RunCtx ctx = make_ctx(p, o);
  set_threads(o);
  std::vector<double> t(values, values + n_values);
  if (use_inverse) std::sort(t.begin(), t.end(), std::greater<double>());
  else std::sort(t.begin(), t.end(), std::less<double>());
  const int R = n_values;
  const int64_t steps = steps_to_run(step_limit, o->steps_per_tick);
  Pcg32 master(o->seed);
  std::vector<Pcg32> rngs;
  for (int r = 0; r < R; ++r) rngs.push_back(master.fork());
  std::vector<Walker> ws(R);
  for (int r = 0; r < R; ++r) {
    ws[r].rng = &rngs[r];
    ws[r].replica = r;
    if (use_inverse) ws[r].set_beta(t[r]); else ws[r].set_temperature(t[r]);
    walker_init(ctx, ws[r], 0);
  }
  // Counter mode: slot r keeps its state; label[k] = slot holding beta index k.
  std::vector<int> label(R);
  std::vector<double> beta_of_index(R);
  for (int k = 0; k < R; ++k) {
    label[k] = k;
    beta_of_index[k] = ws[k].beta;
  }
  std::vector<double> swaps(R, 0.0), tries(R, 0.0);
  // observations per temperature position (parallel_tempering.h:141-188):
  // acceptance rate of the step, cost after the sweeps, direction label
  std::vector<double> o_acc(R, 0.0), o_cost(R, 0.0), o_dir(R, 0.0);
  std::vector<int> dir(R, 0);  // UP 0, DOWN 1, per position
  std::vector<uint64_t> acc0(R, 0);
  const double evals = (double)p->n * sweeps_per_replica;
  for (int64_t s = 0; s < steps; ++s) {
    for (int r = 0; r < R; ++r) acc0[r] = ws[r].accepted;
#pragma omp parallel for schedule(dynamic, 1)
    for (int r = 0; r < R; ++r)
      for (int k = 0; k < sweeps_per_replica; ++k)
        make_sweep(ctx, ws[r], (uint32_t)(s * sweeps_per_replica + k));
    for (int i = 0; i < R; ++i) {
      // the walker at position i: itself in reference mode (states move),
      // the slot holding beta index i in counter mode (labels move)
      const int r = o->rng_mode == ORC_RNG_REFERENCE ? i : label[i];
      o_acc[i] += (double)(ws[r].accepted - acc0[r]) / evals;
      o_cost[i] += ws[r].cost_;
      o_dir[i] += (double)dir[i];
    }
    for (int i = (int)(s & 1); i < R - 1; i += 2) {
      tries[i] += 1;
      if (o->rng_mode == ORC_RNG_REFERENCE) {
        Walker& a = ws[i];
        Walker& b = ws[i + 1];
        if (master.uniform() < std::exp((a.cost_ - b.cost_) * (a.beta - b.beta))) {
          std::swap(a.state, b.state);  // Walker::swap_state (walker.h:256-260)
          std::swap(a.cost_, b.cost_);
          std::swap(dir[i], dir[i + 1]);
          swaps[i] += 1;
        }
      } else {
        Walker& a = ws[label[i]];
        Walker& b = ws[label[i + 1]];
        uint32_t ctr[4] = {(uint32_t)i, (uint32_t)s, 0u, DOM_EXCHANGE}, key[2] = {o->seed, o->stream},
                 w[4];
        philox4x32_10(ctr, key, w);
        double u = canonical2(w[0], w[1]);
        if (u < std::exp((a.cost_ - b.cost_) * (beta_of_index[i] - beta_of_index[i + 1]))) {
          std::swap(label[i], label[i + 1]);
          a.beta = beta_of_index[i + 1];
          b.beta = beta_of_index[i];
          std::swap(dir[i], dir[i + 1]);
          swaps[i] += 1;
        }
      }
    }
    dir[0] = 0;
    dir[R - 1] = 1;
  }
  const int64_t N = p->n;
  for (int r = 0; r < R; ++r) {
    if (slot_beta) slot_beta[r] = ws[r].beta;
    if (lowest) lowest[r] = ws[r].lowest_cost;
    if (lowest_states) copy_spins(ws[r].lowest_spins, lowest_states + (size_t)r * N);
    if (final_costs) final_costs[r] = ws[r].cost_;
    if (final_states) copy_spins(ws[r].state.spins, final_states + (size_t)r * N);
    if (swap_rate) swap_rate[r] = tries[r] > 0 ? swaps[r] / tries[r] : 0.0;
    if (acc_rate) acc_rate[r] = o_acc[r] / (double)steps;
    if (avg_cost) avg_cost[r] = o_cost[r] / (double)steps;
    if (avg_dir) avg_dir[r] = o_dir[r] / (double)steps;
  }
  if (steps_run) *steps_run = steps;
  return 0;
}

// PopulationAnnealing (population_annealing.h:102-616) with the linear or
// geometric beta schedule or an adaptive stepping size (strategy 1 friction
// tensor, 2 energy variance, 3 constant culling, with its constant), and
// Poisson or constant-population resampling (Population::resample,
// population.h:171-218).  Citizens use per-thread RNGs under
// `omp parallel for schedule(static,1)`: citizen i draws from thread i % T,
// threads visiting their citizens in increasing order, which this sequential
// loop reproduces exactly for T = opts.threads (default 1).
// ---------------------------------------------------------------------------
// The reference's test-only toy model (solver/test/test_model.cc:3-38, its
// default constants): the state is a double v, cost |v - 42|, a random state
// uniform() * 100, a random transition v + uniform() * 10 - 5; a sweep is 10
// random transitions (markov/walker.h:168-188, get_sweep_size() = 10), each
// attempted and, if accepted, applied with cost += cost difference and the
// strict-< lowest check (walker.h:209-229, 277-280).  Used only to check the
// oracle's SA and PT loops against the reference's toy-model golden values.
struct ToyWalker {
  Pcg32 rng;
  double v = 0, cost = 0, lowest = 0, beta = 0, T = 0;
  uint64_t evals = 0, accepted = 0;
  void init() {
    v = rng.uniform() * 100.0;
    cost = std::fabs(v - 42.0);
    lowest = cost;
  }
  void set_temperature(double t) {  // metropolis.h:35-61
    T = t > 0 ? t : 0.0;
    beta = t > 0 ? 1.0 / t : std::numeric_limits<double>::infinity();
  }
  void set_beta(double b) {
    beta = b > 0 ? b : 0.0;
    T = b > 0 ? 1.0 / b : std::numeric_limits<double>::infinity();
  }
  void sweep() {
    for (int t = 0; t < 10; ++t) {
      const double nv = v + rng.uniform() * 10.0 - 5.0;
      const double d = std::fabs(nv - 42.0) - std::fabs(v - 42.0);
      ++evals;
      if (metropolis_accept(d, beta, [&](int) { return rng(); })) {
        v = nv;
        ++accepted;
        cost += d;
        if (cost < lowest) lowest = cost;
      }
    }
  }
};

// Milestone (observe/observable.cc:279-305) of the lowest cost after every
// step: (steps so far, lowest) whenever the value changes.
struct Milestones {
  std::vector<double> steps, values;
  double total = 0;
  void record(double value) {
    total += 1.0;
    if (values.empty() || value != values.back()) {
      steps.push_back(total);
      values.push_back(value);
    }
  }
  int write(double* st, double* val, int cap) const {
    const int n = (int)values.size();
    for (int i = 0; i < n && i < cap; ++i) {
      st[i] = steps[i];
      val[i] = values[i];
    }
    return n;
  }
};

// utils::RandomSelector: n weights inserted in order, then k picks without
// replacement with the given uniforms; out[j] = pick j's id (k <= n).
int orc_selector_draw(int n, const double* weights, int k, const double* uniforms, int64_t* out) {
  if (n < 1 || k < 0 || k > n) return -1;
  Selector sel;
  for (int i = 0; i < n; ++i) sel.insert((size_t)i, weights[i]);
  for (int j = 0; j < k; ++j) out[j] = (int64_t)sel.select_and_remove(uniforms[j]);
  return 0;
}

// SimulatedAnnealing on the toy model (simulated_annealing.h:45-110 with
// SteppingSolver::run, stepping_solver.h:112-160): restarts forked from
// PCG(seed) in order, β linear from beta_start to beta_stop stretched to
// step_limit - 1 steps, one sweep per step, the global lowest updated every
// step (a cost_limit is set) and its milestones recorded.  Returns the
// number of milestones.
int orc_toy_sa(uint32_t seed, int64_t step_limit, int restarts, double beta_start, double beta_stop,
               double* ms_steps, double* ms_values, int cap) {
  if (restarts < 1 || step_limit < 2) return -1;
  Pcg32 master(seed);
  std::vector<ToyWalker> w(restarts);
  for (auto& r : w) r.rng = master.fork();
  for (auto& r : w) r.init();
  double lowest = w[0].cost;
  Sched beta;
  beta.type = 0;
  beta.initial = beta_start;
  beta.final_ = beta_stop;
  beta.start = 0;
  beta.stop = (double)std::max<int64_t>(1, step_limit - 1);
  Milestones ms;
  for (int64_t step = 0; step < step_limit; ++step) {
    const double b = beta.value((double)step);
    for (auto& r : w) {
      r.set_beta(b);
      r.sweep();
    }
    for (auto& r : w)
      if (r.lowest < lowest) lowest = r.lowest;
    ms.record(lowest);
  }
  return ms.write(ms_steps, ms_values, cap);
}

// ParallelTempering on the toy model (parallel_tempering.h:68-189): n
// temperatures sorted ascending, replica i forked i-th from PCG(seed), each
// step sweeps_per_replica sweeps per replica, then per replica the observed
// acceptance rate of the step, its cost and its direction label (UP 0 for the
// replica at the lowest temperature, DOWN 1 at the highest, carried along by
// swaps), then swaps of pairs (i, i+1), i = step mod 2, with probability
// min(1, exp((E_i - E_i+1)(β_i - β_i+1))) from the master's uniform.  Outputs
// the step averages the reference's observer reports per replica and the
// lowest-cost milestones; returns the number of milestones.
int orc_toy_pt(uint32_t seed, int64_t step_limit, int n, const double* temps, int sweeps_per_replica,
               double* T_out, double* acc_rate, double* avg_cost, double* avg_dir, double* swap_rate,
               double* ms_steps, double* ms_values, int cap) {
  if (n < 2 || step_limit < 1) return -1;
  std::vector<double> t(temps, temps + n);
  std::sort(t.begin(), t.end());
  Pcg32 master(seed);
  std::vector<ToyWalker> w(n);
  for (auto& r : w) r.rng = master.fork();
  std::vector<int> dir(n, 0);
  for (int i = 0; i < n; ++i) {
    w[i].set_temperature(t[i]);
    w[i].init();
  }
  double lowest = w[0].cost;
  std::vector<double> acc(n, 0), cst(n, 0), dsum(n, 0), sw(n, 0), swn(n, 0);
  Milestones ms;
  for (int64_t step = 0; step < step_limit; ++step) {
    for (auto& r : w)
      for (int k = 0; k < sweeps_per_replica; ++k) r.sweep();
    for (int i = 0; i < n; ++i) {
      acc[i] += (double)w[i].accepted / (double)w[i].evals;
      cst[i] += w[i].cost;
      if (w[i].lowest < lowest) lowest = w[i].lowest;
      w[i].accepted = w[i].evals = 0;
      dsum[i] += (double)dir[i];
    }
    for (int i = (int)(step & 1); i < n - 1; i += 2) {
      ToyWalker &a = w[i], &b = w[i + 1];
      const bool swap = master.uniform() < std::exp((a.cost - b.cost) * (a.beta - b.beta));
      if (swap) {
        std::swap(a.v, b.v);
        std::swap(a.cost, b.cost);
        std::swap(dir[i], dir[i + 1]);
      }
      sw[i] += swap ? 1.0 : 0.0;
      swn[i] += 1.0;
    }
    dir[0] = 0;
    dir[n - 1] = 1;
    ms.record(lowest);
  }
  for (int i = 0; i < n; ++i) {
    T_out[i] = w[i].T;
    acc_rate[i] = acc[i] / (double)step_limit;
    avg_cost[i] = cst[i] / (double)step_limit;
    avg_dir[i] = dsum[i] / (double)step_limit;
    swap_rate[i] = swn[i] > 0 ? sw[i] / swn[i] : -1.0;
  }
  return ms.write(ms_steps, ms_values, cap);
}

// The adaptive stepping size of one step (strategy 1 friction tensor, 2
// energy variance, 3 constant culling) from the costs before and after its
// sweeps.
double orc_adaptive_delta_beta(int strategy, double c, int R, const double* before, const double* after,
                               int target, double scale, int sweeps) {
  return adaptive_delta_beta(strategy, c, before, after, (size_t)R, (size_t)target, scale, sweeps);
}

// The culling estimate over costs already shifted to >= 0.
double orc_culling_estimate(int R, const double* shifted, double epsilon0, int target, double scale,
                            double db) {
  return culling_estimate(std::vector<double>(shifted, shifted + R), epsilon0, (size_t)target, scale, db);
}

void orc_poisson_counter(double lam, uint32_t seed, uint32_t step, int n, uint32_t* out) {
  for (int i = 0; i < n; ++i) out[i] = poisson_counter(lam, (uint32_t)i, step, seed);
}

// One population annealing step in counter mode, after the step's sweeps:
// the decisions of orc_pa's step loop below (population_annealing.h:262-488,
// resample :490-616) for the costs and lowest costs the sweeps left (and,
// for strategies 1 and 2, the costs before them), with offspring counts from
// poisson_counter(citizen, step) or, with constant_population, residual picks
// whose uniforms come from Philox (j, step, 1, DOM_PICK), and the pick of an
// emptied population from Philox (0, step, 0, DOM_PICK).
// Offspring of citizen i fill consecutive slots in citizen order (copies of
// one citizen are identical, so this is orc_pa's order); after a restart a
// population below the target is topped up with copies of its last citizen
// (fresh states follow for everyone, as init_population draws them).
// Outputs: src[0..R_final) = the citizen each new slot copies, fam_out =
// families after the step; iout = {R_final, R_resampled, moved, restart,
// best_src (-1: no new global lowest), target after the step}; dout = {rho,
// dbeta, new best}.  exp_repopulation: a restart doubles the target first.
int orc_pa_step_counter(int R, const double* cost, const double* lowest, const int32_t* family,
                        int64_t step, int64_t base_step, const double* beta_table, int n_table, int target,
                        int cap, double alpha, double scale, uint32_t seed, int have_best, double best,
                        int32_t* src, int32_t* fam_out, int32_t* iout, double* dout, const double* cost_before,
                        int strategy, double strategy_const, int constant_population, int sweeps_per_step,
                        int exp_repopulation) {
  if (R < 1 || n_table < 2 || cap < target || strategy < 0 || strategy > 3) return -1;
  if ((strategy == 1 || strategy == 2) && !cost_before) return -1;
  // update_lowest_cost over the citizens (first minimum)
  int barg = 0;
  for (int i = 1; i < R; ++i)
    if (lowest[i] < lowest[barg]) barg = i;
  const bool improved = !have_best || lowest[barg] < best;
  double avg = 0, mn = std::numeric_limits<double>::max();
  for (int i = 0; i < R; ++i) {
    avg += cost[i];
    mn = cost[i] < mn ? cost[i] : mn;
  }
  avg /= (double)R;
  double var = 0;
  for (int i = 0; i < R; ++i) var += (cost[i] - avg) * (cost[i] - avg);
  var /= (double)R;
  const int64_t k = step - base_step, last = n_table - 1;
  const double db = strategy == 0 ? beta_table[std::min(k + 1, last)] - beta_table[std::min(k, last)]
                                  : adaptive_delta_beta(strategy, strategy_const, cost_before, cost, (size_t)R,
                                                        (size_t)target, scale, sweeps_per_step);
  double rho = 1.0;
  int Rn = R;
  bool moved = false;
  std::vector<int32_t> fam(family, family + R);
  if (db > 0) {
    std::vector<double> w(R);
    double total = 0;
    for (int i = 0; i < R; ++i) {
      w[i] = std::exp(-db * (cost[i] - mn) * scale);
      total += w[i];
    }
    std::vector<size_t> count(R);
    if (constant_population) {
      residual_counts(w, total, (size_t)target, count, [&](size_t j) {
        uint32_t ctr[4] = {(uint32_t)j, (uint32_t)step, 1u, DOM_PICK}, key[2] = {seed, 0u}, wd[4];
        philox4x32_10(ctr, key, wd);
        return canonical2(wd[0], wd[1]);
      });
    } else {
      const double rrq = (double)target / total;
      for (int i = 0; i < R; ++i) {
        const double mean = rrq * w[i];
        count[i] = mean > 0.0 ? poisson_counter(mean, (uint32_t)i, (uint32_t)step, seed) : 0;
      }
    }
    std::vector<int32_t> s2;
    std::vector<int32_t> f2;
    for (int i = 0; i < R; ++i) {
      for (size_t j = 0; j < count[i]; ++j) {
        s2.push_back(i);
        f2.push_back(family[i]);
      }
    }
    if (s2.empty()) {
      uint32_t ctr[4] = {0u, (uint32_t)step, 0u, DOM_PICK}, key[2] = {seed, 0u}, wd[4];
      philox4x32_10(ctr, key, wd);
      const int i = (int)std::floor(canonical2(wd[0], wd[1]) * (double)R);
      s2.push_back(i);
      f2.push_back(family[i]);
    }
    if ((int)s2.size() > cap) return -2;  // the device clamps and flags; not restated
    std::map<int32_t, int64_t> census;
    for (int32_t f : f2) census[f]++;
    double sq = 0;
    for (auto& kv : census) sq += (double)kv.second * (double)kv.second;
    rho = sq / (double)R;
    Rn = (int)s2.size();
    for (int j = 0; j < Rn; ++j) src[j] = s2[j];
    fam = f2;
    moved = true;
  } else {
    for (int j = 0; j < R; ++j) src[j] = j;
  }
  const bool restart = (double)R / rho < alpha || var == 0.0;
  int Rf = Rn;
  int tnew = target;
  if (restart && exp_repopulation) {  // population_annealing.h:463-470
    tnew = 2 * target;
    if (tnew > cap) return -2;  // the device clamps and flags; not restated
  }
  if (restart) {
    if (Rn < tnew) {
      for (int j = Rn; j < tnew; ++j) src[j] = src[Rn - 1];
      fam.resize(tnew);
      Rf = tnew;
    }
    for (int j = 0; j < Rf && j < tnew; ++j) fam[j] = j;
    moved = true;
  }
  for (int j = 0; j < Rf; ++j) fam_out[j] = fam[j];
  iout[0] = Rf;
  iout[1] = Rn;
  iout[2] = moved ? 1 : 0;
  iout[3] = restart ? 1 : 0;
  iout[4] = improved ? barg : -1;
  iout[5] = tnew;
  dout[0] = rho;
  dout[1] = db;
  dout[2] = improved ? lowest[barg] : best;
  return 0;
}

int orc_pa(void* pp, const orc_opts* o, int geometric, double beta_start, double beta_stop,
           int64_t step_limit, int population, int sweeps_per_replica, double alpha,
           int constant_population, double* best_cost, uint8_t* best_state,
           int64_t* final_population, int64_t* restarts_out, int strategy, double strategy_const,
           int exp_repopulation) {
  auto* p = static_cast<Problem*>(pp);
  if (o->rng_mode != ORC_RNG_REFERENCE || population <= 0 || step_limit <= 1) return -1;
  if (strategy < 0 || strategy > 3 || (strategy > 0 && !(strategy_const > 0))) return -1;
  if (strategy > 0) {  // adaptive: beta starts at 0 (:733-735)
    beta_start = beta_stop = 0.0;
    geometric = 0;
  }
  RunCtx ctx = make_ctx(p, o);
  const int T = o->threads > 0 ? o->threads : 1;
  Sched beta;
  beta.type = geometric ? 1 : 0;
  beta.initial = beta_start;
  beta.final_ = beta_stop;
  beta.start = 0;
  beta.stop = (double)std::max<int64_t>(1, step_limit - 1);
  const double b0 = beta.value(0);
  const int64_t steps = steps_to_run(step_limit, o->steps_per_tick);
  const double scale = p->scale_factor();
  Pcg32 master(o->seed);
  std::vector<Pcg32> rngs;
  for (int t = 0; t < T; ++t) rngs.push_back(master.fork());

  struct Citizen {
    Walker w;
    size_t family;
  };
  std::vector<Citizen> pop;
  size_t target = (size_t)population;
  double cur_beta = b0;
  int64_t base_step = 0, n_restarts = 0;
  bool have_best = false;
  double best = 0;
  std::vector<uint8_t> best_spins;
  auto update_best = [&](double c, const std::vector<uint8_t>& s) {
    if (!have_best || c < best) {
      have_best = true;
      best = c;
      best_spins = s;
    }
  };
  // init_population (:141-167)
  auto init_population = [&]() {
    for (size_t i = 0; i < target; ++i) {
      if (pop.size() <= i) {
        Citizen c;
        c.w.set_beta(cur_beta);
        pop.push_back(c);
      }
      pop[i].family = i;
    }
    for (size_t i = 0; i < pop.size(); ++i) {
      pop[i].w.rng = &rngs[i % T];
      walker_init(ctx, pop[i].w, 0);
    }
  };
  init_population();
  update_best(pop[0].w.cost_, pop[0].w.state.spins);

  std::vector<double> before, after;
  for (int64_t step = 0; step < steps; ++step) {
    const size_t R = pop.size();
    const double Rd = (double)R;
    before.resize(R);
    after.resize(R);
    for (size_t i = 0; i < R; ++i) {
      before[i] = pop[i].w.cost_;
      pop[i].w.rng = &rngs[i % T];
      for (int k = 0; k < sweeps_per_replica; ++k) make_sweep(ctx, pop[i].w, 0);
      after[i] = pop[i].w.cost_;
    }
    for (size_t i = 0; i < R; ++i) update_best(pop[i].w.lowest_cost, pop[i].w.lowest_spins);
    double avg_after = 0, min_cost = std::numeric_limits<double>::max();
    for (size_t i = 0; i < R; ++i) {
      avg_after += after[i];
      min_cost = after[i] < min_cost ? after[i] : min_cost;
    }
    avg_after /= Rd;
    double var_after = 0;
    for (size_t i = 0; i < R; ++i) {
      double d = after[i] - avg_after;
      var_after += d * d;
    }
    var_after /= Rd;
    const double delta_beta =
        strategy == 0
            ? beta.value((double)(step + 1 - base_step)) - beta.value((double)(step - base_step))
            : adaptive_delta_beta(strategy, strategy_const, before.data(), after.data(), R, target, scale,
                                  sweeps_per_replica);
    for (size_t i = 0; i < R; ++i) after[i] -= min_cost;
    double rho_t = 1.0;
    if (delta_beta > 0) {
      // resample (:490-616)
      std::vector<double> weight(R);
      double total = 0;
      for (size_t i = 0; i < R; ++i) {
        weight[i] = std::exp(-delta_beta * after[i] * scale);
        total += weight[i];
      }
      std::vector<size_t> count(R, 1);
      uint64_t current = R;
      if (constant_population) {
        residual_counts(weight, total, target, count, [&](size_t) { return master.uniform(); });
        current = 0;
        for (size_t c : count) current += c;
      } else {
        const double RRQ = (double)target / total;
        for (size_t i = 0; i < R; ++i) {
          size_t nc = (RRQ * weight[i]) > 0.0 ? (size_t)std::floor((double)master.poisson(RRQ * weight[i])) : 0;
          if (nc == 0) {
            current -= count[i];
            count[i] = 0;
          } else if (nc > 1) {
            current += nc - 1;
            count[i] += nc - 1;
          }
        }
        if (current == 0) {
          size_t i = (size_t)std::floor(master.uniform() * (double)R);
          count[i] += 1;
        }
      }
      // Population::resample: for each surviving citizen, (count-1) copies
      // then the original.
      std::vector<Citizen> next;
      next.reserve(current + 1);
      for (size_t i = 0; i < R; ++i) {
        if (count[i] == 0) continue;
        for (size_t k = 1; k < count[i]; ++k) next.push_back(pop[i]);
        next.push_back(pop[i]);
      }
      pop.swap(next);
      std::map<size_t, size_t> fam;
      for (auto& c : pop) fam[c.family]++;
      rho_t = 0;
      for (auto& kv : fam) rho_t += (double)kv.second * (double)kv.second;
      rho_t /= Rd;
    }
    if (Rd / rho_t < alpha || var_after == 0) {
      cur_beta = b0;
      base_step = step;
      if (exp_repopulation) target *= 2;  // population_annealing.h:463-470
      init_population();
      n_restarts++;
    } else {
      cur_beta += delta_beta;
      for (auto& c : pop) c.w.set_beta(cur_beta);
    }
  }
  // finalize -> populate_solutions(population): best over citizens' lowest
  for (auto& c : pop) update_best(c.w.lowest_cost, c.w.lowest_spins);
  if (best_cost) *best_cost = best;
  if (best_state) copy_spins(best_spins, best_state);
  if (final_population) *final_population = (int64_t)pop.size();
  if (restarts_out) *restarts_out = n_restarts;
  return 0;
}

}  // extern "C"
