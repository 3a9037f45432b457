// Internal structures shared by the HIP kernels and the C-ABI implementation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/qmc.h"

namespace qmc {

// Device copy of a problem in position space (see qmc_kernels.hip).
struct ProblemDev {
  int32_t N;
  int32_t n_classes;
  int32_t lanes;
  int32_t kind;
  int32_t pairwise;
  int32_t coupling;
  int64_t M;
  const int32_t* class_off;   // n_classes + 1
  const int32_t* var_of_pos;  // N
  const int32_t* pos_of_var;  // N
  const int32_t* inc_off;     // N + 1, incident entries of each position (term order)
  const void* coef;           // int32 or float per entry
  const int32_t* nbr;         // pairwise: the other member's position (N = ghost)
  const int32_t* oth_off;     // entries + 1
  const int32_t* oth_pos;     // other members' positions
  const int64_t* term_off;    // M + 1 (variable ids, for FP64 energies)
  const int32_t* term_var;
  const double* term_c;
};

// One work-group's share: up to 32 replicas of one problem.
struct UnitDev {
  // problem: sliced ELL of the pairwise form for this unit's lane count --
  // slice (class c, round k) = ell_slice_base[c] + k; ell_slice[slice] =
  // {entry offset, width, first position (class start + k*lanes), lane
  // split (log2 lanes per position)}; the table runs through the classes in
  // order and ends with copies of its first two slices (round k past the
  // last class's end is round k of the next sweep's first class); entry e
  // of lane t is ell_ent[offset + e*lanes + t] = {neighbour position,
  // coefficient bits (int32 or fp32)}, padded with the ghost position and a
  // zero coefficient; a slice stores max(width, 8) entries per lane
  const int32_t* ell_slice_base;
  const int4* ell_slice;
  const uint2* ell_ent;
  // general problems whose terms have at most 4 variables: the same slices
  // over 16-byte entries {coefficient bits, the other members' positions,
  // padded with the ghost position N} (null: incidence CSR below)
  const uint4* ell4_ent;
  // pairwise problems with int8 coefficients (kernels with PAIR 2): the same
  // slices over 4-byte entries {neighbour position | coefficient << 24}
  const uint32_t* ell_ent32;
  const int32_t* class_off;
  const int32_t* var_of_pos;
  const int32_t* pos_of_var;
  const int32_t* inc_off;
  const void* coef;
  const int32_t* nbr;
  const int32_t* oth_off;
  const int32_t* oth_pos;
  int32_t N, n_classes, lanes, G;
  int32_t lanes_log2;
  int32_t stride;   // bytes of one unit's spin array, multiple of 4
  int32_t ladder;   // > 0: PT ladder of this many replicas (exchanges in the kernel)
  int32_t slot0;    // ladder slot of the unit's first replica
  int32_t upl;      // work-groups per ladder (> 1: cooperative, per-step barrier)
  int32_t* sync;    // per-ladder barrier counter (upl > 1)
  double* cbuf;     // per-ladder [3][ladder] published energies, NaN = not yet (upl > 1)
  int32_t* err;     // set when a barrier wait times out
  // state (unit-local pointers: element 0 is the unit's first replica)
  uint8_t* spins;
  uint8_t* best;
  double* cost;
  double* best_cost;
  double* beta;
  unsigned long long* accepted;
  int32_t* label;   // the ladder's labels: slot holding beta index k
  const double* ladder_beta;
  const double* sched;   // beta of absolute sweep s = sched[s - sched_base]; null: per-replica betas
  int64_t sched_base;
  int64_t* progress;     // [0] next sweep index, [1] sweeps left (updated by the kernel)
  int32_t spr;           // PT: sweeps per exchange step
  int32_t rbase;         // index of the unit's first replica in its walker set
  uint32_t r0;      // replica id of the unit's first replica (RNG counter)
  uint32_t nrep;
  uint32_t seed, stream;
  // population on the device: replicas rbase + b >= *live are idle (null: all
  // nrep replicas are live)
  const int32_t* live;
};

// Replicas of the unit that take part: nrep, or fewer when the population
// size lives on the device.
__host__ __device__ inline int unit_nrep(const UnitDev& U) {
  int n = (int)U.nrep;
  if (U.live) {
    const int left = *U.live - U.rbase;
    n = left < 0 ? 0 : (left < n ? left : n);
  }
  return n;
}

// Population annealing state kept on the device (one per walker set); the
// step kernel reads and updates it, the host reads it at the end.
struct PaDev {
  int32_t R;          // live population: replicas 0..R-1 of the walker set
  int32_t R_prev;     // population the last step started from
  int32_t R_resampled;// population after its resampling, before a restart's top-up
  int32_t moved;      // the last step copied citizens (resampling or restart)
  int32_t restart;    // the last step restarted the population
  int32_t err;        // the population outgrew the capacity (clamped)
  int32_t restarts;   // restarts so far
  int32_t best_src;   // replica holding a new global lowest after the last step, -1 none
  int32_t have_best;
  uint32_t epoch;     // initial-state epoch of the last restart
  int64_t step;       // next step
  int64_t base_step;  // step of the last restart
  int64_t attempts;   // spin-flip attempts of the steps so far
  double cur_beta;
  double best;        // lowest cost any citizen has reached
  double rho;         // family statistic of the last resampling
  double delta_beta;  // stepping size of the last step
  int32_t target;     // target population (doubles at restarts with exp_repopulation)
  int32_t pad;
  // cost + difference evaluations as make_step accumulates its citizens'
  // counters (population_annealing.h:300-308): the step's sweeps plus the
  // (re)initialisations made before them
  int64_t evals;
  int64_t pending_inits;
};

// Residual resampling keeps its selection tree in LDS up to this many
// citizens (a power of two), in PaArgs::tree above it.
constexpr int kPaTreeLds = 4096;
// Per-step record of a population annealing step (the observations of
// population_annealing.h:262-470): epoch before the step, mean cost after the
// sweeps, population, sigma, tau, zeta (NaN when not defined), the stepping
// size, family count, sum of family sizes, size of the family with the
// largest id, rho, restart flag, resampled flag.
constexpr int kPaLogFields = QMC_PA_LOG_FIELDS;

struct PaArgs {
  PaDev* pa;
  const double* beta_table;  // beta schedule value at k = step - base_step, k = 0..n_table-1
  int32_t n_table;
  int32_t target, cap, n_vars;
  int32_t sweeps_per_step, pad;
  double alpha, scale, beta_start;
  uint32_t seed;
  int32_t strategy;           // QMC_PA_SCHEDULE ... QMC_PA_CONSTANT_CULLING
  int32_t constant_population;// residual resampling instead of Poisson offspring
  int32_t exp_repopulation;   // a restart doubles the target population
  double strategy_const;      // friction constant, culling fraction (initial or constant)
  double* cost0;              // costs before the step's sweeps (friction tensor, energy variance)
  double* tree;               // selection tree above kPaTreeLds citizens (pow2(cap) + 1 entries)
  int32_t* ids;               // its ids (cap entries)
  double* log;                // per-step records (kPaLogFields doubles each), or null
  int32_t log_steps;          // records the log holds
  double* cost;
  double* best_cost;
  double* beta;
  unsigned long long* accepted;
  int32_t* family;
  int32_t* family_next;
  int32_t* src;
  int32_t* hist;             // cap entries (families below the target)
  double* weight;            // cap entries
  int32_t* count;            // cap entries
  double* tcost;             // gathered copies (cap entries each)
  double* tbest;
  double* tbeta;
  unsigned long long* tacc;
  uint8_t* plane_cur;        // plain states, cap x n_vars
  uint8_t* plane_low;
  uint8_t* best_state;       // n_vars
  const UnitDev* units;
  int32_t n_units, B;
};
// One population annealing step after the step's sweeps
// (population_annealing.h:262-488): statistics, the global lowest, weights,
// Poisson offspring and their scan, the resampling copies and a restart's
// fresh citizens -- all on `st`, reading the population size on the device.
hipError_t launch_pa_step(const PaArgs& a, const ProblemDev& P, hipStream_t st);

struct SweepArgs {
  int32_t n_sweeps;      // at most this many sweeps per unit in this launch
  int32_t best_lds;      // lowest states kept in LDS during the launch (else in HBM)
  const double* lut;     // device, 1024 values
  const int32_t* unit_of_block;  // optional block -> unit map (null: identity)
  // Split parallel tempering ladders: dynamic LDS at byte pre_off holds the
  // local fields (NB per thread) and variable ids of pre_rounds rounds of
  // colour class 0, computed while the ladder's energies travel between
  // work-groups (0: no room, fields computed in the class as usual), and
  // with pre_draws set also their draws (NB fp64 table values and a mask
  // per thread)
  int32_t pre_rounds;
  int32_t pre_off;
  int32_t pre_draws;
};
// Bytes of that area for a launch of `block` threads with NB replicas per
// byte lane.
inline size_t sweep_pre_bytes(int rounds, int block, int nb, bool draws) {
  return (size_t)rounds * block * ((nb + 1) * 4 + (draws ? nb * 8 + 4 : 0));
}

// Dynamic LDS of a sweep launch whose largest unit spin array is max_stride
// bytes; *best_lds (optional) tells whether the lowest states fit in it too.
// Dynamic LDS of a sweep work-group of `threads` threads with nb replicas
// per byte lane; allow_best_lds 0: the lowest states stay in global memory
// (two bytes of LDS per position and byte group instead of three)
size_t sweep_smem_bytes(int max_stride, int ordered, int nb, int threads, int* best_lds = nullptr,
                        int allow_best_lds = 1);
// pairwise: 0 general terms, 1 pairwise 8-byte entries, 2 pairwise int8 4-byte entries
// row_e: entries per pairwise row read ahead by the kernel (4, 6 or 8; any
// value is valid for every layout, since slices store at least 8 entries)
hipError_t launch_sweep(int kind, int pairwise, int coupling, int nb, int row_e, const UnitDev* units, int n_units,
                        int block, size_t smem, const SweepArgs& a, int ordered, int cooperative, int ladder,
                        hipStream_t st);
// largest grid of cooperative sweep work-groups that can be co-resident
int max_coresident_blocks(int kind, int pairwise, int coupling, int nb, int row_e, int block, size_t smem);
// Per-step observations of graph-replayed parallel tempering
// (parallel_tempering.h:141-188): for ladder ld, step s and ladder position
// k, log[((ld * log_steps + s) * L + k) * 4 + {0,1,2,3}] = cost after the
// sweeps, transitions accepted during the step, direction label (0 up, 1
// down) before the swaps, swap of pair (k, k+1) (-1: not attempted).
struct PtObs {
  double* log;
  int64_t log_steps;
  const unsigned long long* accepted;  // per replica, cumulative
  unsigned long long* acc_prev;        // per replica, at the last step
  int32_t* dir;                        // per ladder position
};
hipError_t launch_exchange(double* cost, double* beta, int32_t* label, const double* lbeta,
                           int n_ladders, int L, const int64_t* progress, int units_per_ladder,
                           int spr, int64_t* last_step, uint32_t seed, uint32_t stream,
                           const PtObs& obs, hipStream_t st);
hipError_t launch_init(const UnitDev* units, int n_units, int N, uint32_t epoch, hipStream_t st);
// Set-up of one freshly created unit (one work-group per entry): optionally
// clear its counters (accepted, beta, label, the walker set's error flag),
// set its progress, draw the initial states of `epoch` (spins = lowest) and
// store every replica's energy as its cost and lowest cost, summed exactly
// as launch_energy sums.
struct SetupDesc {
  const UnitDev* unit;
  const int64_t* term_off;
  const int32_t* term_var;
  const double* term_c;
  int64_t M;
  int64_t sweep0, n_sweeps;  // progress to set (set_progress != 0)
  int32_t kind;
  uint32_t epoch;
  int32_t clear;
  int32_t set_progress;
  int32_t keep_ladder;  // clearing leaves beta and labels (a ladder plan uploaded them)
  int32_t pad_;
};
// max_stride: largest spin array among the units; fails with
// hipErrorInvalidValue when it does not fit in LDS (setup_fits says so).
bool setup_fits(int max_stride);
hipError_t launch_setup(const SetupDesc* descs, int n, int max_stride, hipStream_t st);
hipError_t launch_unpack(const UnitDev* units, int B, int r0, int R, int N, int which, uint8_t* out,
                         hipStream_t st);
hipError_t launch_pack(const UnitDev* units, int n_units, int B, int R, int N, const uint8_t* in,
                       int which, hipStream_t st);
hipError_t launch_pack_gather(const UnitDev* units, int n_units, int B, int R, int N,
                              const int32_t* src, const uint8_t* cur, const uint8_t* low, hipStream_t st);
hipError_t launch_gather_replicas(int n, const int32_t* src, const double* c0, const double* c1,
                                  const double* c2, const unsigned long long* a0, double* o0,
                                  double* o1, double* o2, unsigned long long* oa, hipStream_t st);
// energies of n plain states into out[] (and out2[] when given)
hipError_t launch_energy(const ProblemDev& P, const uint8_t* states, int n, double* out,
                         hipStream_t st, double* out2 = nullptr);
// Zeroes up to kZeroSpans device byte ranges in one launch; each pointer is
// 16-byte aligned.
constexpr int kZeroSpans = 12;
struct ZeroSpans {
  void* p[kZeroSpans];
  size_t n[kZeroSpans];
};
hipError_t launch_zero(const ZeroSpans& z, int count, hipStream_t st);
hipError_t launch_delta_all(const ProblemDev& P, int n_states, const uint8_t* states, double* out,
                            hipStream_t st);

// One walker set of a batch read-back: its replicas occupy global replica
// slots roff .. roff+R-1 and its states bytes soff .. soff+R*N-1.
struct ReadDesc {
  const UnitDev* units;
  int64_t soff;
  int64_t roff;
  int32_t B, R, N, kind;
  int64_t M;
  const int64_t* term_off;
  const int32_t* term_var;
  const double* term_c;
};
// Lowest states (plain layout), their FP64 energies and accepted counts of
// every replica of every walker set; rep_member maps a global replica to its
// walker set.
hipError_t launch_batch_read(const ReadDesc* desc, const int32_t* rep_member, int64_t n_replicas,
                             uint8_t* states, double* energy, unsigned long long* accepted,
                             hipStream_t st);
// LDS streaming probe: `blocks` work-groups of 1024 threads, each thread
// reading `iters` 16-byte LDS words; returns the bytes read.
double launch_lds_probe(int blocks, int iters, uint32_t* out, hipStream_t st);
hipError_t launch_philox(int n, const uint32_t* ctr, uint32_t k0, uint32_t k1, uint32_t* out,
                         hipStream_t st);
// One staged host-to-device upload: `bytes` bytes from offset `src` of the
// device copy of the staging ring to device address `dst`.
struct ScatterRec {
  uint64_t src;
  uint64_t dst;
  uint64_t bytes;
};
// Applies n such copies (one work-group each, in no particular order: their
// destinations do not overlap).
hipError_t launch_scatter(const uint8_t* stage, const ScatterRec* rec, int n, hipStream_t st);

}  // namespace qmc
