// C-ABI of the sweep engine (include/qmc.h): problem upload (colouring and
// position-space CSR), walker sets, and launches.  Host-side HIP runtime code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "err_buf.h"
#include "qmc_internal.h"

using namespace qmc;

namespace {

thread_local qmc::ErrBuf g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(e_ == hipErrorOutOfMemory ? QMC_MEMORY_LIMITED : QMC_DEVICE_ERROR,   \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                  \
  } while (0)

// Device allocations are recycled through size buckets (powers of two from
// 256 B): hipMalloc/hipFree synchronise and cost far more than a solve's
// setup, and a serving loop must not call them per request.
class DevicePool {
 public:
  static DevicePool& get() {
    static DevicePool* p = new DevicePool();  // lives until process exit
    return *p;
  }
  hipError_t alloc(void** out, size_t bytes) {
    const size_t b = bucket(bytes);
    {
      std::lock_guard<std::mutex> lk(mu_);
      int dev = 0;
      (void)hipGetDevice(&dev);
      auto& fl = free_[{dev, b}];
      if (!fl.empty()) {
        *out = fl.back();
        fl.pop_back();
        return hipSuccess;
      }
    }
    hipError_t e = hipMalloc(out, b);
    if (e == hipErrorOutOfMemory) {
      // give the device's cached free buffers back (hipFree waits for the
      // device first) and try once more
      (void)hipGetLastError();
      trim();
      e = hipMalloc(out, b);
    }
    return e;
  }
  // Frees the cached buffers of the current device.
  void trim() {
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& kv : free_) {
      if (kv.first.first != dev) continue;
      for (void* q : kv.second) (void)hipFree(q);
      kv.second.clear();
    }
  }
  void release(void* p, size_t bytes) {
    if (!p) return;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(mu_);
    free_[{dev, bucket(bytes)}].push_back(p);
  }
  static size_t bucket(size_t bytes) {
    size_t b = 256;
    while (b < bytes) b <<= 1;
    return b;
  }

 private:
  std::mutex mu_;
  std::map<std::pair<int, size_t>, std::vector<void*>> free_;
};

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) DevicePool::get().release(p, n * sizeof(T));
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return hipSuccess;
    return DevicePool::get().alloc((void**)&p, count * sizeof(T));
  }
  // at least `count` elements, keeping the current buffer when it is large enough
  hipError_t reserve(size_t count) { return count <= n && p ? hipSuccess : alloc(count); }
  hipError_t upload(const T* src, size_t count) {
    hipError_t e = alloc(count);
    if (e != hipSuccess || count == 0) return e;
    return hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice);
  }
};

}  // namespace

constexpr int kAuxStreams = 3;
constexpr int kMaxBlock = 512;  // threads of the largest sweep work-group

// Pairwise problems whose coefficients are integers in [-128, 127] and whose
// positions fit 24 bits hold 4-byte entries {position | int8 coefficient <<
// 24}: half the bytes of the 8-byte entries.
constexpr int64_t kInt8PosLimit = (int64_t)1 << 24;
static bool int8_entries_fit(qmc_problem p);


// Handles are reference counted: a problem keeps its context alive, a
// walker set its problem, a batch its context and members.  *_destroy drops
// the caller's reference, so handles may be destroyed in any order.
struct qmc_ctx_s {
  std::atomic<int> refs{1};
  int device = 0;
  int n_cus = 256;  // compute units of the device (launch-shape defaults)
  // The context's stream.  Every use goes through stream_of() or
  // sync_stream(), which first enqueue the uploads staged since the last use.
  hipStream_t raw_stream = nullptr;
  DevBuf<double> lut;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // batches fan their launch groups out over the main stream and these
  // (created on first use, default priority), joined back by events
  hipStream_t aux[kAuxStreams] = {};
  hipEvent_t fork_ev = nullptr, join_ev[kAuxStreams] = {};
  // pinned staging ring for small host-to-device uploads: they are copied
  // into the ring and recorded, and the next use of the stream brings all of
  // them to the device with one copy of the ring's span (into `dstage`, its
  // device mirror) and one scatter launch, instead of one copy per upload
  std::mutex stage_mu;
  uint8_t* stage = nullptr;
  size_t stage_off = 0;
  DevBuf<uint8_t> dstage;
  std::vector<ScatterRec> pending;  // staged, not yet enqueued (src = ring offset)
  std::unordered_map<const void*, size_t> pending_dst;  // destination -> its record in `pending`
  size_t pend_begin = 0;            // ring offset of the first pending byte
  hipError_t sticky = hipSuccess;   // a failed flush, reported by sync_stream()
  // beta schedules on the device, shared by walker sets that plan the same
  // values (many solves of one schedule upload it once)
  struct SchedEntry {
    std::vector<double> values;
    std::shared_ptr<DevBuf<double>> dev;
    uint64_t used = 0;
  };
  std::mutex sched_mu;
  std::vector<SchedEntry> sched_cache;
  uint64_t sched_clock = 0;
  // set-up descriptors of the last set-up launch (reused in stream order)
  DevBuf<SetupDesc> setup_descs;
};

constexpr size_t kSchedCacheEntries = 16;
constexpr size_t kSchedCacheMaxValues = 1 << 17;

constexpr size_t kStageBytes = 32u << 20;

// Enqueues the staged uploads on the stream (caller holds stage_mu): one
// copy of the ring's span and one scatter launch, or direct copies for one
// or two uploads (or when the copy records do not fit behind the data).
static hipError_t flush_locked(qmc_ctx c) {
  if (c->pending.empty()) return hipSuccess;
  hipError_t e = hipSuccess;
  const size_t n = c->pending.size();
  const size_t at = (c->stage_off + 15) & ~(size_t)15;
  const size_t rec_bytes = n * sizeof(ScatterRec);
  if (n <= 2 || at + rec_bytes > kStageBytes) {
    for (const ScatterRec& r : c->pending) {
      e = hipMemcpyAsync((void*)r.dst, c->stage + r.src, r.bytes, hipMemcpyHostToDevice, c->raw_stream);
      if (e != hipSuccess) break;
    }
  } else {
    std::memcpy(c->stage + at, c->pending.data(), rec_bytes);
    c->stage_off = at + rec_bytes;
    e = hipMemcpyAsync(c->dstage.p + c->pend_begin, c->stage + c->pend_begin, c->stage_off - c->pend_begin,
                       hipMemcpyHostToDevice, c->raw_stream);
    if (e == hipSuccess)
      e = launch_scatter(c->dstage.p, reinterpret_cast<const ScatterRec*>(c->dstage.p + at), (int)n,
                         c->raw_stream);
  }
  c->pending.clear();
  c->pending_dst.clear();
  c->pend_begin = c->stage_off;
  return e;
}

// The context's stream, after enqueueing the staged uploads (a failed
// flush is kept and reported by the next sync_stream()).
static hipStream_t stream_of(qmc_ctx c) {
  std::lock_guard<std::mutex> lk(c->stage_mu);
  const hipError_t e = flush_locked(c);
  if (e != hipSuccess && c->sticky == hipSuccess) c->sticky = e;
  return c->raw_stream;
}

// A context's stream for a function body: each use converts through
// stream_of(), so work enqueued after an upload always follows its flush.
struct CtxStream {
  qmc_ctx c;
  operator hipStream_t() const { return stream_of(c); }
};

// Launch helpers that take a raw stream check that nothing is staged (a
// launch before the flush would read device memory the staged uploads have
// not written yet).
static bool uploads_pending(qmc_ctx c) {
  std::lock_guard<std::mutex> lk(c->stage_mu);
  return !c->pending.empty();
}

static hipError_t sync_stream(qmc_ctx c) {
  const const CtxStream st{c};
  if (c->sticky != hipSuccess) return c->sticky;
  return hipStreamSynchronize(st);
}

// Copies `bytes` from host `src` to device `dst` in order on the context's
// stream.  `src` may be reused as soon as this returns: small uploads are
// copied into the pinned ring and enqueued at the next use of the stream
// (the ring wraps after draining the stream); large ones wait for their own
// copy.
static hipError_t upload_async(qmc_ctx c, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return hipSuccess;
  std::lock_guard<std::mutex> lk(c->stage_mu);
  if (!c->stage) {
    hipError_t e = hipHostMalloc((void**)&c->stage, kStageBytes, hipHostMallocDefault);
    if (e != hipSuccess) {
      c->stage = nullptr;
      return e;
    }
    e = c->dstage.alloc(kStageBytes);
    if (e != hipSuccess) {
      (void)hipHostFree(c->stage);
      c->stage = nullptr;
      return e;
    }
  }
  hipError_t e = hipSuccess;
  if (bytes > kStageBytes / 4) {
    if ((e = flush_locked(c)) != hipSuccess) return e;
    e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->raw_stream);
    return e != hipSuccess ? e : hipStreamSynchronize(c->raw_stream);
  }
  // A second upload to the same place: of the same size, it replaces the
  // first one's staged bytes (nothing has been enqueued from them yet);
  // otherwise it waits for the first to be enqueued (the scatter applies one
  // batch's copies in no particular order).
  auto same = c->pending_dst.find(dst);
  if (same != c->pending_dst.end()) {
    const ScatterRec& r = c->pending[same->second];
    if (r.bytes == bytes) {
      std::memcpy(c->stage + r.src, src, bytes);
      return hipSuccess;
    }
    if ((e = flush_locked(c)) != hipSuccess) return e;
  }
  size_t at = (c->stage_off + 15) & ~(size_t)15;
  if (at + bytes > kStageBytes) {
    // every earlier copy out of the ring is enqueued on the stream
    if ((e = flush_locked(c)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(c->raw_stream)) != hipSuccess) return e;
    at = 0;
    c->pend_begin = 0;
  }
  if (c->pending.empty()) c->pend_begin = at;
  std::memcpy(c->stage + at, src, bytes);
  c->stage_off = at + bytes;
  c->pending.push_back(ScatterRec{(uint64_t)at, (uint64_t)(uintptr_t)dst, (uint64_t)bytes});
  c->pending_dst[dst] = c->pending.size() - 1;
  return hipSuccess;
}

struct EllLayout {
  DevBuf<int32_t> slice_base;
  DevBuf<int4> slice;  // {entry offset, width, first position, lane split}
  DevBuf<uint2> ent;   // pairwise problems, 8-byte entries
  // pairwise problems with int8 coefficients: 4-byte entries {position |
  // coefficient << 24} (int8 set) instead of `ent`
  DevBuf<uint32_t> ent32;
  bool int8 = false;
  // pairwise: the smallest of 4, 6, 8 entries covering every slice's width
  // (8 when wider slices exist; their further entries are read per row)
  int row_e = 8;
  DevBuf<uint4> ent4;  // general problems with at most 3 other members per term
};

struct qmc_problem_s {
  std::atomic<int> refs{1};
  qmc_ctx ctx = nullptr;
  std::map<int, std::unique_ptr<EllLayout>> ell;       // per lane count (compact entries where they fit)
  std::map<int, std::unique_ptr<EllLayout>> ell_wide;  // per lane count, 8-byte pairwise entries
  std::vector<int32_t> h_inc_off, h_nbr;          // host copies to build layouts
  std::vector<double> h_coef;
  std::vector<int32_t> h_oth_off, h_oth_pos;      // general problems
  int max_others = 0;                             // most other members of one term
  int kind = QMC_ISING;
  int32_t N = 0;
  int64_t M = 0;
  int coupling = QMC_COUPLING_I32;
  int pairwise = 1;
  int lanes = 64;
  bool lanes_given = false;  // lanes chosen by the caller (else walker sets pick their own)
  int32_t max_class = 0;
  int max_degree = 0;
  std::vector<int32_t> class_off, var_of_pos, pos_of_var;
  DevBuf<int32_t> d_class_off, d_var_of_pos, d_pos_of_var, d_inc_off, d_nbr, d_oth_off, d_oth_pos;
  DevBuf<int32_t> d_coef_i;
  DevBuf<float> d_coef_f;
  DevBuf<int64_t> d_term_off;
  DevBuf<int32_t> d_term_var;
  DevBuf<double> d_term_c;
  int64_t n_inc = 0;
  ProblemDev dev() const {
    ProblemDev P{};
    P.N = N;
    P.n_classes = (int32_t)class_off.size() - 1;
    P.lanes = lanes;
    P.kind = kind;
    P.pairwise = pairwise;
    P.coupling = coupling;
    P.M = M;
    P.class_off = d_class_off.p;
    P.var_of_pos = d_var_of_pos.p;
    P.pos_of_var = d_pos_of_var.p;
    P.inc_off = d_inc_off.p;
    P.coef = coupling == QMC_COUPLING_I32 ? (const void*)d_coef_i.p : (const void*)d_coef_f.p;
    P.nbr = d_nbr.p;
    P.oth_off = d_oth_off.p;
    P.oth_pos = d_oth_pos.p;
    P.term_off = d_term_off.p;
    P.term_var = d_term_var.p;
    P.term_c = d_term_c.p;
    return P;
  }
};

struct qmc_walkers_s {
  std::atomic<int> refs{1};
  qmc_problem prob = nullptr;
  int32_t R = 0, B = 1, G = 1, n_units = 0, stride = 0;
  int32_t cap = 0, cap_units = 0;  // allocated replicas / work-groups
  int32_t lanes = 64;               // positions side by side in a class round
  // work-groups expected to hold live replicas (population mode: the target
  // population's, not the capacity's)
  int32_t live_units = 0;
  int32_t per_seed = 0;
  uint32_t seed = 0, stream = 0;
  int32_t first_replica = 0;
  DevBuf<uint8_t> spins, best, scratch_states;
  DevBuf<double> cost, best_cost, beta, ladder_beta, energies;
  std::shared_ptr<DevBuf<double>> sched;  // beta schedule (possibly shared, never rewritten)
  DevBuf<unsigned long long> accepted;
  DevBuf<int32_t> label;
  DevBuf<int64_t> progress;  // 2 per unit: next sweep, sweeps left
  DevBuf<int64_t> last_step; // per PT ladder: last exchanged step
  DevBuf<int32_t> sync;      // per PT ladder: barrier counter (cooperative ladders)
  DevBuf<double> cbuf;       // per PT ladder: [3][L] published energies (NaN: not yet)
  DevBuf<int32_t> err;       // barrier time-out flag
  int coop = 0;              // ladders split over work-groups, exchanged in-kernel
  int force_graph = 0;       // split ladders through graph-replayed launches instead
  int nb = 8;                // compile-time replicas per byte lane (1, 2, 4, 8)
  int pair = 0;              // kernel form: 0 general terms, 1 pairwise 8-byte, 2 pairwise int8 entries
  int row_e = 8;             // pairwise entries per row read ahead (4, 6, 8)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_key[4] = {0, 0, 0, 0};
  DevBuf<UnitDev> d_units;
  std::vector<UnitDev> units;
  int ladder = 0;
  int spr = 1;
  int64_t sched_base = 0;
  int64_t planned = 0;       // sweeps of the last plan
  uint64_t version = 0;      // bumped whenever `units` changes
  bool units_stale = true;   // `d_units` behind `units` (uploaded before device use)
  bool pend_zero = true;     // device arrays not cleared yet (the first set-up clears them)
  // qmc_walkers_init and a plan record what to do; the set-up launch (alone,
  // or for every pending member of a batch at once) performs it
  bool pend_init = false;
  uint32_t pend_epoch = 0;
  bool pend_prog = false;
  int64_t pend_sweep0 = 0, pend_n = 0;
  bool pend_keep_ladder = false;  // the pending set-up keeps the uploaded ladder labels and betas
  // population annealing with the population size on the device
  int pa_on = 0;
  PaArgs pa_args{};
  DevBuf<PaDev> pa;
  DevBuf<double> pa_table, pa_weight, pa_tcost, pa_tbest, pa_tbeta, pa_cost0, pa_tree, pa_log;
  DevBuf<int32_t> pa_family, pa_family_next, pa_src, pa_hist, pa_count, pa_ids;
  bool pa_stepped = false;  // a step was enqueued since qmc_walkers_pa_begin
  int ordered = 0;           // sweeps visit variables in index order (k_sweep_ordered)
  // parallel tempering observations (graph-replayed steps only)
  DevBuf<double> pt_log;
  DevBuf<unsigned long long> pt_acc_prev;
  DevBuf<int32_t> pt_dir;
  PtObs pt_obs{};
  DevBuf<unsigned long long> pa_tacc;
  DevBuf<uint8_t> pa_plane_cur, pa_plane_low, pa_best_state;
  float last_ms = 0.f;
};

struct qmc_batch_s {
  qmc_ctx ctx = nullptr;
  std::vector<qmc_walkers> members;
  std::vector<uint64_t> seen_version;
  bool dirty = true;
  // one launch group per kernel instantiation
  struct Group {
    int kind, pairwise, coupling, nb, row_e, coop, ladder, ordered;
    int block = 64;
    int stride = 0;
    int64_t live = 0;  // work-groups expected to hold live replicas
    double work = 0;   // attempts planned (for spreading groups over streams)
    std::vector<UnitDev> units;
    DevBuf<UnitDev> d_units;
    // cooperative groups: launches of units [chunk[i], chunk[i+1]) (at most
    // the co-resident work-group count, never splitting a ladder);
    // chunk_err: a ladder exceeds the co-resident limit
    std::vector<int> chunk;
    bool chunk_err = false;
    int pre_rounds = 0;  // early-field rounds of its split ladders (largest)
  };
  std::vector<std::unique_ptr<Group>> groups;
  // where each member's descriptors sit in its group (valid while the
  // member's version equals seen_version)
  std::vector<std::pair<Group*, size_t>> member_at;
  float last_ms = 0.f;
  // chunked runs: launches of `graph_spl` sweeps captured once and replayed
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int64_t graph_key[3] = {0, 0, 0};  // sweeps per launch, launches, rebuild count
  int64_t rebuilds = 0;
  // device time of the runs since the last wait: ev0 at the first run's
  // start, ev1 at every run's end
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool span_open = false;
  ~qmc_batch_s() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }
  // batched read-back
  DevBuf<ReadDesc> d_desc;
  DevBuf<int32_t> d_rep_member;
  DevBuf<uint8_t> d_states;
  DevBuf<double> d_energy;
  DevBuf<unsigned long long> d_acc;
};

// Set by an exit handler registered when the first context is created, i.e.
// after the HIP runtime registered its own teardown, so it runs first: a
// context released after that point (a finaliser during process exit) is
// abandoned instead of calling into a runtime that is shutting down.
static std::atomic<bool> g_exiting{false};
static void mark_exiting() { g_exiting.store(true); }

// Device-side staging of qmc_energy_noalloc: a fixed array in every GPU's
// copy of the code object, so that entry point never allocates.  In host code
// `scratch` names a host-side stand-in; its device address comes from
// hipGetSymbolAddress, looked up once per device (with the device current)
// when the first context on it is created, and only that address is used.
constexpr size_t kScratchBytes = (size_t)QMC_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};
// one caller at a time per device: every call shares that device's array
static std::mutex g_scratch_use[kMaxDevices];

// With `device` current: the device address of its scratch array.
static hipError_t scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = ptr;
  return hipSuccess;
}

extern "C" {

const char* qmc_last_error(void) { return g_err.c_str(); }

int qmc_ctx_create(int device, qmc_ctx* out) {
  if (!out) return fail(QMC_INVALID_ARGUMENT, "null output handle");
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(QMC_INVALID_ARGUMENT, "no such device");
  HIP_TRY(hipSetDevice(device));
  static std::once_flag exit_hook;
  std::call_once(exit_hook, [] { std::atexit(mark_exiting); });
  {
    const hipError_t e = scratch_setup(device);
    if (e != hipSuccess)
      return fail(QMC_DEVICE_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  }
  auto* c = new qmc_ctx_s();
  c->device = device;
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
      c->n_cus = cus;
  }
  if (hipStreamCreateWithFlags(&c->raw_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(QMC_DEVICE_ERROR, "hipStreamCreate failed");
  }
  // utils::uniform_ln (utils/utils.cc:12-32), computed with the host libm as the reference does
  std::vector<double> lut(1024);
  for (int r = 0; r < 4; ++r)
    for (int i = 0; i < 256; ++i) lut[r * 256 + i] = std::log((double)i / std::pow(256.0, (double)(r + 1)));
  if (c->lut.upload(lut.data(), lut.size()) != hipSuccess || hipEventCreate(&c->ev0) != hipSuccess ||
      hipEventCreate(&c->ev1) != hipSuccess) {
    qmc_ctx_destroy(c);
    return fail(QMC_DEVICE_ERROR, "context allocation failed");
  }
  *out = c;
  return QMC_OK;
}

static void ctx_release(qmc_ctx c) {
  if (!c || --c->refs > 0) return;
  if (g_exiting.load()) return;  // the process is ending: nothing left to free
  (void)hipSetDevice(c->device);
  if (c->raw_stream) (void)hipStreamSynchronize(c->raw_stream);
  c->lut.release();
  c->dstage.release();
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
  for (int i = 0; i < kAuxStreams; ++i) {
    if (c->aux[i]) (void)hipStreamSynchronize(c->aux[i]);
    if (c->join_ev[i]) (void)hipEventDestroy(c->join_ev[i]);
    if (c->aux[i]) (void)hipStreamDestroy(c->aux[i]);
  }
  if (c->raw_stream) (void)hipStreamDestroy(c->raw_stream);
  if (c->stage) (void)hipHostFree(c->stage);
  delete c;
}

void qmc_ctx_destroy(qmc_ctx c) { ctx_release(c); }

int qmc_ctx_compute_units(qmc_ctx c, int32_t* compute_units) {
  if (!c || !compute_units) return fail(QMC_INVALID_ARGUMENT, "null argument");
  *compute_units = c->n_cus;
  return QMC_OK;
}

int qmc_sync(qmc_ctx c) {
  if (!c) return fail(QMC_INVALID_ARGUMENT, "null context");
  HIP_TRY(sync_stream(c));
  return QMC_OK;
}

int qmc_problem_create(qmc_ctx ctx, int kind, int32_t n_vars, int64_t n_terms,
                       const int64_t* term_offsets, const int32_t* term_vars, const double* coeffs,
                       int32_t lanes, qmc_problem* out) {
  return qmc_problem_create_visit(ctx, kind, n_vars, n_terms, term_offsets, term_vars, coeffs, lanes,
                                  QMC_VISIT_COLOR_CLASSES, out);
}

int qmc_problem_create_visit(qmc_ctx ctx, int kind, int32_t n_vars, int64_t n_terms,
                             const int64_t* term_offsets, const int32_t* term_vars, const double* coeffs,
                             int32_t lanes, int32_t visit, qmc_problem* out) {
  if (!ctx || !out) return fail(QMC_INVALID_ARGUMENT, "null handle");
  if (visit != QMC_VISIT_COLOR_CLASSES && visit != QMC_VISIT_INDEX_LEVELS)
    return fail(QMC_INVALID_ARGUMENT, "unknown visit order");
  if (kind != QMC_ISING && kind != QMC_PUBO) return fail(QMC_INVALID_ARGUMENT, "unknown model kind");
  if (n_vars <= 0 || n_terms <= 0) return fail(QMC_INVALID_ARGUMENT, "empty problem");
  if (lanes != 0 && lanes != 64 && lanes != 128 && lanes != 256 && lanes != 512)
    return fail(QMC_INVALID_ARGUMENT, "lanes must be 64, 128, 256 or 512");
  const int32_t N = n_vars;
  std::vector<std::vector<int32_t>> node_terms(N);
  bool all_int = true;
  double abs_sum = 0.0;
  int max_term = 0;
  for (int64_t j = 0; j < n_terms; ++j) {
    const int64_t a = term_offsets[j], b = term_offsets[j + 1];
    if (b <= a) return fail(QMC_INVALID_ARGUMENT, "constant terms must be folded into const_cost");
    max_term = std::max<int>(max_term, (int)(b - a));
    for (int64_t k = a; k < b; ++k) {
      const int32_t v = term_vars[k];
      if (v < 0 || v >= N) return fail(QMC_INVALID_ARGUMENT, "variable id out of range");
      if (k > a && term_vars[k - 1] >= v)
        return fail(QMC_DUPLICATED_VARIABLE, "term ids must be strictly increasing");
      node_terms[v].push_back((int32_t)j);
    }
    const double c = coeffs[j];
    if (!std::isfinite(c)) return fail(QMC_INVALID_ARGUMENT, "non-finite coefficient");
    if (c != std::floor(c)) all_int = false;
    abs_sum += std::fabs(c);
  }
  for (int32_t v = 0; v < N; ++v)
    if (node_terms[v].empty()) return fail(QMC_INVALID_ARGUMENT, "variable without terms");
  auto p = std::make_unique<qmc_problem_s>();
  p->ctx = ctx;
  p->kind = kind;
  p->N = N;
  p->M = n_terms;
  p->coupling = (all_int && 2.0 * abs_sum < 2147483647.0) ? QMC_COUPLING_I32 : QMC_COUPLING_F32;
  p->pairwise = max_term <= 2;

  // Classes over "shares a term": members of a class can be updated
  // simultaneously.  Colour classes: greedy colouring in variable order (few
  // classes).  Index levels: class of v = 1 + the largest class of a
  // neighbour u < v (0 without one), so running the classes in turn gives
  // every variable the neighbour values it sees in an index-order sweep --
  // the same decisions and end-of-sweep states.
  std::vector<int32_t> color(N, -1), stamp;
  int n_colors = 0;
  for (int32_t v = 0; v < N; ++v) {
    int level = 0;
    for (int32_t j : node_terms[v])
      for (int64_t k = term_offsets[j]; k < term_offsets[j + 1]; ++k) {
        const int32_t u = term_vars[k];
        if (u < v) {
          const int32_t cu = color[u];
          level = std::max(level, cu + 1);
          if ((int)stamp.size() <= cu) stamp.resize(cu + 1, -1);
          stamp[cu] = v;
        }
      }
    int cv = 0;
    if (visit == QMC_VISIT_INDEX_LEVELS) {
      cv = level;
    } else {
      while (cv < (int)stamp.size() && stamp[cv] == v) ++cv;
    }
    color[v] = cv;
    n_colors = std::max(n_colors, cv + 1);
  }
  std::vector<int32_t> csize(n_colors, 0);
  for (int32_t v = 0; v < N; ++v) csize[color[v]]++;
  p->class_off.assign(n_colors + 1, 0);
  for (int c = 0; c < n_colors; ++c) p->class_off[c + 1] = p->class_off[c] + csize[c];
  p->var_of_pos.assign(N, 0);
  p->pos_of_var.assign(N, 0);
  {
    // inside a class: descending degree, then variable id -- members of a
    // class are independent, so any order is a valid sweep; degree order
    // keeps the padding of the sliced layout small
    std::vector<int32_t> fillp(p->class_off.begin(), p->class_off.end() - 1);
    for (int32_t v = 0; v < N; ++v) p->var_of_pos[fillp[color[v]]++] = v;
    for (int c = 0; c < n_colors; ++c)
      std::stable_sort(p->var_of_pos.begin() + p->class_off[c], p->var_of_pos.begin() + p->class_off[c + 1],
                       [&](int32_t a, int32_t b) { return node_terms[a].size() > node_terms[b].size(); });
    for (int32_t q = 0; q < N; ++q) p->pos_of_var[p->var_of_pos[q]] = q;
  }
  const int32_t nmax = *std::max_element(csize.begin(), csize.end());
  p->max_class = nmax;
  p->lanes_given = lanes != 0;
  if (lanes) {
    p->lanes = lanes;
  } else {
    int l = 64;
    while (l < 512 && l * 4 < nmax) l *= 2;
    p->lanes = l;
  }

  // position-space incidence CSR
  std::vector<int32_t> inc_off(N + 1, 0), nbr, oth_off(1, 0), oth_pos, coef_i;
  std::vector<float> coef_f;
  for (int32_t q = 0; q < N; ++q) {
    const int32_t v = p->var_of_pos[q];
    p->max_degree = std::max<int>(p->max_degree, (int)node_terms[v].size());
    for (int32_t j : node_terms[v]) {
      int32_t other = N;
      int n_oth = 0;
      for (int64_t k = term_offsets[j]; k < term_offsets[j + 1]; ++k) {
        const int32_t u = term_vars[k];
        if (u == v) continue;
        other = p->pos_of_var[u];
        oth_pos.push_back(other);
        ++n_oth;
      }
      oth_off.push_back((int32_t)oth_pos.size());
      nbr.push_back(n_oth <= 1 ? other : N);
      coef_i.push_back(p->coupling == QMC_COUPLING_I32 ? (int32_t)coeffs[j] : 0);
      coef_f.push_back((float)coeffs[j]);
    }
    inc_off[q + 1] = (int32_t)nbr.size();
  }
  p->n_inc = (int64_t)nbr.size();
  p->h_inc_off = inc_off;
  p->h_nbr = nbr;
  for (size_t e = 0; e + 1 < oth_off.size(); ++e) p->max_others = std::max(p->max_others, oth_off[e + 1] - oth_off[e]);
  if (!p->pairwise) {
    p->h_oth_off = oth_off;
    p->h_oth_pos = oth_pos;
  }
  p->h_coef.resize(nbr.size());
  for (size_t e = 0; e < nbr.size(); ++e) p->h_coef[e] = coef_f[e];
  if (p->coupling == QMC_COUPLING_I32)
    for (size_t e = 0; e < nbr.size(); ++e) p->h_coef[e] = coef_i[e];
  std::vector<int64_t> toff(term_offsets, term_offsets + n_terms + 1);
  std::vector<int32_t> tvar(term_vars, term_vars + term_offsets[n_terms]);
  std::vector<double> tc(coeffs, coeffs + n_terms);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(p->d_class_off.upload(p->class_off.data(), p->class_off.size()));
  {
    // padded by two class rounds of the widest lane count, so sweep lanes
    // prefetching past the end of the last class may read it unconditionally
    std::vector<int32_t> vp(p->var_of_pos);
    vp.resize((size_t)N + 1024, 0);
    HIP_TRY(p->d_var_of_pos.upload(vp.data(), vp.size()));
  }
  HIP_TRY(p->d_pos_of_var.upload(p->pos_of_var.data(), N));
  HIP_TRY(p->d_inc_off.upload(inc_off.data(), inc_off.size()));
  HIP_TRY(p->d_nbr.upload(nbr.data(), nbr.size()));
  HIP_TRY(p->d_oth_off.upload(oth_off.data(), oth_off.size()));
  oth_pos.push_back(0);  // never empty
  HIP_TRY(p->d_oth_pos.upload(oth_pos.data(), oth_pos.size()));
  if (p->coupling == QMC_COUPLING_I32) HIP_TRY(p->d_coef_i.upload(coef_i.data(), coef_i.size()));
  else HIP_TRY(p->d_coef_f.upload(coef_f.data(), coef_f.size()));
  HIP_TRY(p->d_term_off.upload(toff.data(), toff.size()));
  HIP_TRY(p->d_term_var.upload(tvar.data(), tvar.size()));
  HIP_TRY(p->d_term_c.upload(tc.data(), tc.size()));
  ctx->refs++;
  *out = p.release();
  return QMC_OK;
}

static void problem_release(qmc_problem p) {
  if (!p || --p->refs > 0) return;
  qmc_ctx c = p->ctx;
  delete p;
  ctx_release(c);
}

void qmc_problem_destroy(qmc_problem p) { problem_release(p); }

int qmc_problem_get_info(qmc_problem p, qmc_problem_info* info) {
  if (!p || !info) return fail(QMC_INVALID_ARGUMENT, "null handle");
  info->n_vars = p->N;
  info->n_terms = p->M;
  info->n_incidences = p->n_inc;
  info->n_classes = (int32_t)p->class_off.size() - 1;
  info->lanes = p->lanes;
  info->coupling = p->coupling;
  info->pairwise = p->pairwise;
  info->max_degree = p->max_degree;
  info->mean_degree = (double)p->n_inc / (double)p->N;
  info->max_class = 0;
  for (size_t c = 0; c + 1 < p->class_off.size(); ++c)
    info->max_class = std::max(info->max_class, p->class_off[c + 1] - p->class_off[c]);
  info->entry_bytes = p->pairwise ? (int8_entries_fit(p) ? 4 : 8) : (p->max_others <= 3 ? 16 : 0);
  return QMC_OK;
}

int qmc_problem_order(qmc_problem p, int32_t lanes, int32_t* order, int32_t* class_offsets) {
  if (!p) return fail(QMC_INVALID_ARGUMENT, "null handle");
  if (lanes == 0) lanes = p->lanes;
  if (lanes != 64 && lanes != 128 && lanes != 256 && lanes != 512)
    return fail(QMC_INVALID_ARGUMENT, "lanes must be 64, 128, 256 or 512");
  const int nclass = (int)p->class_off.size() - 1;
  int32_t k = 0;
  for (int c = 0; c < nclass; ++c) {
    const int32_t c0 = p->class_off[c], nc = p->class_off[c + 1] - c0;
    if (class_offsets) class_offsets[c] = c0;
    for (int32_t tp = 0; tp < lanes; ++tp)
      for (int32_t rho = tp; rho < nc; rho += lanes)
        if (order) order[k++] = p->var_of_pos[c0 + rho];
  }
  if (class_offsets) class_offsets[nclass] = p->N;
  return QMC_OK;
}

// Sliced ELL of the pairwise form for `lanes` lanes: slice (c, k) covers
// positions c0 + k*lanes .. + lanes of class c; width = its max degree.
static bool int8_entries_fit(qmc_problem p) {
  if (!p->pairwise || p->coupling != QMC_COUPLING_I32 || (int64_t)p->N >= kInt8PosLimit) return false;
  for (double c : p->h_coef)
    if (c < -128.0 || c > 127.0) return false;
  return true;
}

static EllLayout* get_layout(qmc_problem p, int lanes, bool wide = false) {
  auto& cache = (wide && p->pairwise) ? p->ell_wide : p->ell;
  auto it = cache.find(lanes);
  if (it != cache.end()) return it->second.get();
  auto L = std::make_unique<EllLayout>();
  const int nclass = (int)p->class_off.size() - 1;
  const bool i32 = p->coupling == QMC_COUPLING_I32;
  auto bits = [&](double c) -> uint32_t {
    if (i32) return (uint32_t)(int32_t)c;
    const float f = (float)c;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
  };
  std::vector<int32_t> sbase(nclass + 1, 0);
  std::vector<int4> slice;
  std::vector<uint2> ent;
  // The sweep fetches at most one round past a class's end as slice
  // sbase[c] + K_c: the table ends with copies of its first two slices (the
  // next sweep's first rounds), or zero-width slices when it has fewer.
  auto close_table = [&](int32_t n_entries) {
    const size_t real = slice.size();
    for (size_t i = 0; i < 2; ++i)
      slice.push_back(i < real ? slice[i] : make_int4(n_entries, 0, p->N, 0));
  };
  if (!p->pairwise) {
    // general terms with at most 3 other members: 16-byte entries
    if (p->max_others > 3) return nullptr;
    const uint32_t ghost = (uint32_t)p->N;
    std::vector<uint4> ent4;
    for (int c = 0; c < nclass; ++c) {
      const int32_t c0 = p->class_off[c], nc = p->class_off[c + 1] - c0;
      sbase[c] = (int32_t)slice.size();
      // small classes of integer problems share a position among 2^split
      // lanes (at most 8, while positions x lanes per position fit the lanes)
      int32_t split = 0;
      if (i32)
        while (split < 3 && ((int64_t)nc << (split + 1)) <= lanes) ++split;
      for (int32_t k = 0; k * lanes < nc; ++k) {
        int32_t w = 0;
        for (int32_t t = 0; t < lanes && k * lanes + t < nc; ++t) {
          const int32_t q = c0 + k * lanes + t;
          w = std::max(w, p->h_inc_off[q + 1] - p->h_inc_off[q]);
        }
        slice.push_back(make_int4((int32_t)ent4.size(), w, c0 + k * lanes, split));
        for (int32_t e = 0; e < w; ++e)
          for (int32_t t = 0; t < lanes; ++t) {
            const int32_t rho = k * lanes + t;
            uint4 x = make_uint4(bits(0.0), ghost, ghost, ghost);
            if (rho < nc) {
              const int32_t q = c0 + rho;
              const int32_t e0 = p->h_inc_off[q], deg = p->h_inc_off[q + 1] - e0;
              if (e < deg) {
                const int32_t i = e0 + e;
                uint32_t o[3] = {ghost, ghost, ghost};
                for (int32_t m = p->h_oth_off[i]; m < p->h_oth_off[i + 1]; ++m)
                  o[m - p->h_oth_off[i]] = (uint32_t)p->h_oth_pos[m];
                x = make_uint4(bits(p->h_coef[i]), o[0], o[1], o[2]);
              }
            }
            ent4.push_back(x);
          }
      }
    }
    sbase[nclass] = (int32_t)slice.size();
    close_table((int32_t)ent4.size());
    ent4.push_back(make_uint4(bits(0.0), ghost, ghost, ghost));
    const bool ok = L->slice_base.upload(sbase.data(), sbase.size()) == hipSuccess &&
                    L->slice.upload(slice.data(), slice.size()) == hipSuccess &&
                    L->ent4.upload(ent4.data(), ent4.size()) == hipSuccess;
    if (!ok) return nullptr;
    EllLayout* raw = L.get();
    cache[lanes] = std::move(L);
    return raw;
  }
  const bool compact = !wide && int8_entries_fit(p);
  std::vector<uint32_t> ent32;
  const uint32_t ghost = (uint32_t)p->N;
  auto put = [&](uint32_t pos, double c) {
    if (compact)
      ent32.push_back(pos | ((uint32_t)(uint8_t)(int8_t)(int32_t)c << 24));
    else
      ent.push_back(make_uint2(pos, bits(c)));
  };
  auto n_ent = [&]() { return (int32_t)(compact ? ent32.size() : ent.size()); };
  int32_t widest = 0;
  for (int c = 0; c < nclass; ++c) {
    const int32_t c0 = p->class_off[c], nc = p->class_off[c + 1] - c0;
    sbase[c] = (int32_t)slice.size();
    for (int32_t k = 0; k * lanes < nc; ++k) {
      int32_t w = 0;
      for (int32_t t = 0; t < lanes && k * lanes + t < nc; ++t) {
        const int32_t q = c0 + k * lanes + t;
        w = std::max(w, p->h_inc_off[q + 1] - p->h_inc_off[q]);
      }
      widest = std::max(widest, w);
      slice.push_back(make_int4(n_ent(), w, c0 + k * lanes, 0));
      // at least 8 entries stored: the kernels read their first 4, 6 or 8
      // without masks, and the entries past the width add zero
      for (int32_t e = 0; e < std::max(w, 8); ++e)
        for (int32_t t = 0; t < lanes; ++t) {
          const int32_t rho = k * lanes + t;
          bool real = false;
          if (rho < nc) {
            const int32_t q = c0 + rho;
            const int32_t e0 = p->h_inc_off[q], deg = p->h_inc_off[q + 1] - e0;
            if (e < deg) {
              put((uint32_t)p->h_nbr[e0 + e], p->h_coef[e0 + e]);
              real = true;
            }
          }
          if (!real) put(ghost, 0.0);
        }
    }
  }
  sbase[nclass] = (int32_t)slice.size();
  L->row_e = widest <= 4 ? 4 : (widest <= 6 ? 6 : 8);
  // the table's closing copies, and 8 rows of tail padding: the sweep
  // reads up to 8 rows of a slice
  close_table(n_ent());
  for (int64_t i = 0; i < (int64_t)8 * lanes + 1; ++i) put(ghost, 0.0);
  bool ok = L->slice_base.upload(sbase.data(), sbase.size()) == hipSuccess &&
            L->slice.upload(slice.data(), slice.size()) == hipSuccess;
  if (compact) {
    ok = ok && L->ent32.upload(ent32.data(), ent32.size()) == hipSuccess;
    L->int8 = true;
  } else {
    ok = ok && L->ent.upload(ent.data(), ent.size()) == hipSuccess;
  }
  if (!ok) return nullptr;
  EllLayout* raw = L.get();
  cache[lanes] = std::move(L);
  return raw;
}

static void build_units(qmc_walkers w) {
  qmc_problem p = w->prob;
  w->units.assign(w->n_units, UnitDev{});
  EllLayout* ell = (p->pairwise || p->max_others <= 3) ? get_layout(p, w->lanes) : nullptr;
  w->pair = !p->pairwise ? 0 : (ell && ell->int8 ? 2 : 1);
  w->row_e = (p->pairwise && ell) ? ell->row_e : 8;
  for (int u = 0; u < w->n_units; ++u) {
    UnitDev& U = w->units[u];
    if (ell) {
      U.ell_slice_base = ell->slice_base.p;
      U.ell_slice = ell->slice.p;
      U.ell_ent = p->pairwise ? ell->ent.p : nullptr;
      U.ell4_ent = p->pairwise ? nullptr : ell->ent4.p;
      U.ell_ent32 = ell->ent32.p;
    }
    U.class_off = p->d_class_off.p;
    U.var_of_pos = p->d_var_of_pos.p;
    U.pos_of_var = p->d_pos_of_var.p;
    U.inc_off = p->d_inc_off.p;
    U.coef = p->coupling == QMC_COUPLING_I32 ? (const void*)p->d_coef_i.p : (const void*)p->d_coef_f.p;
    U.nbr = p->d_nbr.p;
    U.oth_off = p->d_oth_off.p;
    U.oth_pos = p->d_oth_pos.p;
    U.N = p->N;
    U.n_classes = (int32_t)p->class_off.size() - 1;
    U.lanes = w->lanes;
    U.lanes_log2 = 31 - __builtin_clz((unsigned)w->lanes);
    U.G = w->G;
    U.stride = w->stride;
    U.ladder = 0;
    const int32_t r = u * w->B;
    U.spins = w->spins.p + (size_t)u * w->stride;
    U.best = w->best.p + (size_t)u * w->stride;
    U.cost = w->cost.p + r;
    U.best_cost = w->best_cost.p + r;
    U.beta = w->beta.p + r;
    U.accepted = w->accepted.p + r;
    U.label = w->label.p + r;
    U.ladder_beta = w->ladder_beta.p;
    U.ladder = w->ladder;
    U.slot0 = 0;
    U.upl = 1;
    U.sync = nullptr;
    U.cbuf = nullptr;
    U.err = w->err.p;
    if (w->ladder > 0) {
      const int L = w->ladder;
      const int ld = r / L;
      U.slot0 = r % L;
      U.upl = L / w->B;
      U.label = w->label.p + (size_t)ld * L;
      if (U.upl > 1) {
        U.sync = w->sync.p + ld;
        U.cbuf = w->cbuf.p + (size_t)ld * 3 * L;
      }
    }
    U.spr = w->spr;
    U.sched = w->sched ? w->sched->p : nullptr;
    U.sched_base = w->sched_base;
    U.progress = w->progress.p + 2 * (size_t)u;
    U.nrep = (uint32_t)std::min(w->B, w->R - r);
    // replicas are grouped per seed: instance k = r / per_seed uses seed + k
    const int32_t gr = w->first_replica + r;
    U.seed = w->seed + (uint32_t)(gr / w->per_seed);
    U.r0 = (uint32_t)(gr % w->per_seed);
    U.stream = w->stream;
    U.rbase = r;
    U.live = w->pa_on ? &w->pa.p->R : nullptr;
  }
}

// Rebuilds the host descriptors; the device copy follows lazily
// (device_units), since batched runs launch from their own copies.
static int upload_units(qmc_walkers w) {
  std::vector<UnitDev> before;
  before.swap(w->units);
  build_units(w);
  if (before.size() == w->units.size() &&
      std::memcmp(before.data(), w->units.data(), before.size() * sizeof(UnitDev)) == 0)
    return QMC_OK;  // unchanged: batches keep their descriptors
  w->version++;
  w->units_stale = true;
  return QMC_OK;
}

// The walkers' unit descriptors on the device, brought up to date in order
// on the context stream.  Called before every launch that reads them.
static int device_units(qmc_walkers w, const UnitDev** out) {
  if (w->units_stale) {
    HIP_TRY(upload_async(w->prob->ctx, w->d_units.p, w->units.data(), w->units.size() * sizeof(UnitDev)));
    w->units_stale = false;
  }
  *out = w->d_units.p;
  return QMC_OK;
}

// every unit: next sweep = sweep0, sweeps left = n
static int set_progress(qmc_walkers w, int64_t sweep0, int64_t n) {
  std::vector<int64_t> pr(2 * (size_t)w->n_units);
  for (int u = 0; u < w->n_units; ++u) {
    pr[2 * u] = sweep0;
    pr[2 * u + 1] = n;
  }
  HIP_TRY(upload_async(w->prob->ctx, w->progress.p, pr.data(), pr.size() * sizeof(int64_t)));
  w->planned = n;
  return QMC_OK;
}

// Class-round width of a walker set on a problem created without one: from
// 64, lanes double while the set's work-groups hold fewer than 16 waves per
// compute unit, one round does not yet cover the largest colour class (or,
// where small classes share a position among up to 8 lanes -- integer terms
// in sliced entries -- 8 lanes of each of its positions), and lanes x G
// stays within 512.  Few work-groups get wide ones (shorter class rounds,
// more waves to hide latency); many get narrow ones.
static int default_lanes(qmc_problem p, int64_t n_units, int G) {
  const int64_t want = 16 * (int64_t)p->ctx->n_cus;
  const bool shared = !p->pairwise && p->coupling == QMC_COUPLING_I32 && p->max_others <= 3;
  const int64_t cover = (int64_t)p->max_class * (shared ? 8 : 1);
  int l = 64;
  while (l * 2 * G <= 512 && l < cover && n_units * (l * G / 64) < want) l *= 2;
  return l;
}

int qmc_walkers_create(qmc_problem p, int32_t n_replicas, int32_t replicas_per_group,
                        uint32_t seed, uint32_t stream, int32_t first_replica,
                        int32_t replicas_per_seed, qmc_walkers* out) {
  if (!p || !out) return fail(QMC_INVALID_ARGUMENT, "null handle");
  if (n_replicas <= 0) return fail(QMC_INVALID_ARGUMENT, "n_replicas must be positive");
  if (replicas_per_group < 1 || replicas_per_group > 32)
    return fail(QMC_INVALID_ARGUMENT, "replicas_per_group must be in 1..32");
  if (replicas_per_seed <= 0) return fail(QMC_INVALID_ARGUMENT, "replicas_per_seed must be positive");
  auto w = std::make_unique<qmc_walkers_s>();
  w->prob = p;
  w->R = n_replicas;
  w->B = replicas_per_group;
  w->G = (w->B + 7) / 8;
  w->n_units = (w->R + w->B - 1) / w->B;
  w->lanes = p->lanes_given ? p->lanes : default_lanes(p, w->n_units, w->G);
  while (w->lanes * w->G > 512) w->lanes /= 2;
  if (w->lanes < 64) return fail(QMC_INVALID_ARGUMENT, "replica bytes exceed the work-group");
  w->cap = w->R;
  w->cap_units = w->n_units;
  w->live_units = w->n_units;
  w->stride = (int32_t)(((int64_t)(p->N + 1) * w->G + 3) / 4 * 4);
  w->seed = seed;
  w->stream = stream;
  w->first_replica = first_replica;
  w->per_seed = replicas_per_seed;
  w->nb = w->B > 4 ? 8 : (w->B > 2 ? 4 : (w->B > 1 ? 2 : 1));
  const size_t smem = sweep_smem_bytes(w->stride, 0, 8, kMaxBlock);
  if (smem > 160 * 1024) return fail(QMC_MEMORY_LIMITED, "problem too large for LDS residency at this group size");
  HIP_TRY(hipSetDevice(p->ctx->device));
  HIP_TRY(w->spins.alloc((size_t)w->n_units * w->stride));
  HIP_TRY(w->best.alloc((size_t)w->n_units * w->stride));
  HIP_TRY(w->cost.alloc(w->R));
  HIP_TRY(w->best_cost.alloc(w->R));
  HIP_TRY(w->beta.alloc(w->R));
  HIP_TRY(w->accepted.alloc(w->R));
  HIP_TRY(w->label.alloc(w->R));
  HIP_TRY(w->ladder_beta.alloc(32));
  HIP_TRY(w->err.alloc(1));
  HIP_TRY(w->d_units.alloc(w->n_units));
  HIP_TRY(w->progress.alloc(2 * (size_t)w->n_units));
  // everything starts at zero: cleared by the first qmc_walkers_init, or by
  // ensure_cleared before any other device use
  int rc = upload_units(w.get());
  if (rc) return rc;
  p->refs++;
  *out = w.release();
  return QMC_OK;
}

// Clears a new walker set's device arrays in one launch unless that already
// happened; every entry point that touches walker memory on the device calls
// this first (qmc_walkers_init folds it into its set-up launch).
static int ensure_cleared(qmc_walkers w) {
  if (!w->pend_zero) return QMC_OK;
  ZeroSpans z{};
  int nz = 0;
  auto zero = [&](void* ptr, size_t bytes) {
    z.p[nz] = ptr;
    z.n[nz++] = bytes;
  };
  zero(w->spins.p, w->spins.n);
  zero(w->best.p, w->best.n);
  zero(w->cost.p, w->cap * sizeof(double));
  zero(w->best_cost.p, w->cap * sizeof(double));
  zero(w->beta.p, w->cap * sizeof(double));
  zero(w->accepted.p, w->cap * sizeof(unsigned long long));
  zero(w->label.p, w->cap * sizeof(int32_t));
  zero(w->err.p, sizeof(int32_t));
  zero(w->progress.p, w->progress.n * sizeof(int64_t));
  HIP_TRY(launch_zero(z, nz, stream_of(w->prob->ctx)));
  w->pend_zero = false;
  return QMC_OK;
}

static void add_setup_descs(qmc_walkers w, const UnitDev* du, std::vector<SetupDesc>& out) {
  qmc_problem p = w->prob;
  for (int u = 0; u < w->n_units; ++u) {
    SetupDesc d{};
    d.unit = du + u;
    d.term_off = p->d_term_off.p;
    d.term_var = p->d_term_var.p;
    d.term_c = p->d_term_c.p;
    d.M = p->M;
    d.kind = p->kind;
    d.epoch = w->pend_epoch;
    d.clear = w->pend_zero ? 1 : 0;
    d.set_progress = w->pend_prog ? 1 : 0;
    d.keep_ladder = w->pend_keep_ladder ? 1 : 0;
    d.sweep0 = w->pend_sweep0;
    d.n_sweeps = w->pend_n;
    out.push_back(d);
  }
}

// The costs a set-up just computed are also the first population-annealing
// step's costs before its sweeps.
static int setup_done(qmc_walkers w) {
  w->pend_init = false;
  w->pend_zero = false;
  w->pend_prog = false;
  w->pend_keep_ladder = false;
  if (w->pa_on && w->pa_args.cost0)
    HIP_TRY(hipMemcpyAsync(w->pa_args.cost0, w->cost.p, (size_t)w->cap * sizeof(double),
                           hipMemcpyDeviceToDevice, stream_of(w->prob->ctx)));
  return QMC_OK;
}

// One set-up launch for the recorded work of several walker sets.
static int launch_setups(qmc_ctx c, const std::vector<SetupDesc>& d, int max_stride) {
  if (d.empty()) return QMC_OK;
  if (c->setup_descs.n < d.size()) {
    HIP_TRY(sync_stream(c));  // a launch may still read the old array
    HIP_TRY(c->setup_descs.alloc(d.size()));
  }
  HIP_TRY(upload_async(c, c->setup_descs.p, d.data(), d.size() * sizeof(SetupDesc)));
  HIP_TRY(launch_setup(c->setup_descs.p, (int)d.size(), max_stride, stream_of(c)));
  return QMC_OK;
}

static int ensure_scratch(qmc_walkers w);

// Performs what qmc_walkers_init and a plan recorded for this walker set
// (or only the clearing); every entry point that touches walker memory on
// the device calls this first.
static int materialize(qmc_walkers w) {
  if (!w->pend_init) return ensure_cleared(w);
  qmc_problem p = w->prob;
  const CtxStream st{p->ctx};
  const UnitDev* du = nullptr;
  int rc = device_units(w, &du);
  if (rc) return rc;
  if (setup_fits(w->stride)) {
    std::vector<SetupDesc> d;
    add_setup_descs(w, du, d);
    if ((rc = launch_setups(p->ctx, d, w->stride))) return rc;
    return setup_done(w);
  }
  // states too large for the one-launch set-up
  if ((rc = ensure_cleared(w)) || (rc = ensure_scratch(w))) return rc;
  HIP_TRY(launch_init(du, w->n_units, p->N, w->pend_epoch, st));
  HIP_TRY(launch_unpack(du, w->B, 0, w->R, p->N, 0, w->scratch_states.p, st));
  HIP_TRY(launch_energy(p->dev(), w->scratch_states.p, w->R, w->cost.p, st, w->best_cost.p));
  const bool prog = w->pend_prog;
  const int64_t s0 = w->pend_sweep0, n = w->pend_n;
  if ((rc = setup_done(w))) return rc;
  return prog ? set_progress(w, s0, n) : QMC_OK;
}

static void drop_graph(qmc_walkers w) {
  if (w->graph_exec) (void)hipGraphExecDestroy(w->graph_exec);
  if (w->graph) (void)hipGraphDestroy(w->graph);
  w->graph_exec = nullptr;
  w->graph = nullptr;
}

static void walkers_release(qmc_walkers w) {
  if (!w || --w->refs > 0) return;
  qmc_problem p = w->prob;
  (void)sync_stream(p->ctx);
  drop_graph(w);
  delete w;
  problem_release(p);
}

void qmc_walkers_destroy(qmc_walkers w) { walkers_release(w); }

static int ensure_scratch(qmc_walkers w) {
  const size_t need = (size_t)w->cap * w->prob->N;
  if (w->scratch_states.n < need) HIP_TRY(w->scratch_states.alloc(need));
  if (w->energies.n < (size_t)w->cap) HIP_TRY(w->energies.alloc(w->cap));
  return QMC_OK;
}

int qmc_walkers_resample(qmc_walkers w, const int32_t* src, int32_t new_R) {
  if (!w || (!src && new_R > 0)) return fail(QMC_INVALID_ARGUMENT, "null argument");
  if (w->pa_on) return fail(QMC_INVALID_ARGUMENT, "the population is on the device (qmc_walkers_pa_end first)");
  if (int rc0 = materialize(w)) return rc0;
  if (new_R < 1 || new_R > w->cap)
    return fail(QMC_MEMORY_LIMITED, "population exceeds the walker capacity");
  for (int32_t r = 0; r < new_R; ++r)
    if (src[r] < 0 || src[r] >= w->R) return fail(QMC_INVALID_ARGUMENT, "resample source out of range");
  qmc_problem p = w->prob;
  const CtxStream st{p->ctx};
  int rc = ensure_scratch(w);
  if (rc) return rc;
  const size_t n = (size_t)w->R * p->N;
  DevBuf<uint8_t> low;
  DevBuf<int32_t> dsrc;
  DevBuf<double> c0, c1, c2;
  DevBuf<unsigned long long> a0;
  HIP_TRY(low.alloc(std::max<size_t>(n, 1)));
  HIP_TRY(dsrc.upload(src, new_R));
  HIP_TRY(c0.alloc(new_R));
  HIP_TRY(c1.alloc(new_R));
  HIP_TRY(c2.alloc(new_R));
  HIP_TRY(a0.alloc(new_R));
  const UnitDev* du = nullptr;
  if ((rc = device_units(w, &du))) return rc;
  HIP_TRY(launch_unpack(du, w->B, 0, w->R, p->N, 0, w->scratch_states.p, st));
  HIP_TRY(launch_unpack(du, w->B, 0, w->R, p->N, 1, low.p, st));
  HIP_TRY(launch_gather_replicas(new_R, dsrc.p, w->cost.p, w->best_cost.p, w->beta.p, w->accepted.p,
                                 c0.p, c1.p, c2.p, a0.p, st));
  HIP_TRY(hipMemcpyAsync(w->cost.p, c0.p, new_R * sizeof(double), hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(w->best_cost.p, c1.p, new_R * sizeof(double), hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(w->beta.p, c2.p, new_R * sizeof(double), hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(w->accepted.p, a0.p, new_R * sizeof(unsigned long long), hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));
  w->R = new_R;
  w->n_units = (new_R + w->B - 1) / w->B;
  if ((rc = upload_units(w)) || (rc = device_units(w, &du))) return rc;
  HIP_TRY(launch_pack_gather(du, w->n_units, w->B, new_R, p->N, dsrc.p, w->scratch_states.p,
                             low.p, st));
  HIP_TRY(hipStreamSynchronize(st));
  return QMC_OK;
}

int qmc_walkers_count(qmc_walkers w, int32_t* n_replicas) {
  if (!w || !n_replicas) return fail(QMC_INVALID_ARGUMENT, "null argument");
  *n_replicas = w->R;
  return QMC_OK;
}

int qmc_walkers_pa_begin(qmc_walkers w, int32_t population, double alpha, double scale,
                         const double* beta_table, int32_t n_table, uint32_t seed,
                         int32_t sweeps_per_step) {
  if (!w || !beta_table) return fail(QMC_INVALID_ARGUMENT, "null argument");
  if (int rc0 = materialize(w)) return rc0;
  if (w->pa_on) return fail(QMC_INVALID_ARGUMENT, "population mode already on");
  if (w->R != w->cap) return fail(QMC_INVALID_ARGUMENT, "walkers must hold their full capacity");
  if (population < 1 || population > w->cap)
    return fail(QMC_INVALID_ARGUMENT, "population must be in 1..capacity");
  if (n_table < 2) return fail(QMC_INVALID_ARGUMENT, "beta_table needs at least 2 values");
  if (sweeps_per_step < 1) return fail(QMC_INVALID_ARGUMENT, "sweeps_per_step must be >= 1");
  qmc_problem p = w->prob;
  qmc_ctx c = p->ctx;
  // the population's work-groups, not the capacity's, are the ones that run
  w->live_units = std::min<int32_t>(w->n_units, (population + w->B - 1) / w->B);
  const int cap = w->cap;
  const size_t N = (size_t)p->N;
  HIP_TRY(w->pa.alloc(1));
  HIP_TRY(w->pa_table.alloc(n_table));
  HIP_TRY(w->pa_weight.alloc(cap));
  HIP_TRY(w->pa_tcost.alloc(cap));
  HIP_TRY(w->pa_tbest.alloc(cap));
  HIP_TRY(w->pa_tbeta.alloc(cap));
  HIP_TRY(w->pa_tacc.alloc(cap));
  HIP_TRY(w->pa_family.alloc(cap));
  HIP_TRY(w->pa_family_next.alloc(cap));
  HIP_TRY(w->pa_src.alloc(cap));
  HIP_TRY(w->pa_count.alloc(cap));
  HIP_TRY(w->pa_hist.alloc(cap));
  HIP_TRY(w->pa_plane_cur.alloc(std::max<size_t>((size_t)cap * N, 1)));
  HIP_TRY(w->pa_plane_low.alloc(std::max<size_t>((size_t)cap * N, 1)));
  HIP_TRY(w->pa_best_state.alloc(std::max<size_t>(N, 1)));
  PaDev st{};
  st.R = population;
  st.target = population;
  st.pending_inits = population;
  st.best_src = -1;
  st.cur_beta = beta_table[0];
  HIP_TRY(upload_async(c, w->pa.p, &st, sizeof st));
  HIP_TRY(upload_async(c, w->pa_table.p, beta_table, (size_t)n_table * sizeof(double)));
  // family of citizen i = i; every family id stays below the population
  std::vector<int32_t> fam(cap, 0);
  for (int i = 0; i < population; ++i) fam[i] = i;
  HIP_TRY(upload_async(c, w->pa_family.p, fam.data(), fam.size() * sizeof(int32_t)));
  std::vector<double> b(cap, beta_table[0]);
  HIP_TRY(upload_async(c, w->beta.p, b.data(), b.size() * sizeof(double)));
  PaArgs& a = w->pa_args;
  a = PaArgs{};
  a.pa = w->pa.p;
  a.beta_table = w->pa_table.p;
  a.n_table = n_table;
  a.target = population;
  a.cap = cap;
  a.n_vars = p->N;
  a.sweeps_per_step = sweeps_per_step;
  a.alpha = alpha;
  a.scale = scale;
  a.beta_start = beta_table[0];
  a.seed = seed;
  a.cost = w->cost.p;
  a.best_cost = w->best_cost.p;
  a.beta = w->beta.p;
  a.accepted = w->accepted.p;
  a.family = w->pa_family.p;
  a.family_next = w->pa_family_next.p;
  a.src = w->pa_src.p;
  a.hist = w->pa_hist.p;
  a.weight = w->pa_weight.p;
  a.count = w->pa_count.p;
  a.tcost = w->pa_tcost.p;
  a.tbest = w->pa_tbest.p;
  a.tbeta = w->pa_tbeta.p;
  a.tacc = w->pa_tacc.p;
  a.plane_cur = w->pa_plane_cur.p;
  a.plane_low = w->pa_plane_low.p;
  a.best_state = w->pa_best_state.p;
  a.B = w->B;
  a.strategy = QMC_PA_SCHEDULE;
  w->pa_on = 1;
  w->pa_stepped = false;
  return upload_units(w);
}

// The costs before the next step's sweeps kept on the device: now, or when
// a pending set-up computes them.
static int pa_keep_costs_before(qmc_walkers w) {
  if (!w->pa_cost0.p) HIP_TRY(w->pa_cost0.alloc(w->cap));
  if (!w->pend_init)
    HIP_TRY(hipMemcpyAsync(w->pa_cost0.p, w->cost.p, (size_t)w->cap * sizeof(double), hipMemcpyDeviceToDevice,
                           stream_of(w->prob->ctx)));
  w->pa_args.cost0 = w->pa_cost0.p;
  return QMC_OK;
}

int qmc_walkers_pa_observe(qmc_walkers w, int32_t max_steps) {
  if (!w) return fail(QMC_INVALID_ARGUMENT, "null handle");
  if (!w->pa_on) return fail(QMC_INVALID_ARGUMENT, "population mode is off");
  if (w->pa_stepped) return fail(QMC_INVALID_ARGUMENT, "observation starts before the first step");
  if (max_steps < 1) return fail(QMC_INVALID_ARGUMENT, "max_steps must be >= 1");
  if (int rc = pa_keep_costs_before(w)) return rc;
  HIP_TRY(w->pa_log.alloc((size_t)max_steps * kPaLogFields));
  ZeroSpans z{};
  z.p[0] = w->pa_log.p;
  z.n[0] = (size_t)max_steps * kPaLogFields * sizeof(double);
  HIP_TRY(launch_zero(z, 1, stream_of(w->prob->ctx)));
  w->pa_args.log = w->pa_log.p;
  w->pa_args.log_steps = max_steps;
  return QMC_OK;
}

int qmc_walkers_pa_log(qmc_walkers w, int32_t n_steps, double* records) {
  if (!w || (!records && n_steps > 0)) return fail(QMC_INVALID_ARGUMENT, "null argument");
  if (!w->pa_log.p) return fail(QMC_INVALID_ARGUMENT, "observation was not started");
  if (n_steps < 0 || n_steps > w->pa_args.log_steps) return fail(QMC_INVALID_ARGUMENT, "n_steps out of range");
  HIP_TRY(sync_stream(w->prob->ctx));
  if (n_steps > 0)
    HIP_TRY(hipMemcpy(records, w->pa_log.p, (size_t)n_steps * kPaLogFields * sizeof(double), hipMemcpyDeviceToHost));
  return QMC_OK;
}

int qmc_walkers_pa_strategy(qmc_walkers w, int32_t strategy, double constant, int32_t constant_population,
                            int32_t exp_repopulation) {
  if (!w) return fail(QMC_INVALID_ARGUMENT, "null handle");
  if (!w->pa_on) return fail(QMC_INVALID_ARGUMENT, "population mode is off");
  if (w->pa_stepped) return fail(QMC_INVALID_ARGUMENT, "the strategy is set before the first step");
  if (strategy < QMC_PA_SCHEDULE || strategy > QMC_PA_CONSTANT_CULLING)
    return fail(QMC_INVALID_ARGUMENT, "unknown resampling strategy");
  if (strategy != QMC_PA_SCHEDULE && !(constant > 0.0))
    return fail(QMC_INVALID_ARGUMENT, "the strategy constant must be greater than 0");
  PaArgs& a = w->pa_args;
  const int cap = w->cap;
  if (strategy == QMC_PA_FRICTION_TENSOR || strategy == QMC_PA_ENERGY_VARIANCE)
    if (int rc = pa_keep_costs_before(w)) return rc;
  int pow2 = 1;
  while (pow2 < cap) pow2 <<= 1;
  if (constant_population && pow2 > kPaTreeLds && !w->pa_tree.p) {
    HIP_TRY(w->pa_tree.alloc((size_t)pow2 + 1));
    HIP_TRY(w->pa_ids.alloc(cap));
  }
  a.tree = w->pa_tree.p;
  a.ids = w->pa_ids.p;
  a.strategy = strategy;
  a.strategy_const = strategy == QMC_PA_SCHEDULE ? 0.0 : constant;
  a.constant_population = constant_population ? 1 : 0;
  a.exp_repopulation = exp_repopulation ? 1 : 0;
  return QMC_OK;
}

int qmc_walkers_pa_step(qmc_walkers w) {
  if (!w) return fail(QMC_INVALID_ARGUMENT, "null handle");
  if (!w->pa_on) return fail(QMC_INVALID_ARGUMENT, "population mode is off");
  const UnitDev* du = nullptr;
  int rc = device_units(w, &du);
  if (rc) return rc;
  w->pa_args.units = du;
  w->pa_args.n_units = w->n_units;
  HIP_TRY(launch_pa_step(w->pa_args, w->prob->dev(), stream_of(w->prob->ctx)));
  w->pa_stepped = true;
  return QMC_OK;
}

int qmc_walkers_pa_info(qmc_walkers w, qmc_pa_info* info, uint8_t* best_state) {
  if (!w || !info) return fail(QMC_INVALID_ARGUMENT, "null argument");
  if (!w->pa.p) return fail(QMC_INVALID_ARGUMENT, "population mode was never started");
  const CtxStream st{w->prob->ctx};
  HIP_TRY(hipStreamSynchronize(st));
  PaDev d{};
  HIP_TRY(hipMemcpy(&d, w->pa.p, sizeof d, hipMemcpyDeviceToHost));
  info->population = d.R;
  info->restarts = d.restarts;
  info->overflow = d.err;
  info->have_best = d.have_best;
  info->steps = d.step;
  info->attempts = d.attempts;
  info->best = d.best;
  info->beta = d.cur_beta;
  info->rho = d.rho;
  info->delta_beta = d.delta_beta;
  info->target = d.target;
  info->evaluations = d.evals;
  if (best_state && w->prob->N > 0)
    HIP_TRY(hipMemcpy(best_state, w->pa_best_state.p, (size_t)w->prob->N, hipMemcpyDeviceToHost));
  return QMC_OK;
}

int qmc_walkers_pa_end(qmc_walkers w) {
  if (!w) return fail(QMC_INVALID_ARGUMENT, "null handle");
  if (!w->pa_on) return QMC_OK;
  HIP_TRY(sync_stream(w->prob->ctx));
  int32_t R = 0;
  HIP_TRY(hipMemcpy(&R, &w->pa.p->R, sizeof R, hipMemcpyDeviceToHost));
  if (R < 1 || R > w->cap) return fail(QMC_RUNTIME_ERROR, "device population size out of range");
  w->pa_on = 0;
  w->R = R;
  w->n_units = (R + w->B - 1) / w->B;
  drop_graph(w);
  return upload_units(w);
}

int qmc_walkers_init(qmc_walkers w, uint32_t epoch) {
  if (!w) return fail(QMC_INVALID_ARGUMENT, "null handle");
  // recorded; performed by the next launch that needs the state
  w->pend_init = true;
  w->pend_epoch = epoch;
  return QMC_OK;
}
int qmc_walkers_set_states(qmc_walkers w, const uint8_t* states) {
  if (!w || !states) return fail(QMC_INVALID_ARGUMENT, "null argument");
  if (int rc0 = materialize(w)) return rc0;
  qmc_problem p = w->prob;
  const CtxStream st{p->ctx};
  const UnitDev* du = nullptr;
  int rc = device_units(w, &du);
  if (rc || (rc = ensure_scratch(w))) return rc;
  const size_t n = (size_t)w->R * p->N;
  HIP_TRY(upload_async(p->ctx, w->scratch_states.p, states, n));
  HIP_TRY(launch_pack(du, w->n_units, w->B, w->R, p->N, w->scratch_states.p, 2, st));
  HIP_TRY(launch_energy(p->dev(), w->scratch_states.p, w->R, w->cost.p, st, w->best_cost.p));
  return QMC_OK;
}

int qmc_walkers_set_betas(qmc_walkers w, const double* betas) {
  if (!w || !betas) return fail(QMC_INVALID_ARGUMENT, "null argument");
  if (int rc0 = materialize(w)) return rc0;
  HIP_TRY(upload_async(w->prob->ctx, w->beta.p, betas, w->R * sizeof(double)));
  return QMC_OK;
}

// Which kernel a launch runs, for occupancy queries.
struct LaunchKernel {
  int kind, pairwise, coupling, row_e;
};

// Dynamic LDS of a sweep launch, plus room for the early fields of split
// ladders (pre_rounds rounds of colour class 0, see SweepArgs) when they fit
// the 160 KiB; fills the launch arguments that depend on it.  The lowest
// states stay in LDS unless keeping them in global memory lets more
// work-groups share a compute unit and the launch has more work-groups
// (live_units) than fit with them in LDS: large problems, where the third
// byte per position would leave one work-group per compute unit.
// Cooperative launches (pre_rounds, live_units 0) keep the placement their
// co-residency was checked with.
static size_t sweep_launch_smem(int stride, int ordered, int block, int nb, int pre_rounds, SweepArgs& a,
                                const LaunchKernel* k = nullptr, int64_t live_units = 0) {
  size_t smem = sweep_smem_bytes(stride, ordered, nb, block, &a.best_lds);
  if (a.best_lds && k && live_units > 0 && !ordered && pre_rounds == 0) {
    const int fit = max_coresident_blocks(k->kind, k->pairwise, k->coupling, nb, k->row_e, block, smem);
    if (fit > 0 && live_units > fit) {
      const size_t lean = sweep_smem_bytes(stride, ordered, nb, block, nullptr, 0);
      if (max_coresident_blocks(k->kind, k->pairwise, k->coupling, nb, k->row_e, block, lean) > fit) {
        smem = lean;
        a.best_lds = 0;
      }
    }
  }
  a.pre_rounds = 0;
  a.pre_off = 0;
  a.pre_draws = 0;
  if (pre_rounds > 0 && !ordered) {
    const size_t off = (smem + 15) & ~(size_t)15;
    // fields and draws, else fields only, else neither
    for (int draws = 1; draws >= 0; --draws) {
      const size_t bytes = sweep_pre_bytes(pre_rounds, block, nb, draws != 0);
      if (off + bytes <= 160 * 1024) {
        a.pre_rounds = pre_rounds;
        a.pre_off = (int32_t)off;
        a.pre_draws = draws;
        return off + bytes;
      }
    }
  }
  return smem;
}

// Rounds of colour class 0 of a split ladder's sweeps (0: no early fields --
// not a split ladder, or terms of more than two variables).
static int early_rounds(qmc_walkers w) {
  const qmc_problem p = w->prob;
  if (!p->pairwise || w->ladder <= 0 || w->ladder == w->B || p->class_off.size() < 3) return 0;
  return (p->class_off[1] - p->class_off[0] + w->lanes - 1) / w->lanes;
}

// ladder: the units hold parallel tempering ladders (exchange code needed)
static int launch_units(qmc_walkers w, const UnitDev* d_units, int block, int32_t max_sweeps, int ordered,
                        hipStream_t st, int coop, int ladder, int pre_rounds = 0) {
  const qmc_problem p = w->prob;
  if (uploads_pending(p->ctx)) return fail(QMC_DEVICE_ERROR, "internal: sweep launch before its uploads");
  const int nb = w->nb, n_units = w->n_units;
  SweepArgs a{};
  a.n_sweeps = max_sweeps;
  a.lut = p->ctx->lut.p;
  a.unit_of_block = nullptr;
  const LaunchKernel lk{p->kind, ordered ? p->pairwise : w->pair, p->coupling, w->row_e};
  const size_t smem = sweep_launch_smem(w->stride, ordered, block, nb, coop ? pre_rounds : 0, a, &lk,
                                        coop ? 0 : std::min(w->live_units, n_units));
  HIP_TRY(launch_sweep(p->kind, ordered ? p->pairwise : w->pair, p->coupling, nb, w->row_e, d_units, n_units, block,
                       smem, a, ordered, coop, ladder, st));
  return QMC_OK;
}

static int upload_schedule(qmc_walkers w, int64_t sweep0, int32_t n, const double* schedule) {
  qmc_ctx c = w->prob->ctx;
  std::shared_ptr<DevBuf<double>> buf;
  const bool cacheable = n > 0 && (size_t)n <= kSchedCacheMaxValues;
  if (cacheable) {
    std::lock_guard<std::mutex> lk(c->sched_mu);
    for (auto& e : c->sched_cache)
      if (e.values.size() == (size_t)n && std::memcmp(e.values.data(), schedule, n * sizeof(double)) == 0) {
        e.used = ++c->sched_clock;
        buf = e.dev;
        break;
      }
  }
  if (!buf) {
    // a fresh buffer: one a launch may still read is never rewritten
    buf = std::make_shared<DevBuf<double>>();
    HIP_TRY(buf->alloc(std::max<int32_t>(n, 1)));
    HIP_TRY(upload_async(c, buf->p, schedule, (size_t)n * sizeof(double)));
    if (cacheable) {
      std::lock_guard<std::mutex> lk(c->sched_mu);
      if (c->sched_cache.size() >= kSchedCacheEntries) {
        auto old = std::min_element(c->sched_cache.begin(), c->sched_cache.end(),
                                    [](const auto& x, const auto& y) { return x.used < y.used; });
        c->sched_cache.erase(old);
      }
      c->sched_cache.push_back({std::vector<double>(schedule, schedule + n), buf, ++c->sched_clock});
    }
  }
  w->sched = buf;
  w->sched_base = sweep0;
  return QMC_OK;
}

int qmc_walkers_plan(qmc_walkers w, int64_t sweep0, int32_t n_sweeps, const double* schedule) {
  if (!w) return fail(QMC_INVALID_ARGUMENT, "null handle");
  if (n_sweeps < 0) return fail(QMC_INVALID_ARGUMENT, "n_sweeps must be >= 0");
  // a pending set-up stays pending: it sets the progress recorded here
  if (!w->pend_init)
    if (int rc0 = materialize(w)) return rc0;
  int rc;
  w->ladder = 0;
  w->coop = 0;
  w->spr = 1;
  if (schedule) {
    if ((rc = upload_schedule(w, sweep0, n_sweeps, schedule))) return rc;
  } else {
    w->sched.reset();
  }
  if ((rc = upload_units(w))) return rc;
  if (w->pend_init) {
    w->pend_prog = true;
    w->pend_sweep0 = sweep0;
    w->pend_n = n_sweeps;
    w->pend_keep_ladder = false;
    w->planned = n_sweeps;
    return QMC_OK;
  }
  return set_progress(w, sweep0, n_sweeps);
}

static int set_ladder(qmc_walkers w, int32_t ladder, const double* ladder_betas, int64_t step0) {
  if (ladder < 1 || ladder > 32 || w->R % ladder != 0)
    return fail(QMC_INVALID_ARGUMENT, "ladder must divide the replica count and be <= 32");
  if (w->per_seed != ladder)
    return fail(QMC_INVALID_ARGUMENT, "create the walkers with replicas_per_seed == ladder");
  if (step0 == 0 || w->ladder != ladder) {
    // fresh ladder: slot k holds beta index k
    std::vector<int32_t> lab(w->R);
    std::vector<double> be(w->R);
    for (int r = 0; r < w->R; ++r) {
      lab[r] = r % ladder;
      be[r] = ladder_betas[r % ladder];
    }
    HIP_TRY(upload_async(w->prob->ctx, w->label.p, lab.data(), w->R * sizeof(int32_t)));
    HIP_TRY(upload_async(w->prob->ctx, w->beta.p, be.data(), w->R * sizeof(double)));
  }
  HIP_TRY(upload_async(w->prob->ctx, w->ladder_beta.p, ladder_betas, ladder * sizeof(double)));
  return QMC_OK;
}

int qmc_walkers_plan_pt(qmc_walkers w, int32_t ladder, const double* ladder_betas, int64_t step0,
                        int32_t n_steps, int32_t sweeps_per_replica) {
  if (!w || !ladder_betas) return fail(QMC_INVALID_ARGUMENT, "null argument");
  // A pending one-launch set-up stays pending (a batch runs those of all
  // its members in one launch): it keeps the labels and betas uploaded here
  // and sets the progress recorded here.
  const bool defer = w->pend_init && setup_fits(w->stride);
  if (!defer)
    if (int rc0 = materialize(w)) return rc0;
  if (ladder < 1 || ladder % w->B != 0)
    return fail(QMC_INVALID_ARGUMENT, "the ladder must be a multiple of replicas_per_group");
  if (sweeps_per_replica < 1) return fail(QMC_INVALID_ARGUMENT, "sweeps_per_replica must be >= 1");
  int rc = set_ladder(w, ladder, ladder_betas, step0);
  if (rc) return rc;
  const int n_ladders = w->R / ladder;
  const int upl = ladder / w->B;
  w->coop = upl > 1;
  if (w->coop) {
    // ladders split over work-groups exchange inside a cooperative launch;
    // the barrier counter holds steps * work-groups-per-ladder
    if (w->sync.n < (size_t)n_ladders) HIP_TRY(w->sync.alloc(n_ladders));
    if (w->cbuf.n < (size_t)n_ladders * 3 * ladder) HIP_TRY(w->cbuf.alloc((size_t)n_ladders * 3 * ladder));
    // every energy slot starts "not yet published"
    std::vector<double> nan((size_t)n_ladders * 3 * ladder, std::numeric_limits<double>::quiet_NaN());
    HIP_TRY(upload_async(w->prob->ctx, w->cbuf.p, nan.data(), nan.size() * sizeof(double)));
    std::vector<int32_t> ctr(n_ladders, (int32_t)(step0 * upl));
    HIP_TRY(upload_async(w->prob->ctx, w->sync.p, ctr.data(), n_ladders * sizeof(int32_t)));
    const int32_t zero = 0;
    HIP_TRY(upload_async(w->prob->ctx, w->err.p, &zero, sizeof zero));
  }
  w->ladder = ladder;
  w->spr = sweeps_per_replica;
  w->sched.reset();
  if ((rc = upload_units(w))) return rc;
  if (defer) {
    w->pend_keep_ladder = true;
    w->pend_prog = true;
    w->pend_sweep0 = step0 * sweeps_per_replica;
    w->pend_n = (int64_t)n_steps * sweeps_per_replica;
    w->planned = w->pend_n;
    return QMC_OK;
  }
  return set_progress(w, step0 * sweeps_per_replica, (int64_t)n_steps * sweeps_per_replica);
}

static int check_barrier_error(qmc_walkers w) {
  int32_t e = 0;
  HIP_TRY(hipMemcpy(&e, w->err.p, sizeof e, hipMemcpyDeviceToHost));
  if (e) return fail(QMC_RUNTIME_ERROR, "ladder barrier timed out (work-groups not co-resident)");
  return QMC_OK;
}

int qmc_walkers_sweep(qmc_walkers w, int64_t sweep0, int32_t n_sweeps, int beta_mode,
                      const double* schedule, int ordered) {
  if (!w) return fail(QMC_INVALID_ARGUMENT, "null handle");
  if (int rc0 = materialize(w)) return rc0;
  if (n_sweeps <= 0) return QMC_OK;
  if (beta_mode == QMC_BETA_SCHEDULE && !schedule) return fail(QMC_INVALID_ARGUMENT, "schedule required");
  int rc = qmc_walkers_plan(w, sweep0, n_sweeps, beta_mode == QMC_BETA_SCHEDULE ? schedule : nullptr);
  if (rc) return rc;
  qmc_ctx c = w->prob->ctx;
  const UnitDev* du = nullptr;
  if ((rc = device_units(w, &du))) return rc;
  HIP_TRY(hipEventRecord(c->ev0, stream_of(c)));
  ordered = ordered || w->ordered;
  rc = launch_units(w, du, ordered ? 64 : w->lanes * w->G, n_sweeps, ordered, stream_of(c), 0, 0);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(c->ev1, stream_of(c)));
  HIP_TRY(sync_stream(c));
  HIP_TRY(hipEventElapsedTime(&w->last_ms, c->ev0, c->ev1));
  return QMC_OK;
}

int qmc_walkers_pt(qmc_walkers w, int32_t ladder, const double* ladder_betas, int64_t step0,
                   int32_t n_steps, int32_t sweeps_per_replica) {
  if (!w || !ladder_betas) return fail(QMC_INVALID_ARGUMENT, "null argument");
  if (int rc0 = materialize(w)) return rc0;
  if (sweeps_per_replica < 1) return fail(QMC_INVALID_ARGUMENT, "sweeps_per_replica must be >= 1");
  qmc_problem p = w->prob;
  qmc_ctx c = p->ctx;
  const CtxStream st{c};
  const int n_ladders = w->R / std::max(1, ladder);
  int rc;
  SweepArgs sa{};
  const int pre = (p->pairwise && w->B != ladder && p->class_off.size() >= 3)
                      ? (p->class_off[1] - p->class_off[0] + w->lanes - 1) / w->lanes : 0;
  const size_t smem = sweep_launch_smem(w->stride, 0, w->lanes * w->G, w->nb, pre, sa);
  const bool can_coop = w->B != ladder && ladder % w->B == 0 &&
      w->n_units <= max_coresident_blocks(p->kind, w->pair, p->coupling, w->nb, w->row_e, w->lanes * w->G, smem);
  if (!w->pt_obs.log && (w->B == ladder || (can_coop && !w->force_graph))) {
    if ((rc = qmc_walkers_plan_pt(w, ladder, ladder_betas, step0, n_steps, sweeps_per_replica))) return rc;
    const UnitDev* du = nullptr;
    if ((rc = device_units(w, &du))) return rc;
    HIP_TRY(hipEventRecord(c->ev0, st));
    if ((rc = launch_units(w, du, w->lanes * w->G, n_steps * sweeps_per_replica, 0, st, w->coop, 1,
                           early_rounds(w))))
      return rc;
    HIP_TRY(hipEventRecord(c->ev1, st));
    HIP_TRY(hipStreamSynchronize(st));
    if ((rc = check_barrier_error(w))) return rc;
  } else {
    if (ladder % w->B != 0) return fail(QMC_INVALID_ARGUMENT, "ladder must be a multiple of replicas_per_group");
    const bool fresh = (step0 == 0 || w->ladder != ladder);
    if ((rc = set_ladder(w, ladder, ladder_betas, step0))) return rc;
    w->coop = 0;
    w->ladder = 0;  // exchanges run as their own launches
    w->spr = 1;
    w->sched.reset();
    if ((rc = upload_units(w))) return rc;
    w->ladder = ladder;
    if ((rc = set_progress(w, step0 * sweeps_per_replica, (int64_t)n_steps * sweeps_per_replica))) return rc;
    if (w->last_step.n < (size_t)n_ladders) HIP_TRY(w->last_step.alloc(n_ladders));
    {
      std::vector<int64_t> ls(n_ladders, fresh ? -1 : step0 - 1);
      HIP_TRY(upload_async(c, w->last_step.p, ls.data(), n_ladders * sizeof(int64_t)));
    }
    // nothing is uploaded while the stream is being captured
    const UnitDev* du = nullptr;
    if ((rc = device_units(w, &du))) return rc;
    const int upl = ladder / w->B;
    const uint32_t seed0 = w->seed + (uint32_t)(w->first_replica / ladder);
    // Steps are captured once as a graph of K (sweep, exchange) pairs and
    // replayed; the kernels read their step from the device-side progress,
    // and replays past the planned steps do nothing.
    const int K = std::min<int32_t>(n_steps, 64);
    const int key[4] = {K, sweeps_per_replica, w->n_units, ladder};  // observation drops the graph
    if (!w->graph_exec || std::memcmp(key, w->graph_key, sizeof key) != 0) {
      drop_graph(w);
      HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < K; ++k) {
        rc = launch_units(w, du, w->lanes * w->G, sweeps_per_replica, 0, st, 0, 0);
        if (rc) {
          hipGraph_t g;
          (void)hipStreamEndCapture(st, &g);
          return rc;
        }
        HIP_TRY(launch_exchange(w->cost.p, w->beta.p, w->label.p, w->ladder_beta.p, n_ladders, ladder,
                                w->progress.p, upl, sweeps_per_replica, w->last_step.p, seed0,
                                w->stream, w->pt_obs, st));
      }
      HIP_TRY(hipStreamEndCapture(st, &w->graph));
      HIP_TRY(hipGraphInstantiate(&w->graph_exec, w->graph, nullptr, nullptr, 0));
      std::memcpy(w->graph_key, key, sizeof key);
    }
    HIP_TRY(hipEventRecord(c->ev0, st));
    for (int32_t done = 0; done < n_steps; done += K) HIP_TRY(hipGraphLaunch(w->graph_exec, st));
    HIP_TRY(hipEventRecord(c->ev1, st));
  }
  HIP_TRY(hipStreamSynchronize(st));
  HIP_TRY(hipEventElapsedTime(&w->last_ms, c->ev0, c->ev1));
  return QMC_OK;
}

int qmc_walkers_remaining(qmc_walkers w, int64_t* sweeps_left) {
  if (!w || !sweeps_left) return fail(QMC_INVALID_ARGUMENT, "null argument");
  if (int rc0 = materialize(w)) return rc0;
  std::vector<int64_t> pr(w->progress.n);
  HIP_TRY(sync_stream(w->prob->ctx));
  HIP_TRY(hipMemcpy(pr.data(), w->progress.p, pr.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
  int64_t m = 0;
  for (int u = 0; u < w->n_units; ++u) m = std::max(m, pr[2 * u + 1]);
  *sweeps_left = m;
  return QMC_OK;
}

// ---------------------------------------------------------------- batches
int qmc_batch_create(qmc_ctx ctx, qmc_batch* out) {
  if (!ctx || !out) return fail(QMC_INVALID_ARGUMENT, "null handle");
  auto* b = new qmc_batch_s();
  b->ctx = ctx;
  if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    delete b;
    return fail(QMC_DEVICE_ERROR, "hipEventCreate failed");
  }
  ctx->refs++;
  *out = b;
  return QMC_OK;
}

void qmc_batch_destroy(qmc_batch b) {
  if (!b) return;
  qmc_ctx c = b->ctx;
  (void)sync_stream(c);
  for (qmc_walkers w : b->members) walkers_release(w);
  delete b;
  ctx_release(c);
}

int qmc_batch_add(qmc_batch b, qmc_walkers w) {
  if (!b || !w) return fail(QMC_INVALID_ARGUMENT, "null handle");
  if (w->prob->ctx != b->ctx) return fail(QMC_INVALID_ARGUMENT, "walkers belong to another context");
  // pending set-ups of members run in one launch with the batch
  if (!w->pend_init)
    if (int rc0 = materialize(w)) return rc0;
  for (auto* m : b->members)
    if (m == w) return QMC_OK;
  b->members.push_back(w);
  w->refs++;
  b->seen_version.push_back(~0ull);
  b->dirty = true;
  return QMC_OK;
}

int qmc_batch_remove(qmc_batch b, qmc_walkers w) {
  if (!b || !w) return fail(QMC_INVALID_ARGUMENT, "null handle");
  for (size_t i = 0; i < b->members.size(); ++i)
    if (b->members[i] == w) {
      b->members.erase(b->members.begin() + i);
      b->seen_version.erase(b->seen_version.begin() + i);
      b->dirty = true;
      walkers_release(w);
      return QMC_OK;
    }
  return QMC_OK;
}

int qmc_batch_clear(qmc_batch b) {
  if (!b) return fail(QMC_INVALID_ARGUMENT, "null handle");
  for (qmc_walkers w : b->members) walkers_release(w);
  b->members.clear();
  b->seen_version.clear();
  b->dirty = true;
  return QMC_OK;
}

static int batch_rebuild(qmc_batch b) {
  bool stale = b->dirty;
  for (size_t i = 0; i < b->members.size(); ++i)
    if (b->seen_version[i] != b->members[i]->version) stale = true;
  if (!stale) return QMC_OK;
  // launches still running may read the descriptors released below
  HIP_TRY(sync_stream(b->ctx));
  b->groups.clear();
  b->rebuilds++;
  // Largest problems first, so the longest work-groups start earliest.
  std::vector<size_t> order(b->members.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t i, size_t j) {
    const qmc_walkers x = b->members[i], y = b->members[j];
    return (int64_t)x->prob->n_inc * x->B > (int64_t)y->prob->n_inc * y->B;
  });
  b->member_at.assign(b->members.size(), {nullptr, 0});
  for (size_t i : order) {
    qmc_walkers w = b->members[i];
    qmc_problem p = w->prob;
    // one group per kernel instantiation and work-group size, so no launch
    // carries idle threads for its smaller members
    const int block = w->ordered ? 64 : w->lanes * w->G;
    qmc_batch_s::Group* g = nullptr;
    for (auto& gg : b->groups)
      if (gg->kind == p->kind && gg->pairwise == (w->ordered ? p->pairwise : w->pair) && gg->coupling == p->coupling &&
          gg->nb == w->nb && gg->row_e == w->row_e && gg->coop == w->coop && gg->ladder == (w->ladder > 0) &&
          gg->block == block && gg->ordered == w->ordered)
        g = gg.get();
    if (!g) {
      b->groups.push_back(std::make_unique<qmc_batch_s::Group>());
      g = b->groups.back().get();
      g->kind = p->kind;
      g->pairwise = w->ordered ? p->pairwise : w->pair;
      g->coupling = p->coupling;
      g->nb = w->nb;
      g->row_e = w->row_e;
      g->coop = w->coop;
      g->ladder = w->ladder > 0;
      g->ordered = w->ordered;
      g->block = block;
    }
    g->stride = std::max(g->stride, w->stride);
    if (w->coop) g->pre_rounds = std::max(g->pre_rounds, early_rounds(w));
    g->work += (double)p->N * w->R * (double)w->planned;
    g->live += w->live_units;
    b->member_at[i] = {g, g->units.size()};
    g->units.insert(g->units.end(), w->units.begin(), w->units.end());
  }
  for (auto& g : b->groups) {
    HIP_TRY(g->d_units.upload(g->units.data(), g->units.size()));
    if (g->coop) {
      // launch chunks of at most the co-resident number of work-groups,
      // never splitting a ladder (its work-groups are adjacent)
      SweepArgs sa{};
      const size_t smem = sweep_launch_smem(g->stride, g->ordered, g->block, g->nb, g->pre_rounds, sa);
      const int cap = max_coresident_blocks(g->kind, g->pairwise, g->coupling, g->nb, g->row_e, g->block, smem);
      const int n = (int)g->units.size();
      g->chunk.assign(1, 0);
      g->chunk_err = false;
      while (g->chunk.back() < n) {
        const int start = g->chunk.back();
        int end = start;
        while (end < n && end + g->units[end].upl - start <= cap) end += g->units[end].upl;
        if (end == start) {
          g->chunk_err = true;
          break;
        }
        g->chunk.push_back(end);
      }
    }
  }
  for (size_t i = 0; i < b->members.size(); ++i) b->seen_version[i] = b->members[i]->version;
  b->dirty = false;
  return QMC_OK;
}

static int ensure_aux_streams(qmc_ctx c) {
  if (c->fork_ev) return QMC_OK;
  for (int i = 0; i < kAuxStreams; ++i) {
    HIP_TRY(hipStreamCreateWithFlags(&c->aux[i], hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&c->join_ev[i], hipEventDisableTiming));
  }
  HIP_TRY(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
  return QMC_OK;
}

static int launch_group(qmc_batch_s::Group& g, int32_t max_sweeps, const double* lut, hipStream_t st) {
  SweepArgs a{};
  a.n_sweeps = max_sweeps;
  a.lut = lut;
  a.unit_of_block = nullptr;
  const LaunchKernel lk{g.kind, g.pairwise, g.coupling, g.row_e};
  const size_t smem = sweep_launch_smem(g.stride, g.ordered, g.block, g.nb, g.coop ? g.pre_rounds : 0, a, &lk,
                                        g.coop ? 0 : g.live);
  if (smem > 160 * 1024) return fail(QMC_MEMORY_LIMITED, "batch LDS footprint exceeds 160 KiB");
  const int n = (int)g.units.size();
  if (!g.coop) {
    HIP_TRY(launch_sweep(g.kind, g.pairwise, g.coupling, g.nb, g.row_e, g.d_units.p, n, g.block, smem, a, g.ordered,
                         0, g.ladder, st));
    return QMC_OK;
  }
  // Cooperative ladders: the chunks planned when the group was built.
  if (g.chunk_err) return fail(QMC_MEMORY_LIMITED, "a ladder exceeds the co-resident work-group limit");
  for (size_t i = 0; i + 1 < g.chunk.size(); ++i) {
    const int start = g.chunk[i], end = g.chunk[i + 1];
    HIP_TRY(launch_sweep(g.kind, g.pairwise, g.coupling, g.nb, g.row_e, g.d_units.p + start, end - start,
                         g.block, smem, a, 0, 1, 1, st));
  }
  return QMC_OK;
}

// Enqueue one launch per group (max_sweeps sweeps per unit) on the context's
// streams; everything joins back into the main stream.
static int enqueue_groups(qmc_batch b, int32_t max_sweeps) {
  int rc;
  qmc_ctx c = b->ctx;
  std::vector<qmc_batch_s::Group*> gs;
  for (auto& g : b->groups)
    if (!g->units.empty()) gs.push_back(g.get());
  // Groups of different kernels or work-group sizes run concurrently on up
  // to 1 + kAuxStreams streams: largest first, each onto the least loaded
  // stream.  Cooperative groups stay on the main stream, alone, so their
  // co-residency is not shared with other kernels.
  bool coop_any = false;
  for (auto* g : gs) coop_any |= g->coop != 0;
  const int n_streams = (gs.size() > 1 && !coop_any) ? 1 + kAuxStreams : 1;
  if (n_streams > 1) {
    HIP_TRY(hipEventRecord(c->fork_ev, stream_of(c)));
    for (int i = 0; i < kAuxStreams; ++i) HIP_TRY(hipStreamWaitEvent(c->aux[i], c->fork_ev, 0));
  }
  std::stable_sort(gs.begin(), gs.end(), [](auto* x, auto* y) { return x->work > y->work; });
  std::vector<double> load(n_streams, 0.0);
  for (auto* g : gs) {
    const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[s] += g->work;
    const hipStream_t gs = s == 0 ? stream_of(c) : c->aux[s - 1];
    if (uploads_pending(c)) return fail(QMC_DEVICE_ERROR, "internal: batch launch before its uploads");
    if ((rc = launch_group(*g, max_sweeps, c->lut.p, gs))) return rc;
  }
  if (n_streams > 1)
    for (int i = 0; i < kAuxStreams; ++i) {
      HIP_TRY(hipEventRecord(c->join_ev[i], c->aux[i]));
      HIP_TRY(hipStreamWaitEvent(stream_of(c), c->join_ev[i], 0));
    }
  return QMC_OK;
}

// The recorded set-ups of all members in one launch on the main stream,
// through the batch's descriptor copies (current after batch_rebuild).
static int batch_materialize(qmc_batch b) {
  std::vector<SetupDesc> d;
  std::vector<qmc_walkers> done;
  int max_stride = 0;
  for (size_t i = 0; i < b->members.size(); ++i) {
    qmc_walkers w = b->members[i];
    if (!w->pend_init) {
      if (int rc = ensure_cleared(w)) return rc;
      continue;
    }
    const auto& at = b->member_at[i];
    if (!setup_fits(w->stride) || !at.first || b->seen_version[i] != w->version) {
      if (int rc = materialize(w)) return rc;
      continue;
    }
    add_setup_descs(w, at.first->d_units.p + at.second, d);
    max_stride = std::max(max_stride, w->stride);
    done.push_back(w);
  }
  if (int rc = launch_setups(b->ctx, d, max_stride)) return rc;
  for (qmc_walkers w : done)
    if (int rc = setup_done(w)) return rc;
  return QMC_OK;
}

int qmc_batch_run(qmc_batch b, int32_t max_sweeps) {
  if (!b) return fail(QMC_INVALID_ARGUMENT, "null handle");
  int rc = batch_rebuild(b);
  if (rc || (rc = batch_materialize(b))) return rc;
  qmc_ctx c = b->ctx;
  if ((rc = ensure_aux_streams(c))) return rc;
  if (!b->span_open) HIP_TRY(hipEventRecord(b->ev0, stream_of(c)));
  b->span_open = true;
  if ((rc = enqueue_groups(b, max_sweeps))) return rc;
  HIP_TRY(hipEventRecord(b->ev1, stream_of(c)));
  return QMC_OK;
}

static void drop_batch_graph(qmc_batch b) {
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
}

int qmc_batch_run_chunked(qmc_batch b, int32_t sweeps_per_launch) {
  if (!b || sweeps_per_launch < 1) return fail(QMC_INVALID_ARGUMENT, "bad argument");
  int rc = batch_rebuild(b);
  if (rc || (rc = batch_materialize(b))) return rc;
  qmc_ctx c = b->ctx;
  if ((rc = ensure_aux_streams(c))) return rc;
  int64_t longest = 0;
  for (qmc_walkers w : b->members) {
    if (w->coop) return fail(QMC_INVALID_ARGUMENT, "chunked runs take no split PT ladders");
    longest = std::max(longest, w->planned);
  }
  const int64_t launches = (longest + sweeps_per_launch - 1) / sweeps_per_launch;
  const int64_t key[3] = {sweeps_per_launch, launches, b->rebuilds};
  if (!b->graph_exec || !std::equal(key, key + 3, b->graph_key)) {
    drop_batch_graph(b);
    // The kernels read each unit's progress on the device, so one recorded
    // sequence of launches serves every replay without argument updates.
    HIP_TRY(hipStreamBeginCapture(stream_of(c), hipStreamCaptureModeThreadLocal));
    for (int64_t i = 0; i < launches; ++i) {
      rc = enqueue_groups(b, sweeps_per_launch);
      if (rc) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream_of(c), &g);
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
    }
    HIP_TRY(hipStreamEndCapture(stream_of(c), &b->graph));
    HIP_TRY(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    std::copy(key, key + 3, b->graph_key);
  }
  if (!b->span_open) HIP_TRY(hipEventRecord(b->ev0, stream_of(c)));
  b->span_open = true;
  HIP_TRY(hipGraphLaunch(b->graph_exec, stream_of(c)));
  HIP_TRY(hipEventRecord(b->ev1, stream_of(c)));
  return QMC_OK;
}

int qmc_batch_wait(qmc_batch b, double* ms) {
  if (!b) return fail(QMC_INVALID_ARGUMENT, "null handle");
  HIP_TRY(sync_stream(b->ctx));
  if (b->span_open) HIP_TRY(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
  b->span_open = false;
  if (ms) *ms = b->last_ms;
  for (qmc_walkers w : b->members)
    if (w->coop) {
      int rc = check_barrier_error(w);
      if (rc) return rc;
    }
  return QMC_OK;
}

int qmc_batch_read(qmc_batch b, double* lowest_energy, uint64_t* accepted, uint8_t* lowest_states) {
  if (!b) return fail(QMC_INVALID_ARGUMENT, "null handle");
  for (qmc_walkers w : b->members)
    if (int rc = materialize(w)) return rc;
  qmc_ctx c = b->ctx;
  std::vector<ReadDesc> desc;
  std::vector<int32_t> rep_member;
  desc.reserve(b->members.size());
  int64_t roff = 0, soff = 0;
  for (size_t i = 0; i < b->members.size(); ++i) {
    qmc_walkers w = b->members[i];
    qmc_problem p = w->prob;
    ReadDesc d{};
    if (i < b->member_at.size() && b->member_at[i].first && b->seen_version[i] == w->version) {
      d.units = b->member_at[i].first->d_units.p + b->member_at[i].second;
    } else {
      int rc = device_units(w, &d.units);
      if (rc) return rc;
    }
    d.soff = soff;
    d.roff = roff;
    d.B = w->B;
    d.R = w->R;
    d.N = p->N;
    d.kind = p->kind;
    d.M = p->M;
    d.term_off = p->d_term_off.p;
    d.term_var = p->d_term_var.p;
    d.term_c = p->d_term_c.p;
    rep_member.insert(rep_member.end(), (size_t)w->R, (int32_t)desc.size());
    desc.push_back(d);
    roff += w->R;
    soff += (int64_t)w->R * p->N;
  }
  if (roff == 0) return QMC_OK;
  HIP_TRY(sync_stream(c));  // pooled buffers below may be reused
  HIP_TRY(b->d_desc.upload(desc.data(), desc.size()));
  HIP_TRY(b->d_rep_member.upload(rep_member.data(), rep_member.size()));
  HIP_TRY(b->d_states.reserve((size_t)std::max<int64_t>(soff, 1)));
  HIP_TRY(b->d_energy.reserve((size_t)roff));
  HIP_TRY(b->d_acc.reserve((size_t)roff));
  HIP_TRY(launch_batch_read(b->d_desc.p, b->d_rep_member.p, roff, b->d_states.p, b->d_energy.p, b->d_acc.p,
                            stream_of(c)));
  if (lowest_energy)
    HIP_TRY(hipMemcpyAsync(lowest_energy, b->d_energy.p, roff * sizeof(double), hipMemcpyDeviceToHost, stream_of(c)));
  if (accepted)
    HIP_TRY(hipMemcpyAsync(accepted, b->d_acc.p, roff * sizeof(uint64_t), hipMemcpyDeviceToHost, stream_of(c)));
  if (lowest_states && soff)
    HIP_TRY(hipMemcpyAsync(lowest_states, b->d_states.p, (size_t)soff, hipMemcpyDeviceToHost, stream_of(c)));
  HIP_TRY(sync_stream(c));
  return QMC_OK;
}

int qmc_walkers_set_lanes(qmc_walkers w, int32_t lanes) {
  if (!w) return fail(QMC_INVALID_ARGUMENT, "null handle");
  if (lanes == 0) lanes = w->prob->lanes;
  if ((lanes != 64 && lanes != 128 && lanes != 256 && lanes != 512) || lanes * w->G > 512)
    return fail(QMC_INVALID_ARGUMENT, "lanes must be 64, 128, 256 or 512 with lanes * ceil(B/8) <= 512");
  w->lanes = lanes;
  drop_graph(w);
  return upload_units(w);
}

int qmc_walkers_set_sweep_order(qmc_walkers w, int32_t ordered) {
  if (!w) return fail(QMC_INVALID_ARGUMENT, "null handle");
  if (ordered && w->B > 64)
    return fail(QMC_INVALID_ARGUMENT, "index-order sweeps take at most 64 replicas per group");
  if (ordered && (w->ladder > 0 || w->coop))
    return fail(QMC_INVALID_ARGUMENT, "index-order sweeps take no replica-exchange ladder");
  w->ordered = ordered ? 1 : 0;
  drop_graph(w);
  return upload_units(w);
}

int qmc_walkers_pt_observe(qmc_walkers w, int32_t ladder, int64_t max_steps) {
  if (!w) return fail(QMC_INVALID_ARGUMENT, "null handle");
  if (ladder < 2 || ladder > 32 || w->R % ladder != 0)
    return fail(QMC_INVALID_ARGUMENT, "ladder must be 2..32 and divide the replica count");
  if (max_steps < 1) return fail(QMC_INVALID_ARGUMENT, "max_steps must be >= 1");
  if (int rc0 = materialize(w)) return rc0;
  const int n_ladders = w->R / ladder;
  const size_t n = (size_t)n_ladders * (size_t)max_steps * ladder * 4;
  HIP_TRY(w->pt_log.alloc(n));
  HIP_TRY(w->pt_acc_prev.alloc(w->cap));
  HIP_TRY(w->pt_dir.alloc((size_t)n_ladders * ladder));
  // accepted counts so far, direction labels up (0) at the start
  HIP_TRY(hipMemcpyAsync(w->pt_acc_prev.p, w->accepted.p, (size_t)w->cap * sizeof(unsigned long long),
                         hipMemcpyDeviceToDevice, stream_of(w->prob->ctx)));
  ZeroSpans z{};
  z.p[0] = w->pt_dir.p;
  z.n[0] = (size_t)n_ladders * ladder * sizeof(int32_t);
  z.p[1] = w->pt_log.p;
  z.n[1] = n * sizeof(double);
  HIP_TRY(launch_zero(z, 2, stream_of(w->prob->ctx)));
  w->pt_obs.log = w->pt_log.p;
  w->pt_obs.log_steps = max_steps;
  w->pt_obs.accepted = w->accepted.p;
  w->pt_obs.acc_prev = w->pt_acc_prev.p;
  w->pt_obs.dir = w->pt_dir.p;
  drop_graph(w);
  return QMC_OK;
}

int qmc_walkers_pt_log(qmc_walkers w, int64_t n_values, double* out) {
  if (!w || (!out && n_values > 0)) return fail(QMC_INVALID_ARGUMENT, "null argument");
  if (!w->pt_log.p) return fail(QMC_INVALID_ARGUMENT, "observation was not started");
  if (n_values < 0 || (size_t)n_values > w->pt_log.n) return fail(QMC_INVALID_ARGUMENT, "n_values out of range");
  HIP_TRY(sync_stream(w->prob->ctx));
  if (n_values > 0) HIP_TRY(hipMemcpy(out, w->pt_log.p, (size_t)n_values * sizeof(double), hipMemcpyDeviceToHost));
  return QMC_OK;
}

int qmc_walkers_get_lanes(qmc_walkers w, int32_t* lanes) {
  if (!w || !lanes) return fail(QMC_INVALID_ARGUMENT, "null argument");
  *lanes = w->lanes;
  return QMC_OK;
}

int qmc_walkers_set_pt_mode(qmc_walkers w, int mode) {
  if (!w || mode < 0 || mode > 1) return fail(QMC_INVALID_ARGUMENT, "bad argument");
  w->force_graph = mode;
  return QMC_OK;
}

int qmc_walkers_read(qmc_walkers w, double* cost, double* lowest, double* beta, uint64_t* accepted,
                     uint8_t* states, uint8_t* lowest_states) {
  if (!w) return fail(QMC_INVALID_ARGUMENT, "null handle");
  if (int rc0 = materialize(w)) return rc0;
  qmc_problem p = w->prob;
  const CtxStream st{p->ctx};
  HIP_TRY(hipStreamSynchronize(st));
  if (cost) HIP_TRY(hipMemcpy(cost, w->cost.p, w->R * sizeof(double), hipMemcpyDeviceToHost));
  if (lowest) HIP_TRY(hipMemcpy(lowest, w->best_cost.p, w->R * sizeof(double), hipMemcpyDeviceToHost));
  if (beta) HIP_TRY(hipMemcpy(beta, w->beta.p, w->R * sizeof(double), hipMemcpyDeviceToHost));
  if (accepted)
    HIP_TRY(hipMemcpy(accepted, w->accepted.p, w->R * sizeof(uint64_t), hipMemcpyDeviceToHost));
  if (states || lowest_states) {
    const UnitDev* du = nullptr;
    int rc = device_units(w, &du);
    if (rc || (rc = ensure_scratch(w))) return rc;
    const size_t n = (size_t)w->R * p->N;
    if (states) {
      HIP_TRY(launch_unpack(du, w->B, 0, w->R, p->N, 0, w->scratch_states.p, st));
      HIP_TRY(hipMemcpyAsync(states, w->scratch_states.p, n, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
    }
    if (lowest_states) {
      HIP_TRY(launch_unpack(du, w->B, 0, w->R, p->N, 1, w->scratch_states.p, st));
      HIP_TRY(hipMemcpyAsync(lowest_states, w->scratch_states.p, n, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
    }
  }
  return QMC_OK;
}

int qmc_walkers_read_replica(qmc_walkers w, int32_t replica, uint8_t* state, uint8_t* lowest_state) {
  if (!w) return fail(QMC_INVALID_ARGUMENT, "null handle");
  if (replica < 0 || replica >= w->R) return fail(QMC_INVALID_ARGUMENT, "replica out of range");
  if (int rc0 = materialize(w)) return rc0;
  qmc_problem p = w->prob;
  const CtxStream st{p->ctx};
  const UnitDev* du = nullptr;
  int rc = device_units(w, &du);
  if (rc || (rc = ensure_scratch(w))) return rc;
  for (int which = 0; which < 2; ++which) {
    uint8_t* out = which ? lowest_state : state;
    if (!out) continue;
    HIP_TRY(launch_unpack(du, w->B, replica, 1, p->N, which, w->scratch_states.p, st));
    HIP_TRY(hipMemcpyAsync(out, w->scratch_states.p, p->N, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  return QMC_OK;
}

int qmc_walkers_energy(qmc_walkers w, int lowest, double* energies) {
  if (!w || !energies) return fail(QMC_INVALID_ARGUMENT, "null argument");
  if (int rc0 = materialize(w)) return rc0;
  qmc_problem p = w->prob;
  const CtxStream st{p->ctx};
  const UnitDev* du = nullptr;
  int rc = device_units(w, &du);
  if (rc || (rc = ensure_scratch(w))) return rc;
  HIP_TRY(launch_unpack(du, w->B, 0, w->R, p->N, lowest ? 1 : 0, w->scratch_states.p, st));
  HIP_TRY(launch_energy(p->dev(), w->scratch_states.p, w->R, w->energies.p, st));
  HIP_TRY(hipMemcpyAsync(energies, w->energies.p, w->R * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return QMC_OK;
}

int qmc_walkers_labels(qmc_walkers w, int32_t* labels) {
  if (!w || !labels) return fail(QMC_INVALID_ARGUMENT, "null argument");
  if (int rc0 = materialize(w)) return rc0;
  HIP_TRY(sync_stream(w->prob->ctx));
  HIP_TRY(hipMemcpy(labels, w->label.p, w->R * sizeof(int32_t), hipMemcpyDeviceToHost));
  return QMC_OK;
}

int qmc_walkers_last_ms(qmc_walkers w, double* ms) {
  if (!w || !ms) return fail(QMC_INVALID_ARGUMENT, "null argument");
  *ms = w->last_ms;
  return QMC_OK;
}

int qmc_energy(qmc_problem p, int32_t n_states, const uint8_t* states, double* energies) {
  if (!p || !states || !energies || n_states < 0) return fail(QMC_INVALID_ARGUMENT, "bad argument");
  if (n_states == 0) return QMC_OK;
  const CtxStream st{p->ctx};
  DevBuf<uint8_t> ds;
  DevBuf<double> de;
  HIP_TRY(ds.upload(states, (size_t)n_states * p->N));
  HIP_TRY(de.alloc(n_states));
  HIP_TRY(launch_energy(p->dev(), ds.p, n_states, de.p, st));
  HIP_TRY(hipStreamSynchronize(st));
  HIP_TRY(hipMemcpy(energies, de.p, n_states * sizeof(double), hipMemcpyDeviceToHost));
  return QMC_OK;
}

int qmc_energy_noalloc(qmc_problem p, int32_t n_states, const uint8_t* states, double* energies) {
  if (!p || !states || !energies || n_states < 0) return fail(QMC_INVALID_ARGUMENT, "bad argument");
  if (n_states == 0) return QMC_OK;
  // states first (256-byte aligned end), then the energies
  const size_t state_bytes = ((size_t)n_states * (size_t)p->N + 255) & ~(size_t)255;
  const size_t need = state_bytes + (size_t)n_states * sizeof(double);
  if (need > kScratchBytes)
    return fail(QMC_INVALID_ARGUMENT, "qmc_energy_noalloc: " + std::to_string(n_states) + " states of " +
                                          std::to_string(p->N) + " variables need " + std::to_string(need) +
                                          " bytes of device scratch; the limit is " +
                                          std::to_string(kScratchBytes));
  const int dev = p->ctx->device;
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = (dev >= 0 && dev < kMaxDevices) ? g_scratch_ptr[dev] : nullptr;
  }
  if (!base) return fail(QMC_DEVICE_ERROR, "qmc_energy_noalloc: no scratch address for this device");
  std::lock_guard<std::mutex> use(g_scratch_use[dev]);
  HIP_TRY(hipSetDevice(dev));
  const CtxStream st{p->ctx};
  uint8_t* d_states = static_cast<uint8_t*>(base);
  double* d_energy = reinterpret_cast<double*>(d_states + state_bytes);
  // every byte the kernel reads is written by this copy, and every energy by
  // the kernel: nothing needs clearing between calls
  HIP_TRY(hipMemcpyAsync(d_states, states, (size_t)n_states * p->N, hipMemcpyHostToDevice, st));
  HIP_TRY(launch_energy(p->dev(), d_states, n_states, d_energy, st));
  HIP_TRY(hipMemcpyAsync(energies, d_energy, (size_t)n_states * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return QMC_OK;
}

int qmc_delta_all(qmc_problem p, int32_t n_states, const uint8_t* states, double* out) {
  if (!p || !states || !out || n_states < 0) return fail(QMC_INVALID_ARGUMENT, "bad argument");
  if (n_states == 0) return QMC_OK;
  const CtxStream st{p->ctx};
  DevBuf<uint8_t> ds;
  DevBuf<double> d;
  HIP_TRY(ds.upload(states, (size_t)n_states * p->N));
  HIP_TRY(d.alloc((size_t)n_states * p->N));
  HIP_TRY(launch_delta_all(p->dev(), n_states, ds.p, d.p, st));
  HIP_TRY(hipStreamSynchronize(st));
  HIP_TRY(hipMemcpy(out, d.p, (size_t)n_states * p->N * sizeof(double), hipMemcpyDeviceToHost));
  return QMC_OK;
}

int qmc_lds_bandwidth(qmc_ctx ctx, double* bytes_per_s) {
  if (!ctx || !bytes_per_s) return fail(QMC_INVALID_ARGUMENT, "null argument");
  int dev = 0, cus = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = 2 * cus, iters = 1 << 15;
  DevBuf<uint32_t> out;
  HIP_TRY(out.alloc((size_t)blocks * 1024));
  (void)launch_lds_probe(blocks, 64, out.p, stream_of(ctx));  // warm-up
  HIP_TRY(hipEventRecord(ctx->ev0, stream_of(ctx)));
  const double bytes = launch_lds_probe(blocks, iters, out.p, stream_of(ctx));
  HIP_TRY(hipEventRecord(ctx->ev1, stream_of(ctx)));
  HIP_TRY(sync_stream(ctx));
  if (bytes < 0) return fail(QMC_DEVICE_ERROR, "LDS probe launch failed");
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  *bytes_per_s = bytes / (ms * 1e-3);
  return QMC_OK;
}

int qmc_philox(qmc_ctx ctx, int32_t n, const uint32_t* ctr, const uint32_t key[2], uint32_t* out) {
  if (!ctx || !ctr || !key || !out || n < 0) return fail(QMC_INVALID_ARGUMENT, "bad argument");
  if (n == 0) return QMC_OK;
  DevBuf<uint32_t> dc, dout;
  HIP_TRY(dc.upload(ctr, (size_t)4 * n));
  HIP_TRY(dout.alloc((size_t)4 * n));
  HIP_TRY(launch_philox(n, dc.p, key[0], key[1], dout.p, stream_of(ctx)));
  HIP_TRY(sync_stream(ctx));
  HIP_TRY(hipMemcpy(out, dout.p, (size_t)16 * n, hipMemcpyDeviceToHost));
  return QMC_OK;
}

}  // extern "C"
