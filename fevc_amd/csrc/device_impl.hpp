// Internal device-engine state shared by the HIP translation units of the
// library (device.hip: solve path; smoother.hip: learned smoother).  Not part
// of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <chrono>
#include <exception>
#include <cstdio>
#include <string>
#include <type_traits>

#include "common.hpp"
#include "device.hpp"
#include "kernels_engine.hpp"

#define VK_HIP(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      ::vk::fail(::vk::kResourceLimit, "hip", std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace vk {
namespace gpu {


using dev::kBlock;

constexpr uint32_t kChunkEntries = 64;  // default CSR entries or aggregate members per chunk
constexpr int kPollSlots = 8;
constexpr int kPollLag = 2;
// finest rows x lanes from which one PCG iteration outlasts the host's enqueue
// of the next (then the poll lags one iteration)
constexpr uint64_t kLongIteration = 1ull << 20;

// Live and peak bytes of the engine's device buffers (all DBufs of the
// process), for the memory receipts.
struct DeviceBytes {
  static inline std::atomic<uint64_t> live{0}, peak{0};
  static void add(uint64_t b) {
    const uint64_t now = live.fetch_add(b) + b;
    uint64_t prev = peak.load();
    while (now > prev && !peak.compare_exchange_weak(prev, now)) {
    }
  }
  static void sub(uint64_t b) { live.fetch_sub(b); }
};

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf &operator=(DBuf &&o) noexcept { std::swap(p, o.p); std::swap(n, o.n); return *this; }
  ~DBuf() { release(); }
  void release() {
    if (p) {
      (void)hipFree(p);
      DeviceBytes::sub(std::max<size_t>(n, 1) * sizeof(T));
    }
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (p && count <= n) return;
    if (p) {
      VK_HIP(hipFree(p));
      DeviceBytes::sub(std::max<size_t>(n, 1) * sizeof(T));
    }
    p = nullptr;
    n = 0;
    VK_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
    n = count;
    DeviceBytes::add(std::max<size_t>(count, 1) * sizeof(T));
  }
  template <typename V>
  void upload(const std::vector<V> &v) {
    alloc(v.size());
    if (v.empty()) return;
    if constexpr (std::is_same_v<T, V>) {
      VK_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    } else {
      std::vector<T> c(v.begin(), v.end());
      VK_HIP(hipMemcpy(p, c.data(), c.size() * sizeof(T), hipMemcpyHostToDevice));
    }
  }
  // Ordered after the work already queued on `stream`; returns once the copy
  // has landed.
  void download(std::vector<T> &v, size_t count, hipStream_t stream) const {
    v.resize(count);
    if (count) VK_HIP(hipMemcpyAsync(v.data(), p, count * sizeof(T), hipMemcpyDeviceToHost, stream));
    VK_HIP(hipStreamSynchronize(stream));
  }
};

// Grid for `threads` threads, capped, rounded up to a multiple of the 8 XCDs
// (the segment kernels map block ranges per XCD).
inline unsigned grid_for(uint64_t threads, unsigned cap = VK_GRID_CAP) {
  uint64_t g = (threads + kBlock - 1) / kBlock;
  g = std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
  return static_cast<unsigned>((g + 7) / 8 * 8);
}
inline unsigned reduce_grid(uint64_t threads, unsigned cap = dev::kMaxReduceBlocks) { return grid_for(threads, cap); }

template <typename F>
inline void with_k(uint32_t k, F &&f) {
  switch (k) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    case 64: f(std::integral_constant<int, 64>{}); break;
    default: fail(kInvalidInput, "device", "batch lane count must be a power of two in [1, 64]");
  }
}

inline uint32_t lanes_for(uint32_t columns) {
  uint32_t k = 1;
  while (k < columns) k <<= 1;
  if (k > 64) fail(kInvalidInput, "device", "at most 64 right-hand sides per batch");
  return k;
}

struct DevChunks {
  DBuf<uint32_t> desc, mrange, ptr;
  uint32_t count = 0, n_multi = 0;
  void upload(const ChunkTable &t) {
    desc.upload(t.desc);
    mrange.upload(t.mrange);
    ptr.upload(t.ptr);
    count = static_cast<uint32_t>(t.chunks());
    n_multi = t.multi();
  }
  dev::Chunks view() const {
    return {reinterpret_cast<const uint4 *>(desc.p), mrange.p, ptr.p, count, n_multi};
  }
};

struct DevLevel {
  uint32_t n = 0, coarse = 0, repeat = 0;
  uint64_t nnz = 0;
  int terminal = 0;
  DBuf<uint32_t> col, lab;
  DBuf<double> val, inv;
  DBuf<float> val32, inv32;
  DBuf<double> linv;  // learned smoother inverse diagonal omega / d (smoothed levels)
  DBuf<float> linv32;
  DevChunks rows, aggs;
  DevChunks rows_g, aggs_g;  // the same with runs of short segments grouped (kSegmentGroup)
  std::vector<uint32_t> perm;  // host copy: renumbered -> original vertex
  DBuf<uint32_t> dperm;        // the same on the device
  std::vector<uint32_t> lab_host;  // host copy of lab (renumbered coarse vertex)
  // Cycle shape: the next level the tuned cycle visits after this one; when it
  // skips levels, the composite aggregation to it (clab, caggs).
  uint32_t next = 0;
  DBuf<uint32_t> clab;
  DevChunks caggs, caggs_g;
  // coarse column of every CSR entry, lab[col[k]] and clab[col[k]]
  DBuf<uint32_t> pcol, cpcol;
};

// Short segments per grouped chunk.  Operations with at least this many
// slices per chunk (dev::Shape<K>::V) use the grouped tables: each slice then
// walks its own short row, instead of all slices sharing one.
constexpr uint32_t kSegmentGroup = VK_SEG_GROUP;
template <int K>
constexpr bool grouped() { return dev::Shape<K>::V >= (int)kSegmentGroup; }

template <typename T>
struct LevelVecs {
  DBuf<T> braw, b, x, y, r;
  DBuf<double> mean;  // coarse mean of braw, filled by the parent's restriction
};

constexpr uint64_t kTagLeverage = 0x4c45564552414745ULL;
constexpr uint64_t kTagTarget = 0x0054415247455401ULL;

inline std::vector<double> to_renumbered(const std::vector<double> &in, const std::vector<uint32_t> &perm, uint32_t K) {
  std::vector<double> out(in.size());
  for (size_t p = 0; p < perm.size(); ++p)
    for (uint32_t j = 0; j < K; ++j) out[p * K + j] = in[(size_t)perm[p] * K + j];
  return out;
}
inline std::vector<double> from_renumbered(const std::vector<double> &in, const std::vector<uint32_t> &perm, uint32_t K) {
  std::vector<double> out(in.size());
  for (size_t p = 0; p < perm.size(); ++p)
    for (uint32_t j = 0; j < K; ++j) out[(size_t)perm[p] * K + j] = in[p * K + j];
  return out;
}

// Recorded passes for multi-sample launches.  While an engine records (its
// `rec` is set), every pass it would launch is kept instead: the functor's
// bytes, its grid, and the launcher of k_multi for the functor's type.  The
// key is the pass's position in a nested numbering (a segment operation, a
// V-cycle visit of a level and each of its iterations open a nested level),
// so the passes of several samples whose cycles differ in depth or repeats
// still line up: equal keys and types across samples become one k_multi
// launch, and merging in key order keeps every sample's own order.
using MultiLaunch = void (*)(const void *table, const uint32_t *ends, uint32_t count, uint32_t blocks,
                             hipStream_t stream);
template <class F>
void multi_launch(const void *table, const uint32_t *ends, uint32_t count, uint32_t blocks, hipStream_t stream) {
  hipLaunchKernelGGL((dev::k_multi<F>), blocks, F::kThreads, 0, stream, static_cast<const F *>(table), ends, count);
}
struct Record {
  std::vector<uint32_t> key;
  MultiLaunch launch = nullptr;  // also the functor type's identity
  uint32_t grid = 0, threads = 0, size = 0, align = 0;
  std::vector<uint8_t> bytes;
};
struct Recorder {
  std::vector<Record> records;
  std::vector<uint32_t> path{0};
  template <class F>
  void add(const F &f, uint64_t grid) {
    static_assert(std::is_trivially_copyable_v<F>, "recorded functors are copied as bytes");
    Record r;
    r.key = path;
    ++path.back();
    r.launch = &multi_launch<F>;
    r.grid = static_cast<uint32_t>(grid);
    r.threads = F::kThreads;
    r.size = sizeof(F);
    r.align = alignof(F);
    r.bytes.assign(reinterpret_cast<const uint8_t *>(&f), reinterpret_cast<const uint8_t *>(&f) + sizeof(F));
    records.push_back(std::move(r));
  }
  void enter() { path.push_back(0); }
  void leave() {
    path.pop_back();
    ++path.back();
  }
  void at(uint32_t i) { path.back() = i; }
  std::vector<Record> take() {
    std::vector<Record> out = std::move(records);
    records.clear();
    path.assign(1, 0);
    return out;
  }
};

// One engine's PCG call in a lockstep sample batch: its recorded setup (with
// the passes that produce its right-hand sides), the preconditioning step and
// iteration body of both cycle parities, and the tail (the centring of x and
// the passes that read the solution); the driver runs the calls waiting
// together (iteration it runs body[(it - 1) % 2], then pre[it % 2]) until no
// sample has an active lane, then the tails, and returns each call's lane
// states and launch counts.
struct PcgProgram {
  std::vector<Record> setup, pre[2], body[2], tail;
  hipStream_t stream = nullptr;    // the engine's stream (ordered before and after)
  const int *active = nullptr;     // the engine's any-lane-active flag
  const dev::ColState *st = nullptr;
  uint32_t lanes = 0, max_iterations = 0;
  std::vector<dev::ColState> out;
  uint64_t iterations = 0, cycles = 0;
  std::exception_ptr error;
  bool done = false;
};
// Blocks until the batch has run `prog` (rethrows its failure).
void lockstep_submit(Lockstep &l, PcgProgram &prog);

// Learned smoother state (kernels_smoother.hpp): the network's inputs and
// activations for every smoothed level, parameters with Adam moments, and the
// training V-cycle vectors (32 error columns).
struct SmootherLevel {
  uint32_t level = 0, n = 0;
  // activations of the input layer and the message-passing layers (h[0..L])
  // and the messages of each message-passing layer (m[t - 1] for layer t)
  DBuf<double> X, rs, h[kSmootherLayers + 1], m[kSmootherLayers], omega, gomega;
  // the solve path's BF16 inference: activations (two alternating), messages,
  // and the FP32 last layer
  DBuf<uint16_t> hb[2], mb;
  DBuf<float> h32;
};
struct TrainLevel {
  DBuf<double> braw, b, s, t1, t2, t3, y, ly, mean;
};
struct SmootherState {
  std::vector<SmootherLevel> L;
  std::vector<TrainLevel> T;
  DBuf<double> theta, grad, am, av, partial, dz, dh, dpre, dms, dtmp;
  DBuf<uint16_t> wcat;  // BF16 layer weights of the inference path
  DBuf<double> err, ae, u, au, lam, q, uau;
  bool features = false, train_buffers = false;
  int mode = 0;             // 0 plain cycle, 1 learned weights with fallback
  bool calibrated = false;  // learned-vs-plain iteration check done
  bool active = false;      // learned inverse diagonals used by the solve path
  bool breakdown = false;   // a learned solve broke down and fell back
  bool shaped = false;      // the tuned cycle shape (levels skipped) is in use
  std::vector<uint32_t> shape;  // visited levels of the tuned shape
  // calibration iterations: reference cycle, learned weights, tuned shape,
  // both (UINT32_MAX where not tried)
  std::vector<uint32_t> calibration;
  uint32_t plain_iterations = 0, learned_iterations = 0;
  std::vector<double> params;  // parameters of the learned weights in use
};

}  // namespace gpu

using namespace gpu;

struct DeviceEngine::Impl {
  hipStream_t stream = nullptr;
  const Prepared *P = nullptr;
  const HybridSystem *HS = nullptr;
  const Hierarchy *H = nullptr;
  uint32_t chunk = kChunkEntries;
  uint64_t W = 0, F = 0, C = 0, G = 0, S = 0, n0 = 0, aux = 0;
  // units [0, Gm) are match units, [Gm, G) observation units (Go of them, the
  // last Go stored rows, obs_physical physical copies)
  uint64_t Gm = 0, Go = 0, obs_physical = 0;
  double target_total = 0, bound = 0;
  DBuf<uint64_t> worker_ptr, firm_ptr, cell_unit_ptr, cell_stratum_ptr, aux_ptr;
  DBuf<uint32_t> firm_cells, cell_worker, cell_firm, cell_units, unit_cell, aux_nbr;
  // when every cell is one unit (match deletion), else empty: the unit of
  // each cell in firm_cells order
  DBuf<uint32_t> firm_units, cell_unit;  // ... and of each cell in cell order
  // the worker and the weight of each cell in firm_cells order; the worker
  // and the firm of each unit
  DBuf<uint32_t> firm_worker, unit_worker, unit_firm;
  DBuf<double> firm_weight;
  DBuf<double> cell_weight, wdiag, fdiag, stratum_per_copy, cell_target, unit_outcome, aux_w;
  DBuf<uint64_t> unit_rank, unit_count, stratum_rank, stratum_count;
  // observation units (null without them): per unit the first copy in its
  // class (all units; zero for match units); per observation unit its first
  // physical copy (prefix of frequencies) and outcome; per observation unit
  // the probe sums of p^2 and p^4 and per physical copy those of s p and s p^3
  // (sum, correction); the observation adjustment; the per-unit correction
  // weight contributions of both kinds
  DBuf<uint64_t> unit_offset, row_phys;
  DBuf<double> row_outcome, obs_m, obs_corr, obs_adjust, contrib;
  // numerical attachment of observation units: copy pullbacks (q1, q3 per
  // copy) and the Counter-V1 words one replay column reads (over classes)
  DBuf<double> ncopy, nograd, nofolds, noscores;
  uint64_t obs_class_words = 0;
  std::vector<DevLevel> levels;
  DBuf<uint32_t> perm0, pos0;  // level-0 renumbering
  DevChunks firm_chunks;       // firms over their cells (firm_cells order)
  DBuf<double> ginv;
  // reductions and segment scratch
  DBuf<double> rpartial, bpartial, seg_partial;
  uint64_t max_chunks = 0;
  uint32_t max_multi = 0;
  DBuf<int> flags;
  const int *gate = nullptr;
  int *host_flags = nullptr;
  // pinned host copies of a batch's lane states and gate figures (reduced
  // residual, complete residual, right-hand side norm per lane), read with one
  // synchronisation after the extraction
  dev::ColState *host_states = nullptr;
  double *host_gates = nullptr;
  hipEvent_t poll_events[kPollSlots] = {};
  uint32_t cap = 0;
  std::vector<LevelVecs<double>> lv64;
  std::vector<LevelVecs<float>> lv32;
  DBuf<double> b0, proj, x, r, rc, z, p, ap, guess_rhs;
  DBuf<dev::ColState> st;
  DBuf<uint8_t> use;
  DBuf<double> wrhs, frhs, worker, red, fsol, sch, wsum, sout;
  DBuf<double> lane_a, lane_b, lane_c, reduced_res, complete_res, rhs_norm, sums;
  DBuf<int32_t> atoms;
  DBuf<double> moments, adjust, cellw, dir, totals, draws, fitted;
  // numerical MCSE attachment: unit gradients (5 x G), offsets, fold sums and
  // errors (12 x G), state partials, replay scores (6 per lane)
  DBuf<double> ngrad, noffs, nfolds, npart, nscores;
  // cap the first PCG pass of a batch at 0.1 gate ||b|| (see solve_columns)
  bool first_pass_cap = true;
  // controls, uploaded at the first solve that uses them: rows x q controls,
  // frequency, outcome, the row -> cell map and the unit row lists; per solve
  // the FE residuals of the controls and the outcome (rows x (q + 1)), the
  // joint residual, the whitened control rows, per unit (a, b) (units x 2q)
  // and (sum f e, B), per cell [Zc | hc] (cells x 2q)
  bool ctl_ready = false;
  uint32_t ctl_block = 0;           // largest block dimension min(rows of a unit, 1 + q)
  std::vector<double> ctl_gram0;    // sum f z z' (q x q)
  std::vector<double> ctl_s_inv, ctl_m2;
  DBuf<double> ctl_z, ctl_f, ctl_y, ctl_zry, ctl_e, ctl_u, ctl_uab, ctl_umass, ctl_cellmat, ctl_cellz;
  DBuf<double> ctl_small, ctl_part, ctl_out;
  DBuf<uint32_t> ctl_row_cell, ctl_unit_rows;
  DBuf<uint64_t> ctl_unit_ptr;
  // first per-unit gate failure (dev::StatusSink key; all ones when none)
  DBuf<unsigned long long> status;
  uint64_t status_launch = 0;
  bool timing = false;
  // HIP events around the launches of the timed kernels in a timed run:
  // [0] the finest-level direction and A p of the Krylov iteration (pass 1
  // of OpDirectionCurvature), [1] the finest prolong + smooth pass of the
  // V-cycle; with
  // their algorithmic bytes per launch
  struct TimedKernel {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    double bytes = 0;
  };
  TimedKernel timed[2];
  std::pair<hipEvent_t, hipEvent_t> timed_pair(int which, double bytes) {
    if (!timing || rec) return {nullptr, nullptr};
    hipEvent_t a = nullptr, e = nullptr;
    VK_HIP(hipEventCreate(&a));
    VK_HIP(hipEventCreate(&e));
    timed[which].events.push_back({a, e});
    timed[which].bytes = bytes;
    return {a, e};
  }
  void drop_timing() {
    for (auto &t : timed) {
      for (auto &e : t.events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
      }
      t.events.clear();
    }
  }
  uint64_t vcycles = 0;
  // diagnostics: solve batches, refinement rounds, refined columns, PCG iterations launched
  uint64_t batches = 0, refine_rounds = 0, refined_columns = 0, iterations_launched = 0;
  SmootherState sm;
  int precision_hint = 32;  // V-cycle precision of the calibration batches
  // HIP graphs of the PCG iteration: the body (A p .. iteration limit) and
  // the preconditioning sequence (lane means, V-cycle, rho, direction), keyed
  // by a fingerprint of every pointer, count and flag their kernels read.
  bool use_graphs = false;
  struct PcgGraphs {
    hipGraphExec_t body = nullptr, pre = nullptr;
  };
  std::map<uint64_t, PcgGraphs> pcg_graphs;
  uint64_t graph_captures = 0, graph_launches = 0;
  void drop_graphs() {
    for (auto &g : pcg_graphs) {
      if (g.second.body) (void)hipGraphExecDestroy(g.second.body);
      if (g.second.pre) (void)hipGraphExecDestroy(g.second.pre);
    }
    pcg_graphs.clear();
  }

  // recording of passes for a multi-sample launch (null: launch directly),
  // and the sample batch whose PCG calls run in lockstep (null: alone)
  Recorder *rec = nullptr;
  Lockstep *lockstep = nullptr;

  const int *always() const { return flags.p + dev::kFlagOne; }

  dev::SegScratch scratch() const { return {seg_partial.p, bpartial.p, gate}; }

  void ensure(uint32_t lanes) {
    if (lanes <= cap) return;
    cap = lanes;
    lv64.clear();
    lv32.clear();
    lv64.resize(levels.size());
    lv32.resize(levels.size());
    for (size_t l = 0; l < levels.size(); ++l) {
      const size_t m = (size_t)levels[l].n * lanes;
      for (auto *v : {&lv64[l].x, &lv64[l].y, &lv64[l].r}) v->alloc(m);
      for (auto *v : {&lv32[l].braw, &lv32[l].b, &lv32[l].x, &lv32[l].y, &lv32[l].r}) v->alloc(m);
      if (l > 0) { lv64[l].braw.alloc(m); lv64[l].b.alloc(m); }
      lv64[l].mean.alloc(64);
      lv32[l].mean.alloc(64);
    }
    const size_t m0 = (size_t)n0 * lanes;
    for (auto *v : {&b0, &proj, &x, &r, &rc, &z, &p, &ap, &guess_rhs}) v->alloc(m0);
    wrhs.alloc(W * lanes); frhs.alloc(F * lanes); worker.alloc(W * lanes); wsum.alloc(W * lanes);
    red.alloc(F * lanes); fsol.alloc(F * lanes); sch.alloc(F * lanes); sout.alloc(F * lanes);
    seg_partial.alloc(std::max<uint64_t>(max_chunks, 1) * lanes);
  }

  // a sink for one launch of a gated kernel; launches are numbered in issue order
  dev::StatusSink sink() { return {status.p, ++status_launch}; }
  // clears the failure word (each solve starts without a recorded failure)
  void reset_status() {
    VK_HIP(hipMemsetAsync(status.p, 0xff, sizeof(unsigned long long), stream));
    status_launch = 0;
  }
  void check_status(const char *phase) {
    unsigned long long h = ~0ull;
    VK_HIP(hipMemcpyAsync(&h, status.p, sizeof h, hipMemcpyDeviceToHost, stream));
    VK_HIP(hipStreamSynchronize(stream));
    if (h != ~0ull) {
      const int code = (int)(h & 0xff);
      const uint64_t unit = (h >> 8) & 0xffffffffull;
      fail(code, phase, "device gate failed with status " + std::to_string(code) + " at zero-based unit " +
                            std::to_string(unit));
    }
  }
};

namespace gpu {

// Tuned cycle shape: the cycle visits the levels of `shape` (ascending, from 0
// to the terminal); between visited levels a and b > a + 1 it restricts and
// prolongs through the composite aggregation lab_{b-1}(... lab_a(i)), which
// is monotone in the renumbered order, so its member ranges are contiguous.
inline void set_cycle_shape(DeviceEngine::Impl &E, const std::vector<uint32_t> &shape) {
  for (DevLevel &D : E.levels) D.next = 0;
  for (size_t k = 0; k + 1 < shape.size(); ++k) {
    const uint32_t a = shape[k], b = shape[k + 1];
    if (b <= a || b >= E.levels.size()) fail(kInvariant, "cycle_shape", "levels must ascend within the hierarchy");
    DevLevel &D = E.levels[a];
    D.next = b;
    if (b == a + 1) continue;
    std::vector<uint32_t> lab = D.lab_host;
    for (uint32_t l = a + 1; l < b; ++l)
      for (uint32_t &x : lab) x = E.levels[l].lab_host[x];
    const uint32_t nb = E.levels[b].n;
    std::vector<uint32_t> ptr(nb + 1, 0);
    for (uint32_t x : lab) ptr[x + 1]++;
    for (uint32_t c = 0; c < nb; ++c) ptr[c + 1] += ptr[c];
    D.clab.upload(lab);
    D.cpcol.alloc(D.nnz);
    hipLaunchKernelGGL(dev::k_compose, grid_for(D.nnz), kBlock, 0, E.stream, D.nnz, D.col.p, D.clab.p, D.cpcol.p);
    const ChunkTable table = chunk_segments(ptr, E.chunk);
    D.caggs.upload(table);
    D.caggs_g.upload(chunk_segments(ptr, E.chunk, kSegmentGroup));
    E.max_chunks = std::max<uint64_t>(E.max_chunks, table.chunks());
    E.max_multi = std::max(E.max_multi, table.multi());
  }
  E.sm.shape = shape;
  if (E.cap) E.seg_partial.alloc(std::max<uint64_t>(E.max_chunks, 1) * E.cap);
}

// Scope in which the V-cycle uses the reference smoother (0.5 / d).
struct PlainCycle {
  DeviceEngine::Impl &E;
  bool saved;
  bool saved_shape;
  explicit PlainCycle(DeviceEngine::Impl &e) : E(e), saved(e.sm.active), saved_shape(e.sm.shaped) {
    E.sm.active = false;
    E.sm.shaped = false;
  }
  ~PlainCycle() {
    E.sm.active = saved;
    E.sm.shaped = saved_shape;
  }
};

using Impl = DeviceEngine::Impl;

}  // namespace gpu

// Plug-in components of worker and firm effects (jla.rs:286-364) and the
// relative accounting residual of a component vector.
void plugin_components(const Prepared &P, const std::vector<double> &alpha, const std::vector<double> &psi,
                       double out[4]);
double component_identity(const double c[4]);

namespace gpu {

// Launch functor f of kernel `kernel` on `grid` blocks on the engine's stream,
// or record it while the engine records.
template <class F>
void run(Impl &E, void (*kernel)(F), uint64_t grid, const F &f) {
  if (E.rec) {
    E.rec->add(f, grid);
    return;
  }
  hipLaunchKernelGGL(kernel, dim3((uint32_t)grid), dim3(F::kThreads), 0, E.stream, f);
}
// A nested level of the recorded numbering (no-op when not recording).
struct Nest {
  Recorder *r;
  explicit Nest(Impl &E) : r(E.rec) {
    if (r) r->enter();
  }
  ~Nest() {
    if (r) r->leave();
  }
  void at(uint32_t i) const {
    if (r) r->at(i);
  }
};

// Final per-lane reduction of `items` block partials (one block per lane).
template <int K, int NV, class Fin>
void finish(Impl &E, const double *partial, unsigned items, const Fin &fin, const int *gate = nullptr) {
  run(E, dev::k_finish<K, NV, Fin>, K, dev::Finish<K, NV, Fin>{partial, items, fin, gate});
}

// Launch a segment operation over `ch` with operator `op`: pass 1 over all
// chunks, pass 2 for multi-chunk segments, then the final reduction (in a
// recording, positions 0, 1 and 2 of a nested level whether or not pass 2
// runs, so samples with and without multi-chunk segments line up).
template <typename T, int K, int NV, class Op>
void segments_in(Impl &E, const DevChunks &ch, const Op &op, const dev::SegScratch &sc,
                 hipEvent_t before = nullptr, hipEvent_t after = nullptr, unsigned cap = dev::kMaxReduceBlocks) {
  constexpr int G = dev::Shape<K>::G;
  const unsigned grid1 = NV > 0 ? reduce_grid((uint64_t)ch.count * G, cap) : grid_for((uint64_t)ch.count * G);
  const unsigned grid2 = std::max(1u, std::min<unsigned>(ch.n_multi, dev::kMaxReduceBlocks));
  const Nest nest(E);
  if (before) VK_HIP(hipEventRecord(before, E.stream));
  run(E, dev::k_segments<T, K, NV, Op>, grid1, dev::SegPass<T, K, NV, Op>{ch.view(), op, sc});
  if (after) VK_HIP(hipEventRecord(after, E.stream));
  if (ch.n_multi) {
    nest.at(1);
    run(E, dev::k_segments_combine<T, K, NV, Op>, grid2,
        dev::SegCombine<T, K, NV, Op>{ch.view(), op, sc, NV > 0 ? grid1 : 0u});
  }
  if constexpr (NV > 0) {
    nest.at(2);
    finish<K, NV>(E, sc.bpartial, grid1 + (ch.n_multi ? grid2 : 0u), op, E.gate);
  }
}

// The same with the engine's own scratch (chunk partials and block partials).
template <typename T, int K, int NV, class Op>
void segments(Impl &E, const DevChunks &ch, const Op &op, hipEvent_t before = nullptr, hipEvent_t after = nullptr,
              unsigned cap = dev::kMaxReduceBlocks) {
  segments_in<T, K, NV, Op>(E, ch, op, E.scratch(), before, after, cap);
}

}  // namespace gpu
}  // namespace vk
