// Device engine: uploads the prepared problem and CMG hierarchy to HBM and
// runs the per-probe path of the leave-out estimator on the GPU:
//   Counter-V1 atoms -> exact right-hand sides -> Schur right-hand side ->
//   lockstep certified PCG with the CMG V-cycle -> extraction and the
//   complete-residual gate -> moments / target contraction.
// Reference flow: engine.rs:1060-1563, full_cmg.rs:809-1340, pcg.rs:1537-1909,
// preconditioner/cycle.rs:136-478.
//
// Every level is renumbered so that aggregates are contiguous (restriction is
// a contiguous segmented sum, prolongation labels are monotone); the level-0
// renumbering is undone at the firm/auxiliary boundary of the hybrid system.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <numeric>

#include "device_impl.hpp"
#include "kernels_controls.hpp"
#include "numerical.hpp"

namespace vk {
namespace {

// ------------------------------------------------------------------ V-cycle
template <typename T, int K>
struct Cycle {
  // The FP32 segment operations carry four probe lanes per thread (float4:
  // 16-byte loads and stores, four entries of a row walked per wavefront);
  // the scalar per-row values (matrix entries, inverse diagonals) stay float.
  static constexpr bool kPacked = std::is_same_v<T, float> && K >= 16;
  using VT = std::conditional_t<kPacked, float4, T>;
  static constexpr int KV = kPacked ? K / 4 : K;
  static VT *vp(T *p) { return reinterpret_cast<VT *>(p); }
  static const VT *vp(const T *p) { return reinterpret_cast<const VT *>(p); }

  Impl &E;
  std::vector<LevelVecs<T>> &V;
  const T *braw0;  // level-0 right-hand side
  T *b0;           // level-0 centred copy

  const T *val(const DevLevel &L) const {
    if constexpr (sizeof(T) == 4) return L.val32.p; else return L.val.p;
  }
  // smoother inverse diagonal: 0.5 / d, or the learned omega / d on smoothed
  // levels when the learned weights are in use (terminals stay plain)
  const T *inv(const DevLevel &L) const {
    const bool learned = E.sm.active && !L.terminal;
    if constexpr (sizeof(T) == 4) return learned ? L.linv32.p : L.inv32.p;
    else return learned ? L.linv.p : L.inv.p;
  }

  // cycle.rs:136-311 with the coarse centring of cycle.rs:459-478 folded into
  // the first kernel of the child level.
  void level(size_t l, uint32_t iterations, const double *mean) {
    DevLevel &L = E.levels[l];
    LevelVecs<T> &v = V[l];
    const T *braw = l ? v.braw.p : braw0;
    T *b = l ? v.b.p : b0;
    const uint64_t work = (uint64_t)L.n * K;
    if (L.terminal) {
      if (L.terminal == kDirect)
        run(E, dev::k_terminal_direct<T, K>, grid_for(dev::terminal_threads<K>(L.n)),
            dev::TerminalDirect<T, K>{L.n, E.ginv.p, braw, mean, b, v.x.p, E.gate});
      else
        run(E, dev::k_terminal_scale<T, K>, grid_for(work),
            dev::TerminalScale<T, K>{L.n, inv(L), braw, mean, b, v.x.p, E.gate});
      return;
    }
    // next level visited: l + 1, or with the tuned shape a deeper level
    // reached through the composite aggregation
    const uint32_t nx = (E.sm.shaped && L.next) ? L.next : (uint32_t)l + 1;
    const bool composite = nx != l + 1;
    const uint32_t *lab = composite ? L.clab.p : L.lab.p;
    const uint32_t *pcol = composite ? L.cpcol.p : L.pcol.p;
    constexpr bool g = grouped<KV>();
    const DevChunks &rows = g ? L.rows_g : L.rows;
    const DevChunks &aggs = composite ? (g ? L.caggs_g : L.caggs) : (g ? L.aggs_g : L.aggs);
    DevLevel &Lc = E.levels[nx];
    LevelVecs<T> &c = V[nx];
    // The packed FP32 cycle is never centred (mean == nullptr): b is braw
    // itself, and the producer of braw also wrote x0 = D^-1 braw into v.x.
    if (kPacked && mean) fail(kInvariant, "vcycle", "the packed FP32 cycle takes uncentred right-hand sides");
    const T *bb = kPacked ? braw : b;
    for (uint32_t it = 0; it < iterations; ++it) {
      const Nest visit(E);  // the passes of this iteration, numbered apart from other iterations'
      if (it == 0) {
        if constexpr (kPacked)
          segments<VT, KV, 0>(E, rows,
                              dev::OpFirstResidual<VT, KV>{{L.col.p, val(L)}, vp(braw), vp(v.x.p), vp(v.r.p)});
        else
          segments<VT, KV, 0>(E, rows, dev::OpScaleResidual<VT, KV>{{L.col.p, val(L)}, vp(braw), mean, inv(L),
                                                                        vp(b), vp(v.x.p), vp(v.r.p)});
      } else {
        segments<VT, KV, 0>(E, rows,
                            dev::OpSmooth<VT, KV>{{L.col.p, val(L)}, vp(bb), vp(v.x.p), inv(L), vp(v.y.p)});
        std::swap(v.x, v.y);
        segments<VT, KV, 0>(E, rows, dev::OpResidual<VT, KV>{{L.col.p, val(L)}, vp(bb), vp(v.x.p), vp(v.r.p)});
      }
      // The coarse right-hand side sums residuals of a Laplacian (zero row
      // sums) of a centred right-hand side, so its lane means vanish up to
      // rounding: the reference's coarse centring (cycle.rs:459-478) is a
      // rounding-level projection here and is not recomputed.
      const bool x0 = kPacked && !Lc.terminal;  // the child's first iterate, written with its braw
      segments<VT, KV, 0>(E, aggs, dev::OpAggregate<VT, KV>{vp(v.r.p), vp(c.braw.p), nullptr, Lc.n,
                                                            x0 ? inv(Lc) : nullptr, x0 ? vp(c.x.p) : nullptr});
      {
        const Nest child(E);
        level(nx, E.levels[nx - 1].repeat, nullptr);
      }
      if (l == 0) {
        // the finest prolong-smooth pass, timed with HIP events in a timed
        // run; algorithmic bytes: CSR once (values, columns, chunk
        // descriptors), inverse diagonal and labels, r and x read, y
        // written, the coarse correction read
        const double s = sizeof(T);
        const auto ev = E.timed_pair(1, (double)L.nnz * (s + 4) + 16.0 * rows.count + (s + 4) * L.n +
                                            K * s * (3.0 * L.n + Lc.n));
        segments<VT, KV, 0>(E, rows,
                            dev::OpProlongSmooth<VT, KV, true>{{pcol, val(L)}, vp(v.r.p), vp(v.x.p), vp(c.x.p),
                                                               inv(L), lab, vp(v.y.p)},
                            ev.first, ev.second);
      } else {
        segments<VT, KV, 0>(E, rows,
                            dev::OpProlongSmooth<VT, KV>{{pcol, val(L)}, vp(v.r.p), vp(v.x.p), vp(c.x.p), inv(L),
                                                         lab, vp(v.y.p)});
      }
      std::swap(v.x, v.y);
    }
  }
};

// z = M r for K lanes; r is centred by `mean_r` (lane means) into `rc`.
// z is the FP64 level-0 output, or with the FP32 cycle its FP32 output (the
// Krylov kernels read it directly).
struct CycleOut {
  const double *z64 = nullptr;
  const float *z32 = nullptr;
};
template <int K>
CycleOut apply_cycle_raw(Impl &E, const double *r, const double *mean_r, double *rc, bool fp32) {
  E.vcycles++;
  if (!fp32) {
    Cycle<double, K> cyc{E, E.lv64, r, rc};
    const Nest cycle(E);
    cyc.level(0, 1, mean_r);
    return {E.lv64[0].x.p, nullptr};
  }
  const uint64_t m0 = E.n0 * K;
  Cycle<float, K> cyc{E, E.lv32, E.lv32[0].braw.p, E.lv32[0].b.p};
  const bool x0 = Cycle<float, K>::kPacked;  // level 0's first iterate, written with its braw
  run(E, dev::k_centre_split<K>, grid_for(m0),
      dev::CentreSplit<K>{E.n0, r, mean_r, rc, E.lv32[0].braw.p, x0 ? cyc.inv(E.levels[0]) : nullptr,
                          x0 ? E.lv32[0].x.p : nullptr, E.gate});
  {
    const Nest cycle(E);
    cyc.level(0, 1, nullptr);
  }
  return {nullptr, E.lv32[0].x.p};
}
// As above with z always in FP64 (returns its device address).
template <int K>
const double *apply_cycle(Impl &E, const double *r, const double *mean_r, double *rc, bool fp32) {
  const CycleOut o = apply_cycle_raw<K>(E, r, mean_r, rc, fp32);
  if (o.z64) return o.z64;
  const uint64_t m0 = E.n0 * K;
  hipLaunchKernelGGL((dev::k_to_double<K>), grid_for(m0), kBlock, 0, E.stream, m0, o.z32, E.z.p, E.gate);
  return E.z.p;
}

// ------------------------------------------------------------- HIP graphs
// Fingerprint of everything the captured PCG kernels read as arguments or
// launch shape: widths, flags, tolerances, buffer addresses, level tables and
// their counts.  The V-cycle swaps each level's x and y buffers on the host
// while it enqueues, so the pair enters as (min, max): a captured sequence is
// self-contained (every cycle buffer it reads is written inside it), so one
// graph serves either assignment.
struct Fingerprint {
  uint64_t h = 1469598103934665603ULL;
  void add(uint64_t v) {
    for (int k = 0; k < 8; ++k) {
      h ^= (v >> (8 * k)) & 0xff;
      h *= 1099511628211ULL;
    }
  }
  void ptr(const void *p) { add((uint64_t)(uintptr_t)p); }
  void real(double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    add(u);
  }
  void chunks(const DevChunks &c) {
    ptr(c.desc.p);
    ptr(c.mrange.p);
    ptr(c.ptr.p);
    add(c.count);
    add(c.n_multi);
  }
  template <typename T>
  void pair(const DBuf<T> &a, const DBuf<T> &b) {
    ptr(std::min(a.p, b.p));
    ptr(std::max(a.p, b.p));
  }
};

template <int K>
uint64_t pcg_fingerprint(const Impl &E, const double *b, const dev::PcgParams &prm, bool fp32) {
  Fingerprint f;
  f.add(K);
  f.add(fp32);
  f.add(E.sm.active);
  f.add(E.sm.shaped);
  f.add(E.n0);
  f.add(prm.n);
  f.real(prm.rel);
  f.real(prm.abs);
  f.real(prm.cap);
  f.real(prm.bound);
  f.add((uint64_t)(int64_t)prm.max_iterations);
  f.add((uint64_t)(int64_t)prm.interval);
  for (const void *p : {(const void *)b, (const void *)E.proj.p, (const void *)E.x.p, (const void *)E.r.p,
                        (const void *)E.rc.p, (const void *)E.p.p, (const void *)E.ap.p, (const void *)E.z.p,
                        (const void *)E.st.p, (const void *)E.lane_a.p, (const void *)E.lane_b.p,
                        (const void *)E.flags.p, (const void *)E.rpartial.p, (const void *)E.bpartial.p,
                        (const void *)E.seg_partial.p, (const void *)E.use.p, (const void *)E.ginv.p,
                        (const void *)E.gate})
    f.ptr(p);
  for (const DevLevel &L : E.levels) {
    f.add(L.n);
    f.add(L.nnz);
    f.add(L.terminal);
    f.add(L.repeat);
    f.add(L.next);
    for (const void *p : {(const void *)L.col.p, (const void *)L.val.p, (const void *)L.val32.p,
                          (const void *)L.inv.p, (const void *)L.inv32.p, (const void *)L.linv.p,
                          (const void *)L.linv32.p, (const void *)L.lab.p, (const void *)L.clab.p,
                          (const void *)L.pcol.p, (const void *)L.cpcol.p})
      f.ptr(p);
    for (const DevChunks *c : {&L.rows, &L.rows_g, &L.aggs, &L.aggs_g, &L.caggs, &L.caggs_g}) f.chunks(*c);
  }
  for (const auto &v : E.lv32) {
    f.ptr(v.braw.p);
    f.ptr(v.b.p);
    f.ptr(v.r.p);
    f.ptr(v.mean.p);
    f.pair(v.x, v.y);
  }
  for (const auto &v : E.lv64) {
    f.ptr(v.braw.p);
    f.ptr(v.b.p);
    f.ptr(v.r.p);
    f.ptr(v.mean.p);
    f.pair(v.x, v.y);
  }
  return f.h;
}

// Record what `enqueue` launches on the engine stream into an executable graph.
template <class F>
hipGraphExec_t capture(Impl &E, F &&enqueue) {
  hipGraph_t g = nullptr;
  VK_HIP(hipStreamBeginCapture(E.stream, hipStreamCaptureModeThreadLocal));
  try {
    enqueue();
  } catch (...) {
    (void)hipStreamEndCapture(E.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  VK_HIP(hipStreamEndCapture(E.stream, &g));
  hipGraphExec_t x = nullptr;
  const hipError_t rc = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  VK_HIP(rc);
  E.graph_captures++;
  return x;
}

// One lockstep certified PCG call (pcg.rs:1537-1909) for K lanes on the
// finest hybrid Laplacian (renumbered): the setup passes, the
// preconditioning step (lane means, V-cycle, rho) and the iteration body
// (direction and A p, update, centring, fresh-residual gate, limit).  b:
// n0 x K right-hand sides; E.use holds the lanes that take part; with a warm
// start E.x holds the initial guesses.
template <int K>
struct PcgCall {
  Impl &E;
  const double *b;
  dev::PcgParams prm;
  bool fp32;
  const DevLevel &L0;
  const DevChunks &rows0;
  unsigned rg;
  int *flags;
  const int *active;
  // the preconditioner's output of the latest enqueued cycle: the cycle
  // alternates its level-0 buffers, so the body reads the address the last
  // precondition step produced
  CycleOut zcur{};

  PcgCall(Impl &e, const double *rhs, double rel, double abs_tol, uint32_t max_it, uint32_t interval, bool f32,
          double cap)
      : E(e), b(rhs), prm{(uint32_t)e.n0, rel, abs_tol, e.bound, (int32_t)max_it, (int32_t)interval, cap}, fp32(f32),
        L0(e.levels[0]), rows0(grouped<K>() ? e.levels[0].rows_g : e.levels[0].rows),
        rg(reduce_grid((uint64_t)e.n0 * K, dev::kKrylovReduceBlocks)), flags(e.flags.p),
        active(e.flags.p + dev::kFlagActive) {}

  void setup(bool warm) {
    const uint32_t n = E.n0;
    const uint64_t m0 = (uint64_t)n * K;
    run(E, dev::k_pcg_stats<K>, rg, dev::PcgStats<K>{n, b, E.rpartial.p});
    finish<K, 3>(E, E.rpartial.p, rg, dev::FinStats{E.st.p, E.use.p, n});
    E.gate = E.always();
    if (!warm) {
      run(E, dev::k_pcg_project<K>, grid_for(m0), dev::PcgProject<K>{n, b, E.proj.p, E.r.p, E.x.p, E.st.p});
      run(E, dev::k_pcg_start<K>, 1, dev::PcgStart<K>{E.st.p, prm, pub()});
    } else {
      run(E, dev::k_pcg_project<K>, grid_for(m0), dev::PcgProject<K>{n, b, E.proj.p, E.r.p, E.ap.p, E.st.p});
      run(E, dev::k_lane_sum<K>, rg, dev::LaneSum<K>{(uint64_t)n, E.x.p, E.rpartial.p});
      finish<K, 1>(E, E.rpartial.p, rg, dev::FinMean{E.lane_a.p, E.use.p, n});
      run(E, dev::k_center_cols<K>, grid_for(m0), dev::CenterCols<K>{(uint64_t)n, E.x.p, E.lane_a.p, nullptr});
      segments<double, K, 4>(E, rows0,
                             dev::OpWarm<K>{{L0.col.p, L0.val.p}, b, E.proj.p, E.x.p, E.r.p, E.st.p, prm, pub()},
                             nullptr, nullptr, dev::kKrylovReduceBlocks);
      run(E, dev::k_pcg_flags<K>, 1, dev::PcgFlags<K>{E.st.p, flags});
    }
    E.gate = active;
  }

  // the lane means and flags the state-changing passes write
  dev::LanePublish pub() const { return {E.lane_a.p, E.lane_b.p, flags}; }

  void precondition() {
    const uint32_t n = E.n0;
    // the FP32 cycle centres r into its own FP32 input, so no FP64 copy is
    // stored; the FP64 cycle uses it as its level-0 right-hand side
    zcur = apply_cycle_raw<K>(E, E.r.p, E.lane_b.p, fp32 ? nullptr : E.rc.p, fp32);
    if (zcur.z32)
      run(E, dev::k_pcg_rho<K, float>, rg,
          dev::PcgRho<K, float>{n, E.r.p, E.lane_b.p, zcur.z32, E.rpartial.p, active});
    else
      run(E, dev::k_pcg_rho<K, double>, rg,
          dev::PcgRho<K, double>{n, E.r.p, E.lane_b.p, zcur.z64, E.rpartial.p, active});
    finish<K, 3>(E, E.rpartial.p, rg, dev::FinRho{E.st.p, n}, active);
  }

  void body() {
    const uint32_t n = E.n0;
    const uint64_t m0 = (uint64_t)n * K;
    // direction and A p, the dominant kernel of the iteration, timed with HIP
    // events in a timed run; algorithmic bytes: CSR once (values, columns,
    // chunk descriptors), z read once in the cycle's precision, p and A p
    // each read and written once
    const double zb = zcur.z32 ? 4.0 : 8.0;
    const auto ev = E.timed_pair(0, 12.0 * (double)L0.nnz + 16.0 * rows0.count + (zb + 4.0 * 8.0) * (double)m0);
    if (zcur.z32)
      segments<double, K, 1>(E, rows0,
                             dev::OpDirectionCurvature<K, float>{{L0.col.p, L0.val.p}, zcur.z32, E.p.p, E.ap.p,
                                                                 E.st.p},
                             ev.first, ev.second, dev::kKrylovReduceBlocks);
    else
      segments<double, K, 1>(E, rows0,
                             dev::OpDirectionCurvature<K, double>{{L0.col.p, L0.val.p}, zcur.z64, E.p.p, E.ap.p,
                                                                  E.st.p},
                             ev.first, ev.second, dev::kKrylovReduceBlocks);
    // x is centred inside the next update (its mean after this one is kept
    // in the lane state) and on the fly by the fresh residual
    run(E, dev::k_pcg_update<K>, rg,
        dev::PcgUpdate<K>{n, E.p.p, E.ap.p, E.x.p, E.r.p, E.rpartial.p, E.st.p, E.lane_a.p, E.lane_b.p, active});
    finish<K, 4>(E, E.rpartial.p, rg, dev::FinUpdate{E.st.p, prm, pub()}, active);
    segments<double, K, 3>(E, rows0,
                           dev::OpFresh<K>{{L0.col.p, L0.val.p}, b, E.proj.p, E.x.p, E.r.p, E.st.p, prm, flags,
                                           pub()},
                           nullptr, nullptr, dev::kKrylovReduceBlocks);
    run(E, dev::k_pcg_limit<K>, 1, dev::PcgLimit<K>{E.st.p, prm, pub()});
  }

  // after the iterations: every lane's last update leaves x uncentred by its
  // mean, kept in the lane state
  void center() {
    run(E, dev::k_center_state<K>, grid_for((uint64_t)E.n0 * K), dev::CenterState<K>{E.n0, E.x.p, E.st.p});
  }
};

// The calls of a sample batch in lockstep: record this call's passes (the
// setup, then two preconditioning steps and bodies, one for each parity of
// the cycle's alternating buffers: two cycles swap every level's buffers an
// even number of times) and run them with the other samples' waiting calls.
template <int K>
void pcg_lockstep(Impl &E, PcgCall<K> &c, bool warm, uint32_t max_it, std::vector<dev::ColState> &out,
                  const std::function<void()> *head, const std::function<void()> *tail) {
  PcgProgram prog;
  Recorder rec;
  const uint64_t cycles = E.vcycles;
  E.rec = &rec;
  try {
    if (head) (*head)();
    c.setup(warm);
    prog.setup = rec.take();
    for (int parity = 0; parity < 2; ++parity) {
      c.precondition();
      prog.pre[parity] = rec.take();
      c.body();
      prog.body[parity] = rec.take();
    }
    E.gate = E.always();
    c.center();
    if (tail) (*tail)();
    prog.tail = rec.take();
  } catch (...) {
    E.rec = nullptr;
    throw;
  }
  E.rec = nullptr;
  E.vcycles = cycles;
  prog.stream = E.stream;
  prog.active = c.active;
  prog.st = E.st.p;
  prog.lanes = K;
  prog.max_iterations = max_it;
  lockstep_submit(*E.lockstep, prog);
  E.vcycles += prog.cycles;
  E.iterations_launched += prog.iterations;
  E.gate = E.always();
  out = prog.out;
}

// Lockstep certified PCG on the finest hybrid Laplacian: with `warm`, E.x
// holds the initial guesses.  Alone, the host polls the any-lane-active flag
// with a lag of kPollLag iterations and iterations queued after convergence
// return at once; in a lockstep sample batch the call joins the others'.
// `head` enqueues passes that produce b before the call; `tail` passes that
// follow it (after x is centred).
template <int K>
void pcg_lanes(Impl &E, const double *b, double rel, double abs_tol, uint32_t max_it, uint32_t interval, bool warm,
               bool fp32, const std::vector<uint8_t> &use, std::vector<dev::ColState> &out, double cap = 0.0,
               bool states = true, const std::function<void()> *head = nullptr,
               const std::function<void()> *tail = nullptr) {
  VK_HIP(hipMemcpyAsync(E.use.p, use.data(), K, hipMemcpyHostToDevice, E.stream));
  PcgCall<K> c(E, b, rel, abs_tol, max_it, interval, fp32, cap);
  if (E.lockstep) {
    pcg_lockstep<K>(E, c, warm, max_it, out, head, tail);
    return;
  }
  if (head) (*head)();
  c.setup(warm);
  const dev::PcgParams &prm = c.prm;
  auto precondition = [&]() { c.precondition(); };
  auto body = [&]() { c.body(); };
  // With graphs, each iteration launches the two captured sequences; the flag
  // copy and the host poll stay between them, as without graphs.  Timed runs
  // keep individual launches (their per-launch events).
  Impl::PcgGraphs *graphs = nullptr;
  if (E.use_graphs && !E.timing) {
    const uint64_t key = pcg_fingerprint<K>(E, b, prm, fp32);
    auto found = E.pcg_graphs.find(key);
    if (found == E.pcg_graphs.end()) {
      if (E.pcg_graphs.size() >= 16) {
        VK_HIP(hipStreamSynchronize(E.stream));  // no cached graph may still be running
        E.drop_graphs();
      }
      Impl::PcgGraphs g;
      const uint64_t v = E.vcycles;
      g.pre = capture(E, precondition);  // first: the body reads the cycle output it writes
      E.vcycles = v;
      g.body = capture(E, body);
      found = E.pcg_graphs.emplace(key, g).first;
    }
    graphs = &found->second;
  }
  auto precondition_step = [&]() {
    if (graphs) {
      VK_HIP(hipGraphLaunch(graphs->pre, E.stream));
      E.vcycles++;
      E.graph_launches++;
    } else {
      precondition();
    }
  };
  const int *active = c.active;
  // A long iteration covers the host's enqueue of the next one, so the poll
  // can look one iteration back (one queued no-op iteration after the last);
  // short ones keep two in flight.
  const uint32_t lag = (uint64_t)E.n0 * K >= kLongIteration ? 1u : (uint32_t)kPollLag;
  precondition_step();
  for (uint32_t it = 1; it <= max_it; ++it) {
    E.iterations_launched++;
    if (graphs) {
      VK_HIP(hipGraphLaunch(graphs->body, E.stream));
      E.graph_launches++;
    } else {
      body();
    }
    const int slot = it % kPollSlots;
    VK_HIP(hipMemcpyAsync(E.host_flags + slot, active, sizeof(int), hipMemcpyDeviceToHost, E.stream));
    VK_HIP(hipEventRecord(E.poll_events[slot], E.stream));
    if (it > lag) {
      const int seen = (it - lag) % kPollSlots;
      VK_HIP(hipEventSynchronize(E.poll_events[seen]));
      if (E.host_flags[seen] == 0) break;
    }
    precondition_step();
  }
  E.gate = E.always();
  c.center();
  if (tail) (*tail)();
  if (!states) return;  // the caller reads the lane states later (read_batch)
  std::vector<dev::ColState> host(K);
  VK_HIP(hipMemcpyAsync(host.data(), E.st.p, K * sizeof(dev::ColState), hipMemcpyDeviceToHost, E.stream));
  VK_HIP(hipStreamSynchronize(E.stream));
  out = host;
}

// The lane states of the last PCG call and the gate figures of the last
// extraction, with one synchronisation.
template <int K>
void read_batch(Impl &E, std::vector<dev::ColState> &st, std::vector<double> &red, std::vector<double> &comp,
                std::vector<double> &rn) {
  VK_HIP(hipMemcpyAsync(E.host_states, E.st.p, K * sizeof(dev::ColState), hipMemcpyDeviceToHost, E.stream));
  VK_HIP(hipMemcpyAsync(E.host_gates, E.reduced_res.p, K * sizeof(double), hipMemcpyDeviceToHost, E.stream));
  VK_HIP(hipMemcpyAsync(E.host_gates + 64, E.complete_res.p, K * sizeof(double), hipMemcpyDeviceToHost, E.stream));
  VK_HIP(hipMemcpyAsync(E.host_gates + 128, E.rhs_norm.p, K * sizeof(double), hipMemcpyDeviceToHost, E.stream));
  VK_HIP(hipStreamSynchronize(E.stream));
  st.assign(E.host_states, E.host_states + K);
  red.assign(E.host_gates, E.host_gates + K);
  comp.assign(E.host_gates + 64, E.host_gates + 64 + K);
  rn.assign(E.host_gates + 128, E.host_gates + 128 + K);
}

// full_cmg.rs:1158-1340 extraction for K lanes: normalise the firm mean, the
// reduced Schur residual, worker reconstruction and the complete residual.
// fsol: the normalised firm solution in firm order; red: its recentred copy.
template <int K>
void extract(Impl &E) {
  const uint64_t F = E.F, W = E.W;
  run(E, dev::k_gather_rows<K>, grid_for(F * K), dev::GatherRows<K>{F, E.pos0.p, E.x.p, E.fsol.p});
  const unsigned fg = reduce_grid(F * K), wg = reduce_grid(W * dev::Pack<K>::TPR);
  run(E, dev::k_lane_sum<K>, fg, dev::LaneSum<K>{(uint64_t)(F), E.fsol.p, E.rpartial.p});
  finish<K, 1>(E, E.rpartial.p, fg, dev::FinMean{E.lane_a.p, nullptr, F});
  run(E, dev::k_center_cols<K>, grid_for(E.n0 * K), dev::CenterCols<K>{(uint64_t)(E.n0), E.x.p, E.lane_a.p, nullptr});
  run(E, dev::k_center_cols<K>, grid_for(F * K), dev::CenterCols<K>{(uint64_t)(F), E.fsol.p, E.lane_a.p, nullptr});
  run(E, dev::k_copy_doubles, grid_for(F * K), dev::CopyDoubles{F * K, E.fsol.p, E.red.p});
  run(E, dev::k_lane_sum<K>, fg, dev::LaneSum<K>{(uint64_t)(F), E.red.p, E.rpartial.p});
  finish<K, 1>(E, E.rpartial.p, fg, dev::FinMean{E.lane_a.p, nullptr, F});
  run(E, dev::k_center_cols<K>, grid_for(F * K), dev::CenterCols<K>{(uint64_t)(F), E.red.p, E.lane_a.p, nullptr});
  run(E, dev::k_worker_gather<K>, grid_for(W * dev::Pack<K>::TPR),
      dev::WorkerGather<K>{W, E.worker_ptr.p, E.cell_firm.p, E.cell_weight.p, E.wdiag.p, E.red.p, E.wsum.p});
  segments<double, K, 1>(E, E.firm_chunks,
                         dev::OpSchurApply<K>{{E.firm_cells.p, E.cell_worker.p, E.cell_weight.p, E.firm_worker.p, E.firm_weight.p}, E.fdiag.p, E.red.p,
                                              E.wsum.p, E.sout.p, E.lane_b.p, (uint32_t)F});
  run(E, dev::k_reduced_residual<K>, fg, dev::ReducedResidual<K>{(uint32_t)F, E.sch.p, E.sout.p, E.lane_b.p, E.rpartial.p});
  finish<K, 2>(E, E.rpartial.p, fg, dev::FinRelative{E.reduced_res.p});
  run(E, dev::k_worker_solution<K>, wg,
      dev::WorkerSolution<K>{W, E.worker_ptr.p, E.cell_firm.p, E.cell_weight.p, E.wdiag.p, E.wrhs.p, E.fsol.p,
                             E.worker.p, E.rpartial.p});
  finish<K, 2>(E, E.rpartial.p, wg, dev::FinSums<K>{E.sums.p});
  segments<double, K, 2>(E, E.firm_chunks,
                         dev::OpFirmResidual<K>{{E.firm_cells.p, E.cell_worker.p, E.cell_weight.p, E.firm_worker.p, E.firm_weight.p}, E.fdiag.p, E.frhs.p,
                                                E.fsol.p, E.worker.p, E.sums.p, E.complete_res.p, E.rhs_norm.p});
}

// One batch of right-hand sides already in E.wrhs / E.frhs (K lanes, the first
// `columns` in use): the CMG_FULL_V2 batch solve with its complete-residual
// gate and warm-start refinement ladder (full_cmg.rs:809-1111).
template <int K>
void solve_columns(Impl &E, uint32_t columns, double effective_tol, double inner_ratio, const SolveOptions &opt,
                   std::vector<ColumnReceipt> &receipts) {
  const uint64_t F = E.F;
  const uint64_t m0 = E.n0 * K;
  const double gate = std::max(10.0 * effective_tol, 1e-11);
  // the Schur right-hand side into the hybrid system (before the first PCG
  // call), and the extraction after every PCG call: in a lockstep sample
  // batch both join the call's multi-sample launches
  const std::function<void()> head = [&]() {
    segments<double, K, 1>(E, E.firm_chunks,
                           dev::OpSchurRhs<K>{{E.firm_cells.p, E.cell_worker.p, E.cell_weight.p, E.firm_worker.p,
                                               E.firm_weight.p},
                                              E.wdiag.p, E.wrhs.p, E.frhs.p, E.sch.p, E.lane_a.p, (uint32_t)F});
    run(E, dev::k_center_cols<K>, grid_for(F * K), dev::CenterCols<K>{(uint64_t)(F), E.sch.p, E.lane_a.p, nullptr});
    run(E, dev::k_hybrid_rhs<K>, grid_for(m0), dev::HybridRhs<K>{E.n0, (uint32_t)F, E.perm0.p, E.sch.p, E.b0.p});
  };
  const std::function<void()> tail = [&]() { extract<K>(E); };
  std::vector<uint8_t> use(K, 0);
  for (uint32_t j = 0; j < columns; ++j) use[j] = 1;
  std::vector<dev::ColState> st;
  E.batches++;
  // Gate-aware first pass: the reference's operator-scaled stop
  // rel (||b|| + 2 maxdeg ||x||) can accept residuals far above the complete
  // gate when a heavy firm dominates the degree bound, and the batch is then
  // refined from a warm start with a second extraction.  Capping the stop at
  // 0.1 gate ||b|| per lane (the first refinement round's own target) makes
  // the first pass meet the gate in one extraction; the gate and the ladder
  // below are unchanged and still decide.  E.first_pass_cap = false keeps the
  // reference's stop.
  const double cap = E.first_pass_cap ? 0.1 * gate : 0.0;
  // the extraction is queued behind the PCG call before the host reads the
  // lane states, so the batch synchronises once (a failed lane's extraction is
  // discarded with the batch)
  pcg_lanes<K>(E, E.b0.p, effective_tol * inner_ratio, 0.0, opt.max_iterations, opt.recompute_interval, false,
               opt.precision == 32, use, st, cap, false, &head, &tail);
  std::vector<double> red, comp, rn;
  read_batch<K>(E, st, red, comp, rn);
  receipts.assign(columns, ColumnReceipt{});
  std::vector<double> first_norm(K, 0.0);
  for (uint32_t j = 0; j < columns; ++j) {
    first_norm[j] = st[j].rhs_norm;
    if (st[j].status) fail(st[j].status, "cmg_full_v2", "lane " + std::to_string(j) + " PCG failed");
    receipts[j].iterations = st[j].iters;
    receipts[j].restarts = st[j].restarts;
  }
  static const double factors[3] = {0.1, 0.01, 0.001};
  for (int round = 0; round < 3; ++round) {
    std::vector<uint8_t> fail_lane(K, 0);
    bool any = false;
    for (uint32_t j = 0; j < columns; ++j)
      if (!std::isfinite(comp[j]) || comp[j] > gate || !std::isfinite(red[j]) || red[j] > gate) {
        fail_lane[j] = 1;
        any = true;
      }
    if (!any) break;
    opt.checkpoint("cmg_full_v2_refinement");
    // warm start from the lifted firm solution (full_cmg.rs:970-1088)
    if (E.aux)
      hipLaunchKernelGGL((dev::k_lift<K>), grid_for(E.aux * K), kBlock, 0, E.stream, (uint32_t)F, (uint32_t)E.aux,
                         E.aux_ptr.p, E.aux_nbr.p, E.aux_w.p, E.pos0.p, E.fsol.p, E.x.p);
    // smallest Euclidean norm of a failing hybrid right-hand side (the
    // lane norms of the setup pass of the first solve)
    double smallest = INFINITY;
    for (uint32_t j = 0; j < columns; ++j)
      if (fail_lane[j]) smallest = std::min(smallest, first_norm[j]);
    double target = factors[round] * gate * smallest, rel = 0.0, abs_tol = target;
    if (!(std::isfinite(target) && target > 0.0)) {
      rel = effective_tol * inner_ratio * factors[round];
      abs_tol = 0.0;
    }
    E.refine_rounds++;
    for (uint32_t j = 0; j < columns; ++j) E.refined_columns += fail_lane[j];
    pcg_lanes<K>(E, E.b0.p, rel, abs_tol, opt.max_iterations, opt.recompute_interval, true, opt.precision == 32,
                 fail_lane, st, 0.0, false, nullptr, &tail);
    read_batch<K>(E, st, red, comp, rn);
    for (uint32_t j = 0; j < columns; ++j)
      if (fail_lane[j]) {
        if (st[j].status) fail(st[j].status, "cmg_full_v2_refinement", "refinement PCG failed");
        receipts[j].iterations += st[j].iters;
        receipts[j].restarts += st[j].restarts;
      }
  }
  for (uint32_t j = 0; j < columns; ++j) {
    if (!std::isfinite(comp[j]) || comp[j] > gate || !std::isfinite(red[j]) || red[j] > gate)
      fail(kFullResidual, "cmg_full_v2_refinement",
           "zero-based RHS column " + std::to_string(j) + ": reduced residual " + std::to_string(red[j]) +
               " or complete residual " + std::to_string(comp[j]) + " exceeds tolerance " + std::to_string(gate));
    receipts[j].reduced = red[j];
    receipts[j].complete = comp[j];
    receipts[j].rhs_norm = rn[j];
  }
}

}  // namespace

// jla.rs:286-364 on the host (one pass over the rows after the full fit).
void plugin_components(const Prepared &P, const std::vector<double> &alpha, const std::vector<double> &psi,
                       double out[4]) {
  Neumaier wm, fm;
  for (uint64_t r = 0; r < P.rows; ++r) {
    wm.add(P.target[r] * alpha[P.row_worker[r]]);
    fm.add(P.target[r] * psi[P.row_firm[r]]);
  }
  const double wmean = wm.value() / P.target_total, fmean = fm.value() / P.target_total;
  Neumaier a, b, c, d;
  for (uint64_t r = 0; r < P.rows; ++r) {
    const double wt = P.target[r];
    const double cw = alpha[P.row_worker[r]] - wmean, cf = psi[P.row_firm[r]] - fmean;
    a.add(wt * cw * cw);
    b.add(wt * cf * cf);
    c.add(wt * cw * cf);
    const double ct = cw + cf;
    d.add(wt * ct * ct);
  }
  out[0] = a.value() / P.target_total;
  out[1] = b.value() / P.target_total;
  out[2] = c.value() / P.target_total;
  out[3] = d.value() / P.target_total;
  const double implied = out[0] + out[1] + 2.0 * out[2];
  const double scale = std::max({std::fabs(out[0]), std::fabs(out[1]), std::fabs(2.0 * out[2]), std::fabs(out[3]), 1.0});
  if (std::fabs(out[3] - implied) > 1e-11 * scale)
    fail(kAccounting, "jla_components", "plug-in accounting identity failed");
}

double component_identity(const double c[4]) {
  const double implied = c[0] + c[1] + 2.0 * c[2];
  const double scale = std::max({std::fabs(c[0]), std::fabs(c[1]), std::fabs(2.0 * c[2]), std::fabs(c[3]), 1.0});
  return std::fabs(c[3] - implied) / scale;
}

// ------------------------------------------------------------------ controls
namespace {

// Symmetric eigendecomposition of an n x n matrix (row-major) by cyclic
// Jacobi: eigenvalues ascending in `lam`, eigenvectors as columns of `vec`.
void sym_eigen(std::vector<double> a, uint32_t n, std::vector<double> &lam, std::vector<double> &vec) {
  vec.assign((size_t)n * n, 0.0);
  for (uint32_t i = 0; i < n; ++i) vec[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
      diag += a[(size_t)i * n + i] * a[(size_t)i * n + i];
      for (uint32_t j = i + 1; j < n; ++j) off += a[(size_t)i * n + j] * a[(size_t)i * n + j];
    }
    if (!(off > 1e-32 * diag)) break;
    for (uint32_t p = 0; p < n; ++p)
      for (uint32_t r = p + 1; r < n; ++r) {
        const double apr = a[(size_t)p * n + r];
        if (apr == 0.0) continue;
        const double theta = (a[(size_t)r * n + r] - a[(size_t)p * n + p]) / (2.0 * apr);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (uint32_t k = 0; k < n; ++k) {
          const double akp = a[(size_t)k * n + p], akr = a[(size_t)k * n + r];
          a[(size_t)k * n + p] = c * akp - sn * akr;
          a[(size_t)k * n + r] = sn * akp + c * akr;
        }
        for (uint32_t k = 0; k < n; ++k) {
          const double apk = a[(size_t)p * n + k], ark = a[(size_t)r * n + k];
          a[(size_t)p * n + k] = c * apk - sn * ark;
          a[(size_t)r * n + k] = sn * apk + c * ark;
        }
        for (uint32_t k = 0; k < n; ++k) {
          const double vkp = vec[(size_t)k * n + p], vkr = vec[(size_t)k * n + r];
          vec[(size_t)k * n + p] = c * vkp - sn * vkr;
          vec[(size_t)k * n + r] = sn * vkp + c * vkr;
        }
      }
  }
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return a[(size_t)x * n + x] < a[(size_t)y * n + y]; });
  std::vector<double> v2((size_t)n * n);
  lam.resize(n);
  for (uint32_t k = 0; k < n; ++k) {
    lam[k] = a[(size_t)order[k] * n + order[k]];
    for (uint32_t i = 0; i < n; ++i) v2[(size_t)i * n + k] = vec[(size_t)i * n + order[k]];
  }
  vec.swap(v2);
}

// Eigenvalues of D^-1/2 m D^-1/2 (diagonal scaling); false when a diagonal
// entry is not positive and finite.
bool scaled_spectrum(const std::vector<double> &m, uint32_t n, std::vector<double> &lam) {
  std::vector<double> d(n), a(m);
  for (uint32_t i = 0; i < n; ++i) {
    d[i] = m[(size_t)i * n + i];
    if (!(std::isfinite(d[i]) && d[i] > 0.0)) return false;
    d[i] = 1.0 / std::sqrt(d[i]);
  }
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t j = 0; j < n; ++j) a[(size_t)i * n + j] *= d[i] * d[j];
  std::vector<double> vec;
  sym_eigen(a, n, lam, vec);
  return true;
}

// Inverse of a symmetric positive definite matrix from its eigendecomposition
// after diagonal scaling.
std::vector<double> spd_inverse(const std::vector<double> &m, uint32_t n) {
  std::vector<double> d(n), a(m), lam, vec, out((size_t)n * n, 0.0);
  for (uint32_t i = 0; i < n; ++i) d[i] = 1.0 / std::sqrt(m[(size_t)i * n + i]);
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t j = 0; j < n; ++j) a[(size_t)i * n + j] *= d[i] * d[j];
  sym_eigen(a, n, lam, vec);
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t j = 0; j < n; ++j) {
      double s = 0.0;
      for (uint32_t k = 0; k < n; ++k) s += vec[(size_t)i * n + k] * vec[(size_t)j * n + k] / lam[k];
      out[(size_t)i * n + j] = d[i] * s * d[j];
    }
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t j = i + 1; j < n; ++j) {
      const double v = 0.5 * (out[(size_t)i * n + j] + out[(size_t)j * n + i]);
      out[(size_t)i * n + j] = out[(size_t)j * n + i] = v;
    }
  return out;
}

// Lower Cholesky factor of an SPD matrix (row-major); false if not positive.
bool cholesky(const std::vector<double> &m, uint32_t n, std::vector<double> &l) {
  l.assign((size_t)n * n, 0.0);
  for (uint32_t j = 0; j < n; ++j) {
    double s = m[(size_t)j * n + j];
    for (uint32_t k = 0; k < j; ++k) s -= l[(size_t)j * n + k] * l[(size_t)j * n + k];
    if (!(s > 0.0) || !std::isfinite(s)) return false;
    l[(size_t)j * n + j] = std::sqrt(s);
    for (uint32_t i = j + 1; i < n; ++i) {
      double t = m[(size_t)i * n + j];
      for (uint32_t k = 0; k < j; ++k) t -= l[(size_t)i * n + k] * l[(size_t)j * n + k];
      l[(size_t)i * n + j] = t / l[(size_t)j * n + j];
    }
  }
  return true;
}

// Weighted Gram reduction of rows x m (row-major) on the device; returns the
// upper triangle (mode 0) or the cross block (mode 1) summed over blocks.
std::vector<double> device_gram(Impl &E, uint64_t rows, uint32_t m, int mode, const double *x, const double *w) {
  const uint32_t outs = mode == 1 ? (m / 2) * (m / 2) : m * (m + 1) / 2;
  const uint64_t tiles = (rows + dev::kGramTile - 1) / dev::kGramTile;
  const unsigned blocks = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(tiles, 1024));
  E.ctl_part.alloc((size_t)blocks * outs);
  E.ctl_out.alloc(std::max<uint32_t>(outs, 1));
  hipLaunchKernelGGL(dev::k_ctl_gram, blocks, kBlock, 0, E.stream, rows, m, mode, x, w, E.ctl_part.p);
  hipLaunchKernelGGL(dev::k_ctl_finish, grid_for(outs), kBlock, 0, E.stream, blocks, outs, E.ctl_part.p, E.ctl_out.p);
  std::vector<double> out;
  E.ctl_out.download(out, outs, E.stream);
  return out;
}

// Expand an upper triangle (row by row, i <= j) of an m x m matrix.
std::vector<double> from_upper(const std::vector<double> &u, uint32_t m) {
  std::vector<double> a((size_t)m * m);
  size_t o = 0;
  for (uint32_t i = 0; i < m; ++i)
    for (uint32_t j = i; j < m; ++j, ++o) a[(size_t)i * m + j] = a[(size_t)j * m + i] = u[o];
  return a;
}

void ensure_controls(Impl &E) {
  if (E.ctl_ready) return;
  const Prepared &P = *E.P;
  const uint32_t q = P.controls;
  if (q == 0 || q > (uint32_t)dev::kMaxControls) fail(kInvariant, "controls", "no controls to upload");
  E.ctl_z.upload(P.control);
  std::vector<double> f(P.frequency.begin(), P.frequency.end());
  E.ctl_f.upload(f);
  E.ctl_y.upload(P.outcome);
  E.ctl_row_cell.upload(P.row_cell);
  E.ctl_unit_ptr.upload(P.unit_ptr);
  E.ctl_unit_rows.upload(P.unit_rows);
  E.ctl_cellz.upload(P.cell_control);
  uint64_t widest = 0;
  for (uint64_t g = 0; g < P.units; ++g) widest = std::max<uint64_t>(widest, P.unit_ptr[g + 1] - P.unit_ptr[g]);
  E.ctl_block = (uint32_t)std::min<uint64_t>(widest, q + 1);
  E.ctl_gram0 = from_upper(device_gram(E, P.rows, q, 0, E.ctl_z.p, E.ctl_f.p), q);
  E.ctl_ready = true;
}

// The control fit; leaves the FE residuals (zry) and the joint residual (e)
// on the device and returns the joint two-way coefficients.
ControlFit fit_controls(Impl &E, const SolveOptions &opt, std::vector<double> &alpha, std::vector<double> &psi) {
  const Prepared &P = *E.P;
  const uint32_t q = P.controls, cols = q + 1, K = lanes_for(cols);
  ensure_controls(E);
  E.ensure(K);
  ControlFit cf;
  cf.q = q;
  // model_solver.rs:1799-1870: the raw control Gram must be nonsingular
  std::vector<double> lam;
  if (!scaled_spectrum(E.ctl_gram0, q, lam) || !(lam[q - 1] > 0.0) || !(lam[0] > 1e-12 * lam[q - 1]))
    fail(kSingularInformation, "model_rank", "original control Gram is singular or numerically unresolved");
  // the Q + 1 right-hand sides: frequency-weighted control and outcome sums
  std::vector<double> wi(E.W * K, 0.0), fi(E.F * K, 0.0);
  for (uint64_t c = 0; c < E.C; ++c) {
    double *a = &wi[(size_t)P.cell_worker[c] * K], *b = &fi[(size_t)P.cell_firm[c] * K];
    for (uint32_t j = 0; j < q; ++j) {
      a[j] += P.cell_control[c * q + j];
      b[j] += P.cell_control[c * q + j];
    }
    a[q] += P.cell_outcome[c];
    b[q] += P.cell_outcome[c];
  }
  VK_HIP(hipMemcpy(E.wrhs.p, wi.data(), wi.size() * sizeof(double), hipMemcpyHostToDevice));
  VK_HIP(hipMemcpy(E.frhs.p, fi.data(), fi.size() * sizeof(double), hipMemcpyHostToDevice));
  E.gate = E.always();
  E.ctl_zry.alloc(P.rows * cols);
  with_k(K, [&](auto kk) {
    constexpr int KK = decltype(kk)::value;
    solve_columns<KK>(E, cols, opt.fit_tolerance, 0.01, opt, cf.receipts);
    hipLaunchKernelGGL((dev::k_ctl_residualize<KK>), grid_for(P.rows), kBlock, 0, E.stream, P.rows, q,
                       E.ctl_row_cell.p, E.cell_worker.p, E.cell_firm.p, E.worker.p, E.fsol.p, E.ctl_z.p, E.ctl_y.p,
                       E.ctl_zry.p);
  });
  const std::vector<double> full = from_upper(device_gram(E, P.rows, cols, 0, E.ctl_zry.p, E.ctl_f.p), cols);
  cf.S.assign((size_t)q * q, 0.0);
  cf.b.assign(q, 0.0);
  for (uint32_t i = 0; i < q; ++i) {
    cf.b[i] = full[(size_t)i * cols + q];
    for (uint32_t j = 0; j < q; ++j) cf.S[(size_t)i * q + j] = full[(size_t)i * cols + j];
  }
  // generalised eigenvalues of S against the raw Gram (model_solver.rs:1818-1870)
  std::vector<double> l0, g0inv = spd_inverse(E.ctl_gram0, q);
  if (!cholesky(g0inv, q, l0))
    fail(kSingularInformation, "model_rank", "original-control whitening failed");
  std::vector<double> gen((size_t)q * q, 0.0), tmp((size_t)q * q, 0.0), vec;
  for (uint32_t i = 0; i < q; ++i)
    for (uint32_t j = 0; j < q; ++j)
      for (uint32_t k = 0; k < q; ++k) tmp[(size_t)i * q + j] += cf.S[(size_t)i * q + k] * l0[(size_t)k * q + j];
  for (uint32_t i = 0; i < q; ++i)
    for (uint32_t j = 0; j < q; ++j)
      for (uint32_t k = 0; k < q; ++k) gen[(size_t)i * q + j] += l0[(size_t)k * q + i] * tmp[(size_t)k * q + j];
  for (uint32_t i = 0; i < q; ++i)
    for (uint32_t j = i + 1; j < q; ++j)
      gen[(size_t)i * q + j] = gen[(size_t)j * q + i] = 0.5 * (gen[(size_t)i * q + j] + gen[(size_t)j * q + i]);
  sym_eigen(gen, q, lam, vec);
  cf.generalized_rcond = lam[q - 1] > 0.0 ? lam[0] / lam[q - 1] : 0.0;
  if (!(lam[q - 1] > 0.0) || !std::isfinite(cf.generalized_rcond) ||
      cf.generalized_rcond <= std::max(opt.rank_tolerance, 1e-12))
    fail(kSingularInformation, "model_rank",
         "FE-residualized generalized control information is singular or numerically unresolved");
  // control_geometry (generic_jla.rs:4407-4536): the scaled Schur complement
  if (!scaled_spectrum(cf.S, q, lam) || !(lam[q - 1] > 0.0) || lam[0] <= opt.rank_tolerance * lam[q - 1])
    fail(kSingularInformation, "generic_jla_control_schur", "the control Schur complement is singular");
  cf.s_inv = spd_inverse(cf.S, q);
  cf.beta.assign(q, 0.0);
  for (uint32_t i = 0; i < q; ++i)
    for (uint32_t j = 0; j < q; ++j) cf.beta[i] += cf.s_inv[(size_t)i * q + j] * cf.b[j];
  for (double v : cf.beta)
    if (!std::isfinite(v)) fail(kNonfiniteCorrected, "generic_jla_control_schur", "control coefficients are nonfinite");
  // joint residual, weighted RSS and the joint two-way coefficients
  E.ctl_small.alloc(2 * (size_t)q * q + 64);
  VK_HIP(hipMemcpy(E.ctl_small.p, cf.beta.data(), q * sizeof(double), hipMemcpyHostToDevice));
  E.ctl_e.alloc(P.rows);
  const unsigned rb = reduce_grid(P.rows);
  E.ctl_part.alloc(std::max<size_t>(rb, E.W + E.F));
  hipLaunchKernelGGL(dev::k_ctl_resid, rb, kBlock, 0, E.stream, P.rows, q, E.ctl_zry.p, E.ctl_small.p, E.ctl_f.p,
                     E.ctl_e.p, E.ctl_part.p);
  std::vector<double> part;
  E.ctl_part.download(part, rb, E.stream);
  Neumaier rss;
  for (double v : part) rss.add(v);
  cf.rss = rss.value();
  E.ctl_out.alloc(E.W + E.F);
  with_k(K, [&](auto kk) {
    constexpr int KK = decltype(kk)::value;
    hipLaunchKernelGGL((dev::k_ctl_combine<KK>), grid_for(E.W), kBlock, 0, E.stream, E.W, q, E.worker.p,
                       E.ctl_small.p, E.ctl_out.p);
    hipLaunchKernelGGL((dev::k_ctl_combine<KK>), grid_for(E.F), kBlock, 0, E.stream, E.F, q, E.fsol.p,
                       E.ctl_small.p, E.ctl_out.p + E.W);
  });
  std::vector<double> both;
  E.ctl_out.download(both, E.W + E.F, E.stream);
  alpha.assign(both.begin(), both.begin() + E.W);
  psi.assign(both.begin() + E.W, both.end());
  E.ctl_s_inv = cf.s_inv;
  return cf;
}

// The control terms of one target batch's draws (k probes, columns 2l and
// 2l + 1): worker += b_w' H_w + b_w' M2 b_w, firm likewise, total with
// b_w + b_f and H_w + H_f, covariance = (total - worker - firm) / 2.
void control_draws(Impl &E, uint32_t q, const std::vector<double> &m2, uint32_t k2, uint32_t k,
                   std::vector<double> &d) {
  const uint32_t m = 2 * q;
  std::vector<double> out;
  E.ctl_out.download(out, (size_t)m * k2, E.stream);
  const std::vector<double> &si = E.ctl_s_inv;
  auto term = [&](const std::vector<double> &beta, const std::vector<double> &h) {
    double s = 0.0;
    for (uint32_t i = 0; i < q; ++i) {
      s += beta[i] * h[i];
      for (uint32_t j = 0; j < q; ++j) s += beta[i] * m2[(size_t)i * q + j] * beta[j];
    }
    return s;
  };
  std::vector<double> bw(q), bf(q), hw(q), hf(q), bt(q), ht(q);
  for (uint32_t l = 0; l < k; ++l) {
    for (int side = 0; side < 2; ++side) {
      const uint32_t col = 2 * l + side;
      std::vector<double> &beta = side ? bf : bw, &h = side ? hf : hw;
      for (uint32_t i = 0; i < q; ++i) {
        double v = 0.0;
        for (uint32_t j = 0; j < q; ++j) v -= si[(size_t)i * q + j] * out[(size_t)j * k2 + col];
        beta[i] = v;
        h[i] = out[(size_t)(q + i) * k2 + col];
      }
    }
    for (uint32_t i = 0; i < q; ++i) {
      bt[i] = bw[i] + bf[i];
      ht[i] = hw[i] + hf[i];
    }
    double *dd = &d[4 * l];
    const double worker = dd[0] + term(bw, hw), firm = dd[1] + term(bf, hf), total = dd[3] + term(bt, ht);
    dd[0] = worker;
    dd[1] = firm;
    dd[2] = 0.5 * (total - worker - firm);
    dd[3] = total;
  }
}

// Launch the match-block kernel with the smallest per-thread block bound.
void launch_blocks(Impl &E, uint32_t q, const SolveOptions &opt) {
  const dev::BlockParams bp{(double)opt.probes, opt.block_tolerance, opt.rank_tolerance};
  auto go = [&](auto nn) {
    constexpr int N = decltype(nn)::value;
    hipLaunchKernelGGL((dev::k_ctl_blocks<N>), grid_for(E.Gm), kBlock, 0, E.stream, E.Gm, q, E.ctl_unit_ptr.p,
                       E.ctl_unit_rows.p, E.adjust.p, E.ctl_f.p, E.ctl_y.p, E.ctl_e.p, E.ctl_zry.p, E.ctl_u.p, bp,
                       E.contrib.p, E.ctl_uab.p, E.ctl_umass.p, E.sink());
  };
  if (E.ctl_block <= 8) go(std::integral_constant<int, 8>{});
  else if (E.ctl_block <= 16) go(std::integral_constant<int, 16>{});
  else go(std::integral_constant<int, dev::kMaxControls + 1>{});
}

}  // namespace

namespace {

// Host permutation of K-lane vectors between original and renumbered order.
// Numerical MCSE replay (engine/numerical.rs:160-330): the leverage probes
// are solved again (same Counter-V1 draws, same batches as the leverage
// phase) and every probe is scored against the unit linearisation and the
// target folds; the scores' fold cross-covariance plus the target draws'
// covariance give the MCSE.  A replay solve that fails its PCG or complete
// residual gate stops the replay with status 7.
void numerical_replay(Impl &E, const SolveOptions &opt, SolveResult &res) {
  NumericalResult &num = res.numerical;
  const uint64_t R = opt.probes;
  const uint32_t lw = std::max<uint32_t>(1, std::min<uint32_t>(opt.leverage_width, 64));
  const uint32_t KL = lanes_for(lw);
  num.leverage_probes = num.target_probes = (uint32_t)R;
  num.fold_a = (uint32_t)((R + 1) / 2);
  num.fold_b = (uint32_t)(R / 2);
  num.point_rhs = 1 + 3 * R;
  num.maximum_point_complete_residual = res.max_complete_residual;
  std::vector<double> draws(3 * R), scores(6 * R, 0.0);
  for (uint64_t t = 0; t < R; ++t)
    for (int j = 0; j < 3; ++j) draws[3 * t + j] = res.draws[4 * t + j];
  Matrix3 conditional;
  if (!cross_covariance(draws, draws, conditional)) fail(kInvalidInput, "numerical_mc", "nonfinite target draws");
  uint32_t status = num.status;
  if (status == 2) {
    E.nscores.alloc(6 * 64);
    E.noscores.alloc(6 * 64);
    for (uint64_t first = 0; first < R; first += lw) {
      opt.checkpoint("numerical_replay_batch");
      const uint32_t k = (uint32_t)std::min<uint64_t>(lw, R - first);
      run(E, dev::k_atoms, grid_for(E.G * ((k + 3) / 4 + 1)),
          dev::Atoms{E.G, E.unit_rank.p, E.unit_count.p, E.unit_offset.p, opt.seed, kTagLeverage, first, k, KL,
                     E.atoms.p});
      std::vector<ColumnReceipt> rc;
      bool failed = false;
      with_k(KL, [&](auto kk) {
        constexpr int K = decltype(kk)::value;
        run(E, dev::k_unit_rhs<K>, grid_for(E.W * dev::Pack<K>::TPR),
            dev::UnitRhs<K>{E.W, E.worker_ptr.p, nullptr, E.cell_unit_ptr.p, E.cell_units.p, E.atoms.p, E.G, k,
                            E.wrhs.p, E.cell_unit.p});
        segments<double, K, 0>(E, E.firm_chunks,
                               dev::OpFirmUnits<K>{E.firm_cells.p, E.cell_unit_ptr.p, E.cell_units.p, E.atoms.p, E.G,
                                                   k, E.frhs.p, E.firm_units.p});
        num.attempted_replay_rhs += k;
        try {
          solve_columns<K>(E, k, opt.probe_tolerance, 1.0, opt, rc);
        } catch (const Error &e) {
          if (e.code < kPcgCurvature || e.code > kFullResidual) throw;
          failed = true;
          num.failed_replay_probe = (uint32_t)first;
          num.replay_error_code = (uint32_t)e.code;
          return;
        }
        // observation units (generic_jla/numerical.rs:617-641), then the match
        // units (engine/numerical.rs:246-291); the two sums are added per probe
        if (E.Go) {
          const unsigned g = reduce_grid(E.Go * K);
          hipLaunchKernelGGL((dev::k_obs_scores<K>), g, kBlock, 0, E.stream, E.Go, k, first, opt.seed, kTagLeverage,
                             E.unit_cell.p + E.Gm, E.unit_count.p + E.Gm, E.unit_rank.p + E.Gm,
                             E.unit_offset.p + E.Gm, E.row_phys.p, E.obs_physical, E.cell_worker.p, E.cell_firm.p,
                             E.worker.p, E.fsol.p, E.nograd.p, E.ncopy.p, E.nofolds.p, E.rpartial.p);
          finish<K, 6>(E, E.rpartial.p, g, dev::FinScores{E.noscores.p});
        }
        if (E.Gm) {
          const unsigned g = reduce_grid(E.Gm * K);
          hipLaunchKernelGGL((dev::k_numerical_scores<K>), g, kBlock, 0, E.stream, E.Gm, k, E.unit_cell.p,
                             E.unit_count.p, E.cell_worker.p, E.cell_firm.p, E.worker.p, E.fsol.p, E.atoms.p,
                             E.ngrad.p, E.noffs.p, E.nfolds.p, E.rpartial.p);
          finish<K, 6>(E, E.rpartial.p, g, dev::FinScores{E.nscores.p});
        }
      });
      for (uint64_t g = 0; g < E.Gm; ++g) num.replay_generator_words += (E.P->unit_count[g] + 63) / 64 * k;
      num.replay_generator_words += E.obs_class_words * k;  // generic_jla/numerical.rs:490-519
      if (failed) {
        status = 7;
        break;
      }
      std::vector<double> s(6 * (size_t)KL, 0.0), so;
      if (E.Gm) E.nscores.download(s, 6 * (size_t)KL, E.stream);
      if (E.Go) {
        E.noscores.download(so, 6 * (size_t)KL, E.stream);
        for (size_t q = 0; q < s.size(); ++q) s[q] = E.Gm ? so[q] + s[q] : so[q];
      }
      for (uint32_t c = 0; c < k; ++c) {
        const uint64_t probe = first + c;
        for (int f = 0; f < 2; ++f)
          for (int j = 0; j < 3; ++j) scores[(f * R + probe) * 3 + j] = s[c * 6 + f * 3 + j];
        num.replay_probes.push_back((uint32_t)probe);
        num.replay_complete.push_back(rc[c].complete);
        num.maximum_replay_complete_residual = std::max(num.maximum_replay_complete_residual, rc[c].complete);
      }
      num.certified_replay_rhs += k;
    }
  }
  for (int f = 0; f < 2; ++f) {
    const std::vector<double> fold(scores.begin() + f * 3 * R, scores.begin() + (f + 1) * 3 * R);
    const auto m = score_means(fold);
    for (int j = 0; j < 3; ++j) num.score_means[f][j] = m[j];
  }
  Matrix3 leverage;
  leverage.fill(NAN);
  if (status == 2) {
    const std::vector<double> a(scores.begin(), scores.begin() + 3 * R), b(scores.begin() + 3 * R, scores.end());
    if (!cross_covariance(a, b, leverage)) {
      status = 6;
      leverage.fill(NAN);
    }
  }
  const Covariance cov = finalize_covariance(conditional, leverage);
  std::copy(cov.conditional.begin(), cov.conditional.end(), num.conditional);
  std::copy(cov.leverage.begin(), cov.leverage.end(), num.leverage);
  std::copy(cov.raw.begin(), cov.raw.end(), num.raw);
  std::copy(cov.usable.begin(), cov.usable.end(), num.usable);
  std::copy(cov.mcse.begin(), cov.mcse.end(), num.mcse);
  num.psd_adjustment = cov.psd_adjustment;
  num.status = status != 2 ? status : cov.status;
}

// Device layout of the hierarchy and the firm passes for chunks of at most
// `chunk` entries (renumbering, CSR, labels, chunk tables).
void relayout(Impl &E, uint32_t chunk) {
  const Hierarchy &h = *E.H;
  const Prepared &p = *E.P;
  E.chunk = chunk;
  E.max_chunks = 0;
  E.max_multi = 0;
  {
    std::vector<uint32_t> fptr(p.firm_ptr.begin(), p.firm_ptr.end());
    E.firm_chunks.upload(chunk_segments(fptr, chunk));
  }
  std::vector<LevelLayout> layout = device_layout(h, chunk);
  E.levels.resize(h.levels.size());
  for (size_t l = 0; l < h.levels.size(); ++l) {
    const Level &L = h.levels[l];
    LevelLayout &Y = layout[l];
    DevLevel &D = E.levels[l];
    D.n = (uint32_t)L.graph.n;
    D.nnz = Y.col.size();
    D.coarse = (uint32_t)L.coarse;
    D.repeat = (uint32_t)std::min<uint64_t>(L.repeat, 1u << 30);
    D.terminal = L.terminal;
    D.col.upload(Y.col); D.val.upload(Y.val); D.val32.upload(Y.val);
    D.inv.upload(Y.inv); D.inv32.upload(Y.inv);
    D.rows.upload(Y.rows);
    D.rows_g.upload(chunk_segments(Y.row_ptr, chunk, kSegmentGroup));
    if (!Y.lab.empty()) {
      D.lab.upload(Y.lab);
      D.pcol.alloc(D.nnz);
      hipLaunchKernelGGL(dev::k_compose, grid_for(D.nnz), kBlock, 0, E.stream, D.nnz, D.col.p, D.lab.p, D.pcol.p);
      D.aggs.upload(Y.aggs);
      D.aggs_g.upload(chunk_segments(Y.agg_ptr, chunk, kSegmentGroup));
    }
    D.perm = Y.perm;
    D.dperm.upload(Y.perm);
    D.lab_host = Y.lab;
    E.max_chunks = std::max<uint64_t>({E.max_chunks, Y.rows.chunks(), Y.aggs.chunks()});
    E.max_multi = std::max({E.max_multi, Y.rows.multi(), Y.aggs.multi()});
  }
  E.max_chunks = std::max<uint64_t>(E.max_chunks, E.firm_chunks.count);
  if (!E.sm.shape.empty()) set_cycle_shape(E, E.sm.shape);
  E.perm0.upload(layout[0].perm);
  E.pos0.upload(layout[0].pos);
  if (E.cap) E.seg_partial.alloc(std::max<uint64_t>(E.max_chunks, 1) * E.cap);
}

}  // namespace

void device_memory(uint64_t &live, uint64_t &peak) {
  live = DeviceBytes::live.load();
  peak = DeviceBytes::peak.load();
}

bool device_available() {
  int devices = 0;
  return hipGetDeviceCount(&devices) == hipSuccess && devices > 0;
}

// ------------------------------------------------------------ lockstep batches
// The PCG calls of a sample batch that wait at the same time run together on
// one stream: the recorded passes of all calls are merged in key order, and
// passes with the same key and functor type become one k_multi launch over
// their samples (each sample keeps its own grid, buffers, reductions and
// convergence gate, so its results are those of running alone, bit for bit).
// The host polls the OR of the samples' any-lane-active flags with the same
// lag as a single engine.
class Lockstep {
 public:
  Lockstep() {
    VK_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    VK_HIP(hipHostMalloc((void **)&host_flags_, kPollSlots * sizeof(int), hipHostMallocDefault));
    for (auto &e : poll_) VK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    VK_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
    any_.alloc(1);
  }
  ~Lockstep() {
    (void)hipStreamSynchronize(stream_);
    for (auto &e : poll_) (void)hipEventDestroy(e);
    for (auto &e : ready_) (void)hipEventDestroy(e);
    (void)hipEventDestroy(done_);
    if (pinned_) (void)hipHostFree(pinned_);
    if (host_flags_) (void)hipHostFree(host_flags_);
    (void)hipStreamDestroy(stream_);
  }
  void join() {
    std::lock_guard<std::mutex> lk(m_);
    ++running_;
  }
  void leave() {
    std::unique_lock<std::mutex> lk(m_);
    --running_;
    lead(lk);
  }
  void submit(PcgProgram &p) {
    std::unique_lock<std::mutex> lk(m_);
    waiting_.push_back(&p);
    lead(lk);
    cv_.wait(lk, [&] { return p.done; });
    if (p.error) std::rethrow_exception(p.error);
  }

 private:
  struct Launch {
    MultiLaunch fn;
    uint32_t count, blocks;
    size_t fs, ends;
  };
  using Plan = std::vector<Launch>;

  // Runs the waiting calls once every joined thread waits (the caller's
  // thread leads; the others sleep until their call is done).
  void lead(std::unique_lock<std::mutex> &lk) {
    while (!busy_ && !waiting_.empty() && waiting_.size() >= running_) {
      busy_ = true;
      std::vector<PcgProgram *> batch;
      batch.swap(waiting_);
      lk.unlock();
      std::exception_ptr err;
      try {
        execute(batch);
      } catch (...) {
        err = std::current_exception();
      }
      lk.lock();
      busy_ = false;
      for (PcgProgram *p : batch) {
        if (err && !p->error) p->error = err;
        p->done = true;
      }
      cv_.notify_all();
    }
  }

  static void align(std::vector<uint8_t> &t, size_t a) { t.resize((t.size() + a - 1) / a * a, 0); }

  // Merge one recorded segment of every call into launches; their functor
  // arrays and block ends are appended to `table`.
  static Plan merge(const std::vector<const std::vector<Record> *> &segs, std::vector<uint8_t> &table) {
    struct Item {
      const Record *r;
      uint32_t sample;
    };
    std::vector<Item> items;
    for (uint32_t p = 0; p < segs.size(); ++p)
      for (const Record &r : *segs[p]) items.push_back({&r, p});
    std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) {
      if (a.r->key != b.r->key) return a.r->key < b.r->key;
      return reinterpret_cast<uintptr_t>(a.r->launch) < reinterpret_cast<uintptr_t>(b.r->launch);
    });
    Plan plan;
    for (size_t i = 0; i < items.size();) {
      size_t j = i;
      while (j < items.size() && items[j].r->key == items[i].r->key && items[j].r->launch == items[i].r->launch) ++j;
      for (size_t c = i; c < j; c += dev::kMultiMax) {
        const size_t e = std::min(j, c + (size_t)dev::kMultiMax);
        Launch L{items[c].r->launch, (uint32_t)(e - c), 0, 0, 0};
        align(table, 256);
        L.fs = table.size();
        for (size_t k = c; k < e; ++k) {
          const Record &r = *items[k].r;
          if (r.threads != items[c].r->threads || r.size != items[c].r->size)
            fail(kInvariant, "lockstep", "merged passes differ in functor layout");
          table.insert(table.end(), r.bytes.begin(), r.bytes.end());
        }
        align(table, 16);
        L.ends = table.size();
        uint32_t end = 0;
        for (size_t k = c; k < e; ++k) {
          end += items[k].r->grid;
          const uint8_t *b = reinterpret_cast<const uint8_t *>(&end);
          table.insert(table.end(), b, b + sizeof end);
        }
        L.blocks = end;
        plan.push_back(L);
      }
      i = j;
    }
    return plan;
  }

  void launch(const Plan &plan) {
    for (const Launch &L : plan)
      L.fn(table_.p + L.fs, reinterpret_cast<const uint32_t *>(table_.p + L.ends), L.count, L.blocks, stream_);
  }

  void execute(std::vector<PcgProgram *> &batch) {
    const uint32_t P = (uint32_t)batch.size();
    std::vector<uint8_t> table;
    auto segments = [&](auto pick) {
      std::vector<const std::vector<Record> *> v;
      for (PcgProgram *p : batch) v.push_back(&pick(*p));
      return merge(v, table);
    };
    const Plan setup = segments([](PcgProgram &p) -> const std::vector<Record> & { return p.setup; });
    Plan pre[2], body[2];
    for (int q = 0; q < 2; ++q) {
      pre[q] = segments([q](PcgProgram &p) -> const std::vector<Record> & { return p.pre[q]; });
      body[q] = segments([q](PcgProgram &p) -> const std::vector<Record> & { return p.body[q]; });
    }
    const Plan tail = segments([](PcgProgram &p) -> const std::vector<Record> & { return p.tail; });
    align(table, 16);
    const size_t flags_at = table.size();
    for (PcgProgram *p : batch) {
      const uint8_t *b = reinterpret_cast<const uint8_t *>(&p->active);
      table.insert(table.end(), b, b + sizeof p->active);
    }
    if (table.size() > pinned_bytes_) {
      if (pinned_) VK_HIP(hipHostFree(pinned_));
      pinned_ = nullptr;
      pinned_bytes_ = 0;
      VK_HIP(hipHostMalloc((void **)&pinned_, table.size(), hipHostMallocDefault));
      pinned_bytes_ = table.size();
    }
    std::memcpy(pinned_, table.data(), table.size());
    table_.alloc(table.size());
    // after the work each engine queued before its call
    while (ready_.size() < P) {
      hipEvent_t e = nullptr;
      VK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      ready_.push_back(e);
    }
    for (uint32_t i = 0; i < P; ++i) {
      VK_HIP(hipEventRecord(ready_[i], batch[i]->stream));
      VK_HIP(hipStreamWaitEvent(stream_, ready_[i], 0));
    }
    VK_HIP(hipMemcpyAsync(table_.p, pinned_, table.size(), hipMemcpyHostToDevice, stream_));
    const int *const *flags = reinterpret_cast<const int *const *>(table_.p + flags_at);
    uint32_t max_it = 0;
    for (PcgProgram *p : batch) max_it = std::max(max_it, p->max_iterations);
    launch(setup);
    launch(pre[0]);
    uint64_t cycles = 1, iterations = 0;
    for (uint32_t it = 1; it <= max_it; ++it) {
      ++iterations;
      launch(body[(it - 1) % 2]);
      hipLaunchKernelGGL(dev::k_any_active, 1, kBlock, 0, stream_, flags, P, any_.p);
      const int slot = it % kPollSlots;
      VK_HIP(hipMemcpyAsync(host_flags_ + slot, any_.p, sizeof(int), hipMemcpyDeviceToHost, stream_));
      VK_HIP(hipEventRecord(poll_[slot], stream_));
      if (it > (uint32_t)kPollLag) {
        const int seen = (it - kPollLag) % kPollSlots;
        VK_HIP(hipEventSynchronize(poll_[seen]));
        if (host_flags_[seen] == 0) break;
      }
      launch(pre[it % 2]);
      ++cycles;
    }
    launch(tail);
    for (PcgProgram *p : batch) {
      p->out.assign(p->lanes, dev::ColState{});
      if (p->lanes)
        VK_HIP(hipMemcpyAsync(p->out.data(), p->st, p->lanes * sizeof(dev::ColState), hipMemcpyDeviceToHost,
                              stream_));
      p->iterations = iterations;
      p->cycles = cycles;
    }
    VK_HIP(hipEventRecord(done_, stream_));
    for (PcgProgram *p : batch) VK_HIP(hipStreamWaitEvent(p->stream, done_, 0));
    VK_HIP(hipStreamSynchronize(stream_));
  }

  std::mutex m_;
  std::condition_variable cv_;
  uint32_t running_ = 0;
  std::vector<PcgProgram *> waiting_;
  bool busy_ = false;
  hipStream_t stream_ = nullptr;
  int *host_flags_ = nullptr;
  hipEvent_t poll_[kPollSlots] = {};
  hipEvent_t done_ = nullptr;
  std::vector<hipEvent_t> ready_;
  uint8_t *pinned_ = nullptr;
  size_t pinned_bytes_ = 0;
  DBuf<uint8_t> table_;
  DBuf<int> any_;
};

std::shared_ptr<Lockstep> make_lockstep() { return std::make_shared<Lockstep>(); }
void lockstep_join(Lockstep &l) { l.join(); }
void lockstep_leave(Lockstep &l) { l.leave(); }
namespace gpu {
void lockstep_submit(Lockstep &l, PcgProgram &prog) { l.submit(prog); }
}  // namespace gpu

DeviceEngine::DeviceEngine(const Prepared &p, const HybridSystem &hs, const Hierarchy &h) : impl_(new Impl) {
  Impl &E = *impl_;
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0)
    fail(kUnsupported, "device", "no HIP device is available for the MI355X engine");
  // non-blocking: engines of independent samples run concurrently, each on
  // its own stream (the host copies to the device happen while the stream is
  // idle, after its last synchronisation)
  VK_HIP(hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking));
  E.P = &p;
  E.HS = &hs;
  E.W = p.workers; E.F = p.firms; E.C = p.cells; E.G = p.units; E.S = p.strata;
  E.target_total = p.target_total;
  E.aux = hs.aux;
  E.n0 = h.levels[0].graph.n;
  E.bound = h.levels[0].graph.norm_bound;
  if (h.finest_components != 1) fail(kUnsupported, "device", "the hybrid Laplacian must be connected");
  // device atoms are 32-bit integers: 2 popcount - n needs n < 2^31
  for (const auto *counts : {&p.unit_count, &p.stratum_count})
    for (uint64_t c : *counts)
      if (c >= (1ULL << 31)) fail(kUnsupported, "device", "a physical count of 2^31 or more is not supported");
  std::vector<double> wd(E.W, 0.0), fd(E.F, 0.0);
  for (uint64_t c = 0; c < E.C; ++c) {
    wd[p.cell_worker[c]] += p.cell_weight[c];
    fd[p.cell_firm[c]] += p.cell_weight[c];
  }
  E.worker_ptr.upload(p.worker_ptr); E.firm_ptr.upload(p.firm_ptr); E.firm_cells.upload(p.firm_cells);
  E.cell_worker.upload(p.cell_worker); E.cell_firm.upload(p.cell_firm); E.cell_weight.upload(p.cell_weight);
  E.wdiag.upload(wd); E.fdiag.upload(fd);
  E.cell_unit_ptr.upload(p.cell_unit_ptr); E.cell_units.upload(p.cell_units);
  {
    std::vector<uint32_t> fw(E.C);
    std::vector<double> fwt(E.C);
    for (uint64_t k = 0; k < E.C; ++k) {
      fw[k] = p.cell_worker[p.firm_cells[k]];
      fwt[k] = p.cell_weight[p.firm_cells[k]];
    }
    E.firm_worker.upload(fw);
    E.firm_weight.upload(fwt);
    std::vector<uint32_t> uw(p.units), uf(p.units);
    for (uint64_t g = 0; g < p.units; ++g) {
      uw[g] = p.cell_worker[p.unit_cell[g]];
      uf[g] = p.cell_firm[p.unit_cell[g]];
    }
    E.unit_worker.upload(uw);
    E.unit_firm.upload(uf);
  }
  {
    bool single = true;
    for (uint64_t c = 0; c < E.C && single; ++c) single = p.cell_unit_ptr[c + 1] - p.cell_unit_ptr[c] == 1;
    if (single && E.C) {
      std::vector<uint32_t> fu(E.C);
      for (uint64_t k = 0; k < E.C; ++k) fu[k] = p.cell_units[p.cell_unit_ptr[p.firm_cells[k]]];
      E.firm_units.upload(fu);
      for (uint64_t c = 0; c < E.C; ++c) fu[c] = p.cell_units[p.cell_unit_ptr[c]];
      E.cell_unit.upload(fu);
    }
  }
  E.unit_cell.upload(p.unit_cell); E.unit_count.upload(p.unit_count); E.unit_rank.upload(p.unit_rank);
  E.unit_outcome.upload(p.unit_outcome);
  E.Gm = p.match_units;
  E.Go = p.units - p.match_units;
  if (E.Go) {
    const uint64_t first_row = p.rows - E.Go;  // observation units are the last stored rows
    std::vector<uint64_t> phys(E.Go + 1, 0);
    std::vector<double> y(E.Go);
    for (uint64_t i = 0; i < E.Go; ++i) {
      phys[i + 1] = phys[i] + p.frequency[first_row + i];
      y[i] = p.outcome[first_row + i];
    }
    E.obs_physical = phys[E.Go];
    E.unit_offset.upload(p.unit_offset);
    E.row_phys.upload(phys);
    E.row_outcome.upload(y);
  }
  E.stratum_per_copy.upload(p.stratum_per_copy); E.stratum_rank.upload(p.stratum_rank);
  E.stratum_count.upload(p.stratum_count); E.cell_stratum_ptr.upload(p.cell_stratum_ptr);
  E.cell_target.upload(p.cell_target);
  E.aux_ptr.upload(hs.aux_ptr); E.aux_nbr.upload(hs.aux_nbr); E.aux_w.upload(hs.aux_w);

  E.H = &h;
  relayout(E, kChunkEntries);
  if (h.terminal_reason == kDirect) E.ginv.upload(h.terminal_inverse);
  E.rpartial.alloc((size_t)dev::kPartialStride * dev::kMaxReduceValues * 64);
  E.bpartial.alloc((size_t)dev::kPartialStride * dev::kMaxReduceValues * 64);
  E.flags.alloc(4);
  const int init_flags[4] = {1, 0, 1, 0};
  VK_HIP(hipMemcpy(E.flags.p, init_flags, sizeof init_flags, hipMemcpyHostToDevice));
  E.gate = E.always();
  VK_HIP(hipHostMalloc((void **)&E.host_flags, kPollSlots * sizeof(int), hipHostMallocDefault));
  VK_HIP(hipHostMalloc((void **)&E.host_states, 64 * sizeof(dev::ColState), hipHostMallocDefault));
  VK_HIP(hipHostMalloc((void **)&E.host_gates, 3 * 64 * sizeof(double), hipHostMallocDefault));
  for (auto &ev : E.poll_events) VK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  E.st.alloc(64); E.use.alloc(64);
  E.lane_a.alloc(64); E.lane_b.alloc(64); E.lane_c.alloc(64);
  E.reduced_res.alloc(64); E.complete_res.alloc(64); E.rhs_norm.alloc(64); E.sums.alloc(128);
  E.status.alloc(1);
  E.reset_status();
  E.totals.alloc(128); E.draws.alloc(4 * 64);
}

DeviceEngine::~DeviceEngine() {
  if (impl_ && impl_->stream) {
    (void)hipStreamSynchronize(impl_->stream);
    impl_->drop_graphs();
    impl_->drop_timing();
    for (auto &ev : impl_->poll_events)
      if (ev) (void)hipEventDestroy(ev);
    if (impl_->host_flags) (void)hipHostFree(impl_->host_flags);
    if (impl_->host_states) (void)hipHostFree(impl_->host_states);
    if (impl_->host_gates) (void)hipHostFree(impl_->host_gates);
    (void)hipStreamDestroy(impl_->stream);
  }
}

SolveResult DeviceEngine::run_once(const SolveOptions &opt) {
  Impl &E = *impl_;
  const Prepared &P = *E.P;
  if (opt.probes < 2) fail(kInvalidInput, "jla_validate", "JLA probe count must be at least two");
  const uint32_t lw = std::max<uint32_t>(1, std::min<uint32_t>(opt.leverage_width, 64));
  const uint32_t tw = std::max<uint32_t>(1, std::min<uint32_t>(opt.target_width, 32));
  const uint32_t KL = lanes_for(lw), KT = lanes_for(tw), K2 = lanes_for(2 * tw);
  opt.checkpoint("jla_solver_setup");
  E.ensure(std::max({KL, K2, 1u}));
  // a solve starts with no recorded unit failure (a generation may be solved again)
  E.reset_status();
  SolveResult res;
  auto t0 = std::chrono::steady_clock::now();
  const uint64_t R = opt.probes;

  const bool joint = opt.joint_controls;
  if (joint && (P.controls == 0 || E.Go || P.stayers))
    fail(kInvariant, "generic_jla", "joint controls need a match-deletion problem with controls");
  if (joint && opt.numerical)
    fail(kUnsupported, "numerical_mc", "the numerical attachment with controls is not supported by the MI355X engine");
  const uint32_t q = joint ? P.controls : 0;
  // ---- joint fit with the controls (generic_jla.rs:1930-2000): Q + 1
  // two-way solves, the control Schur complement and beta; the plug-in uses
  // the joint two-way coefficients
  if (joint) {
    opt.checkpoint("generic_jla_full_fit");
    std::vector<double> alpha, psi;
    ControlFit cf = fit_controls(E, opt, alpha, psi);
    res.controls = q;
    res.control_beta = cf.beta;
    res.control_rcond = cf.generalized_rcond;
    res.control_fit.assign(cf.receipts.begin(), cf.receipts.begin() + q);
    res.fit = cf.receipts[q];
    res.max_complete_residual = 0.0;
    for (const auto &c : cf.receipts) res.max_complete_residual = std::max(res.max_complete_residual, c.complete);
    res.fitted_cell.resize(E.C);
    for (uint64_t c = 0; c < E.C; ++c) res.fitted_cell[c] = alpha[P.cell_worker[c]] + psi[P.cell_firm[c]];
    res.weighted_rss = cf.rss;
    plugin_components(P, alpha, psi, res.plugin);
    E.fitted.upload(res.fitted_cell);
    // the control rows of the block factors: sqrt(f) Phi' zr, Phi Phi' = S^-1
    std::vector<double> phi;
    if (!cholesky(cf.s_inv, q, phi))
      fail(kInverseResidual, "generic_jla_control_factor", "control inverse square root failed");
    VK_HIP(hipMemcpy(E.ctl_small.p + q, phi.data(), (size_t)q * q * sizeof(double), hipMemcpyHostToDevice));
    E.ctl_u.alloc(P.rows * q);
    hipLaunchKernelGGL(dev::k_ctl_whiten, grid_for(P.rows), kBlock, 0, E.stream, P.rows, q, E.ctl_zry.p, E.ctl_f.p,
                       E.ctl_small.p + q, E.ctl_u.p);
  }
  // ---- full fit (engine.rs:1108-1134)
  if (!joint) {
    opt.checkpoint("jla_full_fit");
    std::vector<double> wr(E.W, 0.0), fr(E.F, 0.0);
    for (uint64_t c = 0; c < E.C; ++c) {
      wr[P.cell_worker[c]] += P.cell_outcome[c];
      fr[P.cell_firm[c]] += P.cell_outcome[c];
    }
    VK_HIP(hipMemcpy(E.wrhs.p, wr.data(), E.W * sizeof(double), hipMemcpyHostToDevice));
    VK_HIP(hipMemcpy(E.frhs.p, fr.data(), E.F * sizeof(double), hipMemcpyHostToDevice));
    std::vector<ColumnReceipt> rc;
    solve_columns<1>(E, 1, opt.fit_tolerance, 0.01, opt, rc);
    res.fit = rc[0];
    std::vector<double> alpha, psi;
    E.worker.download(alpha, E.W, E.stream);
    E.fsol.download(psi, E.F, E.stream);
    res.fitted_cell.resize(E.C);
    for (uint64_t c = 0; c < E.C; ++c) res.fitted_cell[c] = alpha[P.cell_worker[c]] + psi[P.cell_firm[c]];
    Neumaier rss;
    for (uint64_t r = 0; r < P.rows; ++r) {
      const double e = P.outcome[r] - res.fitted_cell[P.row_cell[r]];
      rss.add(static_cast<double>(P.frequency[r]) * e * e);
    }
    res.weighted_rss = rss.value();
    plugin_components(P, alpha, psi, res.plugin);
    res.max_complete_residual = res.fit.complete;
    E.fitted.upload(res.fitted_cell);
  }

  // ---- leverage probes (engine.rs:1150-1246)
  auto tl = std::chrono::steady_clock::now();
  E.moments.alloc(10 * E.G);
  VK_HIP(hipMemsetAsync(E.moments.p, 0, 10 * E.G * sizeof(double), E.stream));
  E.atoms.alloc(std::max<uint64_t>((uint64_t)KL * E.G, (uint64_t)KT * E.S));
  if (E.Go) {
    E.obs_m.alloc(4 * E.Go);
    VK_HIP(hipMemsetAsync(E.obs_m.p, 0, 4 * E.Go * sizeof(double), E.stream));
    E.obs_corr.alloc(4 * E.obs_physical);
    VK_HIP(hipMemsetAsync(E.obs_corr.p, 0, 4 * E.obs_physical * sizeof(double), E.stream));
  }
  for (uint64_t first = 0; first < R; first += lw) {
    opt.checkpoint("jla_leverage_batch");
    const uint32_t k = (uint32_t)std::min<uint64_t>(lw, R - first);
    std::vector<ColumnReceipt> rc;
    with_k(KL, [&](auto kk) {
      constexpr int K = decltype(kk)::value;
      vk::run(E, dev::k_atoms, grid_for(E.G * ((k + 3) / 4 + 1)),
              dev::Atoms{E.G, E.unit_rank.p, E.unit_count.p, E.unit_offset.p, opt.seed, kTagLeverage, first, k, KL,
                         E.atoms.p});
      vk::run(E, dev::k_unit_rhs<K>, grid_for(E.W * dev::Pack<K>::TPR),
              dev::UnitRhs<K>{E.W, E.worker_ptr.p, nullptr, E.cell_unit_ptr.p, E.cell_units.p, E.atoms.p, E.G, k,
                              E.wrhs.p, E.cell_unit.p});
      segments<double, K, 0>(E, E.firm_chunks,
                             dev::OpFirmUnits<K>{E.firm_cells.p, E.cell_unit_ptr.p, E.cell_units.p, E.atoms.p, E.G,
                                                 k, E.frhs.p, E.firm_units.p});
      solve_columns<K>(E, k, opt.probe_tolerance, 1.0, opt, rc);
      if (E.Gm)
        vk::run(E, dev::k_moments<K>, grid_for(E.Gm * dev::Pack<K>::TPR),
                dev::Moments<K>{E.Gm, k, nullptr, E.unit_count.p, E.unit_worker.p, E.unit_firm.p, E.worker.p,
                                E.fsol.p, E.atoms.p, E.moments.p, E.sink()});
      if (E.Go)  // generic_jla.rs:5540-5892, units [Gm, G)
        vk::run(E, dev::k_obs_moments<K>, grid_for(E.Go * dev::Pack<K>::TPR),
                dev::ObsMoments<K>{E.Go, k, first, opt.seed, kTagLeverage, E.unit_cell.p + E.Gm,
                                   E.unit_count.p + E.Gm, E.unit_rank.p + E.Gm, E.unit_offset.p + E.Gm,
                                   E.row_phys.p, E.obs_physical, E.cell_worker.p, E.cell_firm.p, E.worker.p,
                                   E.fsol.p, E.obs_m.p, E.obs_corr.p, E.sink()});
    });
    for (auto &c : rc) {
      res.leverage.push_back(c);
      res.max_complete_residual = std::max(res.max_complete_residual, c.complete);
    }
  }
  E.check_status("jla_leverage");
  res.seconds_leverage = std::chrono::duration<double>(std::chrono::steady_clock::now() - tl).count();

  // ---- leverage adjustment (engine.rs:2269-2410)
  opt.checkpoint("jla_leverage_adjustment");
  E.adjust.alloc(7 * std::max<uint64_t>(E.Gm, 1));
  E.cellw.alloc(E.C);
  const dev::AdjustParams aprm{(double)R, opt.block_tolerance, opt.rank_tolerance};
  if (E.Gm)
    hipLaunchKernelGGL(dev::k_adjust, grid_for(E.Gm), kBlock, 0, E.stream, E.Gm, E.moments.p, E.unit_cell.p,
                       E.unit_count.p, E.unit_outcome.p, E.fitted.p, aprm, E.adjust.p, E.sink());
  if (E.Go) {  // generic_jla.rs:5893-6017
    E.obs_adjust.alloc(7 * E.Go);
    hipLaunchKernelGGL(dev::k_obs_adjust, grid_for(E.Go), kBlock, 0, E.stream, E.Go, E.obs_m.p, E.obs_corr.p,
                       E.obs_physical, E.row_phys.p, E.unit_count.p + E.Gm, E.unit_cell.p + E.Gm, E.row_outcome.p,
                       E.fitted.p, aprm, E.obs_adjust.p, E.sink());
  }
  E.check_status("jla_leverage_adjustment");
  // per-unit correction weight contributions, match units then observation units
  E.contrib.alloc(E.G);
  if (E.Gm)
    VK_HIP(hipMemcpyAsync(E.contrib.p, E.adjust.p + 6 * E.Gm, E.Gm * sizeof(double), hipMemcpyDeviceToDevice,
                          E.stream));
  if (E.Go)
    VK_HIP(hipMemcpyAsync(E.contrib.p + E.Gm, E.obs_adjust.p + 6 * E.Go, E.Go * sizeof(double),
                          hipMemcpyDeviceToDevice, E.stream));
  std::vector<double> m2;  // sum over units of a_g b_g' (q x q)
  if (joint) {
    // match blocks with the exact control part (generic_jla.rs:6053-6273):
    // per unit contrib = A B, (a, b) and (sum f e, B); per cell [Zc | hc]
    E.ctl_uab.alloc(E.Gm * 2 * q);
    E.ctl_umass.alloc(2 * E.Gm);
    launch_blocks(E, q, opt);
    E.check_status("generic_jla_match_adjustment");
    // 64 doubles of zero padding: the target kernel reads rows 32 or 64 wide
    E.ctl_cellmat.alloc(E.C * 2 * q + 64);
    VK_HIP(hipMemsetAsync(E.ctl_cellmat.p + E.C * 2 * q, 0, 64 * sizeof(double), E.stream));
    hipLaunchKernelGGL(dev::k_ctl_cells, grid_for(E.C), kBlock, 0, E.stream, E.C, q, E.cell_unit_ptr.p,
                       E.cell_units.p, E.unit_outcome.p, E.ctl_umass.p, E.ctl_uab.p, E.ctl_cellz.p, E.ctl_cellmat.p);
    const std::vector<double> cross = device_gram(E, E.Gm, 2 * q, 1, E.ctl_uab.p, nullptr);
    m2.assign((size_t)q * q, 0.0);
    for (uint32_t i = 0; i < q; ++i)
      for (uint32_t j = 0; j < q; ++j) m2[(size_t)i * q + j] = 0.5 * (cross[(size_t)i * q + j] + cross[(size_t)j * q + i]);
  }
  hipLaunchKernelGGL(dev::k_cell_weight, grid_for(E.C), kBlock, 0, E.stream, E.C, E.cell_unit_ptr.p,
                     E.cell_units.p, E.contrib.p, E.cellw.p);
  {
    std::vector<double> adj, oadj;
    if (E.Gm) E.adjust.download(adj, 6 * E.Gm, E.stream);
    if (E.Go) E.obs_adjust.download(oadj, 6 * E.Go, E.stream);
    auto slice = [&](int q, std::vector<double> &out) {
      out.assign(adj.begin() + q * E.Gm, adj.begin() + (q + 1) * E.Gm);
      out.insert(out.end(), oadj.begin() + q * E.Go, oadj.begin() + (q + 1) * E.Go);
    };
    slice(0, res.projection_share); slice(1, res.residual_share); slice(2, res.finite_bias);
    slice(3, res.finite_variance); slice(4, res.residual_mass); slice(5, res.deleted_mass);
    E.cellw.download(res.cell_weight, E.C, E.stream);
    if (joint) {
      std::vector<double> um;
      E.ctl_umass.download(um, 2 * E.Gm, E.stream);
      for (uint64_t g = 0; g < E.Gm; ++g) {
        res.residual_mass[g] = um[2 * g];
        res.deleted_mass[g] = um[2 * g + 1];
      }
    }
  }

  // ---- numerical attachment state (engine/numerical.rs:62-122)
  NumericalResult &num = res.numerical;
  if (opt.numerical && E.Gm) {
    res.has_numerical = true;
    E.ngrad.alloc(5 * E.Gm);
    E.noffs.alloc(E.Gm);
    E.nfolds.alloc(12 * E.Gm);
    VK_HIP(hipMemsetAsync(E.nfolds.p, 0, 12 * E.Gm * sizeof(double), E.stream));
    const unsigned sg = reduce_grid(E.Gm);
    E.npart.alloc(5 * (size_t)sg);
    hipLaunchKernelGGL(dev::k_numerical_state, sg, kBlock, 0, E.stream, E.Gm, E.moments.p, (double)R, E.ngrad.p,
                       E.noffs.p, E.npart.p);
    std::vector<double> part;
    E.npart.download(part, 5 * (size_t)sg, E.stream);
    num.minimum_constrained = INFINITY;
    num.minimum_residual_margin = INFINITY;
    num.maximum_sensitivity_ratio = 0.0;
    double last = -1.0;
    bool clipped = false;
    for (unsigned b = 0; b < sg; ++b) {
      num.minimum_constrained = std::min(num.minimum_constrained, part[5 * b]);
      num.minimum_residual_margin = std::min(num.minimum_residual_margin, part[5 * b + 1]);
      num.maximum_sensitivity_ratio = std::max(num.maximum_sensitivity_ratio, part[5 * b + 2]);
      last = std::max(last, part[5 * b + 3]);
      clipped = clipped || part[5 * b + 4] != 0.0;
    }
    // the last unit with an event sets the status (numerical.rs:88-120)
    num.status = last >= 0.0 ? (uint32_t)((uint64_t)last % 8) : (clipped ? 5u : 2u);
    num.allocation_bytes = (5 + 1 + 12) * E.Gm * sizeof(double) + 6 * R * sizeof(double);
  }
  if (opt.numerical && E.Go) {
    // observation units (generic_jla/numerical.rs:172-270): per copy the
    // finite derivative of its inverse, pulled back to the row's p^2, p^4 and
    // the copy's s p, s p^3 (numerical_mc.rs:137-145); units in class order
    // (class rank, then row), so that the last event sets the status
    res.has_numerical = true;
    const uint64_t Go = E.Go, N = E.obs_physical;
    std::vector<double> m4, corr;
    E.obs_m.download(m4, 4 * Go, E.stream);
    E.obs_corr.download(corr, 4 * N, E.stream);
    std::vector<uint64_t> phys(Go + 1, 0);
    for (uint64_t i = 0; i < Go; ++i) phys[i + 1] = phys[i] + P.unit_count[E.Gm + i];
    std::vector<uint64_t> order(Go);
    std::iota(order.begin(), order.end(), 0ull);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint64_t a, uint64_t b) { return P.unit_rank[E.Gm + a] < P.unit_rank[E.Gm + b]; });
    E.obs_class_words = 0;
    for (uint64_t k = 0; k < Go;) {
      uint64_t mass = 0, q = k;
      for (; q < Go && P.unit_rank[E.Gm + order[q]] == P.unit_rank[E.Gm + order[k]]; ++q)
        mass += P.unit_count[E.Gm + order[q]];
      E.obs_class_words += (mass + 63) / 64;
      k = q;
    }
    std::vector<double> grad(3 * Go, 0.0), copy(2 * N, 0.0);
    if (!E.Gm) {
      num.minimum_constrained = INFINITY;
      num.minimum_residual_margin = INFINITY;
      num.maximum_sensitivity_ratio = 0.0;
    }
    bool clipped = false;
    uint32_t event = 0;
    const double r = (double)R;
    for (uint64_t k = 0; k < Go; ++k) {
      const uint64_t i = order[k];
      const double a2 = (m4[i] + m4[Go + i]) / r, a4 = (m4[2 * Go + i] + m4[3 * Go + i]) / r;
      Neumaier g2, g4, off;
      for (uint64_t x = phys[i]; x < phys[i + 1]; ++x) {
        const double c1 = (corr[x] + corr[N + x]) / r, c3 = (corr[2 * N + x] + corr[3 * N + x]) / r;
        const double u[5] = {a2, 1.0 + a2 - 2.0 * c1, a4, 1.0 + 6.0 * a2 + a4 - 4.0 * c1 - 4.0 * c3,
                             a2 + a4 - 2.0 * c3};
        num.minimum_constrained = std::min(num.minimum_constrained, u[0] + u[1]);
        double gr[5], h = 0.0, mm = 0.0, var = 0.0;
        const int code = dev::unit_linearisation(u, r, gr, h, mm, var);
        if (code == 1) {
          event = 6;
          continue;
        }
        num.minimum_residual_margin = std::min(num.minimum_residual_margin, mm);
        num.maximum_sensitivity_ratio = std::max(num.maximum_sensitivity_ratio, std::sqrt(std::max(var, 0.0)) / mm);
        if (var < 0.0) clipped = true;
        if (code == 5) {
          event = 5;
          continue;
        }
        if (code == 2) {
          event = 6;
          continue;
        }
        const double pb[4] = {gr[0] + gr[1] + 6.0 * gr[3] + gr[4], gr[2] + gr[3] + gr[4],
                              -2.0 * gr[1] - 4.0 * gr[3], -4.0 * gr[3] - 2.0 * gr[4]};
        g2.add(pb[0]);
        g4.add(pb[1]);
        off.add(pb[2] * c1);
        off.add(pb[3] * c3);
        copy[x] = pb[2];
        copy[N + x] = pb[3];
      }
      off.add(g2.value() * a2);
      off.add(g4.value() * a4);
      grad[i] = g2.value();
      grad[Go + i] = g4.value();
      grad[2 * Go + i] = off.value();
    }
    // after the match units' state: the observation adjustment's variance
    // clipping (generic_jla/numerical.rs:172-176), then the last event
    const uint32_t base = E.Gm ? num.status : 2u;
    num.status = event ? event : (clipped ? 5u : base);
    E.nograd.upload(grad);
    E.ncopy.upload(copy);
    E.nofolds.alloc(12 * Go);
    VK_HIP(hipMemsetAsync(E.nofolds.p, 0, 12 * Go * sizeof(double), E.stream));
    num.allocation_bytes += (3 * Go + 2 * N + 12 * Go) * sizeof(double) + (E.Gm ? 0 : 6 * R * sizeof(double));
  }


  // ---- target probes (engine.rs:1373-1484)
  auto tt = std::chrono::steady_clock::now();
  E.dir.alloc((uint64_t)KT * E.C);
  res.draws.assign(4 * R, 0.0);
  uint32_t ctl_k2 = 0;
  for (uint64_t first = 0; first < R; first += tw) {
    opt.checkpoint("jla_target_batch");
    const uint32_t k = (uint32_t)std::min<uint64_t>(tw, R - first);
    VK_HIP(hipMemsetAsync(E.atoms.p, 0, (uint64_t)KT * E.S * sizeof(int32_t), E.stream));
    vk::run(E, dev::k_atoms, grid_for(E.S * ((k + 3) / 4 + 1)),
            dev::Atoms{E.S, E.stratum_rank.p, E.stratum_count.p, nullptr, opt.seed, kTagTarget, first, k, KT, E.atoms.p});
    std::vector<ColumnReceipt> rc;
    with_k(KT, [&](auto kt) {
      // target batches hold at most 32 probes (64 columns)
      constexpr int KTc = decltype(kt)::value > 32 ? 32 : decltype(kt)::value;
      constexpr int K2c = 2 * KTc;
      const unsigned cg = reduce_grid(E.C * KTc);
      vk::run(E, dev::k_target_direction<KTc>, cg,
              dev::TargetDirection<KTc>{E.C, E.cell_stratum_ptr.p, E.stratum_per_copy.p, E.target_total, E.atoms.p,
                                        E.dir.p, E.rpartial.p});
      finish<KTc, 2>(E, E.rpartial.p, cg, dev::FinTotals<KTc>{E.totals.p});
      vk::run(E, dev::k_target_scores<KTc>, grid_for(E.W * K2c),
              dev::TargetScores<KTc>{E.W, E.worker_ptr.p, nullptr, E.dir.p, E.cell_target.p, E.target_total,
                                     E.totals.p, 0, E.wrhs.p});
      segments<double, K2c, 0>(E, E.firm_chunks,
                               dev::OpFirmTargetScores<KTc>{E.firm_cells.p, E.dir.p, E.cell_target.p, E.totals.p,
                                                            E.target_total, E.frhs.p});
      const unsigned bw = reduce_grid(E.W * K2c), bf = reduce_grid(E.F * K2c);
      vk::run(E, dev::k_balance_sum<K2c>, bw, dev::BalanceSum<K2c>{E.W, E.wrhs.p, E.rpartial.p});
      finish<K2c, 1>(E, E.rpartial.p, bw, dev::FinBalance<K2c>{E.wrhs.p, E.W, 0, E.totals.p, E.sink()});
      vk::run(E, dev::k_balance_sum<K2c>, bf, dev::BalanceSum<K2c>{E.F, E.frhs.p, E.rpartial.p});
      finish<K2c, 1>(E, E.rpartial.p, bf, dev::FinBalance<K2c>{E.frhs.p, E.F, 1, E.totals.p, E.sink()});
      solve_columns<K2c>(E, 2 * k, opt.probe_tolerance, 1.0, opt, rc);
      if (opt.numerical && E.Go)  // observation folds: y e kappa / fold size (generic_jla/numerical.rs:377-389)
        hipLaunchKernelGGL((dev::k_numerical_folds<K2c>), grid_for(E.Go), kBlock, 0, E.stream, E.Go, k, first, R,
                           E.unit_cell.p + E.Gm, E.cell_worker.p, E.cell_firm.p, E.row_outcome.p,
                           E.obs_adjust.p + 4 * E.Go, E.worker.p, E.fsol.p, E.nofolds.p);
      if (opt.numerical && E.Gm)
        hipLaunchKernelGGL((dev::k_numerical_folds<K2c>), grid_for(E.Gm), kBlock, 0, E.stream, E.Gm, k, first, R,
                           E.unit_cell.p, E.cell_worker.p, E.cell_firm.p, E.unit_outcome.p, E.adjust.p + 4 * E.Gm,
                           E.worker.p, E.fsol.p, E.nfolds.p);
      vk::run(E, dev::k_contract<KTc>, cg,
              dev::Contract<KTc>{E.C, E.cell_worker.p, E.cell_firm.p, E.cellw.p, E.worker.p, E.fsol.p, E.rpartial.p});
      finish<KTc, 3>(E, E.rpartial.p, cg, dev::FinDraws{E.draws.p});
      if (joint) {
        // statistical_batches.rs:259-346 with the joint solve's control
        // coefficients: per column s, Y_s = sum_c P_cs Zc_c and H_s = sum_c
        // P_cs hc_c (the host adds beta_s' H_s + beta_s' M2 beta_s with
        // beta_s = -S^-1 Y_s)
        const uint32_t m = 2 * q;
        const unsigned tb = (unsigned)std::max<uint64_t>(
            1, std::min<uint64_t>(4096, (E.C + kBlock / K2c - 1) / (kBlock / K2c)));
        E.ctl_part.alloc((size_t)tb * m * K2c);
        E.ctl_out.alloc((size_t)m * K2c);
        auto go = [&](auto qm) {
          constexpr int QM = decltype(qm)::value;
          hipLaunchKernelGGL((dev::k_ctl_target<K2c, QM>), tb, kBlock, 0, E.stream, E.C, m, E.cell_worker.p,
                             E.cell_firm.p, E.worker.p, E.fsol.p, E.ctl_cellmat.p, E.ctl_part.p);
        };
        if (m <= 16) go(std::integral_constant<int, 16>{});
        else if (m <= 32) go(std::integral_constant<int, 32>{});
        else go(std::integral_constant<int, 64>{});
        hipLaunchKernelGGL(dev::k_ctl_finish_wave, grid_for((uint64_t)m * K2c * 64), kBlock, 0, E.stream, tb,
                           m * K2c, E.ctl_part.p, E.ctl_out.p);
        ctl_k2 = K2c;
      }
    });
    std::vector<double> d;
    E.draws.download(d, 4 * k, E.stream);
    if (joint) control_draws(E, q, m2, ctl_k2, k, d);
    std::copy(d.begin(), d.end(), res.draws.begin() + 4 * first);
    for (auto &c : rc) {
      res.target.push_back(c);
      res.max_complete_residual = std::max(res.max_complete_residual, c.complete);
    }
  }
  E.check_status("jla_target");
  res.seconds_target = std::chrono::duration<double>(std::chrono::steady_clock::now() - tt).count();

  // ---- finalize (engine.rs:3108-3272)
  opt.checkpoint("jla_finalize");
  Neumaier m[4];
  for (uint64_t t = 0; t < R; ++t)
    for (int j = 0; j < 4; ++j) {
      if (!std::isfinite(res.draws[4 * t + j])) fail(kCorrectionNonFinite, "jla_target", "nonfinite target draw");
      m[j].add(res.draws[4 * t + j]);
    }
  for (int j = 0; j < 4; ++j) res.correction[j] = m[j].value() / (double)R;
  for (int j = 0; j < 4; ++j) res.corrected[j] = res.plugin[j] - res.correction[j];
  Neumaier v[4];
  for (uint64_t t = 0; t < R; ++t)
    for (int j = 0; j < 4; ++j) {
      const double e = res.draws[4 * t + j] - res.correction[j];
      v[j].add(e * e);
    }
  const double denom = (double)(R * (R - 1));
  for (int j = 0; j < 4; ++j) res.mcse[j] = std::sqrt(v[j].value() / denom);
  double acc = std::max({component_identity(res.plugin), component_identity(res.correction),
                         component_identity(res.corrected)});
  for (uint64_t t = 0; t < R; ++t) acc = std::max(acc, component_identity(&res.draws[4 * t]));
  res.accounting_residual = acc;
  if (acc > 4096.0 * 2.220446049250313e-16)
    fail(kTargetIdentity, "jla_accounting", "component accounting residual exceeds the roundoff gate");
  if (opt.numerical) numerical_replay(E, opt, res);
  res.seconds_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  res.vcycles = E.vcycles;
  return res;
}

void DeviceEngine::atoms(uint64_t seed, uint64_t tag, uint64_t first_probe, uint32_t columns, bool target,
                         std::vector<int64_t> &out) {
  Impl &E = *impl_;
  const uint64_t n = target ? E.S : E.G;
  DBuf<int32_t> buf;
  buf.alloc(n * columns);
  vk::run(E, dev::k_atoms, grid_for(n * ((columns + 3) / 4 + 1)),
          dev::Atoms{n, target ? E.stratum_rank.p : E.unit_rank.p, target ? E.stratum_count.p : E.unit_count.p,
                     target ? nullptr : E.unit_offset.p, seed, tag, first_probe, columns, columns, buf.p});
  std::vector<int32_t> h;
  buf.download(h, n * columns, E.stream);
  out.resize(h.size());
  for (uint64_t row = 0; row < n; ++row)
    for (uint32_t col = 0; col < columns; ++col) out[col * n + row] = (int64_t)h[row * columns + col];
}

void DeviceEngine::spmv(uint32_t level, const std::vector<double> &xin, uint32_t columns, std::vector<double> &y) {
  Impl &E = *impl_;
  PlainCycle plain(E);  // diagnostics use the reference cycle
  if (level >= E.levels.size()) fail(kInvalidInput, "device", "level out of range");
  const uint32_t K = lanes_for(columns);
  if (K != columns) fail(kInvalidInput, "device", "columns must be a power of two");
  DevLevel &L = E.levels[level];
  const uint64_t m = (uint64_t)L.n * K;
  if (xin.size() != m) fail(kInvalidInput, "device", "input has the wrong size");
  E.ensure(K);
  DBuf<double> a, b;
  a.upload(to_renumbered(xin, L.perm, K));
  b.alloc(m);
  E.gate = E.always();
  with_k(K, [&](auto kk) {
    constexpr int KK = decltype(kk)::value;
    segments<double, KK, 0>(E, L.rows, dev::OpProduct<double, KK>{{L.col.p, L.val.p}, a.p, b.p});
  });
  std::vector<double> yn;
  b.download(yn, m, E.stream);
  y = from_renumbered(yn, L.perm, K);
}

void DeviceEngine::vcycle(const std::vector<double> &rhs, uint32_t columns, int precision, std::vector<double> &out) {
  Impl &E = *impl_;
  PlainCycle plain(E);  // diagnostics use the reference cycle
  const uint32_t K = lanes_for(columns);
  if (K != columns) fail(kInvalidInput, "device", "columns must be a power of two");
  E.ensure(K);
  const uint64_t m0 = E.n0 * K;
  if (rhs.size() != m0) fail(kInvalidInput, "device", "rhs has the wrong size");
  const std::vector<double> rn = to_renumbered(rhs, E.levels[0].perm, K);
  VK_HIP(hipMemcpy(E.guess_rhs.p, rn.data(), m0 * sizeof(double), hipMemcpyHostToDevice));
  E.gate = E.always();
  const double *zp = nullptr;
  with_k(K, [&](auto kk) {
    constexpr int KK = decltype(kk)::value;
    zp = apply_cycle<KK>(E, E.guess_rhs.p, nullptr, E.rc.p, precision == 32);
  });
  std::vector<double> zn(m0);
  VK_HIP(hipMemcpyAsync(zn.data(), zp, m0 * sizeof(double), hipMemcpyDeviceToHost, E.stream));
  VK_HIP(hipStreamSynchronize(E.stream));
  out = from_renumbered(zn, E.levels[0].perm, K);
}

void DeviceEngine::pcg(const std::vector<double> &rhs, uint32_t columns, double tolerance, int precision,
                       std::vector<double> &xout, std::vector<ColumnReceipt> &receipts) {
  Impl &E = *impl_;
  PlainCycle plain(E);  // diagnostics use the reference cycle
  const uint32_t K = lanes_for(columns);
  if (K != columns) fail(kInvalidInput, "device", "columns must be a power of two");
  E.ensure(K);
  const uint64_t m0 = E.n0 * K;
  if (rhs.size() != m0) fail(kInvalidInput, "device", "rhs has the wrong size");
  const std::vector<double> rn = to_renumbered(rhs, E.levels[0].perm, K);
  VK_HIP(hipMemcpy(E.guess_rhs.p, rn.data(), m0 * sizeof(double), hipMemcpyHostToDevice));
  std::vector<uint8_t> use(K, 1);
  std::vector<dev::ColState> st;
  with_k(K, [&](auto kk) {
    constexpr int KK = decltype(kk)::value;
    pcg_lanes<KK>(E, E.guess_rhs.p, tolerance, 0.0, 1000, 25, false, precision == 32, use, st);
  });
  receipts.assign(K, ColumnReceipt{});
  for (uint32_t j = 0; j < K; ++j) {
    if (st[j].status) fail(st[j].status, "pcg", "lane " + std::to_string(j) + " failed");
    receipts[j].iterations = st[j].iters;
    receipts[j].restarts = st[j].restarts;
    receipts[j].rhs_norm = st[j].rhs_norm;
  }
  std::vector<double> xn;
  E.x.download(xn, m0, E.stream);
  xout = from_renumbered(xn, E.levels[0].perm, K);
}

ControlFit DeviceEngine::control_fit(const SolveOptions &opt) {
  Impl &E = *impl_;
  if (E.P->controls == 0) fail(kInvariant, "controls", "the prepared problem has no controls");
  std::vector<double> alpha, psi;
  return fit_controls(E, opt, alpha, psi);
}

void DeviceEngine::refresh_outcome() {
  Impl &E = *impl_;
  const Prepared &p = *E.P;
  E.unit_outcome.upload(p.unit_outcome);
  if (E.Go) {
    const uint64_t first_row = p.rows - E.Go;
    std::vector<double> y(p.outcome.begin() + first_row, p.outcome.end());
    E.row_outcome.upload(y);
  }
}

void DeviceEngine::solve_batch(const std::vector<double> &wr, const std::vector<double> &fr, uint32_t columns,
                               double tolerance, double inner_ratio, const SolveOptions &opt,
                               std::vector<double> &worker, std::vector<double> &firm,
                               std::vector<ColumnReceipt> &receipts, uint64_t refine[2]) {
  Impl &E = *impl_;
  const uint64_t W = E.W, F = E.F;
  if (columns == 0 || columns > 64) fail(kInvalidInput, "cmg_full_v2", "1 to 64 columns per batch");
  if (wr.size() != W * columns || fr.size() != F * columns)
    fail(kInvalidInput, "cmg_full_v2", "direct hybrid batch RHS arrays have incompatible dimensions");
  const uint32_t K = lanes_for(columns);
  E.ensure(K);
  E.gate = E.always();
  // column-major host blocks -> lane-interleaved device rows (unused lanes zero)
  std::vector<double> wi(W * K, 0.0), fi(F * K, 0.0);
  for (uint32_t j = 0; j < columns; ++j) {
    for (uint64_t w = 0; w < W; ++w) wi[w * K + j] = wr[j * W + w];
    for (uint64_t f = 0; f < F; ++f) fi[f * K + j] = fr[j * F + f];
  }
  VK_HIP(hipMemcpy(E.wrhs.p, wi.data(), wi.size() * sizeof(double), hipMemcpyHostToDevice));
  VK_HIP(hipMemcpy(E.frhs.p, fi.data(), fi.size() * sizeof(double), hipMemcpyHostToDevice));
  const uint64_t rounds0 = E.refine_rounds, refined0 = E.refined_columns;
  with_k(K, [&](auto kk) {
    constexpr int KK = decltype(kk)::value;
    solve_columns<KK>(E, columns, tolerance, inner_ratio, opt, receipts);
  });
  refine[0] = E.refine_rounds - rounds0;
  refine[1] = E.refined_columns - refined0;
  std::vector<double> wo, fo;
  E.worker.download(wo, W * K, E.stream);
  E.fsol.download(fo, F * K, E.stream);
  worker.assign(W * columns, 0.0);
  firm.assign(F * columns, 0.0);
  for (uint32_t j = 0; j < columns; ++j) {
    for (uint64_t w = 0; w < W; ++w) worker[j * W + w] = wo[w * K + j];
    for (uint64_t f = 0; f < F; ++f) firm[j * F + f] = fo[f * K + j];
  }
}

// ------------------------------------------------ no-allocation level product
// Device scratch of vckss_mi355x_spmv_noalloc_v1: the renumbered input and
// output blocks and the chunk partials of rows split over several chunks.
// Every GPU holds its own copy for the life of the process; all calls on one
// GPU share it (it keeps the last call's contents, and calls that overlap on
// one GPU race).
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
double *g_scratch_address[kMaxDevices] = {};
std::mutex g_scratch_mutex;
}  // namespace

// The device address of `scratch` on the current device, looked up once per
// device.  The array's name is a host-side stand-in: kernels and copies take
// this address, never the name.
static double *scratch_address() {
  int device = 0;
  VK_HIP(hipGetDevice(&device));
  if (device < 0 || device >= kMaxDevices) fail(kResourceLimit, "spmv_noalloc", "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_address[device]) {
    double *ptr = nullptr;
    const hipError_t rc = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !ptr)
      fail(kResourceLimit, "spmv_noalloc",
           std::string("hipGetSymbolAddress of the no-allocation scratch failed: ") + hipGetErrorString(rc));
    g_scratch_address[device] = ptr;
  }
  return g_scratch_address[device];
}

void DeviceEngine::spmv_noalloc(uint32_t level, uint32_t columns, const double *x, double *y) {
  Impl &E = *impl_;
  if (level >= E.levels.size()) fail(kInvalidInput, "spmv_noalloc", "level out of range");
  if (!x || !y) fail(kInvalidInput, "spmv_noalloc", "null device pointer");
  const uint32_t K = lanes_for(columns);
  if (K != columns) fail(kInvalidInput, "spmv_noalloc", "columns must be a power of two in [1, 64]");
  DevLevel &L = E.levels[level];
  const uint64_t m = (uint64_t)L.n * K, parts = (uint64_t)std::max<uint32_t>(L.rows.count, 1) * K;
  const uint64_t need = 2 * m + parts;
  if (need > kNoAllocScratchDoubles)
    fail(kResourceLimit, "spmv_noalloc",
         "level " + std::to_string(level) + " with " + std::to_string(columns) + " columns needs " +
             std::to_string(need) + " scratch doubles; the no-allocation scratch holds " +
             std::to_string(kNoAllocScratchDoubles));
  double *base = scratch_address();
  double *xr = base, *yr = base + m, *part = base + 2 * m;
  E.gate = E.always();
  hipLaunchKernelGGL(dev::k_permute_lanes, grid_for(m), kBlock, 0, E.stream, (uint64_t)L.n, K, L.dperm.p, x, xr,
                     1);
  with_k(K, [&](auto kk) {
    constexpr int KK = decltype(kk)::value;
    const dev::SegScratch sc{part, nullptr, E.gate};
    segments_in<double, KK, 0>(E, L.rows, dev::OpProduct<double, KK>{{L.col.p, L.val.p}, xr, yr}, sc);
  });
  hipLaunchKernelGGL(dev::k_permute_lanes, grid_for(m), kBlock, 0, E.stream, (uint64_t)L.n, K, L.dperm.p, yr, y,
                     0);
  VK_HIP(hipStreamSynchronize(E.stream));
}

double DeviceEngine::time_leverage_batches(const SolveOptions &opt, uint32_t batches, uint64_t &solves,
                                           uint64_t &iterations) {
  Impl &E = *impl_;
  const uint32_t lw = std::max<uint32_t>(1, std::min<uint32_t>(opt.leverage_width, 64));
  const uint32_t KL = lanes_for(lw);
  E.ensure(KL);
  E.atoms.alloc((uint64_t)KL * E.G);
  E.moments.alloc(10 * E.G);
  E.fitted.alloc(E.C);
  VK_HIP(hipMemsetAsync(E.moments.p, 0, 10 * E.G * sizeof(double), E.stream));
  VK_HIP(hipStreamSynchronize(E.stream));
  E.drop_timing();
  E.timing = true;
  solves = 0;
  iterations = 0;
  auto t0 = std::chrono::steady_clock::now();
  for (uint32_t b = 0; b < batches; ++b) {
    const uint64_t first = (uint64_t)b * lw;
    vk::run(E, dev::k_atoms, grid_for(E.G * ((lw + 3) / 4 + 1)),
            dev::Atoms{E.G, E.unit_rank.p, E.unit_count.p, E.unit_offset.p, opt.seed, kTagLeverage, first, lw, KL,
                       E.atoms.p});
    std::vector<ColumnReceipt> rc;
    with_k(KL, [&](auto kk) {
      constexpr int K = decltype(kk)::value;
      vk::run(E, dev::k_unit_rhs<K>, grid_for(E.W * dev::Pack<K>::TPR),
              dev::UnitRhs<K>{E.W, E.worker_ptr.p, nullptr, E.cell_unit_ptr.p, E.cell_units.p, E.atoms.p, E.G, lw,
                              E.wrhs.p, E.cell_unit.p});
      segments<double, K, 0>(E, E.firm_chunks,
                             dev::OpFirmUnits<K>{E.firm_cells.p, E.cell_unit_ptr.p, E.cell_units.p, E.atoms.p, E.G, lw,
                                                 E.frhs.p, E.firm_units.p});
      solve_columns<K>(E, lw, opt.probe_tolerance, 1.0, opt, rc);
      vk::run(E, dev::k_moments<K>, grid_for(E.G * dev::Pack<K>::TPR),
              dev::Moments<K>{E.G, lw, nullptr, E.unit_count.p, E.unit_worker.p, E.unit_firm.p, E.worker.p, E.fsol.p,
                              E.atoms.p, E.moments.p, E.sink()});
    });
    for (auto &c : rc) iterations += c.iterations;
    solves += lw;
  }
  VK_HIP(hipStreamSynchronize(E.stream));
  E.timing = false;
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

DeviceEngine::Tuning DeviceEngine::autotune(const SolveOptions &base, uint32_t batches) {
  Impl &E = *impl_;
  Tuning best;
  double default_rate = 0.0;
  for (uint32_t chunk : {32u, 64u, 128u}) {
    relayout(E, chunk);
    for (uint32_t width : {16u, 32u, 64u}) {
      SolveOptions o = base;
      o.leverage_width = width;
      uint64_t solves = 0, iterations = 0;
      (void)time_leverage_batches(o, 1, solves, iterations);  // warm
      const double sec = time_leverage_batches(o, batches, solves, iterations);
      const double rate = sec > 0 ? (double)solves / sec : 0.0;
      best.trials.push_back({width, chunk, rate});
      if (width == 64 && chunk == kChunkEntries) default_rate = rate;
      if (rate > best.rate) {
        best.rate = rate;
        best.width = width;
        best.chunk = chunk;
      }
    }
  }
  // the upload shape (64 probes, 64-entry chunks) stays unless another
  // candidate is clearly faster, so timing noise does not flip the choice
  if (best.rate < 1.03 * default_rate) {
    best.rate = default_rate;
    best.width = 64;
    best.chunk = kChunkEntries;
  }
  relayout(E, best.chunk);
  return best;
}

void DeviceEngine::set_chunk(uint32_t chunk) { relayout(*impl_, chunk); }

void DeviceEngine::set_first_pass_cap(bool enable) { impl_->first_pass_cap = enable; }
void DeviceEngine::set_lockstep(Lockstep *l) { impl_->lockstep = l; }
bool DeviceEngine::first_pass_cap() const { return impl_->first_pass_cap; }

void DeviceEngine::set_graphs(bool enable) {
  Impl &E = *impl_;
  E.use_graphs = enable;
  if (!enable) {
    VK_HIP(hipStreamSynchronize(E.stream));
    E.drop_graphs();
  }
}

void DeviceEngine::graph_counters(uint64_t &captures, uint64_t &launches) const {
  captures = impl_->graph_captures;
  launches = impl_->graph_launches;
}

void DeviceEngine::counters(uint64_t out[6]) const {
  const Impl &E = *impl_;
  const uint64_t v[6] = {E.vcycles, E.batches, E.refine_rounds, E.refined_columns, E.iterations_launched, 0};
  for (int i = 0; i < 6; ++i) out[i] = v[i];
}

double DeviceEngine::stream_copy(uint64_t bytes) {
  Impl &E = *impl_;
  const uint64_t count = std::max<uint64_t>(1, bytes / 32);
  DBuf<double> a, b;
  a.alloc(count * 4);
  b.alloc(count * 4);
  VK_HIP(hipMemsetAsync(a.p, 0, count * 32, E.stream));
  hipEvent_t s, e;
  VK_HIP(hipEventCreate(&s));
  VK_HIP(hipEventCreate(&e));
  VK_HIP(hipEventRecord(s, E.stream));
  hipLaunchKernelGGL(dev::k_stream_copy, grid_for(count), kBlock, 0, E.stream, count,
                     reinterpret_cast<const double4 *>(a.p), reinterpret_cast<double4 *>(b.p));
  VK_HIP(hipEventRecord(e, E.stream));
  VK_HIP(hipEventSynchronize(e));
  float ms = 0.f;
  VK_HIP(hipEventElapsedTime(&ms, s, e));
  (void)hipEventDestroy(s);
  (void)hipEventDestroy(e);
  return ms;
}

DeviceEngine::KernelTiming DeviceEngine::kernel_timing(int which) const {
  const Impl &E = *impl_;
  KernelTiming t;
  if (which < 0 || which > 1) fail(kInvalidInput, "kernel_timing", "timed kernel must be 0 (A p) or 1");
  const auto &events = E.timed[which].events;
  // Launches queued after every lane converged return at the device gate
  // without touching the level (a few microseconds); only launches that did
  // the pass (longer than a quarter of the longest) are averaged.
  std::vector<float> ms(events.size(), 0.f);
  float longest = 0.f;
  for (size_t k = 0; k < events.size(); ++k) {
    if (hipEventElapsedTime(&ms[k], events[k].first, events[k].second) != hipSuccess) ms[k] = 0.f;
    longest = std::max(longest, ms[k]);
  }
  double total = 0.0;
  for (float v : ms)
    if (v > 0.25f * longest) {
      total += v;
      t.launches++;
    }
  t.ms_per_launch = t.launches ? total / (double)t.launches : 0.0;
  t.bytes_per_launch = E.timed[which].bytes;
  return t;
}

}  // namespace vk
