// Device engine interface: the per-probe hot path of the leave-out estimator
// on one MI355X (gfx950).  Host code prepares the problem and the CMG
// hierarchy; everything per probe runs in HIP kernels.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "common.hpp"
#include "host_cmg.hpp"
#include "host_prepare.hpp"

namespace vk {

// Size of the no-allocation scratch (doubles, 32 MiB): 2 n columns + chunks
// x columns of the level must fit.
constexpr uint64_t kNoAllocScratchDoubles = 4ull << 20;
// Parameters of the learned smoother network (8 features, an input layer and
// three message-passing layers of width 64, a scalar output).
constexpr int kSmootherLayers = 3;
constexpr uint32_t kSmootherParams = 64 * 8 + 64 + kSmootherLayers * (2 * 64 * 64 + 64) + 64 + 1;

struct SolveOptions {
  uint64_t seed = 8675309;
  uint32_t probes = 200;
  uint32_t leverage_width = 32;   // probes per leverage batch (<= 64)
  uint32_t target_width = 32;     // probes per target batch (2 columns each, <= 32)
  double fit_tolerance = 1e-10;   // effective full-fit tolerance
  double probe_tolerance = 1e-6;  // effective probe tolerance
  uint32_t max_iterations = 10000;
  uint32_t recompute_interval = 100;
  double rank_tolerance = 1e-10;
  double block_tolerance = 1e-10;
  int precision = 64;             // 64: FP64 V-cycle; 32: FP32 V-cycle, FP64 Krylov
  bool numerical = false;         // numerical MCSE attachment (replayed leverage probes)
  bool exact = false;             // dense exact estimator instead of JLA probes
  bool joint_controls = false;    // the prepared controls enter jointly (joint nuisance)
  bool fixed_offset = false;      // the prepared controls enter as a fixed offset
  Interrupt *interrupt = nullptr; // caller's checkpoints (none when null)
  void checkpoint(const char *phase) const {
    if (interrupt) interrupt->checkpoint(phase);
  }
};

// Receipt of the dense exact estimator (exact_estimator.rs:127-153).
struct ExactReceipt {
  uint64_t parameters = 0, deletion_units = 0, cells = 0;
  double max_leverage = 0, deletion_rank_gap = 0, information_rcond = 0, inverse_relres = 0,
         fit_relres = 0, fit_tolerance = 0, firm_zero_sum_residual = 0;
  double seconds_factor = 0, seconds_cells = 0;
  // the correction's match-block and physical-copy parts (the stayer hybrid's
  // mover and stayer corrections, ffi_engine.rs:8772-8800)
  double mover_correction[4] = {0, 0, 0, 0}, stayer_correction[4] = {0, 0, 0, 0};
};

// Numerical MCSE attachment (engine/numerical.rs:13-32, numerical_api.rs:54-90).
struct NumericalResult {
  uint32_t status = 0;  // 1 exact zero .. 7 replay failed (numerical_api.rs:59-60)
  uint32_t leverage_probes = 0, target_probes = 0, fold_a = 0, fold_b = 0;
  uint64_t certified_replay_rhs = 0, attempted_replay_rhs = 0, replay_generator_words = 0,
           allocation_bytes = 0, point_rhs = 0;
  uint32_t failed_replay_probe = UINT32_MAX, replay_error_code = 0;
  double psd_adjustment = 0, minimum_constrained = 0, minimum_residual_margin = 0,
         maximum_sensitivity_ratio = 0;
  double score_means[2][3] = {{0, 0, 0}, {0, 0, 0}};
  double conditional[9] = {}, leverage[9] = {}, raw[9] = {}, usable[9] = {}, mcse[4] = {};
  double maximum_point_complete_residual = 0, maximum_replay_complete_residual = 0;
  std::vector<uint32_t> replay_probes;        // certified replay right-hand sides, in order
  std::vector<double> replay_complete;        // their complete residuals
};

struct ColumnReceipt {
  uint32_t iterations = 0, restarts = 0;
  double reduced = 0.0, complete = 0.0, rhs_norm = 0.0;
};

struct SolveResult {
  double plugin[4] = {0, 0, 0, 0}, correction[4] = {0, 0, 0, 0}, corrected[4] = {0, 0, 0, 0},
         mcse[4] = {0, 0, 0, 0};
  double weighted_rss = 0.0, accounting_residual = 0.0, max_complete_residual = 0.0;
  std::vector<double> fitted_cell, projection_share, residual_share, finite_bias, finite_variance,
      residual_mass, deleted_mass, cell_weight, draws;
  std::vector<ColumnReceipt> leverage, target;
  ColumnReceipt fit;
  double seconds_total = 0.0, seconds_leverage = 0.0, seconds_target = 0.0;
  uint64_t vcycles = 0;
  bool has_numerical = false;
  NumericalResult numerical;
  bool has_exact = false;
  ExactReceipt exact;
  // controls: the fit receipts of the control columns (then, under the fixed
  // offset, the outcome column of the joint fit), the control coefficients,
  // and the generalised reciprocal condition of the control information
  uint32_t controls = 0;
  bool fixed_offset = false;
  std::vector<ColumnReceipt> control_fit;
  std::vector<double> control_beta;
  double control_rcond = 0.0;
};

// The joint fit with the controls by Frisch-Waugh-Lovell: beta = S^-1 b from
// the control Schur complement S = Zr' F Zr and b = Zr' F yr of the FE
// residuals of the controls and the outcome (generic_jla.rs:1930-2000).
struct ControlFit {
  uint32_t q = 0;
  std::vector<double> beta, S, s_inv, b;
  std::vector<ColumnReceipt> receipts;  // the q control columns, then the outcome column
  double rss = 0.0, generalized_rcond = 0.0;
};

bool device_available();

// A sample batch whose engines run their PCG calls in lockstep: the calls
// waiting at the same time run together, each pass one launch over all their
// samples, every sample with its own convergence gate.  Each sample thread
// joins before its work and leaves after it (also on failure); the batch runs
// the waiting calls whenever every joined thread is waiting.
class Lockstep;
std::shared_ptr<Lockstep> make_lockstep();
void lockstep_join(Lockstep &l);
void lockstep_leave(Lockstep &l);

// Live and peak bytes of the engine's device buffers in this process.
void device_memory(uint64_t &live, uint64_t &peak);

class DeviceEngine {
 public:
  DeviceEngine(const Prepared &p, const HybridSystem &hs, const Hierarchy &h);
  ~DeviceEngine();
  SolveResult run(const SolveOptions &opt);
  // The joint fit with the prepared controls (Q + 1 two-way solves); refuses
  // a singular control block with status 82.
  ControlFit control_fit(const SolveOptions &opt);
  // Upload the unit and row outcome sums again after the prepared outcome
  // changed (the fixed-offset working outcome).
  void refresh_outcome();
  // Dense exact estimator (match deletion, joint nuisance, no controls):
  // exact leverages and target bilinear forms from the inverse information.
  SolveResult run_exact(const SolveOptions &opt);

  // Diagnostics used by the parity tests (no counterpart in the reference).
  void atoms(uint64_t seed, uint64_t tag, uint64_t first_probe, uint32_t columns, bool target,
             std::vector<int64_t> &out);
  void vcycle(const std::vector<double> &rhs, uint32_t columns, int precision, std::vector<double> &out);
  void spmv(uint32_t level, const std::vector<double> &x, uint32_t columns, std::vector<double> &y);
  // The same product on caller-owned device memory (n x columns, original
  // vertex order, lanes interleaved) with no device allocation: the working
  // blocks live in a fixed device scratch of kNoAllocScratchDoubles doubles.
  void spmv_noalloc(uint32_t level, uint32_t columns, const double *x, double *y);
  // Solve hybrid systems (n x columns, vertex-major) to a relative tolerance.
  void pcg(const std::vector<double> &rhs, uint32_t columns, double tolerance, int precision,
           std::vector<double> &x, std::vector<ColumnReceipt> &receipts);
  // One CMG_FULL_V2 batch solve (full_cmg.rs:809-1111) of host right-hand
  // sides (worker W x columns and firm F x columns, column-major): the Schur
  // right-hand side, lockstep PCG, extraction, complete-residual gate and the
  // refinement ladder.  refine: rounds and refined columns of this call.
  void solve_batch(const std::vector<double> &wr, const std::vector<double> &fr, uint32_t columns,
                   double tolerance, double inner_ratio, const SolveOptions &opt, std::vector<double> &worker,
                   std::vector<double> &firm, std::vector<ColumnReceipt> &receipts, uint64_t refine[2]);
  double time_leverage_batches(const SolveOptions &opt, uint32_t batches, uint64_t &solves,
                               uint64_t &iterations);
  // Timed search over probes per batch {16, 32, 64} and segment chunk sizes
  // {32, 64, 128} on leverage batches of this problem; keeps the fastest
  // layout and returns the choice (the caller keeps the width).
  struct Tuning {
    uint32_t width = 64, chunk = 64;
    double rate = 0.0;
    struct Trial { uint32_t width, chunk; double rate; };
    std::vector<Trial> trials;
  };
  Tuning autotune(const SolveOptions &base, uint32_t batches);
  void set_chunk(uint32_t chunk);
  // First PCG pass of every batch: true caps its stop at 0.1 x the complete
  // gate x ||b|| per lane (default); false keeps the reference's
  // operator-scaled stop alone, so gate failures go through the refinement
  // ladder as in full_cmg.rs:970-1088.
  void set_first_pass_cap(bool enable);
  bool first_pass_cap() const;
  // Run this engine's PCG calls in the lockstep batch `l` (null: alone).
  void set_lockstep(Lockstep *l);
  // HIP-graph capture of the PCG iteration (two graphs per configuration,
  // cached); returns the captures and graph launches so far.
  void set_graphs(bool enable);
  void graph_counters(uint64_t &captures, uint64_t &launches) const;
  // Plain device-to-device stream copy of `bytes` (counter calibration);
  // returns milliseconds.
  double stream_copy(uint64_t bytes);
  // V-cycles applied, solve batches, refinement rounds, refined columns,
  // PCG iterations launched (including ones queued after convergence).
  void counters(uint64_t out[6]) const;
  // Average duration of a timed kernel over the launches of the last timed
  // run that did work (launches returning at the convergence gate excluded):
  // 0 the finest A p of the Krylov iteration (the dominant kernel), 1 the
  // finest prolong + smooth pass of the V-cycle.
  struct KernelTiming { double ms_per_launch = 0; uint64_t launches = 0; double bytes_per_launch = 0; };
  KernelTiming kernel_timing(int which) const;

  // Tuned preconditioner (no counterpart in the reference).  mode 0: the
  // reference cycle; mode 1: the network's per-node smoother weights (theta)
  // and/or a cycle shape that skips levels which coarsen little, chosen by a
  // calibration batch (least work among configurations that need no more PCG
  // iterations than the reference cycle), and dropped for the reference
  // cycle if a solve breaks down.
  void smoother_mode(int mode, const std::vector<double> &theta, int precision);
  // Learned weights chosen for this engine, calibrated on the first solve.
  void smoother_default(const std::vector<double> &theta);
  // Host-forced fallback test: every smoothed level uses the constant weight
  // omega (omega >= 1 makes the V-cycle indefinite); no calibration, so the
  // next solve runs with it and falls back to the reference cycle when PCG
  // reports the loss of positive definiteness.
  void smoother_force(double omega);
  struct SmootherStatus {
    int mode = 0;
    bool active = false, breakdown = false, shaped = false;
    uint32_t plain_iterations = 0, learned_iterations = 0;  // calibration: reference cycle, chosen one
    std::vector<uint32_t> shape;  // visited levels when shaped
    std::vector<uint32_t> calibration;  // iterations: reference, learned, shaped, both
  };
  SmootherStatus smoother_status() const;
  // omega of every node of every smoothed level (levels in order, original
  // vertex order within a level).
  // omega of every smoothed level for parameters theta: the FP64 training
  // forward, or with `bf16` the solve path's BF16 inference
  void smoother_omega(const std::vector<double> &theta, std::vector<double> &omega, bool bf16 = false);
  // Training loss and gradient for 32 given error columns (n x 32, original
  // vertex order).
  double smoother_gradient(const std::vector<double> &theta, const std::vector<double> &errors,
                           std::vector<double> &grad);
  // Adam steps on this hierarchy; state = parameters, first and second
  // moments (3 x count); returns the loss of every step.
  void smoother_train(std::vector<double> &state, uint64_t step0, uint32_t steps, double lr, uint64_t seed,
                      uint32_t passes, std::vector<double> &losses);

  struct Impl;

 private:
  SolveResult run_once(const SolveOptions &opt);
  std::unique_ptr<Impl> impl_;
};

}  // namespace vk
