// C ABI of the MI355X leave-out engine: the vckss_rust_* entry points of the
// reference Stata backend for the CMG_FULL_V2 compressed JLA path, plus the
// vckss_mi355x_* extension and diagnostic entry points.
//
// Conventions follow rust/crates/vckss-plugin/src/ffi_engine.rs:
//   * every call returns an int32 status (error.rs:8-47) and leaves a message
//     readable through vckss_rust_last_error / vckss_rust_engine_last_error;
//   * C++ exceptions never cross the boundary (panic containment -> 99);
//   * one live context per process, identified by a nonzero generation;
//     release is idempotent (ffi_engine.rs:7146, context.rs:19-120);
//   * requests are copied and size-checked before use; outputs carry a
//     capacity that is checked before any write.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <atomic>

#include "../../include/vckss_mi355x.h"
#include "../../include/vckss_mi355x_ext.h"
#include "common.hpp"
#include "device.hpp"
#include "host_cmg.hpp"
#include "host_prepare.hpp"
#include "smoother_params.inc"  // generated by fevc_amd/build.py from fevc_amd/smoother_params.json

namespace {

using namespace vk;

enum State : uint32_t { kNone = 0, kPrepared = 1, kSolving = 2, kSolved = 3, kFailed = 4, kPoisoned = 5 };

struct Context {
  uint64_t generation = 0;
  State state = kNone;
  Prepared P;
  HybridSystem HS;
  Hierarchy H;
  std::unique_ptr<DeviceEngine> dev;
  bool has_result = false;
  SolveResult result;
  StayerReceipt stayer_receipt;
  uint64_t stayer_copy_bytes = 0, stayer_forecast = 0;
  uint32_t precision = 64, leverage_width = 32, target_width = 16;
  // receipt inputs: the prepare request's memory fields and phase times
  // (nanoseconds), the last solve's request, route and timing
  uint64_t memory_limit_bytes = 0, caller_copy_bytes = 0, preparation_forecast = 0;
  uint64_t ingest_ns = 0, hierarchy_ns = 0, hybrid_ns = 0, upload_ns = 0, stayer_ns = 0, solve_ns = 0;
  VckssEngineSolveRequestV5 request{};
  uint32_t engine_requested = 0, engine_selected = 0, algorithm_selected = 0;
  bool full_cmg = false;
  uint64_t counters_before[6] = {}, counters_after[6] = {};
};

uint64_t elapsed_ns(std::chrono::steady_clock::time_point t0) {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0)
      .count();
}

std::mutex g_mutex;
std::unique_ptr<Context> g_ctx;
uint64_t g_next_generation = 1, g_last_released = 0;
std::map<uint64_t, std::unique_ptr<Hierarchy>> g_hierarchies;
uint64_t g_next_hierarchy = 1;
thread_local std::string g_last_error;

template <typename F>
int32_t guarded(F &&f) {
  try {
    f();
    g_last_error.clear();
    return kOk;
  } catch (const Error &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_last_error = "allocation failed";
    return kAllocationFailed;
  } catch (const std::exception &e) {
    g_last_error = std::string("internal error: ") + e.what();
    return kPanic;
  } catch (...) {
    g_last_error = "internal error";
    return kPanic;
  }
}

Context &live(uint64_t generation) {
  if (!g_ctx || generation == 0 || g_ctx->generation != generation)
    fail(kStaleContext, "engine_context", "generation " + std::to_string(generation) + " is not live");
  if (g_ctx->state == kPoisoned) fail(kContextPoisoned, "engine_context", "context is poisoned");
  return *g_ctx;
}

template <typename T>
void require_capacity(const void *out, uint64_t capacity, const char *what) {
  if (!out) fail(kInvalidInput, "engine_abi", std::string(what) + " output pointer is null");
  if (capacity < sizeof(T)) fail(kAbiMismatch, "engine_abi", std::string(what) + " output capacity is too small");
}

// ffi_engine.rs:10717-10752: exact binary64 integers, signed ids order-encoded.
std::vector<uint64_t> integer_column(const double *p, uint64_t rows, const char *label, int32_t code,
                                     bool allow_signed) {
  if (!p) fail(kInvalidInput, "engine_ingest", std::string(label) + " column is null");
  const double lim = 9007199254740992.0;
  std::vector<uint64_t> out(rows);
  for (uint64_t r = 0; r < rows; ++r) {
    const double v = p[r];
    const bool in_range = allow_signed ? (v >= -lim && v <= lim) : (v > 0.0 && v <= lim);
    if (!std::isfinite(v) || !in_range || std::trunc(v) != v)
      fail(code, "engine_ingest", std::string(label) + " must be a " + (allow_signed ? "signed" : "positive") +
                                      " exact binary64 integer at zero-based row " + std::to_string(r));
    out[r] = allow_signed ? (static_cast<uint64_t>(static_cast<int64_t>(v)) ^ (1ULL << 63))
                          : static_cast<uint64_t>(v);
  }
  return out;
}

// Admission of a V4 prepare request before its data
// (ffi_engine.rs:3041-3056, 7392-7445): ABI, size, reserved and boolean
// fields, deletion mode, the implicit-match preconditions, then the
// whole-command memory admission of session.rs:202-290 under the legacy
// budget: the caller's copy must be exactly 8 bytes per row and numeric
// column (six base columns, the controls, the probe-order key), and that
// copy plus the preparation charge (768 bytes a row, 32 a row and control,
// 16 a row for a probe-order key or an implicit match key, 4096 fixed) must
// fit the declared limit.  The engine's own restrictions come last:
// probe-order keys and implicit match are not run here.  Returns the
// preparation peak forecast.
uint64_t admit_prepare(const VckssEnginePrepareRequestV4 &o, uint32_t probeorder_supplied) {
  if (probeorder_supplied > 1) fail(kInvalidInput, "engine_prepare", "probeorder_supplied must be zero or one");
  if (o.v3.v2.abi_version != VCKSS_RUST_ABI_VERSION_V1)
    fail(kAbiMismatch, "engine_prepare", "unsupported ABI version");
  if (o.v3.v2.struct_size < sizeof(VckssEnginePrepareRequestV4))
    fail(kAbiMismatch, "engine_prepare", "V4 prepare request reports a short structure size");
  if (o.v3.v2.reserved || o.v3.reserved_2 || o.reserved_3 || o.implicit_match > 1)
    fail(kAbiMismatch, "engine_prepare", "reserved or boolean V4 preparation fields are invalid");
  if (o.v3.v2.cleanup_abandoned > 1) fail(kInvalidInput, "engine_prepare", "cleanup_abandoned must be zero or one");
  if (o.v3.deletion_mode != VCKSS_DELETION_MATCH && o.v3.deletion_mode != VCKSS_DELETION_OBSERVATION)
    fail(kInvalidInput, "engine_prepare", "unknown deletion mode");
  if (o.implicit_match && (o.v3.deletion_mode != VCKSS_DELETION_MATCH || o.v3.controls_count != 0 ||
                           probeorder_supplied != 1))
    fail(kUnsupported, "engine_prepare",
         "implicit-match preparation requires match deletion, no controls, and an explicit probe order");
  const uint64_t rows = o.v3.v2.rows, limit = o.v3.v2.memory_limit_bytes;
  if (rows == 0 || limit == 0)
    fail(kInvalidInput, "engine_memory", "row count and whole-command memory limit must be positive");
  const unsigned __int128 columns = 6u + (unsigned __int128)o.v3.controls_count + probeorder_supplied;
  const unsigned __int128 copy = (unsigned __int128)rows * columns * 8u;
  if (copy != o.v3.v2.caller_copy_bytes)
    fail(kInvalidInput, "engine_memory",
         "declared caller copy is " + std::to_string(o.v3.v2.caller_copy_bytes) + " bytes; " +
             std::to_string((uint64_t)columns) + " columns require " + std::to_string((uint64_t)copy) + " bytes");
  const unsigned __int128 charge = (unsigned __int128)rows * 768u +
                                   (unsigned __int128)rows * o.v3.controls_count * 32u +
                                   (unsigned __int128)rows * 16u * (probeorder_supplied + o.implicit_match) + 4096u;
  const unsigned __int128 peak = copy + charge;
  if (peak > UINT64_MAX) fail(kResourceLimit, "engine_memory", "preparation byte forecast overflow");
  if (peak > limit)
    fail(kResourceLimit, "engine_memory",
         "simultaneous caller and engine preparation forecast " + std::to_string((uint64_t)peak) +
             " bytes exceeds the declared limit " + std::to_string(limit) + " bytes");
  if (o.v3.controls_count > 32) fail(kInvalidInput, "engine_prepare", "at most 32 controls");
  if (probeorder_supplied)
    fail(kUnsupported, "engine_prepare", "probe-order keys are not supported by the MI355X engine");
  if (o.implicit_match)
    fail(kUnsupported, "engine_prepare", "the certified implicit-match route is not supported by the MI355X engine");
  return (uint64_t)peak;
}

void prepare_into(Context &ctx, const VckssEnginePrepareRequestInterruptV3 *request,
                  const VckssEngineColumnsV3 *columns) {
  if (!request || !columns) fail(kInvalidInput, "engine_prepare", "request or columns pointer is null");
  VckssEnginePrepareRequestInterruptV3 req;
  std::memcpy(&req, request, sizeof req);
  const auto &o = req.options;
  if (req.reserved) fail(kAbiMismatch, "engine_prepare", "reserved fields must be zero");
  ctx.preparation_forecast = admit_prepare(o, columns->probeorder_supplied ? 1u : 0u);
  // ffi_engine.rs:3564-3569: the request and the descriptor declare the same controls
  if (columns->v2.controls_count != o.v3.controls_count)
    fail(kInvalidInput, "engine_prepare", "request and descriptor control counts disagree");
  ctx.memory_limit_bytes = o.v3.v2.memory_limit_bytes;
  ctx.caller_copy_bytes = o.v3.v2.caller_copy_bytes;
  const auto t_ingest = std::chrono::steady_clock::now();
  const VckssEngineColumnsV1 &c = columns->v2.v1;
  if (c.rows != o.v3.v2.rows) fail(kInvalidInput, "engine_prepare", "column rows differ from the request");
  const uint64_t rows = c.rows;
  Interrupt interrupt(req.interrupt_poll, req.interrupt_context, req.checkpoint_interval);
  interrupt.checkpoint("engine_prepare");
  auto worker = integer_column(c.worker, rows, "worker identifier", kInvalidIdentifier, true);
  auto firm = integer_column(c.firm, rows, "firm identifier", kInvalidIdentifier, true);
  auto deletion = integer_column(c.deletion, rows, "deletion identifier", kInvalidIdentifier, true);
  auto frequency = integer_column(c.frequency, rows, "frequency weight", kInvalidWeight, false);
  if (!c.outcome || !c.target_weight) fail(kInvalidInput, "engine_ingest", "outcome or target column is null");
  Columns cols{rows, worker.data(), firm.data(), deletion.data(), c.outcome, frequency.data(), c.target_weight};
  ctx.ingest_ns = elapsed_ns(t_ingest);
  ctx.P = prepare(cols, nullptr, o.v3.deletion_mode == VCKSS_DELETION_OBSERVATION);
  attach_controls(ctx.P, columns->v2.controls, columns->v2.controls_count);
  interrupt.checkpoint("engine_prepare_graph");
  auto t = std::chrono::steady_clock::now();
  ctx.HS = build_hybrid(ctx.P);
  ctx.hybrid_ns = elapsed_ns(t);
  interrupt.checkpoint("cmg_full_v2_hierarchy");
  CmgOptions copt;  // full_cmg.rs:467-476: max_levels 128
  t = std::chrono::steady_clock::now();
  ctx.H = build_hierarchy(ctx.HS.graph, copt);
  ctx.hierarchy_ns = elapsed_ns(t);
}

// Options of a point request on the CMG_FULL_V2 path (the V4 prefix shared
// by the V5 solve and the numerical requests).
SolveOptions point_options(const VckssEngineSolveRequestV4 &v4, uint32_t tolerance_supplied, const Context &ctx) {
  const auto &v3 = v4.v3;
  const auto &v2 = v3.v2;
  const auto &v1 = v2.v1;
  // ffi_engine.rs:10335-10425 eligibility, restricted to the compressed path
  // ffi_engine.rs:5712-5728: AUTO takes the exact estimator when the
  // identified dimension W + F - 1 is within exact_estimator_limit.
  // Without a prepared problem (preflight) AUTO stays unresolved.
  const bool prepared = ctx.P.workers > 0 && ctx.P.firms > 0;
  const uint64_t parameters = prepared ? ctx.P.workers + ctx.P.firms - 1 : 0;
  const bool exact = v2.algorithm == VCKSS_ALGORITHM_EXACT ||
                     (prepared && v2.algorithm == VCKSS_ALGORITHM_AUTO && parameters <= v2.exact_estimator_limit);
  // ffi_engine.rs:4840-4861: a stayer-augmented generation takes stayers-all
  // solves only, and stayers-all needs the augmentation
  if (prepared && v3.stayers_mode == VCKSS_STAYERS_MOVERS && ctx.P.stayers)
    fail(kInvalidInput, "engine_solve", "a mover-only solve cannot consume a stayer-augmented generation");
  if (prepared && v3.stayers_mode == VCKSS_STAYERS_ALL && !ctx.P.stayers)
    fail(kInvalidInput, "engine_solve", "a stayers-all solve requires a reconciled stayer augmentation");
  // ffi_engine.rs:5705-5710: the solve's deletion mode is the prepared graph's
  if (prepared && (v1.deletion_mode == VCKSS_DELETION_OBSERVATION) != ctx.P.observation)
    fail(kInvalidInput, "engine_solve", "solve deletion mode differs from the prepared graph mode");
  // observation deletion: the exact route, or the generic engine's JLA route
  // without controls (generic_jla.rs:5470-6017)
  const bool deletion_ok =
      v1.deletion_mode == VCKSS_DELETION_MATCH || v1.deletion_mode == VCKSS_DELETION_OBSERVATION;
  const bool eligible =
      (v2.algorithm == VCKSS_ALGORITHM_JLA || v2.algorithm == VCKSS_ALGORITHM_AUTO ||
       v2.algorithm == VCKSS_ALGORITHM_EXACT) &&
      (v3.engine == VCKSS_ENGINE_AUTO_OR_UNSPECIFIED || v3.engine == VCKSS_ENGINE_GENERIC) &&
      deletion_ok &&
      (v2.nuisance_mode == VCKSS_NUISANCE_JOINT || v2.nuisance_mode == VCKSS_NUISANCE_FIXED_OFFSET) &&
      v1.rng_contract == VCKSS_RNG_COUNTER_V1 &&
      (v1.solver_route == VCKSS_ROUTE_AUTO || v1.solver_route == VCKSS_ROUTE_CMG_PCG) &&
      v3.batch_mode == VCKSS_BATCH_MODE_AUTO && v4.leverage_batch_mode == VCKSS_BATCH_MODE_AUTO &&
      v4.target_batch_mode == VCKSS_BATCH_MODE_AUTO && v1.leverage_batch_width == 0 &&
      v1.target_batch_width == 0 &&
      (v3.stayers_mode == VCKSS_STAYERS_MOVERS || v3.stayers_mode == VCKSS_STAYERS_ALL) &&
      (v3.target_weight_mode == VCKSS_TARGET_WEIGHT_FREQUENCY_DEFAULT ||
       v3.target_weight_mode == VCKSS_TARGET_WEIGHT_STORED_ROW_EXPLICIT) &&
      (v3.deletion_unit_source == VCKSS_DELETION_SOURCE_CELL_DEFAULT ||
       v3.deletion_unit_source == VCKSS_DELETION_SOURCE_MATCH_ID_EXPLICIT) &&
      v3.probeorder_supplied == 0 &&
      (v3.frequency_use == VCKSS_REQUEST_FREQUENCY_UNIT || v3.frequency_use == VCKSS_REQUEST_FREQUENCY_LITERAL);
  if (!eligible)
    fail(kUnsupported, "cmg_full_v2",
         "the MI355X engine requires auto/exact/JLA, auto/generic engine, auto/CMG solver, match or observation "
         "deletion, joint or fixed-offset nuisance, movers, automatic batches and no probe order");
  SolveOptions o;
  o.exact = exact;
  // controls (generic_jla.rs:1930-2140): the JLA route under match deletion,
  // joint or fixed-offset nuisance
  if (prepared && ctx.P.controls) {
    if (exact)
      fail(kUnsupported, "exact_estimator", "the MI355X engine's exact route does not take controls");
    if (ctx.P.observation || ctx.P.stayers)
      fail(kUnsupported, "generic_jla",
           "the MI355X engine takes controls under match deletion without stayers only");
    o.joint_controls = v2.nuisance_mode == VCKSS_NUISANCE_JOINT;
    o.fixed_offset = v2.nuisance_mode == VCKSS_NUISANCE_FIXED_OFFSET;
  }
  if (exact) {
    // ffi_engine.rs:5733-5751: dimension and block-size limits of the exact route
    if (prepared && parameters > v2.exact_estimator_limit)
      fail(kResourceLimit, "exact_estimator", "identified coefficient dimension exceeds exact_limit()");
    for (uint64_t g = 0; g < (ctx.P.observation ? 0 : ctx.P.units); ++g)
      if (ctx.P.unit_ptr[g + 1] - ctx.P.unit_ptr[g] > v2.blocksize_limit)
        fail(kResourceLimit, "exact_estimator", "a deletion block exceeds blocksize_limit()");
  } else if (v1.probes < 2) {
    fail(kInvalidInput, "jla_validate", "JLA probe count must be at least two");
  }
  o.seed = v1.seed;
  o.probes = v1.probes;
  o.fit_tolerance = v1.pcg_tolerance;
  o.probe_tolerance = tolerance_supplied ? v1.pcg_tolerance : 1.0e-6;  // full_cmg.rs:29
  for (double t : {o.fit_tolerance, o.probe_tolerance})
    if (!std::isfinite(t) || t < 1e-15 || t > 1e-4) fail(kInvalidInput, "cmg_full_v2", "tolerance must lie in [1e-15, 1e-4]");
  o.max_iterations = v1.maximum_iterations;
  o.recompute_interval = v1.residual_replacement_interval;
  if (o.max_iterations == 0 || o.recompute_interval == 0)
    fail(kInvalidInput, "pcg_options", "iteration limit and replacement interval must be positive");
  o.rank_tolerance = v1.rank_tolerance;
  o.block_tolerance = v1.block_tolerance;
  if (!std::isfinite(o.rank_tolerance) || o.rank_tolerance < 1e-14 || o.rank_tolerance >= 0.1)
    fail(kInvalidInput, "jla_validate", "rank tolerance must be finite and lie in [1e-14, 0.1)");
  if (!std::isfinite(o.block_tolerance) || o.block_tolerance < 1e-14 || o.block_tolerance >= 1.0)
    fail(kInvalidInput, "jla_validate", "block tolerance must be finite and lie in [1e-14, 1)");
  o.precision = static_cast<int>(ctx.precision);
  o.leverage_width = ctx.leverage_width;
  o.target_width = ctx.target_width;
  return o;
}

SolveOptions solve_options(const VckssEngineSolveRequestV5 &q, const Context &ctx) {
  const auto &v1 = q.v4.v3.v2.v1;
  if (v1.abi_version != VCKSS_RUST_ABI_VERSION_V1) fail(kAbiMismatch, "engine_solve", "unsupported ABI version");
  if (v1.struct_size < sizeof(VckssEngineSolveRequestV5))
    fail(kAbiMismatch, "engine_solve", "V5 solve request reports a short structure size");
  if (q.reserved_5 != 0 || q.tolerance_supplied > 1 || q.full_cmg_v2 > 1)
    fail(kAbiMismatch, "engine_solve", "invalid V5 solve request flags or reserved field");
  const auto &algorithm = q.v4.v3.v2.algorithm;
  if (q.full_cmg_v2 == 1) {
    if (q.threads == 0)
      fail(kInvalidInput, "cmg_full_v2", "CMG_FULL_V2 requires a positive caller-declared thread count");
    // ffi_engine.rs:10366-10410: CMG_FULL_V2 takes the JLA algorithm only
    if (algorithm != VCKSS_ALGORITHM_JLA)
      fail(kUnsupported, "cmg_full_v2",
           "CMG_FULL_V2 requires JLA, auto/generic engine, auto/CMG solver and joint/fixed-offset point estimation");
    return point_options(q.v4, q.tolerance_supplied, ctx);
  }
  // full_cmg_v2 = 0: the dense exact estimator (EXACT, or AUTO within the
  // exact limit); the legacy direct-solver JLA route is not implemented
  SolveOptions o = point_options(q.v4, q.tolerance_supplied, ctx);
  if (!o.exact)
    fail(kUnsupported, "engine_solve", "the MI355X engine runs JLA on the CMG_FULL_V2 route only (full_cmg_v2 = 1)");
  return o;
}

// numerical_api.rs:113-196 (validate_point) for the path's requests: schema,
// reserved fields and the point prefix; controls, projection and component
// attachments are outside the path.
void validate_numerical_v1(const VckssNumericalRequestV1 &r, uint32_t schema) {
  if (r.abi_version != VCKSS_RUST_ABI_VERSION_V1 || r.point.v3.v2.v1.abi_version != VCKSS_RUST_ABI_VERSION_V1)
    fail(kAbiMismatch, "numerical_preflight", "unsupported ABI version");
  if (r.schema_version != schema || r.reserved != 0 || r.reserved_2 != 0 || r.tolerance_supplied > 1 ||
      r.projection_requested > 1 || r.component_requested > 1)
    fail(kAbiMismatch, "numerical_preflight", "invalid numerical V1 schema or reserved fields");
  if (r.point.v3.v2.v1.struct_size < sizeof(VckssEngineSolveRequestV4))
    fail(kAbiMismatch, "numerical_preflight", "numerical point prefix reports a short structure size");
  if (r.point.v3.reserved_3 != 0 || r.point.reserved_4 != 0)
    fail(kAbiMismatch, "numerical_preflight", "reserved numerical point prefix fields must be zero");
  if (r.threads == 0 || r.point.v3.v2.v1.probes < 2)
    fail(kInvalidInput, "numerical_preflight", "positive threads and at least two probes required");
  if (r.projection_requested || r.component_requested || r.controls_count)
    fail(kUnsupported, "numerical_preflight",
         "the MI355X engine attaches numerical MCSE to the no-control compressed path only");
}

void validate_numerical_v2(const VckssNumericalRequestV2 &r) {
  if (r.v1.schema_version != 2 || r.reserved != 0 || r.full_cmg > 1 || r.component_batch_mode > 1)
    fail(kAbiMismatch, "numerical_preflight", "invalid numerical V2 schema or executor intent");
  validate_numerical_v1(r.v1, 2);
  if (r.full_cmg != 1 || r.execution_mode != 0)
    fail(kUnsupported, "numerical_preflight", "the MI355X engine runs the numerical attachment on CMG_FULL_V2 only");
}

// Point options also checked without a context (preflight).
void preflight_point(VckssEngineSolveRequestV4 point, uint32_t tolerance_supplied) {
  if (point.v3.engine == VCKSS_ENGINE_COMPRESSED) point.v3.engine = VCKSS_ENGINE_AUTO_OR_UNSPECIFIED;
  Context probe;
  (void)point_options(point, tolerance_supplied, probe);
}

std::vector<double> builtin_smoother() {
  return std::vector<double>(kBuiltinSmoother, kBuiltinSmoother + kBuiltinSmootherCount);
}

DeviceEngine &device_of(Context &ctx) {
  if (!ctx.dev) {
    ctx.dev.reset(new DeviceEngine(ctx.P, ctx.HS, ctx.H));
    // the built-in trained smoother weights, used where they help
    if (kBuiltinSmootherCount == kSmootherParams) ctx.dev->smoother_default(builtin_smoother());
  }
  return *ctx.dev;
}

void run_options(Context &ctx, SolveOptions o, Interrupt &interrupt);

// What the receipts report about a solve: the request, the engine the path
// resolves to (the compressed engine for match deletion without stayers, the
// generic engine for observation deletion and the stayer hybrid, none for the
// dense exact estimator; engine_plan.rs:243-290) and the algorithm.
void record_request(Context &ctx, const VckssEngineSolveRequestV5 &q, const SolveOptions &o) {
  ctx.request = q;
  ctx.engine_requested = q.v4.v3.engine;
  ctx.algorithm_selected = o.exact ? VCKSS_ALGORITHM_EXACT : VCKSS_ALGORITHM_JLA;
  ctx.engine_selected = o.exact ? VCKSS_ENGINE_NOT_APPLICABLE
                                : ((ctx.P.observation || ctx.P.stayers || ctx.P.controls) ? VCKSS_ENGINE_GENERIC
                                                                                          : VCKSS_ENGINE_COMPRESSED);
  ctx.full_cmg = !o.exact;
}

void run_solve(Context &ctx, const VckssEngineSolveRequestV5 &q, Interrupt &interrupt) {
  const SolveOptions o = solve_options(q, ctx);
  record_request(ctx, q, o);
  run_options(ctx, o, interrupt);
}

// The fixed-offset nuisance with controls (generic_jla.rs:2012-2045): the
// joint fit's control coefficients give the working outcome y - Z beta (a
// compensated sum per row), with which the two-way route runs (its own
// working fit and the leverage and target probes, whose plan keeps the
// prepared ranks); the prepared outcome is restored afterwards.
SolveResult run_fixed_offset(Context &ctx, DeviceEngine &dev, const SolveOptions &o) {
  Prepared &P = ctx.P;
  const uint32_t q = P.controls;
  ControlFit cf = dev.control_fit(o);
  std::vector<double> working(P.rows);
  for (uint64_t r = 0; r < P.rows; ++r) {
    Neumaier v;
    v.add(P.outcome[r]);
    for (uint32_t j = 0; j < q; ++j) v.add(-P.control[r * q + j] * cf.beta[j]);
    working[r] = v.value();
    if (!std::isfinite(working[r])) fail(kInvalidInput, "generic_jla_fixedoffset", "working outcome is nonfinite");
  }
  const std::vector<double> outcome = P.outcome, cell_outcome = P.cell_outcome, unit_outcome = P.unit_outcome;
  auto restore = [&] {
    P.outcome = outcome;
    P.cell_outcome = cell_outcome;
    P.unit_outcome = unit_outcome;
    dev.refresh_outcome();
  };
  set_outcome(P, working);
  dev.refresh_outcome();
  SolveResult r;
  try {
    r = dev.run(o);
  } catch (...) {
    restore();
    throw;
  }
  restore();
  r.controls = q;
  r.fixed_offset = true;
  r.control_fit = cf.receipts;
  r.control_beta = cf.beta;
  r.control_rcond = cf.generalized_rcond;
  for (const auto &c : cf.receipts) r.max_complete_residual = std::max(r.max_complete_residual, c.complete);
  return r;
}

// context.rs:150-205: a solve needs a prepared generation; a failed solve
// leaves the generation Failed and later solves are refused.  Solving a
// Solved generation again (another seed or tolerance on the same prepared
// problem) is kept as an extension of the reference lifecycle.
void run_options(Context &ctx, SolveOptions o, Interrupt &interrupt) {
  if (ctx.state != kPrepared && ctx.state != kSolved)
    fail(kContextPoisoned, "context_solve",
         "generation " + std::to_string(ctx.generation) + " is in state " +
             (ctx.state == kFailed ? "Failed" : ctx.state == kSolving ? "Solving" : "Poisoned") +
             "; state Prepared is required");
  o.interrupt = &interrupt;
  DeviceEngine &dev = device_of(ctx);
  ctx.state = kSolving;
  ctx.has_result = false;
  const auto t = std::chrono::steady_clock::now();
  dev.counters(ctx.counters_before);
  try {
    o.checkpoint("engine_solve");
    if (o.fixed_offset) ctx.result = run_fixed_offset(ctx, dev, o);
    else ctx.result = dev.run(o);
  } catch (...) {
    ctx.state = kFailed;
    throw;
  }
  dev.counters(ctx.counters_after);
  ctx.solve_ns = elapsed_ns(t);
  ctx.has_result = true;
  ctx.state = kSolved;
}

VckssEngineSolveRequestV5 default_solve_v5() {
  VckssEngineSolveRequestV5 q;
  std::memset(&q, 0, sizeof q);
  auto &v1 = q.v4.v3.v2.v1;
  v1.abi_version = VCKSS_RUST_ABI_VERSION_V1;
  v1.struct_size = sizeof(VckssEngineSolveRequestV5);
  v1.seed = 8675309;  // engine.rs:83-103
  v1.probes = 200;
  v1.deletion_mode = VCKSS_DELETION_MATCH;
  v1.rng_contract = VCKSS_RNG_COUNTER_V1;
  v1.solver_route = VCKSS_ROUTE_AUTO;
  v1.exact_dimension_limit = 500;  // solver.rs:42-58
  v1.cmg_minimum_dimension = 2000;
  v1.pcg_tolerance = 1e-10;
  v1.maximum_iterations = 10000;
  v1.residual_replacement_interval = 100;
  v1.rank_tolerance = 1e-10;
  v1.block_tolerance = 1e-10;
  v1.cmg_terminal_vertices = 128;  // vckss-core/src/cmg/hierarchy.rs:38-55
  v1.cmg_dense_vertex_cap = 6144;
  v1.cmg_maximum_levels = 96;
  v1.cmg_aggregate_cap = 8;
  v1.cmg_minimum_reduction = 0.20;
  v1.cmg_jacobi_weight = 2.0 / 3.0;
  v1.cmg_pre_sweeps = 1;
  v1.cmg_post_sweeps = 1;
  v1.cmg_maximum_edge_complexity = 12.0;
  v1.cmg_maximum_vertex_complexity = 5.0;
  v1.cmg_memory_limit_bytes = 2ULL << 30;
  auto &v2 = q.v4.v3.v2;
  v2.algorithm = VCKSS_ALGORITHM_AUTO;
  v2.nuisance_mode = VCKSS_NUISANCE_JOINT;
  v2.exact_estimator_limit = 500;
  v2.blocksize_limit = 5000;
  auto &v3 = q.v4.v3;
  v3.engine = VCKSS_ENGINE_AUTO_OR_UNSPECIFIED;
  v3.batch_mode = VCKSS_BATCH_MODE_AUTO;
  v3.stayers_mode = VCKSS_STAYERS_MOVERS;
  v3.target_weight_mode = VCKSS_TARGET_WEIGHT_FREQUENCY_DEFAULT;
  v3.deletion_unit_source = VCKSS_DELETION_SOURCE_CELL_DEFAULT;
  v3.capability_schema = VCKSS_REQUEST_CAPABILITY_SCHEMA_V3;
  v3.capability_profile = VCKSS_REQUEST_PROFILE_PLANNED_V1;
  v3.frequency_use = VCKSS_REQUEST_FREQUENCY_UNIT;
  v3.physical_limit = 50000000;
  q.v4.leverage_batch_mode = VCKSS_BATCH_MODE_AUTO;
  q.v4.target_batch_mode = VCKSS_BATCH_MODE_AUTO;
  q.threads = 1;
  return q;
}

// ---------------------------------------------------- capability receipts
// FNV-1a identity of a request (ffi_engine.rs:9114-9138, 9269-9302,
// 9492-9530): a domain string, then every field's little-endian bytes.
struct RequestHash {
  uint64_t h = 0xcbf29ce484222325ULL;
  void byte(uint8_t b) {
    h ^= b;
    h *= 0x00000100000001b3ULL;
  }
  void text(const char *s) {
    for (; *s; ++s) byte((uint8_t)*s);
  }
  void u32(uint32_t v) {
    for (int k = 0; k < 4; ++k) byte((uint8_t)(v >> (8 * k)));
  }
  void u64(uint64_t v) {
    for (int k = 0; k < 8; ++k) byte((uint8_t)(v >> (8 * k)));
  }
};

uint64_t signature_v1(const VckssBackendRequestCapabilityRequestV1 &r, const char *domain) {
  RequestHash f;
  f.text(domain);
  for (uint32_t v : {r.request_schema, r.algorithm, r.deletion_mode, r.nuisance_mode, r.solver_route, r.rng_contract,
                     r.controls_count, r.frequency_use})
    f.u32(v);
  return f.h;
}
uint64_t signature_v2_fields(RequestHash &f, const VckssBackendRequestCapabilityRequestV2 &r) {
  for (uint32_t v : {r.v1.request_schema, r.v1.algorithm, r.v1.deletion_mode, r.v1.nuisance_mode, r.v1.solver_route,
                     r.v1.rng_contract, r.v1.controls_count, r.v1.frequency_use, r.engine, r.batch_mode,
                     r.stayers_mode, r.target_weight_mode, r.deletion_unit_source, r.probeorder_supplied,
                     r.wallseconds_supplied})
    f.u32(v);
  return f.h;
}
uint64_t signature_v2(const VckssBackendRequestCapabilityRequestV2 &r) {
  RequestHash f;
  f.text("VCKSS-REQUEST-CAPABILITY-V2");
  signature_v2_fields(f, r);
  f.u64(r.physical_limit);
  return f.h;
}
uint64_t signature_v3(const VckssBackendRequestCapabilityRequestV3 &r) {
  RequestHash f;
  f.text("VCKSS-REQUEST-CAPABILITY-V3");
  signature_v2_fields(f, r.v2);
  for (uint32_t v : {r.leverage_batch_mode, r.target_batch_mode, r.allow_automatic_cmg_setup_fallback}) f.u32(v);
  f.u64(r.v2.physical_limit);
  f.u64(bits_of(r.wallseconds));
  return f.h;
}

constexpr uint64_t kMaxExactInteger = 1ULL << 53;

bool known(uint32_t v, std::initializer_list<uint32_t> set) {
  for (uint32_t s : set)
    if (v == s) return true;
  return false;
}

// The reference's schema-independent unknown-value checks
// (ffi_engine.rs:9040-9075, 9140-9200, 9304-9400).
uint32_t unknown_v1_fields(const VckssBackendRequestCapabilityRequestV1 &v) {
  if (!known(v.algorithm, {VCKSS_ALGORITHM_AUTO, VCKSS_ALGORITHM_EXACT, VCKSS_ALGORITHM_JLA}))
    return VCKSS_REQUEST_REASON_UNKNOWN_ALGORITHM;
  if (!known(v.deletion_mode, {VCKSS_DELETION_MATCH, VCKSS_DELETION_OBSERVATION}))
    return VCKSS_REQUEST_REASON_UNKNOWN_DELETION;
  if (!known(v.nuisance_mode, {VCKSS_NUISANCE_JOINT, VCKSS_NUISANCE_FIXED_OFFSET}))
    return VCKSS_REQUEST_REASON_UNKNOWN_NUISANCE;
  if (!known(v.solver_route, {VCKSS_ROUTE_AUTO, VCKSS_ROUTE_EXACT, VCKSS_ROUTE_DIAGONAL_PCG, VCKSS_ROUTE_CMG_PCG}))
    return VCKSS_REQUEST_REASON_UNKNOWN_SOLVER_ROUTE;
  if (!known(v.rng_contract, {VCKSS_RNG_NONE, VCKSS_RNG_COUNTER_V1})) return VCKSS_REQUEST_REASON_UNKNOWN_RNG_CONTRACT;
  if (!known(v.frequency_use, {VCKSS_REQUEST_FREQUENCY_UNIT, VCKSS_REQUEST_FREQUENCY_LITERAL}))
    return VCKSS_REQUEST_REASON_UNKNOWN_FREQUENCY_USE;
  return VCKSS_REQUEST_REASON_SUPPORTED;
}

// What this engine adds to a request the reference would support: controls
// on the JLA route under match deletion only (the exact route takes none),
// JLA on the automatic or CMG route only (CMG_FULL_V2; no dense-exact or
// diagonal-PCG probe solves), no probe-order keys, and it chooses its own
// batch widths.
uint32_t engine_restriction(uint32_t algorithm, uint32_t deletion, uint32_t route, uint32_t controls,
                            uint32_t probeorder, bool explicit_batches) {
  if (controls != 0 && algorithm == VCKSS_ALGORITHM_EXACT) return VCKSS_REQUEST_REASON_CONTROLS_LIMIT;
  if (controls != 0 && deletion != VCKSS_DELETION_MATCH) return VCKSS_REQUEST_REASON_JLA_CONTROLS;
  if (algorithm != VCKSS_ALGORITHM_EXACT && route != VCKSS_ROUTE_AUTO && route != VCKSS_ROUTE_CMG_PCG)
    return VCKSS_REQUEST_REASON_JLA_GENERIC_SOLVER_ROUTE;
  if (probeorder != 0) return VCKSS_REQUEST_REASON_PROBEORDER_UNSUPPORTED;
  if (explicit_batches) return VCKSS_REQUEST_REASON_BATCH_MODE_UNSUPPORTED;
  return VCKSS_REQUEST_REASON_SUPPORTED;
}

// request_capability_classification (ffi_engine.rs:9040-9112) then the
// engine's restrictions; the profile is EXACT_V1 or JLA_COUNTER_V1.
std::pair<uint32_t, uint32_t> classify_v1(const VckssBackendRequestCapabilityRequestV1 &v) {
  uint32_t reason = v.request_schema != VCKSS_REQUEST_CAPABILITY_SCHEMA_V1 ? VCKSS_REQUEST_REASON_UNKNOWN_SCHEMA
                                                                           : unknown_v1_fields(v);
  if (reason == VCKSS_REQUEST_REASON_SUPPORTED) {
    if (v.algorithm == VCKSS_ALGORITHM_AUTO) reason = VCKSS_REQUEST_REASON_ALGORITHM_AUTO_UNRESOLVED;
    else if (v.algorithm == VCKSS_ALGORITHM_EXACT && v.controls_count > 32) reason = VCKSS_REQUEST_REASON_CONTROLS_LIMIT;
    else if (v.algorithm == VCKSS_ALGORITHM_EXACT && v.rng_contract != VCKSS_RNG_NONE) reason = VCKSS_REQUEST_REASON_EXACT_RNG;
    else if (v.algorithm == VCKSS_ALGORITHM_EXACT && !known(v.solver_route, {VCKSS_ROUTE_AUTO, VCKSS_ROUTE_EXACT}))
      reason = VCKSS_REQUEST_REASON_EXACT_SOLVER_ROUTE;
    else if (v.algorithm == VCKSS_ALGORITHM_JLA && v.deletion_mode != VCKSS_DELETION_MATCH)
      reason = VCKSS_REQUEST_REASON_JLA_DELETION;
    else if (v.algorithm == VCKSS_ALGORITHM_JLA && v.nuisance_mode != VCKSS_NUISANCE_JOINT)
      reason = VCKSS_REQUEST_REASON_JLA_NUISANCE;
    else if (v.algorithm == VCKSS_ALGORITHM_JLA && v.controls_count != 0) reason = VCKSS_REQUEST_REASON_JLA_CONTROLS;
    else if (v.algorithm == VCKSS_ALGORITHM_JLA && v.rng_contract != VCKSS_RNG_COUNTER_V1)
      reason = VCKSS_REQUEST_REASON_JLA_RNG;
    else
      reason = engine_restriction(v.algorithm, v.deletion_mode, v.solver_route, v.controls_count, 0, false);
  }
  const uint32_t profile = reason != VCKSS_REQUEST_REASON_SUPPORTED ? VCKSS_REQUEST_PROFILE_NONE
                           : v.algorithm == VCKSS_ALGORITHM_EXACT  ? VCKSS_REQUEST_PROFILE_EXACT_V1
                                                                   : VCKSS_REQUEST_PROFILE_JLA_COUNTER_V1;
  return {reason, profile};
}

// request_capability_classification_v2 (ffi_engine.rs:9140-9267), then the
// engine's restrictions.  The reference's generic engine runs JLA on the
// diagonal route only; this engine's generic JLA (observation deletion)
// runs on CMG_FULL_V2, so the generic branch takes the CMG route here.
std::pair<uint32_t, uint32_t> classify_v2(const VckssBackendRequestCapabilityRequestV2 &r) {
  const auto &v = r.v1;
  uint32_t reason = v.request_schema != VCKSS_REQUEST_CAPABILITY_SCHEMA_V2 ? VCKSS_REQUEST_REASON_UNKNOWN_SCHEMA
                                                                           : unknown_v1_fields(v);
  if (reason != VCKSS_REQUEST_REASON_SUPPORTED) {
  } else if (!known(r.engine, {VCKSS_ENGINE_AUTO_OR_UNSPECIFIED, VCKSS_ENGINE_COMPRESSED, VCKSS_ENGINE_GENERIC})) {
    reason = VCKSS_REQUEST_REASON_UNKNOWN_ENGINE;
  } else if (!known(r.batch_mode, {VCKSS_BATCH_MODE_AUTO, VCKSS_BATCH_MODE_EXPLICIT})) {
    reason = VCKSS_REQUEST_REASON_UNKNOWN_BATCH_MODE;
  } else if (!known(r.stayers_mode, {VCKSS_STAYERS_MOVERS, VCKSS_STAYERS_ALL})) {
    reason = VCKSS_REQUEST_REASON_UNKNOWN_STAYERS_MODE;
  } else if (!known(r.target_weight_mode,
                    {VCKSS_TARGET_WEIGHT_FREQUENCY_DEFAULT, VCKSS_TARGET_WEIGHT_STORED_ROW_EXPLICIT})) {
    reason = VCKSS_REQUEST_REASON_UNKNOWN_TARGET_WEIGHT_MODE;
  } else if (!known(r.deletion_unit_source, {VCKSS_DELETION_SOURCE_CELL_DEFAULT, VCKSS_DELETION_SOURCE_MATCH_ID_EXPLICIT,
                                             VCKSS_DELETION_SOURCE_OBSERVATION_ROW})) {
    reason = VCKSS_REQUEST_REASON_UNKNOWN_DELETION_UNIT_SOURCE;
  } else if (r.probeorder_supplied > 1) {
    reason = VCKSS_REQUEST_REASON_PROBEORDER_UNSUPPORTED;
  } else if (r.wallseconds_supplied > 1) {
    reason = VCKSS_REQUEST_REASON_WALLSECONDS_UNSUPPORTED;
  } else if (r.physical_limit == 0 || r.physical_limit > kMaxExactInteger) {
    reason = VCKSS_REQUEST_REASON_PHYSICAL_LIMIT;
  } else if (v.algorithm == VCKSS_ALGORITHM_AUTO) {
    reason = VCKSS_REQUEST_REASON_ALGORITHM_AUTO_UNRESOLVED;
  } else if (v.controls_count > 32) {
    reason = VCKSS_REQUEST_REASON_CONTROLS_LIMIT;
  } else if (v.algorithm == VCKSS_ALGORITHM_EXACT) {
    if (r.engine == VCKSS_ENGINE_COMPRESSED) reason = VCKSS_REQUEST_REASON_EXACT_ENGINE;
    else if (v.rng_contract != VCKSS_RNG_NONE) reason = VCKSS_REQUEST_REASON_EXACT_RNG;
    else if (!known(v.solver_route, {VCKSS_ROUTE_AUTO, VCKSS_ROUTE_EXACT})) reason = VCKSS_REQUEST_REASON_EXACT_SOLVER_ROUTE;
  } else if (r.engine == VCKSS_ENGINE_GENERIC) {
    const bool source_ok = (v.deletion_mode == VCKSS_DELETION_MATCH &&
                            known(r.deletion_unit_source,
                                  {VCKSS_DELETION_SOURCE_CELL_DEFAULT, VCKSS_DELETION_SOURCE_MATCH_ID_EXPLICIT})) ||
                           (v.deletion_mode == VCKSS_DELETION_OBSERVATION &&
                            r.deletion_unit_source == VCKSS_DELETION_SOURCE_OBSERVATION_ROW);
    if (v.rng_contract != VCKSS_RNG_COUNTER_V1) reason = VCKSS_REQUEST_REASON_JLA_RNG;
    else if (!known(v.solver_route, {VCKSS_ROUTE_AUTO, VCKSS_ROUTE_CMG_PCG}))
      reason = VCKSS_REQUEST_REASON_JLA_GENERIC_SOLVER_ROUTE;
    else if (r.probeorder_supplied != 0) reason = VCKSS_REQUEST_REASON_PROBEORDER_UNSUPPORTED;
    else if (r.wallseconds_supplied != 0) reason = VCKSS_REQUEST_REASON_WALLSECONDS_UNSUPPORTED;
    else if (!source_ok) reason = VCKSS_REQUEST_REASON_DELETION_UNIT_SOURCE_MISMATCH;
  } else {
    VckssBackendRequestCapabilityRequestV1 legacy = v;
    legacy.request_schema = VCKSS_REQUEST_CAPABILITY_SCHEMA_V1;
    const uint32_t lr = classify_v1(legacy).first;
    reason = (r.engine == VCKSS_ENGINE_AUTO_OR_UNSPECIFIED && lr != VCKSS_REQUEST_REASON_SUPPORTED)
                 ? VCKSS_REQUEST_REASON_JLA_ENGINE_AUTO_UNRESOLVED
                 : lr;
  }
  if (reason == VCKSS_REQUEST_REASON_SUPPORTED)
    reason = engine_restriction(v.algorithm, v.deletion_mode, v.solver_route, v.controls_count, r.probeorder_supplied,
                                r.batch_mode != VCKSS_BATCH_MODE_AUTO);
  const uint32_t profile = reason != VCKSS_REQUEST_REASON_SUPPORTED     ? VCKSS_REQUEST_PROFILE_NONE
                           : v.algorithm == VCKSS_ALGORITHM_EXACT      ? VCKSS_REQUEST_PROFILE_EXACT_V1
                           : r.engine == VCKSS_ENGINE_GENERIC          ? VCKSS_REQUEST_PROFILE_JLA_GENERIC_COUNTER_V1
                                                                       : VCKSS_REQUEST_PROFILE_JLA_COUNTER_V1;
  return {reason, profile};
}

// request_capability_classification_v3 (ffi_engine.rs:9304-9490), then the
// engine's restrictions; a supported request has the planned profile.
std::pair<uint32_t, uint32_t> classify_v3(const VckssBackendRequestCapabilityRequestV3 &r) {
  const auto &v = r.v2.v1;
  const auto &q = r.v2;
  const bool phases_known = known(r.leverage_batch_mode, {VCKSS_BATCH_MODE_AUTO, VCKSS_BATCH_MODE_EXPLICIT}) &&
                            known(r.target_batch_mode, {VCKSS_BATCH_MODE_AUTO, VCKSS_BATCH_MODE_EXPLICIT});
  const uint32_t summary =
      (r.leverage_batch_mode == VCKSS_BATCH_MODE_AUTO && r.target_batch_mode == VCKSS_BATCH_MODE_AUTO)
          ? VCKSS_BATCH_MODE_AUTO
      : (r.leverage_batch_mode == VCKSS_BATCH_MODE_EXPLICIT && r.target_batch_mode == VCKSS_BATCH_MODE_EXPLICIT)
          ? VCKSS_BATCH_MODE_EXPLICIT
          : VCKSS_BATCH_MODE_INDEPENDENT;
  const bool source_ok =
      (v.deletion_mode == VCKSS_DELETION_MATCH &&
       known(q.deletion_unit_source, {VCKSS_DELETION_SOURCE_CELL_DEFAULT, VCKSS_DELETION_SOURCE_MATCH_ID_EXPLICIT})) ||
      (v.deletion_mode == VCKSS_DELETION_OBSERVATION && q.deletion_unit_source == VCKSS_DELETION_SOURCE_OBSERVATION_ROW);
  const bool all = q.stayers_mode == VCKSS_STAYERS_ALL;
  uint32_t reason = v.request_schema != VCKSS_REQUEST_CAPABILITY_SCHEMA_V3 ? VCKSS_REQUEST_REASON_UNKNOWN_SCHEMA
                                                                           : unknown_v1_fields(v);
  if (reason != VCKSS_REQUEST_REASON_SUPPORTED) {
  } else if (!known(q.engine, {VCKSS_ENGINE_AUTO_OR_UNSPECIFIED, VCKSS_ENGINE_COMPRESSED, VCKSS_ENGINE_GENERIC})) {
    reason = VCKSS_REQUEST_REASON_UNKNOWN_ENGINE;
  } else if (!phases_known) {
    reason = VCKSS_REQUEST_REASON_UNKNOWN_BATCH_MODE;
  } else if (!known(q.batch_mode, {VCKSS_BATCH_MODE_AUTO, VCKSS_BATCH_MODE_EXPLICIT, VCKSS_BATCH_MODE_INDEPENDENT}) ||
             q.batch_mode != summary) {
    reason = VCKSS_REQUEST_REASON_BATCH_SUMMARY_MISMATCH;
  } else if (!known(q.stayers_mode, {VCKSS_STAYERS_MOVERS, VCKSS_STAYERS_ALL})) {
    reason = VCKSS_REQUEST_REASON_UNKNOWN_STAYERS_MODE;
  } else if (!known(q.target_weight_mode,
                    {VCKSS_TARGET_WEIGHT_FREQUENCY_DEFAULT, VCKSS_TARGET_WEIGHT_STORED_ROW_EXPLICIT})) {
    reason = VCKSS_REQUEST_REASON_UNKNOWN_TARGET_WEIGHT_MODE;
  } else if (!known(q.deletion_unit_source, {VCKSS_DELETION_SOURCE_CELL_DEFAULT, VCKSS_DELETION_SOURCE_MATCH_ID_EXPLICIT,
                                             VCKSS_DELETION_SOURCE_OBSERVATION_ROW})) {
    reason = VCKSS_REQUEST_REASON_UNKNOWN_DELETION_UNIT_SOURCE;
  } else if (!source_ok) {
    reason = VCKSS_REQUEST_REASON_DELETION_UNIT_SOURCE_MISMATCH;
  } else if (q.probeorder_supplied > 1) {
    reason = VCKSS_REQUEST_REASON_PROBEORDER_UNSUPPORTED;
  } else if (q.wallseconds_supplied > 1 || (q.wallseconds_supplied == 0 && r.wallseconds != 0.0) ||
             (q.wallseconds_supplied == 1 && (!std::isfinite(r.wallseconds) || r.wallseconds <= 0.0))) {
    reason = VCKSS_REQUEST_REASON_WALLSECONDS_VALUE;
  } else if (r.allow_automatic_cmg_setup_fallback > 1 ||
             (r.allow_automatic_cmg_setup_fallback == 1 && v.solver_route != VCKSS_ROUTE_AUTO)) {
    reason = VCKSS_REQUEST_REASON_FALLBACK_ROUTE_MISMATCH;
  } else if (q.physical_limit == 0 || q.physical_limit > kMaxExactInteger) {
    reason = VCKSS_REQUEST_REASON_PHYSICAL_LIMIT;
  } else if (v.controls_count > 32) {
    reason = VCKSS_REQUEST_REASON_CONTROLS_LIMIT;
  } else if (v.algorithm == VCKSS_ALGORITHM_EXACT) {
    if (q.engine == VCKSS_ENGINE_COMPRESSED) reason = VCKSS_REQUEST_REASON_EXACT_ENGINE;
    else if (all && v.deletion_mode != VCKSS_DELETION_MATCH) reason = VCKSS_REQUEST_REASON_STAYERS_MODE_UNSUPPORTED;
    else if (all && q.probeorder_supplied != 0) reason = VCKSS_REQUEST_REASON_PROBEORDER_UNSUPPORTED;
    else if (all && q.wallseconds_supplied != 0) reason = VCKSS_REQUEST_REASON_WALLSECONDS_UNSUPPORTED;
    else if (v.rng_contract != VCKSS_RNG_NONE) reason = VCKSS_REQUEST_REASON_EXACT_RNG;
    else if (!known(v.solver_route, {VCKSS_ROUTE_AUTO, VCKSS_ROUTE_EXACT})) reason = VCKSS_REQUEST_REASON_EXACT_SOLVER_ROUTE;
  } else if (v.algorithm == VCKSS_ALGORITHM_AUTO) {
    if (all && v.deletion_mode != VCKSS_DELETION_MATCH) reason = VCKSS_REQUEST_REASON_STAYERS_MODE_UNSUPPORTED;
    else if (all && q.engine == VCKSS_ENGINE_COMPRESSED) reason = VCKSS_REQUEST_REASON_COMPRESSED_SCIENTIFIC_INELIGIBILITY;
    else if (all && q.probeorder_supplied != 0) reason = VCKSS_REQUEST_REASON_PROBEORDER_UNSUPPORTED;
    else if (all && q.wallseconds_supplied != 0) reason = VCKSS_REQUEST_REASON_WALLSECONDS_UNSUPPORTED;
    else if (v.rng_contract != VCKSS_RNG_COUNTER_V1) reason = VCKSS_REQUEST_REASON_AUTO_RNG;
    else if (v.solver_route != VCKSS_ROUTE_AUTO) reason = VCKSS_REQUEST_REASON_EXACT_SOLVER_ROUTE;
    else if (q.engine == VCKSS_ENGINE_COMPRESSED) reason = VCKSS_REQUEST_REASON_AUTO_ENGINE_COMPRESSED;
  } else if (all && v.deletion_mode != VCKSS_DELETION_MATCH) {
    reason = VCKSS_REQUEST_REASON_STAYERS_MODE_UNSUPPORTED;
  } else if (all && q.engine == VCKSS_ENGINE_COMPRESSED) {
    reason = VCKSS_REQUEST_REASON_COMPRESSED_SCIENTIFIC_INELIGIBILITY;
  } else if (all && q.probeorder_supplied != 0) {
    reason = VCKSS_REQUEST_REASON_PROBEORDER_UNSUPPORTED;
  } else if (all && q.wallseconds_supplied != 0) {
    reason = VCKSS_REQUEST_REASON_WALLSECONDS_UNSUPPORTED;
  } else if (v.rng_contract != VCKSS_RNG_COUNTER_V1) {
    reason = VCKSS_REQUEST_REASON_JLA_RNG;
  } else if (v.solver_route == VCKSS_ROUTE_EXACT) {
    reason = VCKSS_REQUEST_REASON_JLA_GENERIC_SOLVER_ROUTE;
  } else if (q.engine == VCKSS_ENGINE_COMPRESSED &&
             (v.deletion_mode != VCKSS_DELETION_MATCH || v.controls_count != 0)) {
    reason = VCKSS_REQUEST_REASON_COMPRESSED_SCIENTIFIC_INELIGIBILITY;
  }
  if (reason == VCKSS_REQUEST_REASON_SUPPORTED)
    reason = engine_restriction(v.algorithm, v.deletion_mode, v.solver_route, v.controls_count, q.probeorder_supplied,
                                r.leverage_batch_mode != VCKSS_BATCH_MODE_AUTO ||
                                    r.target_batch_mode != VCKSS_BATCH_MODE_AUTO);
  return {reason, reason == VCKSS_REQUEST_REASON_SUPPORTED ? VCKSS_REQUEST_PROFILE_PLANNED_V1 : VCKSS_REQUEST_PROFILE_NONE};
}

// ------------------------------------------------------------- receipts
// Byte model of a prepared generation (host arrays, hybrid graph, CMG
// hierarchy with its device layout) and of one solve batch's vectors, for
// the memory, preparation and full-CMG receipts.
struct MemoryModel {
  uint64_t host_problem = 0, graph = 0, hierarchy = 0, plan = 0, workspace = 0, device_live = 0, device_peak = 0;
};

MemoryModel memory_model(const Context &ctx) {
  MemoryModel m;
  const Prepared &P = ctx.P;
  m.host_problem = P.rows * (4 * 4 + 8 * 3 + 8) + P.cells * (4 * 2 + 8 * 3 + 8 * 2) + P.units * (4 + 8 * 4 + 8 * 2) +
                   P.strata * (4 + 8 * 4) + P.input_rows;
  m.graph = (uint64_t)ctx.HS.graph.u.size() * 16 + (uint64_t)ctx.HS.graph.n * 8;
  for (const Level &L : ctx.H.levels) {
    const uint64_t nnz = 2 * (uint64_t)L.graph.u.size() + L.graph.n;
    m.hierarchy += nnz * (4 + 8 + 4) + (uint64_t)L.graph.n * (8 + 4 + 4 + 4);
    m.plan += (nnz / 16 + L.graph.n / 16 + 2) * 16 * 2;
  }
  m.hierarchy += (uint64_t)ctx.H.terminal_inverse.size() * 8;
  const uint64_t K = 64, n0 = ctx.H.levels.empty() ? 0 : ctx.H.levels[0].graph.n;
  m.workspace = K * 8 * (12 * n0 + 6 * (P.workers + P.firms)) + K * 4 * P.units;
  device_memory(m.device_live, m.device_peak);
  return m;
}

// One row per solved right-hand side: the fit, then every leverage probe,
// then the worker and firm side of every target probe (ffi_engine.rs:8865-8879,
// 9005-9025).
std::vector<VckssEngineRhsReceiptV1> rhs_rows(const Context &ctx) {
  std::vector<VckssEngineRhsReceiptV1> rows;
  if (!ctx.has_result || !ctx.full_cmg) return rows;
  const SolveResult &r = ctx.result;
  auto row = [&](uint32_t phase, uint32_t side, int64_t probe, const ColumnReceipt &c) {
    VckssEngineRhsReceiptV1 v;
    std::memset(&v, 0, sizeof v);
    v.phase = phase;
    v.side = side;
    v.probe = probe;
    v.route = VCKSS_ROUTE_CMG_PCG;
    v.iterations = c.iterations;
    v.zero_rhs = c.rhs_norm == 0.0 ? 1u : 0u;
    v.reduced_residual = c.reduced;
    v.complete_residual = c.complete;
    rows.push_back(v);
  };
  // ffi_engine.rs:8880-8890 phases: control projections (5), the full joint
  // fit (1), the fixed-offset working fit (4), leverage (2), target (3)
  for (uint32_t j = 0; j < r.controls && j < r.control_fit.size(); ++j) row(5, 0, (int64_t)j, r.control_fit[j]);
  if (r.fixed_offset && r.control_fit.size() > r.controls) {
    row(1, 0, -1, r.control_fit[r.controls]);
    row(4, 0, -1, r.fit);
  } else {
    row(1, 0, -1, r.fit);
  }
  for (size_t t = 0; t < r.leverage.size(); ++t) row(2, 0, (int64_t)t, r.leverage[t]);
  for (size_t c = 0; c < r.target.size(); ++c) row(3, 1 + (uint32_t)(c % 2), (int64_t)(c / 2), r.target[c]);
  return rows;
}

const Context &solved(uint64_t generation) {
  const Context &ctx = live(generation);
  if (!ctx.has_result) fail(kStaleContext, "engine_receipt", "no solved result is available");
  return ctx;
}

VckssEngineDetailedReceiptV6 detailed_v6(uint64_t generation, const Context &ctx) {
  const SolveResult &r = ctx.result;
  const auto &q = ctx.request;
  const auto &v1 = q.v4.v3.v2.v1;
  const MemoryModel m = memory_model(ctx);
  VckssEngineDetailedReceiptV6 o;
  std::memset(&o, 0, sizeof o);
  VckssEngineDetailedReceiptV2 &d = o.v5.v4.v3.v2;
  d.struct_size = sizeof(VckssEngineDetailedReceiptV6);
  d.generation = generation;
  d.seed = v1.seed;
  d.probes_requested = v1.probes;
  d.leverage_probes_accepted = (uint32_t)r.leverage.size();
  d.target_probes_accepted = (uint32_t)(r.target.size() / 2);
  d.solver_requested = v1.solver_route;
  d.solver_selected = ctx.full_cmg ? VCKSS_ROUTE_CMG_PCG : VCKSS_ROUTE_EXACT;
  d.solver_dimension = ctx.full_cmg ? ctx.HS.graph.n : ctx.P.workers + ctx.P.firms - 1;
  d.leverage_batch_width = ctx.full_cmg ? ctx.leverage_width : 0;
  d.target_batch_width = ctx.full_cmg ? ctx.target_width : 0;
  d.rank_tolerance = v1.rank_tolerance;
  d.block_tolerance = v1.block_tolerance;
  const double probe_tol = q.tolerance_supplied ? v1.pcg_tolerance : 1e-6;
  d.full_residual_tolerance = std::max(10.0 * std::max(v1.pcg_tolerance, probe_tol), 1e-11);
  d.full_fit_route = d.solver_selected;
  d.full_fit_iterations = r.fit.iterations;
  d.full_fit_reduced_residual = r.fit.reduced;
  d.full_fit_complete_residual = r.fit.complete;
  d.full_fit_zero_rhs = r.fit.rhs_norm == 0.0 && ctx.full_cmg;
  d.leverage_rhs_count = r.leverage.size();
  d.target_rhs_count = r.target.size();
  double max_red = r.fit.reduced, max_comp = r.max_complete_residual;
  for (const auto *set : {&r.leverage, &r.target})
    for (const ColumnReceipt &c : *set) max_red = std::max(max_red, c.reduced);
  d.max_reduced_residual = max_red;
  d.max_complete_residual = max_comp;
  double max_lev = r.has_exact ? r.exact.max_leverage : 0.0, max_recip = 0.0;
  for (double p : r.projection_share) max_lev = std::max(max_lev, p);
  for (double m2 : r.residual_share)
    if (m2 > 0.0) max_recip = std::max(max_recip, 1.0 / m2);
  d.max_leverage = max_lev;
  d.max_reciprocal_residual = max_recip;
  d.accounting_residual = r.accounting_residual;
  d.topology_checksum = ctx.P.topology_checksum;
  d.cmg_levels = ctx.H.levels.size();
  if (!ctx.H.levels.empty()) {
    d.cmg_fine_vertices = ctx.H.levels[0].graph.n;
    d.cmg_fine_edges = ctx.H.levels[0].graph.u.size();
    d.cmg_terminal_vertices = ctx.H.levels.back().graph.n;
    double e = 0.0, n = 0.0;
    for (const Level &L : ctx.H.levels) {
      e += (double)L.graph.u.size();
      n += (double)L.graph.n;
    }
    d.cmg_edge_complexity = e / std::max<double>(1.0, (double)d.cmg_fine_edges);
    d.cmg_vertex_complexity = n / std::max<double>(1.0, (double)d.cmg_fine_vertices);
  }
  d.cmg_structural_bytes = m.graph + m.hierarchy;
  d.cmg_workspace_bytes = m.workspace;
  d.cmg_dense_factor_bytes = (uint64_t)ctx.H.terminal_inverse.size() * 8;
  d.full_fit_weighted_rss = r.weighted_rss;
  d.memory_limit_bytes = ctx.memory_limit_bytes;
  d.caller_copy_bytes = ctx.caller_copy_bytes;
  d.preparation_peak_forecast_bytes = ctx.preparation_forecast;
  d.prepared_resident_bytes = m.host_problem + m.graph + m.hierarchy + m.plan;
  d.solver_setup_forecast_bytes = m.graph + m.hierarchy + m.plan;
  d.leverage_phase_forecast_bytes = m.workspace;
  d.target_phase_forecast_bytes = m.workspace;
  d.result_forecast_bytes = (ctx.P.cells * 2 + ctx.P.units * 6) * 8 + r.draws.size() * 8;
  d.solve_peak_forecast_bytes = std::max(m.device_peak, d.solver_setup_forecast_bytes + m.workspace);
  d.command_peak_forecast_bytes = d.prepared_resident_bytes + d.solve_peak_forecast_bytes + d.result_forecast_bytes;
  VckssEngineDetailedReceiptV3 &d3 = o.v5.v4.v3;
  d3.rng_contract = v1.rng_contract;
  const uint64_t rows = rhs_rows(ctx).size();
  d3.rhs_receipt_rows = rows;
  d3.caller_result_copy_bytes = rows * (sizeof(VckssEngineRhsReceiptV1) + 8 * sizeof(double));
  VckssEngineDetailedReceiptV4 &d4 = o.v5.v4;
  d4.algorithm_requested = q.v4.v3.v2.algorithm;
  d4.algorithm_selected = ctx.algorithm_selected;
  d4.deletion_mode = v1.deletion_mode;
  d4.nuisance_mode = q.v4.v3.v2.nuisance_mode;
  d4.parameters = d4.full_parameters = d4.correction_parameters = ctx.P.workers + ctx.P.firms - 1;
  if (r.has_exact) {
    d4.information_rcond = r.exact.information_rcond;
    d4.inverse_relres = r.exact.inverse_relres;
  }
  VckssEngineDetailedReceiptV5 &d5 = o.v5;
  d5.working_fit_complete_residual = r.fit.complete;
  if (r.has_exact) {
    d5.deletion_rank_gap = r.exact.deletion_rank_gap;
    d5.firm_zero_sum_residual = r.exact.firm_zero_sum_residual;
  }
  d5.fit_peak_forecast_bytes = d.solver_setup_forecast_bytes + m.workspace;
  d5.correction_peak_forecast_bytes = d.solve_peak_forecast_bytes;
  d5.actual_accounting_residual = r.accounting_residual;
  o.engine_requested = ctx.engine_requested;
  o.engine_selected = ctx.engine_selected;
  o.controls_count = r.controls;
  o.rhs_receipt_schema = rows ? 1u : 0u;
  // the full joint fit: its outcome column (joint nuisance), or the outcome
  // column of the control fit before the working fit (fixed offset)
  o.full_joint_fit_complete_residual =
      r.fixed_offset && r.control_fit.size() > r.controls ? r.control_fit[r.controls].complete : r.fit.complete;
  o.generic_working_fit_complete_residual = r.fit.complete;
  o.capability_schema = q.v4.v3.capability_schema;
  o.capability_profile = q.v4.v3.capability_profile;
  o.batch_mode = q.v4.v3.batch_mode;
  o.stayers_mode = q.v4.v3.stayers_mode;
  o.target_weight_mode = q.v4.v3.target_weight_mode;
  o.deletion_unit_source = q.v4.v3.deletion_unit_source;
  o.probeorder_supplied = q.v4.v3.probeorder_supplied;
  o.wallseconds_supplied = q.v4.v3.wallseconds_supplied;
  o.frequency_use = q.v4.v3.frequency_use;
  o.physical_limit = q.v4.v3.physical_limit;
  o.request_signature = q.v4.v3.request_signature;
  return o;
}

VckssExecutionPlanReceiptV1 execution_plan(uint64_t generation, const Context &ctx) {
  const SolveResult &r = ctx.result;
  const auto &q = ctx.request;
  const auto &v1 = q.v4.v3.v2.v1;
  const MemoryModel m = memory_model(ctx);
  VckssExecutionPlanReceiptV1 e;
  std::memset(&e, 0, sizeof e);
  e.struct_size = sizeof e;
  e.schema_version = 1;
  e.generation = generation;
  e.request_signature = q.v4.v3.request_signature;
  auto &res = e.resolution;
  res.algorithm_schema = res.engine_schema = 1;
  res.algorithm_requested = q.v4.v3.v2.algorithm;
  res.algorithm_selected = ctx.algorithm_selected;
  res.engine_requested = ctx.engine_requested;
  res.engine_selected = ctx.engine_selected;
  res.compressed_eligibility = (!ctx.P.observation && !ctx.P.stayers) ? 1u : 0u;
  res.resolved_before_rng = 1;
  res.identified_complexity = ctx.P.workers + ctx.P.firms - 1;
  res.exact_limit = q.v4.v3.v2.exact_estimator_limit;
  auto &sol = e.solver;
  sol.schema_version = 1;
  sol.requested_route = v1.solver_route;
  sol.selected_route = ctx.full_cmg ? VCKSS_ROUTE_CMG_PCG : VCKSS_ROUTE_EXACT;
  sol.full_setup_complete = 1;
  sol.plan_frozen_before_rng = 1;
  sol.threads_requested = sol.threads_used = std::max<uint32_t>(q.threads, 1);
  sol.planned_rhs = 1 + r.control_fit.size() + r.leverage.size() + r.target.size();
  sol.full_solver_dimension = ctx.full_cmg ? ctx.HS.graph.n : ctx.P.workers + ctx.P.firms - 1;
  auto &b = e.batch;
  b.schema_version = 1;
  b.deterministic = 1;
  b.width_invariance_required = 1;
  b.whole_command_admitted = 1;
  b.applicability = ctx.full_cmg ? VCKSS_PLAN_APPLICABILITY_COMPRESSED : VCKSS_PLAN_APPLICABILITY_EXACT;
  b.non_batched_peak_bytes = m.graph + m.hierarchy + m.plan;
  b.selected_command_peak_bytes = b.non_batched_peak_bytes + m.workspace;
  for (int side = 0; side < 2; ++side) {
    auto &ph = side ? b.target : b.leverage;
    ph.request_mode = side ? q.v4.target_batch_mode : q.v4.leverage_batch_mode;
    ph.selection_reason = ctx.full_cmg ? VCKSS_BATCH_SELECTION_AUTO : VCKSS_BATCH_SELECTION_NOT_APPLICABLE;
    ph.applicability = b.applicability;
    ph.selected_width = ctx.full_cmg ? (side ? ctx.target_width : ctx.leverage_width) : 0;
    ph.probe_width_cap = v1.probes;
    ph.declared_threads = sol.threads_requested;
    ph.route_width_cap = ph.effective_width_cap = side ? 32 : 64;
    ph.hard_memory_bytes = ctx.memory_limit_bytes;
    ph.selected_forecast_bytes = m.workspace;
  }
  e.wall.schema_version = 1;
  e.wall.requested_applicable = q.v4.v3.wallseconds_supplied;
  e.wall.requested_seconds = q.v4.wallseconds;
  auto &c = e.counter;
  c.schema_version = 1;
  c.rng_contract = v1.rng_contract;
  c.completed = 1;
  c.leverage.planned_logical_atoms = c.leverage.actual_logical_atoms = ctx.P.units * (uint64_t)r.leverage.size();
  c.target.planned_logical_atoms = c.target.actual_logical_atoms = ctx.P.strata * (uint64_t)(r.target.size() / 2);
  c.total.planned_logical_atoms = c.total.actual_logical_atoms =
      c.leverage.actual_logical_atoms + c.target.actual_logical_atoms;
  auto &mem = e.memory;
  mem.schema_version = 1;
  mem.applicability = 1;
  mem.hard_limit_bytes = ctx.memory_limit_bytes;
  mem.prepared_persistent_bytes = m.host_problem;
  mem.setup_peak_bytes = m.graph + m.hierarchy + m.plan;
  mem.leverage_peak_bytes = mem.target_peak_bytes = mem.setup_peak_bytes + m.workspace;
  mem.fit_peak_bytes = mem.leverage_peak_bytes;
  mem.command_peak_bytes = m.host_problem + mem.leverage_peak_bytes;
  mem.non_batched_peak_bytes = mem.setup_peak_bytes;
  mem.shared_cmg_persistent_bytes = m.hierarchy;
  mem.cmg_preconditioner_workspace_bytes = m.workspace;
  mem.cmg_hybrid_graph_bytes = m.graph;
  return e;
}

}  // namespace

namespace vk {
void record_unsupported(const char *name) {
  g_last_error = std::string(name) +
                 ": not supported by the MI355X engine (it implements the no-control leave-out path: "
                 "match or observation deletion, movers or the stayer hybrid, JLA on route CMG_FULL_V2 "
                 "and the dense exact estimator)";
}
const char *thread_last_error() { return g_last_error.c_str(); }
}  // namespace vk

// ------------------------------------------------------------ sample batches
// Independent samples (e.g. region-year subsamples) prepared together and
// estimated concurrently: each sample is its own prepared problem, hierarchy
// and device engine, and up to `streams` of them run at once, each driven by
// a host thread on its own non-blocking HIP stream.  No counterpart in the
// reference, whose caller runs one sample per command.
struct BatchSample {
  uint64_t id = 0;
  Prepared P;
  HybridSystem HS;
  Hierarchy H;
  std::unique_ptr<DeviceEngine> dev;
  SolveResult result;
  bool solved = false;
  int32_t status = 0;
  std::string error;
  double seconds = 0.0;
};
struct Batch {
  bool observation = false;
  std::vector<std::unique_ptr<BatchSample>> samples;
};
std::map<uint64_t, std::unique_ptr<Batch>> g_batches;
uint64_t g_next_batch = 1;

Batch &live_batch(uint64_t handle) {
  auto it = g_batches.find(handle);
  if (it == g_batches.end()) fail(kStaleContext, "batch", "batch handle " + std::to_string(handle) + " is not live");
  return *it->second;
}

// Run f(i) for i in [0, n) on `threads` host threads, in the given order;
// the first failure (lowest i) is rethrown after every thread has joined.
template <typename F>
void parallel_for(const std::vector<size_t> &order, uint32_t threads, F &&f) {
  const size_t n = order.size();
  std::vector<std::exception_ptr> errors(n);
  std::atomic<size_t> next{0};
  auto work = [&] {
    for (size_t k = next++; k < n; k = next++) {
      try {
        f(order[k]);
      } catch (...) {
        errors[order[k]] = std::current_exception();
      }
    }
  };
  const uint32_t t = std::max<uint32_t>(1, std::min<uint32_t>(threads, (uint32_t)n));
  if (t == 1) {
    work();
  } else {
    std::vector<std::thread> pool;
    for (uint32_t i = 0; i < t; ++i) pool.emplace_back(work);
    for (auto &th : pool) th.join();
  }
  for (auto &e : errors)
    if (e) std::rethrow_exception(e);
}

extern "C" {

uint32_t vckss_rust_abi_version(void) { return VCKSS_RUST_ABI_VERSION_V1; }
const char *vckss_rust_backend_version(void) { return "vckss-mi355x 0.1.0 (gfx950, CMG_FULL_V2)"; }
const char *vckss_rust_capabilities_json(void) {
  return "{\"backend\":\"vckss-mi355x\",\"device\":\"gfx950\",\"routes\":[\"CMG_FULL_V2\"],"
         "\"deletion\":[\"match\",\"observation\"],\"stayers\":[\"movers\",\"all\"],\"rng\":[\"VCKSS-COUNTER-V1\"],\"controls\":32,"
         "\"precision\":[64,32],\"max_batch_rhs\":64}";
}
const char *vckss_rust_last_error(void) { return g_last_error.c_str(); }
const char *vckss_rust_engine_last_error(void) { return g_last_error.c_str(); }

int32_t vckss_rust_selftest(void) {
  return guarded([] {
    // A two-vertex hierarchy exercises the host setup; a visible HIP device
    // is required for every solve.
    const uint32_t u[1] = {0}, v[1] = {1};
    const double w[1] = {1.0};
    Hierarchy h = build_hierarchy(laplacian_from_edges(2, 1, u, v, w), CmgOptions{});
    if (h.levels.size() != 1 || h.ldl_dim != 1) fail(kInvariant, "selftest", "host hierarchy self-test failed");
    if (!device_available()) fail(kUnsupported, "selftest", "no HIP device is visible");
  });
}

int32_t vckss_rust_engine_default_prepare_request_interrupt_v3(VckssEnginePrepareRequestInterruptV3 *output,
                                                               uint32_t output_capacity_bytes) {
  return guarded([&] {
    require_capacity<VckssEnginePrepareRequestInterruptV3>(output, output_capacity_bytes, "prepare request");
    VckssEnginePrepareRequestInterruptV3 d;
    std::memset(&d, 0, sizeof d);
    d.options.v3.v2.abi_version = VCKSS_RUST_ABI_VERSION_V1;
    d.options.v3.v2.struct_size = sizeof(VckssEnginePrepareRequestInterruptV3);  // ffi_engine.rs:1023-1036
    d.options.v3.deletion_mode = VCKSS_DELETION_MATCH;
    d.checkpoint_interval = 0;  // no callback (ffi_engine.rs:520-540)
    std::memcpy(output, &d, sizeof d);
  });
}

int32_t vckss_rust_engine_default_solve_request_interrupt_v5(VckssEngineSolveRequestInterruptV5 *output,
                                                             uint32_t output_capacity_bytes) {
  return guarded([&] {
    require_capacity<VckssEngineSolveRequestInterruptV5>(output, output_capacity_bytes, "solve request");
    VckssEngineSolveRequestInterruptV5 d;
    std::memset(&d, 0, sizeof d);
    d.options = default_solve_v5();
    d.checkpoint_interval = 0;  // no callback (ffi_engine.rs:520-540)
    std::memcpy(output, &d, sizeof d);
  });
}

static int32_t prepare_common(const VckssEnginePrepareRequestInterruptV3 *request,
                              const VckssEngineColumnsV3 *columns, uint64_t *output_handle, bool upload) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (!output_handle) fail(kInvalidInput, "engine_prepare", "output handle pointer is null");
    *output_handle = 0;
    if (g_ctx) {
      if (request && request->options.v3.v2.cleanup_abandoned) {
        g_last_released = g_ctx->generation;
        g_ctx.reset();
      } else {
        fail(kStaleContext, "engine_prepare", "a prepared context is still live; release it first");
      }
    }
    std::unique_ptr<Context> ctx(new Context);
    prepare_into(*ctx, request, columns);
    const auto t = std::chrono::steady_clock::now();
    if (upload) ctx->dev.reset(new DeviceEngine(ctx->P, ctx->HS, ctx->H));
    ctx->upload_ns = elapsed_ns(t);
    ctx->generation = g_next_generation++;
    ctx->state = kPrepared;
    *output_handle = ctx->generation;
    g_ctx = std::move(ctx);
  });
}

// ---- stayer hybrid: ffi_engine.rs:2867-2879, 3667-3676, 4171-4262, 9672-9760;
// session_retained.rs:850-943; context.rs:385-440 (a Prepared generation only;
// a failed augmentation leaves it Failed)
int32_t vckss_rust_engine_default_stayer_augmentation_request_interrupt_v1(
    VckssStayerAugmentationRequestInterruptV1 *output, uint32_t output_capacity_bytes) {
  return guarded([&] {
    require_capacity<VckssStayerAugmentationRequestInterruptV1>(output, output_capacity_bytes,
                                                                "default interrupt stayer augmentation request");
    VckssStayerAugmentationRequestInterruptV1 d;
    std::memset(&d, 0, sizeof d);
    d.options.abi_version = VCKSS_RUST_ABI_VERSION_V1;
    d.options.struct_size = sizeof(VckssStayerAugmentationRequestInterruptV1);
    std::memcpy(output, &d, sizeof d);
  });
}

static void augment_stayers_value(uint64_t generation, const VckssStayerAugmentationRequestV1 &r,
                                  const VckssStayerAugmentationColumnsV1 *columns, Interrupt &interrupt) {
  if (r.abi_version != VCKSS_RUST_ABI_VERSION_V1) fail(kAbiMismatch, "engine_abi", "unsupported ABI version");
  if (r.struct_size < sizeof(VckssStayerAugmentationRequestV1) || r.reserved != 0)
    fail(kAbiMismatch, "engine_abi", "stayer augmentation request is short or reserves nonzero fields");
  if (!columns) fail(kInvalidInput, "engine_stayer_augmentation", "stayer augmentation column descriptor is null");
  VckssStayerAugmentationColumnsV1 c;
  std::memcpy(&c, columns, sizeof c);
  if (c.struct_size < sizeof(VckssStayerAugmentationColumnsV1))
    fail(kAbiMismatch, "engine_abi", "stayer augmentation column descriptor reports a short structure size");
  if (c.reserved != 0 || c.reserved_2 != 0)
    fail(kAbiMismatch, "engine_abi", "reserved stayer augmentation column fields must be zero");
  if (r.rows != c.rows || r.controls_count != c.controls_count)
    fail(kInvalidInput, "engine_stayer_augmentation", "stayer request and column dimensions disagree");
  if (c.rows == 0) {
    if (c.firm || c.worker || c.outcome || c.frequency || c.target_weight || c.controls)
      fail(kInvalidInput, "engine_stayer_augmentation", "zero stayer rows require null numeric and control pointers");
  } else if (c.controls_count == 0 && c.controls) {
    fail(kInvalidInput, "engine_stayer_augmentation", "zero controls require a null control-pointer array");
  }
  if (c.controls_count != 0)
    fail(kUnsupported, "engine_stayer_augmentation", "the MI355X engine supports the no-control route only");
  const uint64_t rows = c.rows;
  std::vector<uint64_t> firm, worker, frequency;
  std::vector<double> outcome, target;
  if (rows) {
    firm = integer_column(c.firm, rows, "stayer firm identifier", kInvalidIdentifier, false);
    worker = integer_column(c.worker, rows, "stayer worker identifier", kInvalidIdentifier, false);
    frequency = integer_column(c.frequency, rows, "stayer frequency weight", kInvalidWeight, false);
    if (!c.outcome || !c.target_weight)
      fail(kInvalidInput, "engine_stayer_augmentation", "stayer outcome or target column is null");
    outcome.assign(c.outcome, c.outcome + rows);
    target.assign(c.target_weight, c.target_weight + rows);
  }
  Context &ctx = live(generation);
  if (ctx.P.controls)
    fail(kUnsupported, "engine_stayer_augmentation",
         "the MI355X engine does not augment a generation prepared with controls");
  if (ctx.state != kPrepared)
    fail(kContextPoisoned, "context_augment", "generation " + std::to_string(generation) +
                                                  " is not in the Prepared state");
  // session.rs:309-372 under the legacy budget: the stayer copy is exactly 8
  // bytes per row and column (five columns), and the prepared generation's
  // resident bytes plus that copy plus 768 bytes a combined row and 4096
  // fixed must fit the limit the generation was prepared under.
  if (r.caller_copy_bytes != rows * 5u * 8u)
    fail(kInvalidInput, "stayer_augmentation_memory",
         "declared stayer caller copy is " + std::to_string(r.caller_copy_bytes) + " bytes; 5 columns require " +
             std::to_string(rows * 40u) + " bytes");
  {
    const MemoryModel m = memory_model(ctx);
    const uint64_t resident = m.host_problem + m.graph + m.hierarchy;
    const uint64_t peak = resident + r.caller_copy_bytes + (ctx.P.input_rows + rows) * 768u + 4096u;
    if (peak > ctx.memory_limit_bytes)
      fail(kResourceLimit, "stayer_augmentation_memory",
           "stayer augmentation forecast " + std::to_string(peak) + " bytes exceeds the declared limit " +
               std::to_string(ctx.memory_limit_bytes) + " bytes");
    ctx.stayer_forecast = peak;
  }
  interrupt.checkpoint("engine_stayer_augmentation");
  StayerColumns sc{rows, firm.data(), worker.data(), frequency.data(), outcome.data(), target.data()};
  try {
    StayerReceipt rec;
    const auto t = std::chrono::steady_clock::now();
    Prepared P = augment_stayers(ctx.P, sc, rec);
    HybridSystem HS = build_hybrid(P);
    CmgOptions copt;  // full_cmg.rs:467-476: max_levels 128
    Hierarchy H = build_hierarchy(HS.graph, copt);
    const bool upload = static_cast<bool>(ctx.dev);
    ctx.dev.reset();
    ctx.P = std::move(P);
    ctx.HS = std::move(HS);
    ctx.H = std::move(H);
    if (upload) ctx.dev.reset(new DeviceEngine(ctx.P, ctx.HS, ctx.H));
    ctx.stayer_receipt = rec;
    ctx.stayer_copy_bytes = r.caller_copy_bytes;
    ctx.stayer_ns = elapsed_ns(t);
  } catch (...) {
    ctx.state = kFailed;
    throw;
  }
}

int32_t vckss_rust_engine_augment_stayers_v1(uint64_t generation, const VckssStayerAugmentationRequestV1 *request,
                                             const VckssStayerAugmentationColumnsV1 *columns) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (!request) fail(kInvalidInput, "engine_stayer_augmentation", "request pointer is null");
    VckssStayerAugmentationRequestV1 r;
    std::memcpy(&r, request, sizeof r);
    Interrupt none;
    augment_stayers_value(generation, r, columns, none);
  });
}

int32_t vckss_rust_engine_augment_stayers_interrupt_v1(uint64_t generation,
                                                       const VckssStayerAugmentationRequestInterruptV1 *request,
                                                       const VckssStayerAugmentationColumnsV1 *columns) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (!request) fail(kInvalidInput, "engine_stayer_augmentation", "request pointer is null");
    VckssStayerAugmentationRequestInterruptV1 r;
    std::memcpy(&r, request, sizeof r);
    if (r.reserved != 0) fail(kAbiMismatch, "engine_abi", "reserved interrupt fields must be zero");
    Interrupt interrupt(r.interrupt_poll, r.interrupt_context, r.checkpoint_interval);
    augment_stayers_value(generation, r.options, columns, interrupt);
  });
}

// ffi_engine.rs:8737-8770
int32_t vckss_rust_engine_stayer_augmentation_receipt_v1(uint64_t generation, VckssStayerAugmentationReceiptV1 *output,
                                                         uint32_t output_capacity_bytes) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    require_capacity<VckssStayerAugmentationReceiptV1>(output, output_capacity_bytes, "stayer augmentation receipt");
    Context &ctx = live(generation);
    if (!ctx.P.stayers)
      fail(kInvalidInput, "engine_stayer_augmentation", "the generation owns no stayer augmentation");
    const StayerReceipt &s = ctx.stayer_receipt;
    VckssStayerAugmentationReceiptV1 o;
    std::memset(&o, 0, sizeof o);
    o.struct_size = sizeof o;
    o.schema_version = 1;
    o.generation = generation;
    o.mover_stored_rows = s.mover_stored_rows; o.stayer_stored_rows = s.stayer_stored_rows;
    o.combined_stored_rows = s.combined_stored_rows;
    o.mover_physical_mass = s.mover_physical_mass; o.stayer_physical_mass = s.stayer_physical_mass;
    o.combined_physical_mass = s.combined_physical_mass;
    o.mover_workers = s.mover_workers; o.stayer_workers = s.stayer_workers; o.combined_workers = s.combined_workers;
    o.firms = s.firms;
    o.mover_deletion_units = s.mover_deletion_units; o.stayer_deletion_units = s.stayer_deletion_units;
    o.combined_deletion_units = s.combined_deletion_units;
    o.mover_target_mass = s.mover_target_mass; o.stayer_target_mass = s.stayer_target_mass;
    o.combined_target_mass = s.combined_target_mass;
    o.topology_checksum = s.topology_checksum;
    o.caller_copy_bytes = ctx.stayer_copy_bytes;
    std::memcpy(output, &o, sizeof o);
  });
}

// ffi_engine.rs:6122-6145, 8772-8840: after an exact solve of a stayer hybrid
int32_t vckss_rust_engine_stayer_hybrid_result_v1(uint64_t generation, VckssStayerHybridResultV1 *output,
                                                  uint32_t output_capacity_bytes) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    require_capacity<VckssStayerHybridResultV1>(output, output_capacity_bytes, "stayer hybrid result");
    Context &ctx = live(generation);
    if (!ctx.has_result) fail(kInvalidInput, "engine_result", "the generation has no solved result");
    const SolveResult &r = ctx.result;
    if (!r.has_exact || !ctx.P.stayers)
      fail(kUnsupported, "engine_stayer_hybrid_result", "the solved generation has no stayer-hybrid result");
    const ExactReceipt &x = r.exact;
    auto identity = [](const double (&v)[4]) {
      const double scale = std::max({std::fabs(v[0]), std::fabs(v[1]), std::fabs(2.0 * v[2]), std::fabs(v[3]), 1.0});
      return std::fabs(v[3] - (v[0] + v[1] + 2.0 * v[2])) / scale;
    };
    // jla.rs:232-269 verify_accounting(1e-10) on the five vectors, then the
    // source identity mover + stayer = total
    for (const double(*v)[4] : {&r.plugin, &r.correction, &r.corrected, &x.mover_correction, &x.stayer_correction})
      if (!(identity(*v) <= 1e-10)) fail(kAccounting, "engine_stayer_hybrid_result", "accounting identity failed");
    double source = 0.0;
    for (int k = 0; k < 4; ++k)
      source = std::max(source, std::fabs(x.mover_correction[k] + x.stayer_correction[k] - r.correction[k]));
    if (source > 1e-10)
      fail(kInvariant, "engine_stayer_hybrid_result",
           "mixed-deletion source corrections do not sum to the total correction");
    auto vec = [](const double (&v)[4]) {
      VckssComponentVectorV1 c;
      c.worker = v[0]; c.firm = v[1]; c.covariance = v[2]; c.total = v[3];
      return c;
    };
    VckssStayerHybridResultV1 o;
    std::memset(&o, 0, sizeof o);
    o.struct_size = sizeof o;
    o.schema_version = 1;
    o.generation = generation;
    o.plugin = vec(r.plugin);
    o.correction = vec(r.correction);
    o.corrected = vec(r.corrected);
    o.mover_correction = vec(x.mover_correction);
    o.stayer_correction = vec(x.stayer_correction);
    o.weighted_rss = r.weighted_rss;
    o.parameters = o.full_parameters = o.correction_parameters = x.parameters;
    o.deletion_units = x.deletion_units;
    o.max_leverage = x.max_leverage;
    o.information_rcond = x.information_rcond;
    o.inverse_relres = o.inverse_original_relres = x.inverse_relres;
    o.full_fit_relres = o.working_fit_relres = x.fit_relres;
    o.fit_residual_tolerance = x.fit_tolerance;
    o.deletion_rank_gap = x.deletion_rank_gap;
    o.firm_zero_sum_residual = x.firm_zero_sum_residual;
    o.topology_checksum = ctx.P.topology_checksum;
    // ffi_engine.rs:8691-8693: the receipt reports the unscaled identity residuals
    auto raw = [](const double (&v)[4]) { return std::fabs(v[3] - v[0] - v[1] - 2.0 * v[2]); };
    o.accounting_residual = std::max({raw(r.plugin), raw(r.correction), raw(r.corrected), source});
    std::memcpy(output, &o, sizeof o);
  });
}

int32_t vckss_rust_engine_prepare_interrupt_v4(const VckssEnginePrepareRequestInterruptV3 *request,
                                               const VckssEngineColumnsV3 *columns, uint64_t *output_handle,
                                               uint32_t output_capacity_bytes) {
  if (output_capacity_bytes < sizeof(uint64_t)) {
    g_last_error = "engine_prepare: output handle capacity is too small";
    return kAbiMismatch;
  }
  return prepare_common(request, columns, output_handle, true);
}

int32_t vckss_mi355x_prepare_host_v1(const VckssEnginePrepareRequestInterruptV3 *request,
                                     const VckssEngineColumnsV3 *columns, uint64_t *output_handle) {
  return prepare_common(request, columns, output_handle, false);
}

int32_t vckss_rust_engine_retained_mask_v1(uint64_t generation, uint8_t *output, uint64_t output_capacity_bytes) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (!output) fail(kInvalidInput, "engine_mask", "output pointer is null");
    if (output_capacity_bytes < ctx.P.input_rows) fail(kAbiMismatch, "engine_mask", "output capacity is too small");
    std::memcpy(output, ctx.P.mask.data(), ctx.P.input_rows);
  });
}

int32_t vckss_rust_engine_solve_v5(uint64_t generation, const VckssEngineSolveRequestV5 *request) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (!request) fail(kInvalidInput, "engine_solve", "request pointer is null");
    VckssEngineSolveRequestV5 q;
    std::memcpy(&q, request, sizeof q);
    Interrupt none;
    run_solve(live(generation), q, none);
  });
}

int32_t vckss_rust_engine_solve_interrupt_v5(uint64_t generation, const VckssEngineSolveRequestInterruptV5 *request) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (!request) fail(kInvalidInput, "engine_solve", "request pointer is null");
    VckssEngineSolveRequestInterruptV5 q;
    std::memcpy(&q, request, sizeof q);
    if (q.reserved) fail(kAbiMismatch, "engine_solve", "reserved fields must be zero");
    Interrupt interrupt(q.interrupt_poll, q.interrupt_context, q.checkpoint_interval);
    run_solve(live(generation), q.options, interrupt);
  });
}

int32_t vckss_rust_engine_result_v1(uint64_t generation, VckssEngineResultV1 *output, uint32_t output_capacity_bytes) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    require_capacity<VckssEngineResultV1>(output, output_capacity_bytes, "result");
    if (!ctx.has_result) fail(kStaleContext, "engine_result", "no solved result is available");
    const SolveResult &r = ctx.result;
    // ffi_engine.rs:6009-6011: accounting re-verified before export
    if (r.accounting_residual > 1e-10) fail(kAccounting, "engine_result", "accounting identity failed");
    VckssEngineResultV1 o;
    std::memset(&o, 0, sizeof o);
    o.struct_size = sizeof o;
    o.generation = generation;
    auto put = [](VckssComponentVectorV1 &v, const double *c) {
      v.worker = c[0]; v.firm = c[1]; v.covariance = c[2]; v.total = c[3];
    };
    put(o.plugin, r.plugin);
    put(o.correction, r.correction);
    put(o.corrected, r.corrected);
    put(o.numerical_mcse, r.mcse);
    std::memcpy(output, &o, sizeof o);
  });
}

// ---- numerical MCSE attachment (numerical_api.rs:262-664)
uint32_t vckss_rust_numerical_schema_v1(void) { return 1; }
uint32_t vckss_rust_numerical_schema_v2(void) { return 2; }

int32_t vckss_rust_engine_numerical_preflight_v1(const VckssNumericalRequestV1 *request) {
  return guarded([&] {
    if (!request) fail(kInvalidInput, "numerical_preflight", "request pointer is null");
    VckssNumericalRequestV1 r;
    std::memcpy(&r, request, sizeof r);
    validate_numerical_v1(r, 1);
    preflight_point(r.point, r.tolerance_supplied);
  });
}

int32_t vckss_rust_engine_numerical_preflight_v2(const VckssNumericalRequestV2 *request) {
  return guarded([&] {
    if (!request) fail(kInvalidInput, "numerical_preflight", "request pointer is null");
    VckssNumericalRequestV2 r;
    std::memcpy(&r, request, sizeof r);
    validate_numerical_v2(r);
    preflight_point(r.v1.point, r.v1.tolerance_supplied);
  });
}

static void solve_numerical(uint64_t generation, const VckssNumericalRequestV1 &r, Interrupt &interrupt) {
  Context &ctx = live(generation);
  VckssEngineSolveRequestV4 point = r.point;
  // numerical_api.rs:157-161 admits the compressed engine for the attachment
  if (point.v3.engine == VCKSS_ENGINE_COMPRESSED) point.v3.engine = VCKSS_ENGINE_AUTO_OR_UNSPECIFIED;
  SolveOptions o = point_options(point, r.tolerance_supplied, ctx);
  if (o.exact)
    fail(kUnsupported, "numerical_solve", "the MI355X engine attaches numerical MCSE to the JLA route only");
  if (ctx.P.controls)
    fail(kUnsupported, "numerical_solve", "the MI355X engine attaches numerical MCSE without controls only");
  o.numerical = true;
  VckssEngineSolveRequestV5 q;
  std::memset(&q, 0, sizeof q);
  q.v4 = r.point;
  q.threads = r.threads;
  q.tolerance_supplied = r.tolerance_supplied;
  q.full_cmg_v2 = 1;
  record_request(ctx, q, o);
  run_options(ctx, o, interrupt);
}

int32_t vckss_rust_engine_solve_numerical_v1(uint64_t generation, const VckssNumericalRequestV1 *request) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (!request) fail(kInvalidInput, "numerical_solve", "request pointer is null");
    VckssNumericalRequestV1 r;
    std::memcpy(&r, request, sizeof r);
    validate_numerical_v1(r, 1);
    Interrupt none;
    solve_numerical(generation, r, none);
  });
}

int32_t vckss_rust_engine_solve_numerical_interrupt_v1(uint64_t generation, const VckssNumericalRequestV1 *request,
                                                       VckssInterruptPollV1 callback, void *context,
                                                       uint32_t interval) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (!request) fail(kInvalidInput, "numerical_solve", "request pointer is null");
    VckssNumericalRequestV1 r;
    std::memcpy(&r, request, sizeof r);
    validate_numerical_v1(r, 1);
    Interrupt interrupt(callback, context, interval);
    solve_numerical(generation, r, interrupt);
  });
}

int32_t vckss_rust_engine_solve_numerical_interrupt_v2(uint64_t generation, const VckssNumericalRequestV2 *request,
                                                       VckssInterruptPollV1 callback, void *context,
                                                       uint32_t interval) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (!request) fail(kInvalidInput, "numerical_solve", "request pointer is null");
    VckssNumericalRequestV2 r;
    std::memcpy(&r, request, sizeof r);
    validate_numerical_v2(r);
    Interrupt interrupt(callback, context, interval);
    solve_numerical(generation, r.v1, interrupt);
  });
}

static const NumericalResult &numerical_of(Context &ctx) {
  if (!ctx.has_result) fail(kStaleContext, "numerical_result", "no solved result is available");
  if (!ctx.result.has_numerical) fail(kUnsupported, "numerical_result", "result has no numerical V1 attachment");
  return ctx.result.numerical;
}

int32_t vckss_rust_engine_numerical_result_v1(uint64_t generation, VckssNumericalResultV1 *output,
                                              uint32_t output_capacity_bytes) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    require_capacity<VckssNumericalResultV1>(output, output_capacity_bytes, "numerical result");
    const NumericalResult &n = numerical_of(ctx);
    VckssNumericalResultV1 o;
    std::memset(&o, 0, sizeof o);
    o.struct_size = sizeof o;
    o.schema_version = 1;
    o.generation = generation;
    o.status = n.status;
    o.engine = VCKSS_ENGINE_COMPRESSED;
    o.leverage_probes = n.leverage_probes;
    o.target_probes = n.target_probes;
    o.target_fold_a = n.fold_a;
    o.target_fold_b = n.fold_b;
    o.certified_replay_rhs = n.certified_replay_rhs;
    o.executed_replay_rhs = n.certified_replay_rhs;
    o.attempted_replay_rhs = n.attempted_replay_rhs;
    o.replay_generator_words = n.replay_generator_words;
    o.allocation_bound_bytes = n.allocation_bytes;
    o.failed_replay_probe = n.failed_replay_probe;
    o.replay_error_code = n.replay_error_code;
    o.psd_adjustment = n.psd_adjustment;
    o.minimum_constrained = n.minimum_constrained;
    o.minimum_residual_margin = n.minimum_residual_margin;
    o.maximum_sensitivity_ratio = n.maximum_sensitivity_ratio;
    std::memcpy(o.score_means, n.score_means, sizeof o.score_means);
    std::memcpy(o.conditional, n.conditional, sizeof o.conditional);
    std::memcpy(o.leverage, n.leverage, sizeof o.leverage);
    std::memcpy(o.raw, n.raw, sizeof o.raw);
    std::memcpy(o.usable, n.usable, sizeof o.usable);
    std::memcpy(o.mcse, n.mcse, sizeof o.mcse);
    o.point_rhs = n.point_rhs;
    o.maximum_point_complete_residual = n.maximum_point_complete_residual;
    o.maximum_replay_complete_residual = n.maximum_replay_complete_residual;
    std::memcpy(output, &o, sizeof o);
  });
}

int32_t vckss_rust_engine_numerical_rhs_v1(uint64_t generation, VckssNumericalRhsV1 *output, uint64_t capacity,
                                           uint64_t *written) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (!written) fail(kAbiMismatch, "numerical_rhs", "numerical RHS count destination is null");
    Context &ctx = live(generation);
    const NumericalResult &n = numerical_of(ctx);
    const uint64_t count = n.replay_probes.size();
    if (capacity < count || (count && !output)) fail(kAbiMismatch, "numerical_rhs", "invalid or short numerical RHS destination");
    for (uint64_t i = 0; i < count; ++i) {
      VckssNumericalRhsV1 r;
      std::memset(&r, 0, sizeof r);
      r.schema_version = 1;
      r.phase = 1;  // leverage replay
      r.probe = n.replay_probes[i];
      r.complete_residual = n.replay_complete[i];
      std::memcpy(output + i, &r, sizeof r);
    }
    *written = count;
  });
}

int32_t vckss_rust_engine_numerical_cmg_work_v1(uint64_t generation, VckssFullCmgModelReceiptV1 *output,
                                                uint32_t output_capacity_bytes) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    require_capacity<VckssFullCmgModelReceiptV1>(output, output_capacity_bytes, "numerical CMG work");
    const NumericalResult &n = numerical_of(ctx);
    VckssFullCmgModelReceiptV1 o;
    std::memset(&o, 0, sizeof o);
    o.struct_size = sizeof o;
    o.schema_version = 1;
    o.generation = generation;
    o.nuisance_mode = VCKSS_NUISANCE_JOINT;
    o.logical_rhs_count = n.point_rhs + n.certified_replay_rhs;
    std::memcpy(output, &o, sizeof o);
  });
}

int32_t vckss_rust_engine_release_v1(uint64_t generation) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (generation == 0) fail(kStaleContext, "engine_release", "generation 0 is never live");
    if (g_ctx && g_ctx->generation == generation) {
      g_last_released = generation;
      g_ctx.reset();
      return;
    }
    if (generation <= g_last_released || generation < g_next_generation) return;  // idempotent
    fail(kStaleContext, "engine_release", "unknown generation");
  });
}

int32_t vckss_rust_engine_clear_abandoned_v1(void) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (g_ctx) {
      g_last_released = g_ctx->generation;
      g_ctx.reset();
    }
  });
}

int32_t vckss_rust_engine_snapshot_v1(VckssEngineSnapshotV1 *output, uint32_t output_capacity_bytes) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    require_capacity<VckssEngineSnapshotV1>(output, output_capacity_bytes, "snapshot");
    VckssEngineSnapshotV1 s;
    s.struct_size = sizeof s;
    s.state = g_ctx ? g_ctx->state : kNone;
    s.generation = g_ctx ? g_ctx->generation : 0;
    s.last_released_generation = g_last_released;
    std::memcpy(output, &s, sizeof s);
  });
}

// ------------------------------------------------------------------ extensions

int32_t vckss_mi355x_configure_v1(uint64_t generation, uint32_t precision, uint32_t leverage_width,
                                  uint32_t target_width) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (precision != 64 && precision != 32) fail(kInvalidInput, "configure", "precision must be 64 or 32");
    if (leverage_width < 1 || leverage_width > 64) fail(kInvalidInput, "configure", "leverage width must lie in [1, 64]");
    if (target_width < 1 || target_width > 32) fail(kInvalidInput, "configure", "target width must lie in [1, 32]");
    ctx.precision = precision;
    ctx.leverage_width = leverage_width;
    ctx.target_width = target_width;
  });
}

int32_t vckss_mi355x_dims_v1(uint64_t generation, uint64_t *dims, uint32_t capacity) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (!dims || capacity < 15) fail(kAbiMismatch, "dims", "capacity must be at least 15");
    const Prepared &P = ctx.P;
    uint64_t v[15] = {P.input_rows, P.rows, P.physical_total, P.workers, P.firms, P.cells, P.units,
                      P.exact_strata, P.strata, ctx.HS.graph.n, ctx.HS.graph.u.size(), ctx.HS.aux,
                      ctx.H.levels.size(), ctx.H.ldl_dim, P.topology_checksum};
    std::memcpy(dims, v, sizeof v);
  });
}

int32_t vckss_mi355x_array_v1(uint64_t generation, uint32_t which, double *out, uint64_t capacity) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    std::vector<double> tmp;
    const std::vector<double> *src = &tmp;
    const Prepared &P = ctx.P;
    if (which < 20) {
      if (!ctx.has_result) fail(kStaleContext, "array", "no solved result is available");
      const SolveResult &r = ctx.result;
      switch (which) {
        case 0: src = &r.fitted_cell; break;
        case 1: src = &r.projection_share; break;
        case 2: src = &r.residual_share; break;
        case 3: src = &r.finite_bias; break;
        case 4: src = &r.finite_variance; break;
        case 5: src = &r.residual_mass; break;
        case 6: src = &r.deleted_mass; break;
        case 7: src = &r.cell_weight; break;
        case 8: src = &r.draws; break;
        case 9: for (auto &c : r.leverage) tmp.push_back(c.iterations); break;
        case 10: for (auto &c : r.target) tmp.push_back(c.iterations); break;
        case 11: for (auto &c : r.leverage) tmp.push_back(c.complete); break;
        case 12: tmp = {r.seconds_total, r.seconds_leverage, r.seconds_target, (double)r.vcycles}; break;
        default: fail(kInvalidInput, "array", "unknown array");
      }
    } else {
      switch (which) {
        case 20: tmp.assign(P.unit_rank.begin(), P.unit_rank.end()); break;
        case 21: tmp.assign(P.unit_count.begin(), P.unit_count.end()); break;
        case 22: tmp.assign(P.stratum_rank.begin(), P.stratum_rank.end()); break;
        case 23: tmp.assign(P.stratum_count.begin(), P.stratum_count.end()); break;
        case 24: src = &P.cell_weight; break;
        case 25: tmp.assign(P.unit_cell.begin(), P.unit_cell.end()); break;
        case 26: tmp.assign(P.unit_offset.begin(), P.unit_offset.end()); tmp.resize(P.units, 0.0); break;
        case 27: tmp.assign(P.cell_units.begin(), P.cell_units.end()); break;
        case 28:
          for (const auto &L : ctx.H.levels) tmp.push_back((double)L.graph.n);
          break;
        default: fail(kInvalidInput, "array", "unknown array");
      }
    }
    if (!out || capacity < src->size()) fail(kAbiMismatch, "array", "capacity " + std::to_string(capacity) +
                                                                        " < " + std::to_string(src->size()));
    std::memcpy(out, src->data(), src->size() * sizeof(double));
  });
}

int32_t vckss_mi355x_atoms_v1(uint64_t generation, uint64_t seed, uint32_t domain, uint64_t first_probe,
                              uint32_t columns, int64_t *out, uint64_t capacity) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (domain > 1 || columns == 0) fail(kInvalidInput, "atoms", "invalid domain or column count");
    std::vector<int64_t> a;
    device_of(ctx).atoms(seed, domain ? 0x0054415247455401ULL : 0x4c45564552414745ULL, first_probe, columns,
                         domain == 1, a);
    if (!out || capacity < a.size()) fail(kAbiMismatch, "atoms", "capacity is too small");
    std::memcpy(out, a.data(), a.size() * sizeof(int64_t));
  });
}

int32_t vckss_mi355x_spmv_v1(uint64_t generation, uint32_t level, uint32_t columns, const double *x, double *y) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (level >= ctx.H.levels.size()) fail(kInvalidInput, "spmv", "level out of range");
    const uint64_t m = ctx.H.levels[level].graph.n * (uint64_t)columns;
    std::vector<double> in(x, x + m), outv;
    device_of(ctx).spmv(level, in, columns, outv);
    std::memcpy(y, outv.data(), m * sizeof(double));
  });
}

int32_t vckss_mi355x_spmv_noalloc_v1(uint64_t generation, uint32_t level, uint32_t columns, const double *x,
                                     double *y) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    // no device engine is created here: that would allocate
    if (!ctx.dev)
      fail(kInvalidInput, "spmv_noalloc",
           "the generation has no device engine; prepare it with vckss_rust_engine_prepare_interrupt_v4");
    ctx.dev->spmv_noalloc(level, columns, x, y);
  });
}

int32_t vckss_mi355x_vcycle_v1(uint64_t generation, uint32_t columns, uint32_t precision, const double *r, double *z) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    const uint64_t m = ctx.H.levels[0].graph.n * (uint64_t)columns;
    std::vector<double> in(r, r + m), outv;
    device_of(ctx).vcycle(in, columns, (int)precision, outv);
    std::memcpy(z, outv.data(), m * sizeof(double));
  });
}

int32_t vckss_mi355x_pcg_v1(uint64_t generation, uint32_t columns, double tolerance, uint32_t precision,
                            const double *rhs, double *x, uint32_t *iterations) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    const uint64_t m = ctx.H.levels[0].graph.n * (uint64_t)columns;
    std::vector<double> in(rhs, rhs + m), outv;
    std::vector<ColumnReceipt> rc;
    device_of(ctx).pcg(in, columns, tolerance, (int)precision, outv, rc);
    std::memcpy(x, outv.data(), m * sizeof(double));
    for (uint32_t j = 0; j < columns; ++j) iterations[j] = rc[j].iterations;
  });
}

int32_t vckss_mi355x_kernel_timing_v1(uint64_t generation, uint32_t which, double *milliseconds, double *bytes,
                                      uint64_t *launches) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (!milliseconds || !bytes || !launches) fail(kInvalidInput, "kernel_timing", "null output pointer");
    const auto t = device_of(ctx).kernel_timing((int)which);
    *milliseconds = t.ms_per_launch;
    *bytes = t.bytes_per_launch;
    *launches = t.launches;
  });
}

int32_t vckss_mi355x_first_pass_v1(uint64_t generation, uint32_t mode, uint32_t *previous) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (mode > 2) fail(kInvalidInput, "first_pass", "mode must be 0 (reference stop), 1 (gate-aware) or 2 (query)");
    DeviceEngine &d = device_of(ctx);
    if (previous) *previous = d.first_pass_cap() ? 1u : 0u;
    if (mode < 2) d.set_first_pass_cap(mode == 1);
  });
}

int32_t vckss_mi355x_solve_columns_v1(uint64_t generation, uint32_t columns, double tolerance, double inner_ratio,
                                      uint32_t precision, uint32_t maximum_iterations, uint32_t replacement_interval,
                                      const double *worker_rhs, const double *firm_rhs, double *worker_out,
                                      double *firm_out, double *receipts, uint64_t *refine) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (!worker_rhs || !firm_rhs || !worker_out || !firm_out || !receipts || !refine)
      fail(kInvalidInput, "cmg_full_v2", "null input or output pointer");
    if (!std::isfinite(tolerance) || tolerance < 1e-15 || tolerance > 1e-4 || !(inner_ratio > 0.0) ||
        inner_ratio > 1.0)
      fail(kInvalidInput, "cmg_full_v2", "tolerance must lie in [1e-15, 1e-4] and the inner ratio in (0, 1]");
    if (precision != 32 && precision != 64) fail(kInvalidInput, "cmg_full_v2", "precision must be 32 or 64");
    if (maximum_iterations == 0 || replacement_interval == 0)
      fail(kInvalidInput, "pcg_options", "iteration limit and replacement interval must be positive");
    const uint64_t W = ctx.P.workers, F = ctx.P.firms;
    std::vector<double> wr(worker_rhs, worker_rhs + W * (uint64_t)columns),
        fr(firm_rhs, firm_rhs + F * (uint64_t)columns), wo, fo;
    SolveOptions o;
    o.precision = (int)precision;
    o.max_iterations = maximum_iterations;
    o.recompute_interval = replacement_interval;
    std::vector<ColumnReceipt> rc;
    uint64_t rf[2] = {0, 0};
    device_of(ctx).solve_batch(wr, fr, columns, tolerance, inner_ratio, o, wo, fo, rc, rf);
    std::memcpy(worker_out, wo.data(), wo.size() * sizeof(double));
    std::memcpy(firm_out, fo.data(), fo.size() * sizeof(double));
    for (uint32_t j = 0; j < columns; ++j) {
      const double v[5] = {(double)rc[j].iterations, (double)rc[j].restarts, rc[j].reduced, rc[j].complete,
                           rc[j].rhs_norm};
      std::memcpy(receipts + 5 * (uint64_t)j, v, sizeof v);
    }
    refine[0] = rf[0];
    refine[1] = rf[1];
  });
}

int32_t vckss_mi355x_time_leverage_v1(uint64_t generation, uint32_t width, uint32_t precision, double tolerance,
                                      uint32_t batches, double *seconds, uint64_t *solves, uint64_t *iterations,
                                      double *kernel_ms, double *kernel_bytes, uint64_t *kernel_launches) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    SolveOptions o;
    o.leverage_width = width;
    o.precision = (int)precision;
    o.probe_tolerance = tolerance;
    DeviceEngine &d = device_of(ctx);
    *seconds = d.time_leverage_batches(o, batches, *solves, *iterations);
    auto t = d.kernel_timing(0);
    *kernel_ms = t.ms_per_launch;
    *kernel_bytes = t.bytes_per_launch;
    *kernel_launches = t.launches;
  });
}

int32_t vckss_mi355x_autotune_v1(uint64_t generation, uint32_t precision, double tolerance, uint32_t batches,
                                 uint32_t *width, uint32_t *chunk, double *rate, double *trials,
                                 uint32_t trial_capacity) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (precision != 64 && precision != 32) fail(kInvalidInput, "autotune", "precision must be 64 or 32");
    if (!width || !chunk || !rate || batches == 0) fail(kInvalidInput, "autotune", "invalid arguments");
    SolveOptions o;
    o.precision = (int)precision;
    o.probe_tolerance = tolerance;
    auto t = device_of(ctx).autotune(o, batches);
    *width = t.width;
    *chunk = t.chunk;
    *rate = t.rate;
    if (trials)
      for (size_t i = 0; i < t.trials.size() && i < trial_capacity; ++i) {
        trials[3 * i] = t.trials[i].width;
        trials[3 * i + 1] = t.trials[i].chunk;
        trials[3 * i + 2] = t.trials[i].rate;
      }
    ctx.precision = precision;
    ctx.leverage_width = t.width;
    ctx.target_width = std::max<uint32_t>(1, std::min<uint32_t>(t.width / 2, 32));
  });
}

int32_t vckss_mi355x_set_chunk_v1(uint64_t generation, uint32_t chunk) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (chunk < 16 || chunk > 1024) fail(kInvalidInput, "set_chunk", "chunk must lie in [16, 1024]");
    device_of(ctx).set_chunk(chunk);
  });
}

int32_t vckss_mi355x_smoother_v1(uint64_t generation, uint32_t mode, const double *params, uint32_t count) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (mode > 1) fail(kInvalidInput, "smoother", "mode must be 0 (plain) or 1 (learned)");
    std::vector<double> theta;
    if (mode == 1) {
      if (params) {
        theta.assign(params, params + count);
      } else {
        if (kBuiltinSmootherCount != kSmootherParams)
          fail(kUnsupported, "smoother", "the library carries no trained smoother weights for this network");
        theta = builtin_smoother();
      }
    }
    device_of(ctx).smoother_mode((int)mode, theta, (int)ctx.precision);
  });
}

int32_t vckss_mi355x_smoother_force_v1(uint64_t generation, double omega) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] { device_of(live(generation)).smoother_force(omega); });
}

int32_t vckss_mi355x_smoother_status_v1(uint64_t generation, uint32_t *out, uint32_t capacity) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (!out || capacity < 10) fail(kAbiMismatch, "smoother_status", "capacity must be at least 10");
    const auto s = device_of(ctx).smoother_status();
    uint32_t v[10] = {(uint32_t)s.mode, s.active ? 1u : 0u, s.breakdown ? 1u : 0u, s.plain_iterations,
                      s.learned_iterations, s.shaped ? 1u : 0u, UINT32_MAX, UINT32_MAX, UINT32_MAX, UINT32_MAX};
    for (size_t k = 0; k < s.calibration.size() && k < 4; ++k) v[6 + k] = s.calibration[k];
    std::memcpy(out, v, sizeof v);
  });
}

int32_t vckss_mi355x_graphs_v1(uint64_t generation, uint32_t enable, uint64_t *captures, uint64_t *launches) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (enable > 2) fail(kInvalidInput, "graphs", "enable must be 0, 1 or 2 (query)");
    DeviceEngine &dev = device_of(ctx);
    if (enable < 2) dev.set_graphs(enable == 1);
    uint64_t c = 0, l = 0;
    dev.graph_counters(c, l);
    if (captures) *captures = c;
    if (launches) *launches = l;
  });
}

int32_t vckss_mi355x_exact_receipt_v1(uint64_t generation, double *out, uint32_t capacity) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (!out || capacity < 13) fail(kAbiMismatch, "exact_receipt", "capacity must be at least 13");
    if (!ctx.has_result) fail(kStaleContext, "exact_receipt", "no solved result is available");
    if (!ctx.result.has_exact) fail(kUnsupported, "exact_receipt", "the result is not an exact estimate");
    const ExactReceipt &x = ctx.result.exact;
    const double v[13] = {(double)x.parameters, (double)x.deletion_units, (double)x.cells, x.max_leverage,
                          x.deletion_rank_gap, x.information_rcond, x.inverse_relres, x.fit_relres,
                          x.fit_tolerance, x.firm_zero_sum_residual, ctx.result.weighted_rss, x.seconds_factor,
                          x.seconds_cells};
    std::memcpy(out, v, sizeof v);
  });
}

int32_t vckss_mi355x_batch_prepare_v1(uint64_t rows, const double *sample, const double *worker,
                                      const double *firm, const double *deletion, const double *outcome,
                                      const double *frequency, const double *target, uint32_t deletion_mode,
                                      uint32_t threads, uint64_t *handle, uint32_t *samples) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (!handle || !samples) fail(kInvalidInput, "batch_prepare", "handle and sample-count outputs are required");
    if (rows == 0) fail(kInvalidInput, "batch_prepare", "row count must be positive");
    if (deletion_mode != VCKSS_DELETION_MATCH && deletion_mode != VCKSS_DELETION_OBSERVATION)
      fail(kInvalidInput, "batch_prepare", "unknown deletion mode");
    if (!outcome || !target) fail(kInvalidInput, "engine_ingest", "outcome or target column is null");
    auto sid = integer_column(sample, rows, "sample identifier", kInvalidIdentifier, true);
    auto w = integer_column(worker, rows, "worker identifier", kInvalidIdentifier, true);
    auto f = integer_column(firm, rows, "firm identifier", kInvalidIdentifier, true);
    auto d = integer_column(deletion, rows, "deletion identifier", kInvalidIdentifier, true);
    auto fr = integer_column(frequency, rows, "frequency weight", kInvalidWeight, false);
    // rows of each sample, in input order; samples ascend by identifier
    std::vector<uint64_t> order(rows);
    std::iota(order.begin(), order.end(), 0ull);
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return sid[a] < sid[b]; });
    auto batch = std::make_unique<Batch>();
    batch->observation = deletion_mode == VCKSS_DELETION_OBSERVATION;
    std::vector<std::pair<uint64_t, uint64_t>> ranges;
    for (uint64_t i = 0; i < rows;) {
      uint64_t j = i;
      while (j < rows && sid[order[j]] == sid[order[i]]) ++j;
      ranges.push_back({i, j});
      auto bs = std::make_unique<BatchSample>();
      bs->id = sid[order[i]];
      batch->samples.push_back(std::move(bs));
      i = j;
    }
    std::vector<size_t> idx(ranges.size());
    std::iota(idx.begin(), idx.end(), (size_t)0);
    std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
      return ranges[a].second - ranges[a].first > ranges[b].second - ranges[b].first;
    });
    parallel_for(idx, std::max<uint32_t>(threads, 1), [&](size_t k) {
      const uint64_t b = ranges[k].first, e = ranges[k].second, n = e - b;
      std::vector<uint64_t> cw(n), cf(n), cd(n), cfr(n);
      std::vector<double> y(n), t(n);
      for (uint64_t i = 0; i < n; ++i) {
        const uint64_t r = order[b + i];
        cw[i] = w[r]; cf[i] = f[r]; cd[i] = d[r]; cfr[i] = fr[r];
        y[i] = outcome[r]; t[i] = target[r];
      }
      Columns cols{n, cw.data(), cf.data(), cd.data(), y.data(), cfr.data(), t.data()};
      BatchSample &bs = *batch->samples[k];
      try {
        bs.P = prepare(cols, nullptr, batch->observation);
        bs.HS = build_hybrid(bs.P);
        CmgOptions copt;
        bs.H = build_hierarchy(bs.HS.graph, copt);
      } catch (const Error &err) {
        fail(err.code, "batch_prepare", "sample " + std::to_string(k) + ": " + err.what());
      }
    });
    for (auto &bs : batch->samples) {
      bs->dev.reset(new DeviceEngine(bs->P, bs->HS, bs->H));
      // as for a single problem: the built-in trained smoother weights and
      // cycle shape, kept for a sample only where its calibration says they help
      if (kBuiltinSmootherCount == kSmootherParams) bs->dev->smoother_default(builtin_smoother());
    }
    *samples = (uint32_t)batch->samples.size();
    *handle = g_next_batch++;
    g_batches[*handle] = std::move(batch);
  });
}

int32_t vckss_mi355x_batch_solve_v1(uint64_t handle, uint64_t seed, uint32_t probes, double tolerance,
                                    double fit_tolerance, uint32_t precision, uint32_t streams) {
  return vckss_mi355x_batch_solve_v2(handle, seed, probes, tolerance, fit_tolerance, precision, streams, 0);
}

int32_t vckss_mi355x_batch_solve_v2(uint64_t handle, uint64_t seed, uint32_t probes, double tolerance,
                                    double fit_tolerance, uint32_t precision, uint32_t streams, uint32_t mode) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Batch &B = live_batch(handle);
    if (mode > 1) fail(kInvalidInput, "batch_solve", "mode must be 0 (streams) or 1 (lockstep)");
    if (probes < 2) fail(kInvalidInput, "jla_validate", "JLA probe count must be at least two");
    if (precision != 32 && precision != 64) fail(kInvalidInput, "batch_solve", "precision must be 32 or 64");
    for (double t : {tolerance, fit_tolerance})
      if (!std::isfinite(t) || t < 1e-15 || t > 1e-4) fail(kInvalidInput, "cmg_full_v2", "tolerance must lie in [1e-15, 1e-4]");
    SolveOptions o;
    o.seed = seed;
    o.probes = probes;
    o.probe_tolerance = tolerance;
    o.fit_tolerance = fit_tolerance;
    o.precision = (int)precision;
    o.leverage_width = std::min<uint32_t>(64, probes);
    o.target_width = std::min<uint32_t>(32, probes);
    std::vector<size_t> idx(B.samples.size());
    std::iota(idx.begin(), idx.end(), (size_t)0);
    std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return B.samples[a]->P.rows > B.samples[b]->P.rows; });
    for (auto &bs : B.samples) {
      bs->solved = false;
      bs->status = 0;
      bs->error.clear();
    }
    // lockstep: the sample threads' PCG calls meet in one batch (a thread
    // joins before its sample and leaves after it, also on failure, so the
    // others never wait for it)
    std::shared_ptr<Lockstep> ls = mode == 1 ? make_lockstep() : nullptr;
    struct Member {
      Lockstep *l;
      explicit Member(Lockstep *x) : l(x) {
        if (l) lockstep_join(*l);
      }
      ~Member() {
        if (l) lockstep_leave(*l);
      }
    };
    for (auto &bs : B.samples) bs->dev->set_lockstep(ls.get());
    try {
      parallel_for(idx, std::max<uint32_t>(streams, 1), [&](size_t k) {
        BatchSample &bs = *B.samples[k];
        const Member member(ls.get());
        const auto t0 = std::chrono::steady_clock::now();
        try {
          bs.result = bs.dev->run(o);
          bs.solved = true;
        } catch (const Error &err) {
          bs.status = err.code;
          bs.error = err.what();
        }
        bs.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      });
    } catch (...) {
      for (auto &bs : B.samples) bs->dev->set_lockstep(nullptr);
      throw;
    }
    for (auto &bs : B.samples) bs->dev->set_lockstep(nullptr);
    for (size_t k = 0; k < B.samples.size(); ++k)
      if (B.samples[k]->status)
        fail(B.samples[k]->status, "batch_solve", "sample " + std::to_string(k) + ": " + B.samples[k]->error);
  });
}

int32_t vckss_mi355x_batch_result_v1(uint64_t handle, uint32_t sample, double *out, uint32_t capacity) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Batch &B = live_batch(handle);
    if (sample >= B.samples.size()) fail(kInvalidInput, "batch_result", "sample index out of range");
    if (!out || capacity < 24) fail(kAbiMismatch, "batch_result", "capacity must be at least 24");
    const BatchSample &bs = *B.samples[sample];
    if (!bs.solved) fail(bs.status ? bs.status : kStaleContext, "batch_result", "the sample has no solved result");
    const SolveResult &r = bs.result;
    for (int j = 0; j < 4; ++j) {
      out[j] = r.plugin[j];
      out[4 + j] = r.correction[j];
      out[8 + j] = r.corrected[j];
      out[12 + j] = r.mcse[j];
    }
    uint64_t it = 0;
    for (const auto *set : {&r.leverage, &r.target})
      for (const ColumnReceipt &c : *set) it += c.iterations;
    // identifiers are held order-encoded ((i64) ^ 2^63), as the column ids are
    const double extra[8] = {(double)(int64_t)(bs.id ^ (1ULL << 63)), (double)bs.P.rows, (double)bs.P.workers, (double)bs.P.firms,
                             (double)bs.P.units, bs.seconds, (double)it, r.max_complete_residual};
    std::memcpy(out + 16, extra, sizeof extra);
  });
}

int32_t vckss_mi355x_batch_release_v1(uint64_t handle) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] { g_batches.erase(handle); });
}

int32_t vckss_mi355x_controls_v1(uint64_t generation, double *out, uint32_t capacity) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (!ctx.has_result) fail(kStaleContext, "controls_receipt", "no solved result is available");
    const SolveResult &r = ctx.result;
    if (!r.controls) fail(kUnsupported, "controls_receipt", "the result has no controls");
    if (!out || capacity < 3 + r.controls) fail(kAbiMismatch, "controls_receipt", "capacity must be at least 3 + q");
    out[0] = r.controls;
    out[1] = r.fixed_offset ? 1.0 : 0.0;
    out[2] = r.control_rcond;
    for (uint32_t j = 0; j < r.controls; ++j) out[3 + j] = r.control_beta[j];
  });
}

int32_t vckss_mi355x_cycle_shape_v1(uint64_t generation, uint32_t *levels, uint32_t capacity, uint32_t *count) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (!count) fail(kInvalidInput, "cycle_shape", "count is required");
    const auto s = device_of(ctx).smoother_status();
    std::vector<uint32_t> shape = s.shape;
    if (!s.shaped) {
      shape.resize(ctx.H.levels.size());
      for (uint32_t l = 0; l < shape.size(); ++l) shape[l] = l;
    }
    *count = (uint32_t)shape.size();
    if (levels) {
      if (capacity < shape.size()) fail(kAbiMismatch, "cycle_shape", "capacity too small");
      std::memcpy(levels, shape.data(), shape.size() * sizeof(uint32_t));
    }
  });
}

int32_t vckss_mi355x_smoother_params_v1(double *out, uint32_t capacity, uint32_t *count) {
  return guarded([&] {
    if (!count) fail(kInvalidInput, "smoother_params", "count is required");
    // parameters of another network layout are not the library's weights
    *count = kBuiltinSmootherCount == kSmootherParams ? (uint32_t)kBuiltinSmootherCount : 0u;
    if (out && *count) {
      if (capacity < kBuiltinSmootherCount) fail(kAbiMismatch, "smoother_params", "capacity too small");
      std::memcpy(out, kBuiltinSmoother, kBuiltinSmootherCount * sizeof(double));
    }
  });
}

static int32_t smoother_omega_entry(uint64_t generation, const double *params, uint32_t count, double *omega,
                                    uint64_t capacity, uint64_t *written, bool bf16) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (!params || !written) fail(kInvalidInput, "smoother_omega", "params and written are required");
    std::vector<double> w;
    device_of(ctx).smoother_omega(std::vector<double>(params, params + count), w, bf16);
    *written = w.size();
    if (omega) {
      if (capacity < w.size()) fail(kAbiMismatch, "smoother_omega", "capacity too small");
      std::memcpy(omega, w.data(), w.size() * sizeof(double));
    }
  });
}

int32_t vckss_mi355x_smoother_omega_v1(uint64_t generation, const double *params, uint32_t count, double *omega,
                                       uint64_t capacity, uint64_t *written) {
  return smoother_omega_entry(generation, params, count, omega, capacity, written, false);
}

int32_t vckss_mi355x_smoother_infer_v1(uint64_t generation, const double *params, uint32_t count, double *omega,
                                       uint64_t capacity, uint64_t *written) {
  return smoother_omega_entry(generation, params, count, omega, capacity, written, true);
}

int32_t vckss_mi355x_smoother_gradient_v1(uint64_t generation, const double *params, uint32_t count,
                                          const double *errors, uint64_t rows, double *loss, double *grad) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (!params || !errors || !loss || !grad) fail(kInvalidInput, "smoother_gradient", "null argument");
    if (rows != ctx.HS.graph.n) fail(kInvalidInput, "smoother_gradient", "errors must have one row per vertex");
    std::vector<double> g;
    *loss = device_of(ctx).smoother_gradient(std::vector<double>(params, params + count),
                                             std::vector<double>(errors, errors + rows * 32), g);
    std::memcpy(grad, g.data(), g.size() * sizeof(double));
  });
}

int32_t vckss_mi355x_smoother_train_v1(uint64_t generation, double *state, uint32_t count, uint64_t step0,
                                       uint32_t steps, double learning_rate, uint64_t seed, uint32_t passes,
                                       double *losses) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (!state || (steps && !losses)) fail(kInvalidInput, "smoother_train", "null argument");
    std::vector<double> s(state, state + count), l;
    device_of(ctx).smoother_train(s, step0, steps, learning_rate, seed, passes, l);
    std::memcpy(state, s.data(), s.size() * sizeof(double));
    if (steps) std::memcpy(losses, l.data(), l.size() * sizeof(double));
  });
}

int32_t vckss_mi355x_counters_v1(uint64_t generation, uint64_t *out, uint32_t capacity) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (!out || capacity < 6) fail(kAbiMismatch, "counters", "capacity must be at least 6");
    device_of(ctx).counters(out);
  });
}

int32_t vckss_mi355x_stream_copy_v1(uint64_t generation, uint64_t bytes, double *milliseconds) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (!milliseconds || bytes == 0) fail(kInvalidInput, "stream_copy", "invalid arguments");
    *milliseconds = device_of(ctx).stream_copy(bytes);
  });
}

int32_t vckss_mi355x_hierarchy_build_v1(uint64_t vertices, uint64_t edges, const uint32_t *u, const uint32_t *v,
                                        const double *w, uint64_t direct_threshold, uint64_t max_levels,
                                        uint64_t *handle) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (!handle || !u || !v || !w) fail(kInvalidInput, "hierarchy", "null pointer");
    CmgOptions o;
    o.direct_threshold = direct_threshold;
    o.max_levels = max_levels;
    if (!o.direct_threshold || !o.max_levels) fail(kInvalidInput, "hierarchy", "invalid options");
    Laplacian g = laplacian_from_edges(vertices, edges, u, v, w);
    std::unique_ptr<Hierarchy> h(new Hierarchy(build_hierarchy(g, o)));
    *handle = g_next_hierarchy++;
    g_hierarchies[*handle] = std::move(h);
  });
}

int32_t vckss_mi355x_hierarchy_levels_v1(uint64_t handle, uint32_t *levels, uint64_t *factor_nnz) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    auto it = g_hierarchies.find(handle);
    if (it == g_hierarchies.end()) fail(kStaleContext, "hierarchy", "unknown handle");
    *levels = (uint32_t)it->second->levels.size();
    *factor_nnz = it->second->factor_nnz;
  });
}

int32_t vckss_mi355x_hierarchy_level_v1(uint64_t handle, uint32_t level, uint64_t *info, uint32_t *labels,
                                        uint32_t *eu, uint32_t *ev, double *ew, double *diag) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    auto it = g_hierarchies.find(handle);
    if (it == g_hierarchies.end()) fail(kStaleContext, "hierarchy", "unknown handle");
    const Hierarchy &H = *it->second;
    if (level >= H.levels.size()) fail(kInvalidInput, "hierarchy", "level out of range");
    const Level &L = H.levels[level];
    uint64_t v6[6] = {L.graph.n, L.graph.u.size(), L.graph.nnz, L.repeat, (uint64_t)L.terminal, L.coarse};
    std::memcpy(info, v6, sizeof v6);
    if (labels && !L.labels.empty()) std::memcpy(labels, L.labels.data(), L.labels.size() * 4);
    if (eu) std::memcpy(eu, L.graph.u.data(), L.graph.u.size() * 4);
    if (ev) std::memcpy(ev, L.graph.v.data(), L.graph.v.size() * 4);
    if (ew) std::memcpy(ew, L.graph.w.data(), L.graph.w.size() * 8);
    if (diag) std::memcpy(diag, L.graph.diag.data(), L.graph.diag.size() * 8);
  });
}

int32_t vckss_mi355x_layout_v1(uint64_t handle, uint32_t level, uint32_t chunk, uint64_t *sizes, uint32_t *perm,
                               uint32_t *row_ptr, uint32_t *col, double *val, double *inv, uint32_t *lab,
                               uint32_t *agg_ptr, uint32_t *row_chunks, uint32_t *row_multi, uint32_t *agg_chunks,
                               uint32_t *agg_multi, double *terminal_inverse) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    auto it = g_hierarchies.find(handle);
    if (it == g_hierarchies.end()) fail(kStaleContext, "layout", "unknown handle");
    const Hierarchy &H = *it->second;
    if (level >= H.levels.size()) fail(kInvalidInput, "layout", "level out of range");
    if (!sizes) fail(kInvalidInput, "layout", "sizes pointer is null");
    const std::vector<LevelLayout> all = device_layout(H, chunk);
    const LevelLayout &L = all[level];
    const uint64_t s[8] = {L.perm.size(), L.col.size(), L.agg_ptr.empty() ? 0 : L.agg_ptr.size() - 1,
                           L.rows.chunks(), L.rows.multi(), L.aggs.chunks(), L.aggs.multi(),
                           level + 1 == H.levels.size() && H.terminal_reason == kDirect ? L.perm.size() : 0};
    std::memcpy(sizes, s, sizeof s);
    auto put = [](auto *dst, const auto &src) {
      if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
    };
    put(perm, L.perm); put(row_ptr, L.row_ptr); put(col, L.col); put(val, L.val); put(inv, L.inv);
    put(lab, L.lab); put(agg_ptr, L.agg_ptr); put(row_chunks, L.rows.desc); put(row_multi, L.rows.mrange);
    put(agg_chunks, L.aggs.desc); put(agg_multi, L.aggs.mrange);
    if (s[7] && H.terminal_inverse.size() != s[7] * s[7]) fail(kInvariant, "layout", "terminal inverse size");
    if (s[7]) put(terminal_inverse, H.terminal_inverse);
  });
}

int32_t vckss_mi355x_chunk_table_v1(const uint32_t *ptr, uint64_t segments, uint32_t chunk, uint32_t group,
                                    uint64_t *sizes, uint32_t *desc, uint32_t *mrange) {
  return guarded([&] {
    if (!sizes || (!ptr && segments)) fail(kInvalidInput, "chunk_table", "null pointer");
    std::vector<uint32_t> p(segments + 1, 0);
    if (segments) std::memcpy(p.data(), ptr, p.size() * 4);
    for (uint64_t s = 0; s < segments; ++s)
      if (p[s + 1] < p[s]) fail(kInvalidInput, "chunk_table", "segment pointers must not decrease");
    const ChunkTable t = chunk_segments(p, chunk, group);
    sizes[0] = t.chunks();
    sizes[1] = t.multi();
    if (desc && !t.desc.empty()) std::memcpy(desc, t.desc.data(), t.desc.size() * 4);
    if (mrange && !t.mrange.empty()) std::memcpy(mrange, t.mrange.data(), t.mrange.size() * 4);
  });
}

int32_t vckss_mi355x_hierarchy_free_v1(uint64_t handle) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] { g_hierarchies.erase(handle); });
}

// ---- backend and request capabilities (ffi_engine.rs:2501-2720, vckss_rust.h:1647-1665)
int32_t vckss_rust_backend_capabilities_v1(VckssBackendCapabilitiesV1 *output, uint32_t output_capacity_bytes) {
  return guarded([&] {
    require_capacity<VckssBackendCapabilitiesV1>(output, output_capacity_bytes, "backend capabilities");
    VckssBackendCapabilitiesV1 o;
    std::memset(&o, 0, sizeof o);
    o.struct_size = sizeof o;
    o.abi_version = VCKSS_RUST_ABI_VERSION_V1;
    // what this library implements: match and observation graphs, the dense
    // exact estimator, batched PCG on the CMG hybrid graph with Counter-V1
    // probes (CMG_FULL_V2 with its point routing and model receipt), movers
    // and the stayer hybrid
    o.core_ready_flags = VCKSS_CORE_MATCH_GRAPH_READY | VCKSS_CORE_EXACT_READY | VCKSS_CORE_BATCHED_PCG_READY |
                         VCKSS_CORE_CMG_GRAPH_READY | VCKSS_CORE_COUNTER_RNG_READY | VCKSS_CORE_JLA_PLAN_READY |
                         VCKSS_CORE_FULL_CMG_V2_READY | VCKSS_CORE_FULL_CMG_POINT_ROUTING_V1_READY |
                         VCKSS_CORE_FULL_CMG_MODEL_V1_READY | VCKSS_CORE_DELETION_UNIT_MOVERS_V1_READY;
    o.support_flags = VCKSS_SUPPORT_EXACT | VCKSS_SUPPORT_JLA | VCKSS_SUPPORT_MATCH_DELETION |
                      VCKSS_SUPPORT_OBSERVATION_DELETION | VCKSS_SUPPORT_CMG;
    o.deterministic_parallelism = 1;
    std::memcpy(output, &o, sizeof o);
  });
}

int32_t vckss_rust_backend_request_capability_v1(const VckssBackendRequestCapabilityRequestV1 *request,
                                                 VckssBackendRequestCapabilityReceiptV1 *output,
                                                 uint32_t output_capacity_bytes) {
  return guarded([&] {
    require_capacity<VckssBackendRequestCapabilityReceiptV1>(output, output_capacity_bytes,
                                                             "request capability receipt");
    if (!request) fail(kInvalidInput, "engine_abi", "request capability pointer is null");
    VckssBackendRequestCapabilityRequestV1 r;
    std::memcpy(&r, request, sizeof r);
    if (r.abi_version != VCKSS_RUST_ABI_VERSION_V1) fail(kAbiMismatch, "engine_abi", "unsupported ABI version");
    if (r.reserved != 0) fail(kAbiMismatch, "engine_abi", "reserved request capability fields must be zero");
    const auto c = classify_v1(r);
    VckssBackendRequestCapabilityReceiptV1 o;
    std::memset(&o, 0, sizeof o);
    o.struct_size = sizeof o;
    o.abi_version = VCKSS_RUST_ABI_VERSION_V1;
    o.request_schema = r.request_schema;
    o.supported = c.first == VCKSS_REQUEST_REASON_SUPPORTED;
    o.reason_code = c.first;
    o.profile_code = c.second;
    o.algorithm = r.algorithm; o.deletion_mode = r.deletion_mode; o.nuisance_mode = r.nuisance_mode;
    o.solver_route = r.solver_route; o.rng_contract = r.rng_contract; o.controls_count = r.controls_count;
    o.frequency_use = r.frequency_use;
    o.request_signature = signature_v1(r, "VCKSS-REQUEST-CAPABILITY-V1");
    std::memcpy(output, &o, sizeof o);
  });
}

static VckssBackendRequestCapabilityReceiptV2 capability_receipt_v2(const VckssBackendRequestCapabilityRequestV2 &r,
                                                                    std::pair<uint32_t, uint32_t> c, uint64_t sig,
                                                                    uint32_t size) {
  VckssBackendRequestCapabilityReceiptV2 o;
  std::memset(&o, 0, sizeof o);
  o.v1.struct_size = size;
  o.v1.abi_version = VCKSS_RUST_ABI_VERSION_V1;
  o.v1.request_schema = r.v1.request_schema;
  o.v1.supported = c.first == VCKSS_REQUEST_REASON_SUPPORTED;
  o.v1.reason_code = c.first;
  o.v1.profile_code = c.second;
  o.v1.algorithm = r.v1.algorithm; o.v1.deletion_mode = r.v1.deletion_mode; o.v1.nuisance_mode = r.v1.nuisance_mode;
  o.v1.solver_route = r.v1.solver_route; o.v1.rng_contract = r.v1.rng_contract;
  o.v1.controls_count = r.v1.controls_count; o.v1.frequency_use = r.v1.frequency_use;
  o.v1.request_signature = sig;
  o.engine = r.engine; o.batch_mode = r.batch_mode; o.stayers_mode = r.stayers_mode;
  o.target_weight_mode = r.target_weight_mode; o.deletion_unit_source = r.deletion_unit_source;
  o.probeorder_supplied = r.probeorder_supplied; o.wallseconds_supplied = r.wallseconds_supplied;
  o.physical_limit = r.physical_limit;
  return o;
}

int32_t vckss_rust_backend_request_capability_v2(const VckssBackendRequestCapabilityRequestV2 *request,
                                                 VckssBackendRequestCapabilityReceiptV2 *output,
                                                 uint32_t output_capacity_bytes) {
  return guarded([&] {
    require_capacity<VckssBackendRequestCapabilityReceiptV2>(output, output_capacity_bytes,
                                                             "V2 request capability receipt");
    if (!request) fail(kInvalidInput, "engine_abi", "request capability pointer is null");
    VckssBackendRequestCapabilityRequestV2 r;
    std::memcpy(&r, request, sizeof r);
    if (r.v1.abi_version != VCKSS_RUST_ABI_VERSION_V1) fail(kAbiMismatch, "engine_abi", "unsupported ABI version");
    if (r.v1.struct_size < sizeof r) fail(kAbiMismatch, "engine_abi", "V2 request capability reports a short structure size");
    if (r.v1.reserved != 0 || r.reserved_2 != 0)
      fail(kAbiMismatch, "engine_abi", "reserved V2 request capability fields must be zero");
    const auto o = capability_receipt_v2(r, classify_v2(r), signature_v2(r), sizeof(VckssBackendRequestCapabilityReceiptV2));
    std::memcpy(output, &o, sizeof o);
  });
}

int32_t vckss_rust_backend_request_capability_v3(const VckssBackendRequestCapabilityRequestV3 *request,
                                                 VckssBackendRequestCapabilityReceiptV3 *output,
                                                 uint32_t output_capacity_bytes) {
  return guarded([&] {
    require_capacity<VckssBackendRequestCapabilityReceiptV3>(output, output_capacity_bytes,
                                                             "V3 request capability receipt");
    if (!request) fail(kInvalidInput, "engine_abi", "request capability pointer is null");
    VckssBackendRequestCapabilityRequestV3 r;
    std::memcpy(&r, request, sizeof r);
    if (r.v2.v1.abi_version != VCKSS_RUST_ABI_VERSION_V1) fail(kAbiMismatch, "engine_abi", "unsupported ABI version");
    if (r.v2.v1.struct_size < sizeof r)
      fail(kAbiMismatch, "engine_abi", "V3 request capability reports a short structure size");
    if (r.v2.v1.reserved != 0 || r.v2.reserved_2 != 0 || r.reserved_3 != 0 || r.reserved_4 != 0)
      fail(kAbiMismatch, "engine_abi", "reserved V3 request capability fields must be zero");
    VckssBackendRequestCapabilityReceiptV3 o;
    std::memset(&o, 0, sizeof o);
    o.v2 = capability_receipt_v2(r.v2, classify_v3(r), signature_v3(r), sizeof(VckssBackendRequestCapabilityReceiptV3));
    o.leverage_batch_mode = r.leverage_batch_mode;
    o.target_batch_mode = r.target_batch_mode;
    o.allow_automatic_cmg_setup_fallback = r.allow_automatic_cmg_setup_fallback;
    o.wallseconds = r.wallseconds;
    // ffi_engine.rs:2657-2671: what is resolved only with the data
    o.algorithm_resolution_deferred = r.v2.v1.algorithm == VCKSS_ALGORITHM_AUTO;
    o.engine_resolution_deferred = r.v2.v1.algorithm == VCKSS_ALGORITHM_AUTO ||
                                   (r.v2.v1.algorithm == VCKSS_ALGORITHM_JLA &&
                                    r.v2.engine == VCKSS_ENGINE_AUTO_OR_UNSPECIFIED);
    o.route_resolution_deferred = r.v2.v1.algorithm != VCKSS_ALGORITHM_EXACT && r.v2.v1.solver_route == VCKSS_ROUTE_AUTO;
    o.leverage_batch_resolution_deferred = r.leverage_batch_mode == VCKSS_BATCH_MODE_AUTO;
    o.target_batch_resolution_deferred = r.target_batch_mode == VCKSS_BATCH_MODE_AUTO;
    o.wall_advisory_only = 1;
    std::memcpy(output, &o, sizeof o);
  });
}

// ---- admission and preparation receipt (vckss_rust.h:1765-1769, 1923-1927)
int32_t vckss_rust_engine_admit_prepare_v4(const VckssEnginePrepareRequestV4 *request, uint32_t probeorder_supplied) {
  return guarded([&] {
    if (!request) fail(kInvalidInput, "engine_prepare", "request pointer is null");
    VckssEnginePrepareRequestV4 r;
    std::memcpy(&r, request, sizeof r);
    (void)admit_prepare(r, probeorder_supplied);
  });
}

int32_t vckss_rust_engine_preparation_receipt_v4(uint64_t generation, VckssEnginePreparationReceiptV4 *output,
                                                 uint32_t output_capacity_bytes) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    require_capacity<VckssEnginePreparationReceiptV4>(output, output_capacity_bytes, "V4 preparation receipt");
    const Context &ctx = live(generation);
    const Prepared &P = ctx.P;
    const MemoryModel m = memory_model(ctx);
    VckssEnginePreparationReceiptV4 o;
    std::memset(&o, 0, sizeof o);
    VckssEnginePreparationReceiptV2 &d = o.v3.v2;
    d.struct_size = sizeof o;
    d.generation = generation;
    d.input_rows = P.input_rows;
    d.retained_rows = P.rows;
    d.workers = P.workers;
    d.firms = P.firms;
    d.cells = P.cells;
    d.deletion_units = P.units;
    d.target_strata = P.strata;
    d.memory_limit_bytes = ctx.memory_limit_bytes;
    d.caller_copy_bytes = ctx.caller_copy_bytes;
    d.preparation_peak_forecast_bytes = ctx.preparation_forecast;
    d.prepared_resident_bytes = m.host_problem + m.graph + m.hierarchy + m.plan;
    const SelectionReceipt &g = P.selection;
    d.graph_input_rows = g.input_rows;
    d.graph_retained_rows = g.retained_rows;
    d.graph_input_physical_mass = g.input_physical_mass;
    d.graph_retained_physical_mass = g.retained_physical_mass;
    d.graph_initial_components = g.initial_components;
    d.graph_maximum_components = g.maximum_components;
    d.graph_initial_component_rows = g.initial_component_rows;
    d.graph_mover_input_rows = g.mover_input_rows;
    d.graph_initial_deletion_edges = g.initial_deletion_edges;
    d.graph_retained_deletion_edges = g.retained_deletion_edges;
    d.graph_insufficient_workers_removed = g.insufficient_workers_removed;
    d.graph_articulation_workers_removed = g.articulation_workers_removed;
    d.graph_bridge_units_removed = g.bridge_units_removed;
    d.graph_bridge_rows_removed = g.bridge_rows_removed;
    d.graph_degree_iterations = g.degree_iterations;
    d.graph_articulation_iterations = g.articulation_iterations;
    d.graph_bridge_iterations = g.bridge_iterations;
    d.graph_fixed_point_iterations = g.fixed_point_iterations;
    o.v3.target_weight_sum = P.target_total;
    o.controls_count = P.controls;
    o.deletion_mode = P.observation ? VCKSS_DELETION_OBSERVATION : VCKSS_DELETION_MATCH;
    std::memcpy(output, &o, sizeof o);
  });
}

// ---- memory forecast (vckss_rust.h:2322-2332): the generation's host
// arrays and hierarchy plus the peak of the engine's device buffers, with a
// 20% admission allowance (expected <= admission, reserve = the difference)
int32_t vckss_rust_engine_memory_forecast_v1(uint64_t generation, VckssMemoryForecastV1 *output,
                                             uint32_t output_capacity_bytes) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    require_capacity<VckssMemoryForecastV1>(output, output_capacity_bytes, "memory forecast");
    const Context &ctx = live(generation);
    const MemoryModel m = memory_model(ctx);
    VckssMemoryForecastV1 o;
    std::memset(&o, 0, sizeof o);
    o.struct_size = sizeof o;
    o.schema_version = 1;
    // memory_api.rs:215-287: a prepared generation reports its preparation
    // forecast; a solved one the larger of that and the solve's peak, where
    // the expected peak counts the device high-water mark measured by the
    // allocator and the admission peak the modelled hierarchy, plan and
    // K = 64 workspace.
    const uint64_t host = m.host_problem + m.graph + m.hierarchy;
    const uint64_t prep = std::max(ctx.preparation_forecast, ctx.stayer_forecast);
    uint64_t expected = prep, admission = prep;
    if (ctx.has_result) {
      expected = std::max(expected, host + m.device_peak);
      admission = std::max(expected, host + m.hierarchy + m.plan + m.workspace);
    }
    o.expected_peak_bytes = expected;
    o.admission_peak_bytes = admission;
    o.conditional_reserve_bytes = admission - expected;
    std::memcpy(output, &o, sizeof o);
  });
}

// ---- performance receipt (vckss_rust.h:1441-1456, 2116-2120)
int32_t vckss_rust_engine_performance_receipt_v1(uint64_t generation, VckssEnginePerformanceReceiptV1 *output,
                                                 uint32_t output_capacity_bytes) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    require_capacity<VckssEnginePerformanceReceiptV1>(output, output_capacity_bytes, "performance receipt");
    const Context &ctx = solved(generation);
    VckssEnginePerformanceReceiptV1 o;
    std::memset(&o, 0, sizeof o);
    o.struct_size = sizeof o;
    o.schema_version = 1;
    o.generation = generation;
    o.applicability_flags = 3;  // preparation phases and the solve are timed
    o.algorithm_selected = ctx.algorithm_selected;
    o.engine_selected = ctx.engine_selected;
    o.ingest_ns = ctx.ingest_ns;
    o.canonicalize_ns = ctx.P.canonicalize_ns;
    o.graph_ns = ctx.P.graph_ns;
    o.compress_ns = ctx.P.compress_ns;
    o.plan_ns = ctx.P.plan_ns + ctx.hybrid_ns + ctx.hierarchy_ns + ctx.upload_ns;
    o.stayer_augmentation_ns = ctx.stayer_ns;
    o.solve_ns = ctx.solve_ns;
    o.native_total_ns = o.ingest_ns + o.canonicalize_ns + o.graph_ns + o.compress_ns + o.plan_ns +
                        o.stayer_augmentation_ns + o.solve_ns;
    std::memcpy(output, &o, sizeof o);
  });
}

// ---- detailed receipts V6 and V7 (vckss_rust.h:2101-2115)
int32_t vckss_rust_engine_detailed_receipt_v6(uint64_t generation, VckssEngineDetailedReceiptV6 *output,
                                              uint32_t output_capacity_bytes) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    require_capacity<VckssEngineDetailedReceiptV6>(output, output_capacity_bytes, "V6 detailed receipt");
    const Context &ctx = solved(generation);
    const VckssEngineDetailedReceiptV6 o = detailed_v6(generation, ctx);
    std::memcpy(output, &o, sizeof o);
  });
}

int32_t vckss_rust_engine_detailed_receipt_v7(uint64_t generation, VckssEngineDetailedReceiptV7 *output,
                                              uint32_t output_capacity_bytes) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    require_capacity<VckssEngineDetailedReceiptV7>(output, output_capacity_bytes, "V7 detailed receipt");
    const Context &ctx = solved(generation);
    VckssEngineDetailedReceiptV7 o;
    std::memset(&o, 0, sizeof o);
    o.v6 = detailed_v6(generation, ctx);
    o.v6.v5.v4.v3.v2.struct_size = sizeof o;
    o.execution = execution_plan(generation, ctx);
    std::memcpy(output, &o, sizeof o);
  });
}

// ---- RHS receipts (vckss_rust.h:2126-2130): fit, leverage, target rows
int32_t vckss_rust_engine_rhs_receipts_v1(uint64_t generation, VckssEngineRhsReceiptV1 *output,
                                          uint64_t output_capacity_rows) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    const Context &ctx = solved(generation);
    const std::vector<VckssEngineRhsReceiptV1> rows = rhs_rows(ctx);
    if (rows.empty()) fail(kUnsupported, "engine_rhs_receipts", "the solved route keeps no RHS receipts");
    if (!output) fail(kInvalidInput, "engine_rhs_receipts", "output pointer is null");
    if (output_capacity_rows < rows.size())
      fail(kAbiMismatch, "engine_rhs_receipts", "output capacity is smaller than the RHS receipt rows");
    std::memcpy(output, rows.data(), rows.size() * sizeof(VckssEngineRhsReceiptV1));
  });
}

// ---- full-CMG receipt (vckss_rust.h:1458-1509, 2121-2125; ffi_engine.rs:8598-8689)
int32_t vckss_rust_engine_full_cmg_receipt_v1(uint64_t generation, VckssFullCmgReceiptV1 *output,
                                              uint32_t output_capacity_bytes) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    require_capacity<VckssFullCmgReceiptV1>(output, output_capacity_bytes, "full-CMG receipt");
    const Context &ctx = solved(generation);
    if (!ctx.full_cmg) fail(kInvalidInput, "engine_full_cmg_receipt", "the solved generation did not use CMG_FULL_V2");
    const SolveResult &r = ctx.result;
    const auto &v1 = ctx.request.v4.v3.v2.v1;
    const MemoryModel m = memory_model(ctx);
    VckssFullCmgReceiptV1 o;
    std::memset(&o, 0, sizeof o);
    o.struct_size = sizeof o;
    o.schema_version = 1;
    o.generation = generation;
    // the CMG hierarchy is that of the vendored cmg crate at this revision,
    // reproduced bit for bit; identity 2 as the reference reports it
    o.backend_identity = 2;
    o.platform_os = 2;    // Linux
    o.platform_arch = 2;  // x86_64 host
    static const char kCommit[] = "e5613e81fb3b03c78de98658b372646fa802aabb";
    std::memcpy(o.cmg_source_commit, kCommit, 40);
    o.threads_requested = o.threads_used = std::max<uint32_t>(ctx.request.threads, 1);
    o.maximum_concurrency = 1;  // one lockstep batch on the device at a time
    o.vertices = ctx.HS.graph.n;
    o.edges = ctx.HS.graph.u.size();
    o.hierarchy_levels = ctx.H.levels.size();
    o.terminal_vertices = ctx.H.levels.empty() ? 0 : ctx.H.levels.back().graph.n;
    o.graph_copy_bytes = m.graph;
    o.hierarchy_bytes = m.hierarchy;
    o.plan_bytes = m.plan;
    o.workspace_bytes_each = o.workspace_pool_bytes = m.workspace;
    o.fit_effective_tolerance = v1.pcg_tolerance;
    o.probe_effective_tolerance = ctx.request.tolerance_supplied ? v1.pcg_tolerance : 1e-6;
    o.fit_initial_inner_tolerance = 0.01 * o.fit_effective_tolerance;  // full_cmg.rs:399-412
    o.probe_initial_inner_tolerance = o.probe_effective_tolerance;
    const uint64_t *a = ctx.counters_before, *b = ctx.counters_after;
    o.refinement_attempts = b[2] - a[2];
    o.refined_columns = b[3] - a[3];
    o.batch_calls = b[1] - a[1];
    o.across_rhs_batches = o.batch_calls;  // every batch is a lockstep multi-RHS solve
    o.batch_strategy_mask = o.across_rhs_batches > 0 ? 4u : 0u;  // ffi_engine.rs:8607-8609
    o.rhs_count = 1 + r.control_fit.size() + r.leverage.size() + r.target.size();
    uint64_t it = r.fit.iterations;
    double red = r.fit.reduced;
    for (const auto *set : {&r.control_fit, &r.leverage, &r.target})
      for (const ColumnReceipt &c : *set) {
        it += c.iterations;
        red = std::max(red, c.reduced);
      }
    o.total_iterations = it;
    o.total_operator_applications = it + o.rhs_count;
    o.total_preconditioner_applications = b[0] - a[0];
    o.maximum_reduced_residual = red;
    o.maximum_complete_residual = r.max_complete_residual;
    o.graph_ns = ctx.hybrid_ns;
    o.hierarchy_plan_ns = ctx.hierarchy_ns + ctx.upload_ns;
    o.solve_ns = ctx.solve_ns;
    o.preparation_peak_bytes = std::max(ctx.preparation_forecast, ctx.stayer_forecast);
    o.prepared_persistent_bytes = m.host_problem;
    o.actual_retained_bytes = m.graph + m.hierarchy + m.plan + m.workspace;
    o.allocator_allowance_bytes = o.actual_retained_bytes / 5;
    o.non_cmg_command_peak_bytes = m.host_problem + (ctx.P.cells * 2 + ctx.P.units * 6) * 8;
    o.admitted_peak_bytes = std::max(o.preparation_peak_bytes, o.non_cmg_command_peak_bytes + o.actual_retained_bytes +
                                                                   o.allocator_allowance_bytes);
    o.pre_rng_forecast_bytes = o.admitted_peak_bytes;
    // a leverage batch holds one column a probe, a target batch two
    o.maximum_batch_rhs = std::max<uint64_t>(ctx.leverage_width, 2 * (uint64_t)ctx.target_width);
    o.workspace_count = 1;
    std::memcpy(output, &o, sizeof o);
  });
}

// ---- V6 and V7 solve requests (vckss_rust.h:1584-1598;
// generic_execution_api.rs:20-160, 285-520): the generic engine's executor
// selection.  DIRECT_ATTACHMENTS on the CMG route with automatic batches is
// this engine's CMG_FULL_V2 path; the diagonal queue is refused (11).
static VckssEngineSolveRequestV6 default_solve_v6(uint32_t size) {
  VckssEngineSolveRequestV6 r;
  std::memset(&r, 0, sizeof r);
  r.v4 = default_solve_v5().v4;
  r.v4.v3.v2.v1.struct_size = size;
  r.v4.v3.v2.algorithm = VCKSS_ALGORITHM_JLA;
  r.v4.v3.engine = VCKSS_ENGINE_GENERIC;
  r.v4.v3.v2.v1.solver_route = VCKSS_ROUTE_DIAGONAL_PCG;
  r.v4.v3.v2.v1.allow_automatic_cmg_setup_fallback = 0;
  r.execution_mode = VCKSS_GENERIC_EXECUTION_DIAGONAL_QUEUE;
  r.threads = 1;
  return r;
}

int32_t vckss_rust_engine_default_solve_request_v6(VckssEngineSolveRequestV6 *output, uint32_t output_capacity_bytes) {
  return guarded([&] {
    require_capacity<VckssEngineSolveRequestV6>(output, output_capacity_bytes, "V6 solve default");
    const VckssEngineSolveRequestV6 d = default_solve_v6(sizeof(VckssEngineSolveRequestV6));
    std::memcpy(output, &d, sizeof d);
  });
}

int32_t vckss_rust_engine_default_solve_request_interrupt_v6(VckssEngineSolveRequestInterruptV6 *output,
                                                             uint32_t output_capacity_bytes) {
  return guarded([&] {
    require_capacity<VckssEngineSolveRequestInterruptV6>(output, output_capacity_bytes, "V6 interrupt solve default");
    VckssEngineSolveRequestInterruptV6 d;
    std::memset(&d, 0, sizeof d);
    d.options = default_solve_v6(sizeof(VckssEngineSolveRequestInterruptV6));
    std::memcpy(output, &d, sizeof d);
  });
}

int32_t vckss_rust_engine_default_solve_request_v7(VckssEngineSolveRequestV7 *output, uint32_t output_capacity_bytes) {
  return guarded([&] {
    require_capacity<VckssEngineSolveRequestV7>(output, output_capacity_bytes, "V7 solve default");
    VckssEngineSolveRequestV7 d;
    std::memset(&d, 0, sizeof d);
    d.v6 = default_solve_v6(sizeof(VckssEngineSolveRequestV7));
    d.component_batch_mode = 1;
    std::memcpy(output, &d, sizeof d);
  });
}

int32_t vckss_rust_engine_default_solve_request_interrupt_v7(VckssEngineSolveRequestInterruptV7 *output,
                                                             uint32_t output_capacity_bytes) {
  return guarded([&] {
    require_capacity<VckssEngineSolveRequestInterruptV7>(output, output_capacity_bytes, "V7 interrupt solve default");
    VckssEngineSolveRequestInterruptV7 d;
    std::memset(&d, 0, sizeof d);
    d.options.v6 = default_solve_v6(sizeof(VckssEngineSolveRequestInterruptV7));
    d.options.component_batch_mode = 1;
    std::memcpy(output, &d, sizeof d);
  });
}

// generic_execution_api.rs:329-372 (validate) for this engine's executor
static VckssEngineSolveRequestV5 v6_as_v5(const VckssEngineSolveRequestV6 &r, uint32_t minimum_size) {
  const auto &v1 = r.v4.v3.v2.v1;
  if (v1.abi_version != VCKSS_RUST_ABI_VERSION_V1) fail(kAbiMismatch, "engine_solve", "unsupported ABI version");
  if (v1.struct_size < minimum_size || r.reserved_6 != 0)
    fail(kAbiMismatch, "engine_solve", "invalid V6 solve structure size or reserved field");
  if (r.threads == 0) fail(kInvalidInput, "generic_execution", "V6 requires positive permitted threads");
  if (r.v4.v3.v2.algorithm != VCKSS_ALGORITHM_JLA || r.v4.v3.engine != VCKSS_ENGINE_GENERIC ||
      v1.rng_contract != VCKSS_RNG_COUNTER_V1 || v1.allow_automatic_cmg_setup_fallback != 0)
    fail(kUnsupported, "generic_execution", "V6 requires explicit generic JLA, Counter-V1 and no fallback");
  const bool direct = r.execution_mode == VCKSS_GENERIC_EXECUTION_DIRECT_ATTACHMENTS &&
                      v1.solver_route == VCKSS_ROUTE_CMG_PCG && r.v4.v3.batch_mode == VCKSS_BATCH_MODE_AUTO &&
                      r.v4.leverage_batch_mode == VCKSS_BATCH_MODE_AUTO &&
                      r.v4.target_batch_mode == VCKSS_BATCH_MODE_AUTO && v1.leverage_batch_width == 0 &&
                      v1.target_batch_width == 0;
  if (!direct)
    fail(kUnsupported, "generic_execution",
         "the MI355X engine runs the generic engine's direct CMG executor only (CMG route, automatic batches)");
  VckssEngineSolveRequestV5 q;
  std::memset(&q, 0, sizeof q);
  q.v4 = r.v4;
  q.v4.v3.v2.v1.struct_size = sizeof(VckssEngineSolveRequestV5);
  q.threads = r.threads;
  q.tolerance_supplied = 1;  // the executor plans the probes at the request tolerance
  q.full_cmg_v2 = 1;
  return q;
}

static void solve_v6_common(uint64_t generation, const VckssEngineSolveRequestV5 &q, Interrupt &interrupt) {
  Context &ctx = live(generation);
  const SolveOptions o = point_options(q.v4, q.tolerance_supplied, ctx);
  record_request(ctx, q, o);
  run_options(ctx, o, interrupt);
}

int32_t vckss_rust_engine_solve_v6(uint64_t generation, const VckssEngineSolveRequestV6 *request) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (!request) fail(kInvalidInput, "engine_solve", "request pointer is null");
    VckssEngineSolveRequestV6 r;
    std::memcpy(&r, request, sizeof r);
    Interrupt none;
    solve_v6_common(generation, v6_as_v5(r, sizeof(VckssEngineSolveRequestV6)), none);
  });
}

int32_t vckss_rust_engine_solve_interrupt_v6(uint64_t generation, const VckssEngineSolveRequestInterruptV6 *request) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (!request) fail(kInvalidInput, "engine_solve", "request pointer is null");
    VckssEngineSolveRequestInterruptV6 r;
    std::memcpy(&r, request, sizeof r);
    if (r.reserved) fail(kAbiMismatch, "engine_solve", "reserved interrupt solve fields must be zero");
    const VckssEngineSolveRequestV5 q = v6_as_v5(r.options, sizeof(VckssEngineSolveRequestV6));
    Interrupt interrupt(r.interrupt_poll, r.interrupt_context, r.checkpoint_interval);
    solve_v6_common(generation, q, interrupt);
  });
}

// generic_execution_api.rs:473-500 (validate_v7)
static VckssEngineSolveRequestV5 v7_as_v5(const VckssEngineSolveRequestV7 &r) {
  if (r.v6.v4.v3.v2.v1.struct_size < sizeof(VckssEngineSolveRequestV7) || r.component_batch_mode != 1 ||
      r.reserved_7 != 0)
    fail(kAbiMismatch, "engine_solve", "invalid V7 solve structure or component batch mode");
  return v6_as_v5(r.v6, sizeof(VckssEngineSolveRequestV7));
}

int32_t vckss_rust_engine_solve_v7(uint64_t generation, const VckssEngineSolveRequestV7 *request) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (!request) fail(kInvalidInput, "engine_solve", "request pointer is null");
    VckssEngineSolveRequestV7 r;
    std::memcpy(&r, request, sizeof r);
    Interrupt none;
    solve_v6_common(generation, v7_as_v5(r), none);
  });
}

int32_t vckss_rust_engine_solve_interrupt_v7(uint64_t generation, const VckssEngineSolveRequestInterruptV7 *request) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    if (!request) fail(kInvalidInput, "engine_solve", "request pointer is null");
    VckssEngineSolveRequestInterruptV7 r;
    std::memcpy(&r, request, sizeof r);
    if (r.reserved) fail(kAbiMismatch, "engine_solve", "reserved interrupt solve fields must be zero");
    const VckssEngineSolveRequestV5 q = v7_as_v5(r.options);
    Interrupt interrupt(r.interrupt_poll, r.interrupt_context, r.checkpoint_interval);
    solve_v6_common(generation, q, interrupt);
  });
}

}  // extern "C"
