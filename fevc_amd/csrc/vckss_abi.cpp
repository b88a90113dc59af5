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
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>

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
  uint64_t stayer_copy_bytes = 0;
  uint32_t precision = 64, leverage_width = 32, target_width = 16;
};

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

void prepare_into(Context &ctx, const VckssEnginePrepareRequestInterruptV3 *request,
                  const VckssEngineColumnsV3 *columns) {
  if (!request || !columns) fail(kInvalidInput, "engine_prepare", "request or columns pointer is null");
  VckssEnginePrepareRequestInterruptV3 req;
  std::memcpy(&req, request, sizeof req);
  const auto &o = req.options;
  if (o.v3.v2.abi_version != VCKSS_RUST_ABI_VERSION_V1)
    fail(kAbiMismatch, "engine_prepare", "unsupported ABI version");
  if (o.v3.v2.struct_size < sizeof(VckssEnginePrepareRequestV4))
    fail(kAbiMismatch, "engine_prepare", "prepare request reports a short structure size");
  if (o.v3.v2.reserved || o.v3.reserved_2 || o.reserved_3 || req.reserved)
    fail(kAbiMismatch, "engine_prepare", "reserved fields must be zero");
  if (o.v3.deletion_mode != VCKSS_DELETION_MATCH && o.v3.deletion_mode != VCKSS_DELETION_OBSERVATION)
    fail(kInvalidInput, "engine_prepare", "unknown deletion mode");
  if (o.v3.controls_count != 0 || columns->v2.controls_count != 0)
    fail(kUnsupported, "engine_prepare", "the MI355X engine supports the no-control route only");
  if (columns->probeorder_supplied)
    fail(kUnsupported, "engine_prepare", "probe-order keys are not supported by the MI355X engine");
  if (o.implicit_match > 1) fail(kInvalidInput, "engine_prepare", "implicit_match must be 0 or 1");
  if (o.implicit_match)
    fail(kUnsupported, "engine_prepare", "the certified implicit-match route is not supported by the MI355X engine");
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
  ctx.P = prepare(cols, nullptr, o.v3.deletion_mode == VCKSS_DELETION_OBSERVATION);
  interrupt.checkpoint("engine_prepare_graph");
  ctx.HS = build_hybrid(ctx.P);
  interrupt.checkpoint("cmg_full_v2_hierarchy");
  CmgOptions copt;  // full_cmg.rs:467-476: max_levels 128
  ctx.H = build_hierarchy(ctx.HS.graph, copt);
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
      deletion_ok && v2.nuisance_mode == VCKSS_NUISANCE_JOINT &&
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
         "deletion, joint nuisance, movers, automatic batches and no probe order");
  SolveOptions o;
  o.exact = exact;
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
    if (kBuiltinSmootherCount) ctx.dev->smoother_default(builtin_smoother());
  }
  return *ctx.dev;
}

void run_options(Context &ctx, SolveOptions o, Interrupt &interrupt);

void run_solve(Context &ctx, const VckssEngineSolveRequestV5 &q, Interrupt &interrupt) {
  run_options(ctx, solve_options(q, ctx), interrupt);
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
  try {
    o.checkpoint("engine_solve");
    ctx.result = dev.run(o);
  } catch (...) {
    ctx.state = kFailed;
    throw;
  }
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

extern "C" {

uint32_t vckss_rust_abi_version(void) { return VCKSS_RUST_ABI_VERSION_V1; }
const char *vckss_rust_backend_version(void) { return "vckss-mi355x 0.1.0 (gfx950, CMG_FULL_V2)"; }
const char *vckss_rust_capabilities_json(void) {
  return "{\"backend\":\"vckss-mi355x\",\"device\":\"gfx950\",\"routes\":[\"CMG_FULL_V2\"],"
         "\"deletion\":[\"match\",\"observation\"],\"stayers\":[\"movers\",\"all\"],\"rng\":[\"VCKSS-COUNTER-V1\"],\"controls\":0,"
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
    d.options.v3.v2.struct_size = sizeof(VckssEnginePrepareRequestV4);
    d.options.v3.v2.memory_limit_bytes = 4ULL << 30;
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
    if (upload) ctx->dev.reset(new DeviceEngine(ctx->P, ctx->HS, ctx->H));
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
  if (ctx.state != kPrepared)
    fail(kContextPoisoned, "context_augment", "generation " + std::to_string(generation) +
                                                  " is not in the Prepared state");
  interrupt.checkpoint("engine_stayer_augmentation");
  StayerColumns sc{rows, firm.data(), worker.data(), frequency.data(), outcome.data(), target.data()};
  try {
    StayerReceipt rec;
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
  o.numerical = true;
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
        if (!kBuiltinSmootherCount) fail(kUnsupported, "smoother", "the library carries no trained smoother weights");
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
    *count = (uint32_t)kBuiltinSmootherCount;
    if (out) {
      if (capacity < kBuiltinSmootherCount) fail(kAbiMismatch, "smoother_params", "capacity too small");
      std::memcpy(out, kBuiltinSmoother, kBuiltinSmootherCount * sizeof(double));
    }
  });
}

int32_t vckss_mi355x_smoother_omega_v1(uint64_t generation, const double *params, uint32_t count, double *omega,
                                       uint64_t capacity, uint64_t *written) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return guarded([&] {
    Context &ctx = live(generation);
    if (!params || !written) fail(kInvalidInput, "smoother_omega", "params and written are required");
    std::vector<double> w;
    device_of(ctx).smoother_omega(std::vector<double>(params, params + count), w);
    *written = w.size();
    if (omega) {
      if (capacity < w.size()) fail(kAbiMismatch, "smoother_omega", "capacity too small");
      std::memcpy(omega, w.data(), w.size() * sizeof(double));
    }
  });
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

}  // extern "C"
