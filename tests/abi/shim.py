"""The reconciliation checks the reference Stata shim applies to the
receipts of this boundary (rust/stata_backend/cshim/stata_entry.c), as
Python predicates over fevc_amd.abi_types structures.  Each returns a list
of the checks that failed (empty when the receipt reconciles)."""
from fevc_amd import abi_types

K = abi_types.CONSTANTS
STRUCTS = abi_types.STRUCTS


def size(name):
    import ctypes

    return ctypes.sizeof(STRUCTS[name])


def capability_v3(request, receipt):
    """stata_entry.c:1197-1229."""
    bad = []
    r1, q1 = receipt.v2.v1, request.v2.v1
    if r1.struct_size != size("VckssBackendRequestCapabilityReceiptV3"):
        bad.append("struct_size")
    if r1.abi_version != K["VCKSS_RUST_ABI_VERSION_V1"]:
        bad.append("abi_version")
    for f in ("request_schema", "algorithm", "deletion_mode", "nuisance_mode", "solver_route", "rng_contract",
              "controls_count", "frequency_use"):
        if getattr(r1, f) != getattr(q1, f):
            bad.append(f)
    for f in ("engine", "batch_mode", "stayers_mode", "target_weight_mode", "deletion_unit_source",
              "probeorder_supplied", "wallseconds_supplied", "physical_limit"):
        if getattr(receipt.v2, f) != getattr(request.v2, f):
            bad.append(f)
    for f in ("leverage_batch_mode", "target_batch_mode", "allow_automatic_cmg_setup_fallback", "wallseconds"):
        if getattr(receipt, f) != getattr(request, f):
            bad.append(f)
    if r1.reserved or receipt.v2.reserved_2 or receipt.reserved_3 or receipt.reserved_4:
        bad.append("reserved")
    if r1.supported > 1 or receipt.wall_advisory_only != 1:
        bad.append("supported/wall_advisory_only")
    if (r1.supported == 1) != (r1.reason_code == K["VCKSS_REQUEST_REASON_SUPPORTED"]):
        bad.append("supported/reason_code")
    return bad


def memory_forecast(f):
    """stata_entry.c:2923-2932."""
    bad = []
    if f.struct_size != size("VckssMemoryForecastV1") or f.schema_version != 1:
        bad.append("struct_size/schema_version")
    if f.expected_peak_bytes > f.admission_peak_bytes:
        bad.append("expected > admission")
    elif f.conditional_reserve_bytes != f.admission_peak_bytes - f.expected_peak_bytes:
        bad.append("conditional_reserve_bytes")
    return bad


def detailed_v7(generation, v6, v7):
    """stata_entry.c:2960-2982."""
    bad = []
    e = v7.execution
    if e.struct_size != size("VckssExecutionPlanReceiptV1") or e.schema_version != 1:
        bad.append("execution struct_size/schema_version")
    if e.generation != generation:
        bad.append("execution generation")
    if e.request_signature != v6.request_signature or v7.v6.request_signature != v6.request_signature:
        bad.append("request_signature")
    for f in ("capability_schema", "capability_profile", "engine_requested", "engine_selected"):
        if getattr(v7.v6, f) != getattr(v6, f):
            bad.append(f)
    if (e.resolution.reserved or e.solver.reserved_1 or e.solver.reserved_2 or e.batch.leverage.reserved
            or e.batch.target.reserved or e.memory.reserved):
        bad.append("reserved")
    return bad


def performance(generation, perf, v4, v6):
    """stata_entry.c:3040-3060."""
    bad = []
    if perf.struct_size != size("VckssEnginePerformanceReceiptV1") or perf.schema_version != 1:
        bad.append("struct_size/schema_version")
    if perf.generation != generation or (perf.applicability_flags & 3) != 3:
        bad.append("generation/applicability")
    if perf.algorithm_selected != v4.algorithm_selected or perf.engine_selected != v6.engine_selected:
        bad.append("algorithm/engine")
    for f in ("ingest_ns", "canonicalize_ns", "graph_ns", "compress_ns", "plan_ns", "stayer_augmentation_ns",
              "solve_ns"):
        if getattr(perf, f) > perf.native_total_ns:
            bad.append(f)
    return bad


def rhs_accounting(v6):
    """stata_entry.c:3071-3138: the RHS export schema and its caller copy."""
    v3 = v6.v5.v4.v3
    bad = []
    if v6.reserved_6 != 0:
        bad.append("reserved_6")
    if v6.rhs_receipt_schema == 0:
        if v3.rhs_receipt_rows or v3.caller_result_copy_bytes or v6.rhs_v2_caller_copy_bytes:
            bad.append("schema 0")
    elif v6.rhs_receipt_schema == 1:
        per_row = size("VckssEngineRhsReceiptV1") + 8 * 8
        if (v3.rhs_receipt_rows == 0 or v3.caller_result_copy_bytes != v3.rhs_receipt_rows * per_row
                or v6.rhs_v2_caller_copy_bytes != 0):
            bad.append("schema 1")
    else:
        bad.append("schema")
    return bad


def full_cmg(generation, r):
    """stata_entry.c:489-601."""
    import math

    bad = []
    checks = {
        "struct": r.struct_size == size("VckssFullCmgReceiptV1") and r.schema_version == 1,
        "generation": r.generation == generation,
        "identity": r.backend_identity == 2 and bytes(r.cmg_source_commit) ==
        b"e5613e81fb3b03c78de98658b372646fa802aabb",
        "platform": r.platform_os in (1, 2, 3) and r.platform_arch in (1, 2),
        "strategy": r.batch_strategy_mask != 0 and (r.batch_strategy_mask & ~7) == 0,
        "threads": r.threads_requested != 0 and r.threads_used == r.threads_requested,
        "concurrency": 0 < r.maximum_concurrency <= r.threads_used,
        "graph": r.vertices and r.edges and r.hierarchy_levels and 0 < r.terminal_vertices <= r.vertices,
        "workspace": (r.workspace_pool_bytes >= r.workspace_bytes_each and r.maximum_batch_rhs
                      and 0 < r.workspace_count <= min(r.threads_used, r.maximum_batch_rhs)
                      and r.maximum_concurrency <= r.workspace_count),
        "persistent": r.prepared_persistent_bytes <= r.non_cmg_command_peak_bytes,
        "tolerances": all(math.isfinite(x) and x > 0 for x in (
            r.fit_effective_tolerance, r.probe_effective_tolerance, r.fit_initial_inner_tolerance,
            r.probe_initial_inner_tolerance)) and r.fit_initial_inner_tolerance <= r.fit_effective_tolerance
        and r.probe_initial_inner_tolerance <= r.probe_effective_tolerance,
        "counts": r.batch_calls and r.rhs_count,
        "residuals": (math.isfinite(r.maximum_reduced_residual) and math.isfinite(r.maximum_complete_residual)
                      and r.maximum_reduced_residual >= 0 and r.maximum_complete_residual >= 0),
        "batches": r.serial_batches + r.planned_batches + r.across_rhs_batches == r.batch_calls,
        "retained": r.actual_retained_bytes >= (r.graph_copy_bytes + r.hierarchy_bytes + r.plan_bytes
                                                + r.workspace_pool_bytes),
        "allowance": r.allocator_allowance_bytes == r.actual_retained_bytes // 5,
        "admitted": r.admitted_peak_bytes == max(r.preparation_peak_bytes, r.non_cmg_command_peak_bytes
                                                 + r.actual_retained_bytes + r.allocator_allowance_bytes),
        "pre_rng": r.pre_rng_forecast_bytes >= r.admitted_peak_bytes,
        "complete": r.maximum_complete_residual <= max(1e-11, 10 * max(r.fit_effective_tolerance,
                                                                       r.probe_effective_tolerance)),
    }
    bad += [k for k, ok in checks.items() if not ok]
    return bad
