"""The reference Stata shim's own call sequence for a full-CMG JLA command,
replayed through the C ABI with the header's structures, on a 4000-worker
panel (rust/stata_backend/cshim/stata_entry.c; fevc/fevc.ado:1512-1530,
2474-2610):

  probe -> requestcapability (V3) -> admit_prepare_v4 -> clear_abandoned ->
  prepare_interrupt_v4 -> preparation_receipt_v4 + retained_mask ->
  solvefull (default V5 request + capability fields, full_cmg_v2 = 1) ->
  memory_forecast -> result -> performance receipt -> detailed V6 / V7 ->
  full-CMG receipt -> RHS receipts -> release.

Every step must return 0 and every receipt must pass the shim's (and, for
the full-CMG receipt, the ado's) reconciliation; the estimate must match
the oracle engine's on the same Counter-V1 draws within 1e-8 relative.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import fevc_amd as F
from fevc_amd import abi_types, synth

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "abi"))
import shim  # noqa: E402

pytestmark = pytest.mark.gpu
K = abi_types.CONSTANTS
COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")
SPEC = synth.PanelSpec(workers=4000, firms=400, periods=5, mobility=0.2, seed=21, name="shim-sequence")
PROBES, SEED, TOL, THREADS = 8, 90210, 1e-10, 4


def new(name):
    return abi_types.new(name)


def raw_library():
    """The library under a second handle whose functions carry no Python
    prototypes, called as a C caller would (64-bit arguments passed as
    c_uint64, structures by pointer)."""
    lib = C.CDLL(F.library()._name)
    lib.vckss_rust_last_error.restype = C.c_char_p
    return lib


def ok(status, step):
    assert status == 0, (step, status, F.library().vckss_rust_last_error())


def test_shim_full_cmg_sequence(oracle):
    lib = raw_library()
    d = synth.generate(SPEC)
    cols = [np.ascontiguousarray(d[k], np.float64) for k in COLS]
    rows = len(cols[0])
    polls = []
    poll = F.Poll(lambda _ctx: polls.append(1) or 0)

    print("shim: probe and V3 capability", flush=True)
    assert lib.vckss_rust_abi_version() == 1
    caps = new("VckssBackendCapabilitiesV1")
    ok(lib.vckss_rust_backend_capabilities_v1(C.byref(caps), C.sizeof(caps)), "capabilities")
    q = new("VckssBackendRequestCapabilityRequestV3")
    q.v2.v1.abi_version, q.v2.v1.struct_size = 1, C.sizeof(q)
    q.v2.v1.request_schema = K["VCKSS_REQUEST_CAPABILITY_SCHEMA_V3"]
    q.v2.v1.algorithm = K["VCKSS_ALGORITHM_JLA"]
    q.v2.v1.deletion_mode = K["VCKSS_DELETION_MATCH"]
    q.v2.v1.nuisance_mode = K["VCKSS_NUISANCE_JOINT"]
    q.v2.v1.solver_route = K["VCKSS_ROUTE_CMG_PCG"]
    q.v2.v1.rng_contract = K["VCKSS_RNG_COUNTER_V1"]
    q.v2.v1.frequency_use = K["VCKSS_REQUEST_FREQUENCY_LITERAL"]
    q.v2.engine = K["VCKSS_ENGINE_AUTO_OR_UNSPECIFIED"]
    q.v2.batch_mode = K["VCKSS_BATCH_MODE_AUTO"]
    q.v2.stayers_mode = K["VCKSS_STAYERS_MOVERS"]
    q.v2.target_weight_mode = K["VCKSS_TARGET_WEIGHT_FREQUENCY_DEFAULT"]
    q.v2.deletion_unit_source = K["VCKSS_DELETION_SOURCE_CELL_DEFAULT"]
    q.v2.physical_limit = 50_000_000
    q.leverage_batch_mode = q.target_batch_mode = K["VCKSS_BATCH_MODE_AUTO"]
    cap = new("VckssBackendRequestCapabilityReceiptV3")
    ok(lib.vckss_rust_backend_request_capability_v3(C.byref(q), C.byref(cap), C.sizeof(cap)), "capability")
    assert shim.capability_v3(q, cap) == []
    assert cap.v2.v1.supported == 1, cap.v2.v1.reason_code

    print("shim: admit, clear, prepare", flush=True)
    pr = new("VckssEnginePrepareRequestInterruptV3")
    ok(lib.vckss_rust_engine_default_prepare_request_interrupt_v3(C.byref(pr), C.sizeof(pr)), "default prepare")
    o = pr.options
    o.v3.v2.abi_version = 1
    o.v3.v2.rows = rows
    o.v3.v2.cleanup_abandoned = 1
    o.v3.v2.memory_limit_bytes = 8 << 30
    o.v3.v2.caller_copy_bytes = rows * 6 * 8
    o.v3.deletion_mode = K["VCKSS_DELETION_MATCH"]
    pr.interrupt_poll = C.cast(poll, C.c_void_p)
    pr.checkpoint_interval = 1
    ok(lib.vckss_rust_engine_admit_prepare_v4(C.byref(o), C.c_uint32(0)), "admit")
    ok(lib.vckss_rust_engine_clear_abandoned_v1(), "clear abandoned")
    o.v3.v2.cleanup_abandoned = 0
    c = new("VckssEngineColumnsV3")
    c.v2.v1.struct_size, c.v2.v1.rows = C.sizeof(c), rows
    (c.v2.v1.worker, c.v2.v1.firm, c.v2.v1.deletion, c.v2.v1.outcome, c.v2.v1.frequency,
     c.v2.v1.target_weight) = [a.ctypes.data for a in cols]
    h = C.c_uint64()
    ok(lib.vckss_rust_engine_prepare_interrupt_v4(C.byref(pr), C.byref(c), C.byref(h), C.c_uint32(8)), "prepare")
    generation = h.value
    g64 = C.c_uint64(generation)
    try:
        prep = new("VckssEnginePreparationReceiptV4")
        ok(lib.vckss_rust_engine_preparation_receipt_v4(g64, C.byref(prep), C.sizeof(prep)), "prep receipt")
        p2 = prep.v3.v2
        assert p2.generation == generation and p2.input_rows == rows
        assert p2.memory_limit_bytes == 8 << 30 and p2.caller_copy_bytes == rows * 48
        assert p2.preparation_peak_forecast_bytes == rows * (48 + 768) + 4096
        mask = np.zeros(rows, np.uint8)
        ok(lib.vckss_rust_engine_retained_mask_v1(g64, C.c_void_p(mask.ctypes.data), C.c_uint64(rows)), "retained mask")
        assert int(mask.sum()) == p2.retained_rows

        print("shim: solvefull", flush=True)
        sr = new("VckssEngineSolveRequestInterruptV5")
        ok(lib.vckss_rust_engine_default_solve_request_interrupt_v5(C.byref(sr), C.sizeof(sr)), "default solve")
        s = sr.options
        v1 = s.v4.v3.v2.v1
        v1.abi_version = 1
        v1.seed, v1.probes = SEED, PROBES
        v1.solver_route = cap.v2.v1.solver_route
        v1.pcg_tolerance = TOL
        v1.maximum_iterations = 10000
        v1.deletion_mode = cap.v2.v1.deletion_mode
        v1.rng_contract = K["VCKSS_RNG_COUNTER_V1"]
        v1.allow_automatic_cmg_setup_fallback = cap.allow_automatic_cmg_setup_fallback
        s.v4.v3.v2.algorithm = cap.v2.v1.algorithm
        s.v4.v3.v2.nuisance_mode = cap.v2.v1.nuisance_mode
        s.v4.v3.engine = cap.v2.engine
        s.v4.v3.batch_mode = cap.v2.batch_mode
        s.v4.v3.stayers_mode = cap.v2.stayers_mode
        s.v4.v3.target_weight_mode = cap.v2.target_weight_mode
        s.v4.v3.deletion_unit_source = cap.v2.deletion_unit_source
        s.v4.v3.probeorder_supplied = cap.v2.probeorder_supplied
        s.v4.v3.wallseconds_supplied = cap.v2.wallseconds_supplied
        s.v4.v3.physical_limit = cap.v2.physical_limit
        s.v4.v3.capability_schema = cap.v2.v1.request_schema
        s.v4.v3.capability_profile = cap.v2.v1.profile_code
        s.v4.v3.frequency_use = cap.v2.v1.frequency_use
        s.v4.v3.request_signature = cap.v2.v1.request_signature
        s.v4.leverage_batch_mode = cap.leverage_batch_mode
        s.v4.target_batch_mode = cap.target_batch_mode
        s.v4.wallseconds = cap.wallseconds
        s.threads = THREADS
        s.tolerance_supplied = 1
        s.full_cmg_v2 = 1
        sr.interrupt_poll = C.cast(poll, C.c_void_p)
        sr.checkpoint_interval = 1
        before = len(polls)
        ok(lib.vckss_rust_engine_solve_interrupt_v5(g64, C.byref(sr)), "solve")
        assert len(polls) > before  # the shim's poll ran at the solve's checkpoints

        print("shim: result sequence", flush=True)
        mf = new("VckssMemoryForecastV1")
        ok(lib.vckss_rust_engine_memory_forecast_v1(g64, C.byref(mf), C.sizeof(mf)), "memory forecast")
        assert shim.memory_forecast(mf) == []
        res = new("VckssEngineResultV1")
        ok(lib.vckss_rust_engine_result_v1(g64, C.byref(res), C.sizeof(res)), "result")
        perf = new("VckssEnginePerformanceReceiptV1")
        ok(lib.vckss_rust_engine_performance_receipt_v1(g64, C.byref(perf), C.sizeof(perf)), "performance")
        v6 = new("VckssEngineDetailedReceiptV6")
        ok(lib.vckss_rust_engine_detailed_receipt_v6(g64, C.byref(v6), C.sizeof(v6)), "detailed v6")
        assert shim.performance(generation, perf, v6.v5.v4, v6) == []
        assert shim.rhs_accounting(v6) == []
        assert v6.capability_schema == K["VCKSS_REQUEST_CAPABILITY_SCHEMA_V3"]
        assert v6.request_signature == cap.v2.v1.request_signature
        v7 = new("VckssEngineDetailedReceiptV7")
        ok(lib.vckss_rust_engine_detailed_receipt_v7(g64, C.byref(v7), C.sizeof(v7)), "detailed v7")
        assert shim.detailed_v7(generation, v6, v7) == []

        fc = new("VckssFullCmgReceiptV1")
        ok(lib.vckss_rust_engine_full_cmg_receipt_v1(g64, C.byref(fc), C.sizeof(fc)), "full-CMG receipt")
        assert shim.full_cmg(generation, fc) == []
        d2 = v6.v5.v4.v3.v2
        # fevc.ado:2577-2602, the checks that depend on the request
        assert fc.threads_requested == fc.threads_used == THREADS
        assert fc.fit_effective_tolerance == TOL and fc.probe_effective_tolerance == TOL
        assert fc.fit_initial_inner_tolerance == 0.01 * TOL and fc.probe_initial_inner_tolerance == TOL
        assert fc.maximum_batch_rhs == max(d2.leverage_batch_width, 2 * d2.target_batch_width)
        assert fc.admitted_peak_bytes <= fc.pre_rng_forecast_bytes
        assert fc.maximum_complete_residual <= max(1e-11, 10 * TOL)

        n_rhs = v6.v5.v4.v3.rhs_receipt_rows
        rr = (abi_types.STRUCTS["VckssEngineRhsReceiptV1"] * n_rhs)()
        ok(lib.vckss_rust_engine_rhs_receipts_v1(g64, rr, C.c_uint64(n_rhs)), "rhs receipts")
        assert n_rhs == 1 + PROBES + 2 * PROBES
        assert all(r.route == K["VCKSS_ROUTE_CMG_PCG"] for r in rr)
        assert max(r.complete_residual for r in rr) <= max(1e-11, 10 * TOL)
        got = {k: [getattr(getattr(res, k), f) for f in ("worker", "firm", "covariance", "total")]
               for k in ("plugin", "correction", "corrected")}
    finally:
        ok(lib.vckss_rust_engine_release_v1(g64), "release")

    print("shim: oracle", flush=True)
    want = oracle.Problem(*cols).run(seed=SEED, probes=PROBES, threads=4, fit_tolerance=TOL,
                                     probe_tolerance=TOL)
    for key in ("plugin", "correction", "corrected"):
        g, w = np.asarray(got[key]), np.asarray(want[key])
        err = np.abs(g - w) / np.maximum(np.abs(w), 1e-12)
        print(f"shim: {key} max rel error {err.max():.2e}")
        assert np.all(err <= 1e-8), (key, g, w)
