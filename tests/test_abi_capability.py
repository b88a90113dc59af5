"""Capability queries and prepare admission of the C ABI, on the CPU.

Known answers and cases from the reference's own tests:
  * the V1 request signature of engine_ffi.rs:858-910 (0x511d0bcb0a6fe389)
    and its sensitivity to every field;
  * the reason codes of the unknown-value and algorithm rules
    (engine_ffi.rs:682-857, 971-1130), which precede this engine's own
    restrictions; reserved request fields are ABI errors (10);
  * 96 distinct V2 signatures over the reference's generic matrix;
  * admission: engine_ffi.rs:610-653 (exact memory forecast of
    caller copy + 768 bytes a row + 4096, one byte short is status 40).
What this engine adds (no controls, joint nuisance only, JLA on the
automatic or CMG route) is checked against its own stated rules, and the
receipts must pass the reconciliation the Stata shim applies
(stata_entry.c:1197-1229).
"""
import ctypes as C
import os
import sys

import pytest

import fevc_amd as F
from fevc_amd import abi_types

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "abi"))
import shim  # noqa: E402

K = abi_types.CONSTANTS


def v1_request(**changes):
    r = abi_types.new("VckssBackendRequestCapabilityRequestV1")
    r.abi_version, r.struct_size = 1, C.sizeof(r)
    r.request_schema = K["VCKSS_REQUEST_CAPABILITY_SCHEMA_V1"]
    r.algorithm = K["VCKSS_ALGORITHM_EXACT"]
    r.deletion_mode = K["VCKSS_DELETION_MATCH"]
    r.nuisance_mode = K["VCKSS_NUISANCE_JOINT"]
    r.solver_route = K["VCKSS_ROUTE_AUTO"]
    r.rng_contract = K["VCKSS_RNG_NONE"]
    r.frequency_use = K["VCKSS_REQUEST_FREQUENCY_UNIT"]
    for k, v in changes.items():
        setattr(r, k, v)
    return r


def receipt_v1(r):
    o = F.request_capability(r)
    assert o.struct_size == 64 and o.abi_version == 1 and o.reserved == 0
    for f in ("request_schema", "algorithm", "deletion_mode", "nuisance_mode", "solver_route", "rng_contract",
              "controls_count", "frequency_use"):
        assert getattr(o, f) == getattr(r, f), f
    assert (o.supported == 1) == (o.reason_code == 0)
    assert (o.profile_code == 0) == (o.supported == 0)
    return o


def test_v1_signature_known_answer_and_every_field():
    base = dict(deletion_mode=K["VCKSS_DELETION_OBSERVATION"], nuisance_mode=K["VCKSS_NUISANCE_FIXED_OFFSET"],
                solver_route=K["VCKSS_ROUTE_EXACT"], controls_count=17,
                frequency_use=K["VCKSS_REQUEST_FREQUENCY_LITERAL"])
    sig = receipt_v1(v1_request(**base)).request_signature
    assert sig == 0x511D0BCB0A6FE389
    assert receipt_v1(v1_request(**base)).request_signature == sig
    for k, v in (("request_schema", 2), ("algorithm", K["VCKSS_ALGORITHM_JLA"]),
                 ("deletion_mode", K["VCKSS_DELETION_MATCH"]), ("nuisance_mode", K["VCKSS_NUISANCE_JOINT"]),
                 ("solver_route", K["VCKSS_ROUTE_AUTO"]), ("rng_contract", K["VCKSS_RNG_COUNTER_V1"]),
                 ("controls_count", 18), ("frequency_use", 0)):
        assert receipt_v1(v1_request(**{**base, k: v})).request_signature != sig, k


def test_v1_reason_codes_and_engine_restrictions():
    R = {k[len("VCKSS_REQUEST_REASON_"):]: v for k, v in K.items() if k.startswith("VCKSS_REQUEST_REASON_")}
    jla = dict(algorithm=K["VCKSS_ALGORITHM_JLA"], rng_contract=K["VCKSS_RNG_COUNTER_V1"])
    cases = [
        (dict(request_schema=99), R["UNKNOWN_SCHEMA"]),
        (dict(algorithm=99), R["UNKNOWN_ALGORITHM"]),
        (dict(algorithm=K["VCKSS_ALGORITHM_AUTO"]), R["ALGORITHM_AUTO_UNRESOLVED"]),
        (dict(deletion_mode=99), R["UNKNOWN_DELETION"]),
        (dict(nuisance_mode=99), R["UNKNOWN_NUISANCE"]),
        (dict(solver_route=99), R["UNKNOWN_SOLVER_ROUTE"]),
        (dict(rng_contract=99), R["UNKNOWN_RNG_CONTRACT"]),
        (dict(frequency_use=99), R["UNKNOWN_FREQUENCY_USE"]),
        (dict(controls_count=33), R["CONTROLS_LIMIT"]),
        (dict(rng_contract=K["VCKSS_RNG_COUNTER_V1"]), R["EXACT_RNG"]),
        (dict(solver_route=K["VCKSS_ROUTE_DIAGONAL_PCG"]), R["EXACT_SOLVER_ROUTE"]),
        (dict(jla, deletion_mode=K["VCKSS_DELETION_OBSERVATION"]), R["JLA_DELETION"]),
        (dict(jla, nuisance_mode=K["VCKSS_NUISANCE_FIXED_OFFSET"]), R["JLA_NUISANCE"]),
        (dict(jla, controls_count=1), R["JLA_CONTROLS"]),
        (dict(jla, rng_contract=K["VCKSS_RNG_NONE"]), R["JLA_RNG"]),
        # this engine: no controls on the exact route, JLA on the automatic or CMG route
        (dict(controls_count=32), R["CONTROLS_LIMIT"]),
        (dict(jla, solver_route=K["VCKSS_ROUTE_DIAGONAL_PCG"]), R["JLA_GENERIC_SOLVER_ROUTE"]),
        (dict(jla, solver_route=K["VCKSS_ROUTE_EXACT"]), R["JLA_GENERIC_SOLVER_ROUTE"]),
    ]
    for changes, reason in cases:
        o = receipt_v1(v1_request(**changes))
        assert (o.supported, o.reason_code, o.profile_code) == (0, reason, 0), changes
    for deletion in (K["VCKSS_DELETION_MATCH"], K["VCKSS_DELETION_OBSERVATION"]):
        for route in (K["VCKSS_ROUTE_AUTO"], K["VCKSS_ROUTE_EXACT"]):
            for freq in (0, 1):
                for nuisance in (K["VCKSS_NUISANCE_JOINT"], K["VCKSS_NUISANCE_FIXED_OFFSET"]):
                    o = receipt_v1(v1_request(deletion_mode=deletion, solver_route=route, frequency_use=freq,
                                              nuisance_mode=nuisance))
                    assert (o.supported, o.profile_code) == (1, K["VCKSS_REQUEST_PROFILE_EXACT_V1"])
    for route in (K["VCKSS_ROUTE_AUTO"], K["VCKSS_ROUTE_CMG_PCG"]):
        o = receipt_v1(v1_request(**jla, solver_route=route))
        assert (o.supported, o.profile_code) == (1, K["VCKSS_REQUEST_PROFILE_JLA_COUNTER_V1"])
    bad = v1_request(reserved=1)
    out = abi_types.new("VckssBackendRequestCapabilityReceiptV1")
    assert F.library().vckss_rust_backend_request_capability_v1(C.byref(bad), C.byref(out), C.sizeof(out)) == 10
    assert F.library().vckss_rust_backend_request_capability_v1(C.byref(v1_request()), C.byref(out), 63) == 10


def v2_generic(**changes):
    r = abi_types.new("VckssBackendRequestCapabilityRequestV2")
    r.v1 = v1_request(struct_size=C.sizeof(r), request_schema=K["VCKSS_REQUEST_CAPABILITY_SCHEMA_V2"],
                      algorithm=K["VCKSS_ALGORITHM_JLA"], solver_route=K["VCKSS_ROUTE_CMG_PCG"],
                      rng_contract=K["VCKSS_RNG_COUNTER_V1"], frequency_use=1)
    r.engine = K["VCKSS_ENGINE_GENERIC"]
    r.batch_mode = K["VCKSS_BATCH_MODE_AUTO"]
    r.stayers_mode = K["VCKSS_STAYERS_MOVERS"]
    r.target_weight_mode = K["VCKSS_TARGET_WEIGHT_FREQUENCY_DEFAULT"]
    r.deletion_unit_source = K["VCKSS_DELETION_SOURCE_CELL_DEFAULT"]
    r.physical_limit = 50_000_000
    for k, v in changes.items():
        if k in ("deletion_mode", "nuisance_mode", "controls_count", "frequency_use", "solver_route"):
            setattr(r.v1, k, v)
        else:
            setattr(r, k, v)
    return r


def test_v2_signatures_distinct_and_reasons():
    R = {k[len("VCKSS_REQUEST_REASON_"):]: v for k, v in K.items() if k.startswith("VCKSS_REQUEST_REASON_")}
    sigs = set()
    S = {k[len("VCKSS_DELETION_SOURCE_"):]: v for k, v in K.items() if k.startswith("VCKSS_DELETION_SOURCE_")}
    for deletion, sources in ((K["VCKSS_DELETION_MATCH"], (S["CELL_DEFAULT"], S["MATCH_ID_EXPLICIT"])),
                              (K["VCKSS_DELETION_OBSERVATION"], (S["OBSERVATION_ROW"],))):
        for source in sources:
            for nuisance in (1, 2):
                for controls in (0, 32):
                    for freq in (0, 1):
                        for tw in (0, 1):
                            for limit in (1, 1 << 53):
                                o = F.request_capability(v2_generic(
                                    deletion_mode=deletion, deletion_unit_source=source, nuisance_mode=nuisance,
                                    controls_count=controls, frequency_use=freq, target_weight_mode=tw,
                                    physical_limit=limit))
                                assert o.v1.struct_size == 104
                                # controls run on the match route only
                                ok = controls == 0 or deletion == K["VCKSS_DELETION_MATCH"]
                                assert o.v1.supported == int(ok), (deletion, source, nuisance, controls)
                                if ok:
                                    assert o.v1.profile_code == K["VCKSS_REQUEST_PROFILE_JLA_GENERIC_COUNTER_V1"]
                                sigs.add(o.v1.request_signature)
    assert len(sigs) == 96
    cases = [
        (dict(engine=99), R["UNKNOWN_ENGINE"]),
        (dict(batch_mode=99), R["UNKNOWN_BATCH_MODE"]),
        (dict(stayers_mode=99), R["UNKNOWN_STAYERS_MODE"]),
        (dict(target_weight_mode=99), R["UNKNOWN_TARGET_WEIGHT_MODE"]),
        (dict(deletion_unit_source=99), R["UNKNOWN_DELETION_UNIT_SOURCE"]),
        (dict(probeorder_supplied=1), R["PROBEORDER_UNSUPPORTED"]),
        (dict(wallseconds_supplied=1), R["WALLSECONDS_UNSUPPORTED"]),
        (dict(physical_limit=0), R["PHYSICAL_LIMIT"]),
        (dict(physical_limit=(1 << 53) + 1), R["PHYSICAL_LIMIT"]),
        (dict(solver_route=K["VCKSS_ROUTE_EXACT"]), R["JLA_GENERIC_SOLVER_ROUTE"]),
        (dict(deletion_unit_source=K["VCKSS_DELETION_SOURCE_OBSERVATION_ROW"]),
         R["DELETION_UNIT_SOURCE_MISMATCH"]),
        (dict(deletion_mode=K["VCKSS_DELETION_OBSERVATION"]), R["DELETION_UNIT_SOURCE_MISMATCH"]),
        (dict(batch_mode=K["VCKSS_BATCH_MODE_EXPLICIT"]), R["BATCH_MODE_UNSUPPORTED"]),
    ]
    for changes, reason in cases:
        o = F.request_capability(v2_generic(**changes))
        assert (o.v1.supported, o.v1.reason_code, o.v1.profile_code) == (0, reason, 0), changes
    base = F.request_capability(v2_generic()).v1.request_signature
    for changes in (dict(engine=K["VCKSS_ENGINE_COMPRESSED"]), dict(batch_mode=K["VCKSS_BATCH_MODE_EXPLICIT"]),
                    dict(stayers_mode=K["VCKSS_STAYERS_ALL"]), dict(target_weight_mode=1),
                    dict(deletion_unit_source=K["VCKSS_DELETION_SOURCE_MATCH_ID_EXPLICIT"]),
                    dict(probeorder_supplied=1), dict(wallseconds_supplied=1),
                    dict(physical_limit=50_000_001)):
        assert F.request_capability(v2_generic(**changes)).v1.request_signature != base, changes


def ado_default_v3(**changes):
    """The V3 tuple fevc.ado:1512-1521 sends for a default JLA command:
    automatic algorithm, engine, route and batches, Counter-V1, movers,
    automatic CMG set-up fallback, no wall budget."""
    r = abi_types.new("VckssBackendRequestCapabilityRequestV3")
    r.v2 = v2_generic(engine=K["VCKSS_ENGINE_AUTO_OR_UNSPECIFIED"], solver_route=K["VCKSS_ROUTE_AUTO"])
    r.v2.v1.struct_size = C.sizeof(r)
    r.v2.v1.request_schema = K["VCKSS_REQUEST_CAPABILITY_SCHEMA_V3"]
    r.v2.v1.algorithm = K["VCKSS_ALGORITHM_AUTO"]
    r.leverage_batch_mode = r.target_batch_mode = K["VCKSS_BATCH_MODE_AUTO"]
    r.allow_automatic_cmg_setup_fallback = 1
    for k, v in changes.items():
        setattr(r, k, v)
    return r


def test_v3_ado_default_is_supported_and_reconciles():
    r = ado_default_v3()
    o = F.request_capability(r)
    assert shim.capability_v3(r, o) == []
    assert o.v2.v1.supported == 1 and o.v2.v1.profile_code == K["VCKSS_REQUEST_PROFILE_PLANNED_V1"]
    assert o.algorithm_resolution_deferred == 1 and o.route_resolution_deferred == 1
    assert o.leverage_batch_resolution_deferred == 1 and o.target_batch_resolution_deferred == 1
    # unsupported tuples still reconcile (the shim then reports the reason)
    for changes in (dict(leverage_batch_mode=K["VCKSS_BATCH_MODE_EXPLICIT"]), dict(wallseconds=3.0),
                    dict(allow_automatic_cmg_setup_fallback=2)):
        q = ado_default_v3(**changes)
        o = F.request_capability(q)
        assert shim.capability_v3(q, o) == [] and o.v2.v1.supported == 0, changes
    sigs = {F.request_capability(ado_default_v3(**c)).v2.v1.request_signature
            for c in (dict(), dict(allow_automatic_cmg_setup_fallback=0), dict(target_batch_mode=2),
                      dict(wallseconds=1.0))}
    assert len(sigs) == 4


def prepare_v4(rows, **changes):
    r = abi_types.new("VckssEnginePrepareRequestV4")
    r.v3.v2.abi_version, r.v3.v2.struct_size = 1, C.sizeof(r)
    r.v3.v2.rows = rows
    r.v3.v2.caller_copy_bytes = rows * 6 * 8
    r.v3.v2.memory_limit_bytes = rows * (48 + 768) + 4096
    r.v3.deletion_mode = K["VCKSS_DELETION_MATCH"]
    for k, v in changes.items():
        obj = r if k in ("implicit_match", "reserved_3") else (r.v3 if k in ("deletion_mode", "controls_count")
                                                                else r.v3.v2)
        setattr(obj, k, v)
    return r


def test_prepare_admission():
    assert F.admit_prepare(prepare_v4(10)) == 0
    assert F.admit_prepare(prepare_v4(10, memory_limit_bytes=10 * 816 + 4095)) == 40
    assert F.admit_prepare(prepare_v4(10, memory_limit_bytes=0)) == 20
    assert F.admit_prepare(prepare_v4(0, memory_limit_bytes=1 << 30)) == 20
    assert F.admit_prepare(prepare_v4(10, caller_copy_bytes=10 * 48 + 8)) == 20
    assert F.admit_prepare(prepare_v4(10, reserved=1)) == 10
    assert F.admit_prepare(prepare_v4(10, implicit_match=2)) == 10
    assert F.admit_prepare(prepare_v4(10, abi_version=2)) == 10
    assert F.admit_prepare(prepare_v4(10, struct_size=40)) == 10
    assert F.admit_prepare(prepare_v4(10, cleanup_abandoned=2)) == 20
    assert F.admit_prepare(prepare_v4(10, deletion_mode=9)) == 20
    assert F.admit_prepare(prepare_v4(10), probeorder_supplied=2) == 20
    # implicit match needs match deletion, no controls and a probe order (11)
    assert F.admit_prepare(prepare_v4(10, implicit_match=1)) == 11
    # controls: 8 bytes a row and control in the caller's copy, 32 in the
    # preparation charge (session.rs:202-290)
    r = prepare_v4(10, controls_count=2, caller_copy_bytes=10 * 8 * 8,
                   memory_limit_bytes=10 * (64 + 768 + 64) + 4096)
    assert F.admit_prepare(r) == 0
    r.v3.v2.memory_limit_bytes -= 1
    assert F.admit_prepare(r) == 40
    assert F.admit_prepare(prepare_v4(10, controls_count=33, caller_copy_bytes=10 * 39 * 8,
                                      memory_limit_bytes=1 << 30)) == 20
    # the engine's restrictions, after a passing admission (11)
    r = prepare_v4(10, caller_copy_bytes=10 * 7 * 8, memory_limit_bytes=10 * (56 + 768 + 16) + 4096)
    assert F.admit_prepare(r, probeorder_supplied=1) == 11
    assert F.admit_prepare(prepare_v4(10, deletion_mode=K["VCKSS_DELETION_OBSERVATION"])) == 0


def test_backend_capabilities_receipt():
    o = F.backend_capabilities()
    assert o.struct_size == 32 and o.abi_version == 1 and o.reserved == 0
    assert o.core_ready_flags & K["VCKSS_CORE_FULL_CMG_V2_READY"]
    assert o.core_ready_flags & K["VCKSS_CORE_COUNTER_RNG_READY"]
    assert o.deterministic_parallelism == 1


_PEAK_SCRIPT = r"""
import gc, sys
import numpy as np
import fevc_amd as F
from fevc_amd import synth
spec = synth.PanelSpec(workers=200000, firms=10000, periods=5, mobility=0.1, seed=3, name="admission")
d = synth.generate(spec)
cols = [np.ascontiguousarray(d[k], np.float64) for k in ("worker", "firm", "deletion", "outcome", "frequency",
                                                         "target")]
del d
gc.collect()
def status():
    s = dict(line.split(":", 1) for line in open("/proc/self/status"))
    return int(s["VmHWM"].split()[0]) * 1024, int(s["VmRSS"].split()[0]) * 1024
with open("/proc/self/clear_refs", "w") as f:
    f.write("5")
_, before = status()
e = F.Engine(*cols, upload=False)
peak, _ = status()
print(len(cols[0]), peak - before)
"""


def test_preparation_peak_within_admission_charge(tmp_path):
    """The admission charges 768 bytes a row beyond the caller's copy (the
    reference's figure); the host preparation of this engine must stay
    inside it.  Measured as the rise of the process's peak resident set
    (VmHWM, reset through /proc/self/clear_refs) over a 1M-row panel."""
    import subprocess

    if not os.access("/proc/self/clear_refs", os.W_OK):
        pytest.skip("cannot reset the peak resident set here")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, PYTHONPATH=root + os.pathsep + os.environ.get("PYTHONPATH", ""))
    out = subprocess.run([sys.executable, "-c", _PEAK_SCRIPT], capture_output=True, text=True, env=env,
                         timeout=600, check=True).stdout.split()
    rows, rise = int(out[0]), int(out[1])
    print(f"preparation peak rise {rise / rows:.0f} bytes a row over {rows} rows")
    assert rows == 1_000_000
    assert rise <= rows * 768 + 4096
