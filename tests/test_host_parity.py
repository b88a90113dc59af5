"""Host side of the product against the oracle (no GPU needed).

* every entry point declared in include/*.h is exported by the library;
* preparation (selection, compression, JLA plan) reproduces the oracle's
  dimensions, retained mask and plan arrays exactly, on the reference KAT
  (rust/crates/vckss-core/tests/rust_mata_shared_atoms.rs:39-294) and on
  synthetic panels;
* the host CMG hierarchy is bit-identical to the oracle's, level by level
  (graph, diagonal, aggregation labels, repeat counts, terminal reason), on
  chains, grids, random graphs and the degree-4 hybrid graphs of panels.
"""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, load_golden

F = pytest.importorskip("fevc_amd")


def _declared():
    names = []
    for h in ("vckss_mi355x.h", "vckss_mi355x_ext.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(v\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    lib = F.library()
    declared = _declared()
    assert len(declared) >= 29
    bound = {name for name, _, _ in F.SIGNATURES}
    for name in declared:
        assert hasattr(lib, name), name
        assert name in bound, f"{name} has no Python binding"


def test_abi_identity():
    lib = F.library()
    assert lib.vckss_rust_abi_version() == 1
    assert b"CMG_FULL_V2" in lib.vckss_rust_capabilities_json()
    req = F.default_solve_request()
    assert req.options.probes == 200 and req.options.seed == 8675309
    assert req.options.pcg_tolerance == 1e-10
    assert req.options.maximum_iterations == 10000
    assert req.options.residual_replacement_interval == 100


def kat_columns():
    K = load_golden("shared_atoms_kat.json")
    inp = K["input"]
    return K, [np.array([inp[k][r] for r in inp["order"]]) for k in
               ("worker", "firm", "deletion", "outcome", "frequency", "target")]


def _compare_prepared(oracle, cols, observation=False):
    p = oracle.Problem(*cols, observation=observation)
    e = F.Engine(*cols, upload=False, deletion_mode="observation" if observation else "match")
    try:
        d = e.dims
        assert (d["input_rows"], d["rows"], d["physical"], d["workers"], d["firms"], d["cells"],
                d["units"], d["exact_strata"], d["strata"]) == (
            p.input_rows, p.rows, p.physical, p.workers, p.firms, p.cells, p.units,
            p.exact_strata, p.strata)
        assert d["topology_checksum"] == p.topology_checksum
        np.testing.assert_array_equal(e.mask(), p.mask())
        for which, name, n in [(20, "unit_rank", p.units), (21, "unit_count", p.units),
                               (22, "stratum_rank", p.strata), (23, "stratum_count", p.strata),
                               (24, "cell_weight", p.cells), (25, "unit_cell", p.units),
                               (26, "unit_offset", p.units)]:
            got = e.array(which, n)
            want = p.array(name).astype(np.float64)
            assert np.array_equal(got, want), name
        g, aux = p.hybrid_graph()
        assert (d["hybrid_vertices"], d["hybrid_edges"], d["auxiliary"]) == (g.n, g.m, aux)
        return p, e
    except Exception:
        e.release()
        raise


def test_prepare_kat(oracle):
    K, cols = kat_columns()
    p, e = _compare_prepared(oracle, cols)
    e.release()


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_prepare_synthetic_panels(oracle, seed):
    from fevc_amd import synth
    spec = synth.PanelSpec(workers=3000, firms=150, periods=4, mobility=0.15, seed=seed,
                           name=f"p{seed}")
    data = synth.generate(spec)
    cols = [data[k] for k in ("worker", "firm", "deletion", "outcome", "frequency", "target")]
    p, e = _compare_prepared(oracle, cols)
    assert e.dims["cells"] > 1000
    e.release()


@pytest.mark.parametrize("seed", [1, 2])
def test_prepare_observation_plan(oracle, seed):
    """Observation deletion: one unit per stored row with its class rank
    (generic_jla.rs:4267-4303) and copy offset; strata keyed by the least
    observation rank (:3902-3995).  Duplicated rows share a class and take
    consecutive copies."""
    from fevc_amd import synth
    spec = synth.PanelSpec(workers=800, firms=60, periods=4, mobility=0.3, seed=seed, name=f"o{seed}")
    data = synth.generate(spec)
    cols = [np.asarray(data[k], float) for k in ("worker", "firm", "deletion", "outcome", "frequency", "target")]
    rng = np.random.default_rng(seed)
    cols[4] = rng.integers(1, 4, cols[0].size).astype(float)
    cols[5] = cols[4] * rng.uniform(0.5, 1.5, cols[0].size)
    # duplicate some rows (same worker, firm, outcome and per-copy target mass)
    dup = rng.choice(cols[0].size, 40, replace=False)
    cols = [np.concatenate((c, c[dup])) for c in cols]
    cols[2][-40:] = cols[2].max() + 1 + np.arange(40)  # their own deletion identifiers
    p, e = _compare_prepared(oracle, cols, observation=True)
    try:
        assert e.dims["units"] == e.dims["rows"]
        # the device's cell -> unit gather lists every stored row once, cells ascending
        gather = e.array(27, e.dims["units"]).astype(np.int64)
        assert np.array_equal(np.sort(gather), np.arange(e.dims["units"]))
        assert np.all(np.diff(e.array(25, e.dims["units"])[gather]) >= 0)
        offsets = e.array(26, e.dims["units"])
        assert offsets.max() >= 1  # a duplicated row continues its class's copies
    finally:
        e.release()


def test_stayer_hybrid_plan(oracle):
    """Stayer augmentation (stayer_hybrid.rs:71-348): the host's combined
    problem and hybrid plan (mover match units with their semantic entities,
    stayer rows as observation units with class ranks and copy offsets,
    strata) equal the oracle's."""
    from fevc_amd import synth
    spec = synth.PanelSpec(workers=800, firms=60, periods=4, mobility=0.3, seed=4, name="stayers")
    data = synth.generate(spec)
    cols = [np.asarray(data[k], float) for k in ("worker", "firm", "deletion", "outcome", "frequency", "target")]
    cols[4] = np.random.default_rng(4).integers(1, 4, cols[0].size).astype(float)
    p = oracle.Problem(*cols)
    st = synth.eligible_stayers(cols[0], cols[1], cols[3], cols[4], cols[5], p.mask())
    h = p.with_stayers(*st)
    e = F.Engine(*cols, upload=False)
    try:
        rec = e.augment_stayers(*st)
        d = e.dims
        assert (d["rows"], d["physical"], d["workers"], d["firms"], d["cells"], d["units"], d["strata"]) == (
            h.rows, h.physical, h.workers, h.firms, h.cells, h.units, h.strata)
        assert d["topology_checksum"] == h.topology_checksum
        for which, name, n in [(20, "unit_rank", h.units), (21, "unit_count", h.units),
                               (22, "stratum_rank", h.strata), (23, "stratum_count", h.strata),
                               (24, "cell_weight", h.cells), (25, "unit_cell", h.units),
                               (26, "unit_offset", h.units)]:
            assert np.array_equal(e.array(which, n), h.array(name).astype(np.float64)), name
        gather = e.array(27, h.units).astype(np.int64)
        assert np.array_equal(np.sort(gather), np.arange(h.units))
        assert (rec["stayer_stored_rows"], rec["stayer_workers"], rec["combined_workers"]) == (
            st[0].size, int(st[1].max()), h.workers)
        assert rec["stayer_physical_mass"] == int(st[3].sum())
        assert rec["combined_deletion_units"] == p.units + int(st[3].sum())
        assert abs(rec["combined_target_mass"] - rec["mover_target_mass"] - rec["stayer_target_mass"]) < 1e-9
        # ffi_engine.rs:4840-4861: mover-only solves are refused once augmented
        e.stayers_mode = 1
        with pytest.raises(F.EngineError) as ei:
            e.solve(probes=8)
        assert ei.value.code == 20
        # context.rs:385-404: a second augmentation needs the Prepared state
        with pytest.raises(F.EngineError) as ei:
            e.augment_stayers(*st)
        assert ei.value.code in (81,)
    finally:
        e.release()


def test_stayer_augmentation_refusals():
    from fevc_amd import synth
    spec = synth.PanelSpec(workers=400, firms=40, periods=4, mobility=0.3, seed=6, name="stayers-bad")
    data = synth.generate(spec)
    cols = [np.asarray(data[k], float) for k in ("worker", "firm", "deletion", "outcome", "frequency", "target")]
    good = [np.array([1.0, 1.0]), np.array([1.0, 1.0]), np.array([0.5, 0.7]), np.array([1.0, 1.0]),
            np.array([1.0, 1.0])]
    cases = [
        ([np.array([0.0, 1.0])] + good[1:], 21),                     # firm not a retained mover firm
        ([np.array([1.0, 2.0])] + good[1:], 21),                     # one stayer in two firms
        (good[:1] + [np.array([1.0, 3.0])] + good[2:], 21),          # stayer workers not dense
        ([good[0][:1], good[1][:1], good[2][:1], good[3][:1], good[4][:1]], 20),  # one physical observation
    ]
    for stayers, code in cases:
        e = F.Engine(*cols, upload=False)
        try:
            with pytest.raises(F.EngineError) as ei:
                e.augment_stayers(*stayers)
            assert ei.value.code == code, (stayers, ei.value)
        finally:
            e.release()
    # stayers-all without an augmentation (ffi_engine.rs:4850-4855)
    e = F.Engine(*cols, upload=False)
    try:
        e.stayers_mode = 2
        with pytest.raises(F.EngineError) as ei:
            e.solve(probes=8)
        assert ei.value.code == 20
    finally:
        e.release()
    # the hybrid needs a match-deletion mover generation (session_retained.rs:866-872)
    e = F.Engine(*cols, upload=False, deletion_mode="observation")
    try:
        with pytest.raises(F.EngineError) as ei:
            e.augment_stayers(*good)
        assert ei.value.code == 11
    finally:
        e.release()


def test_prepare_refuses_bad_input():
    K, cols = kat_columns()
    bad = [c.astype(np.float64) for c in cols]
    bad[0][3] = 0.5  # non-integer worker id
    with pytest.raises(F.EngineError) as ei:
        F.Engine(*bad, upload=False)
    assert ei.value.code == 21  # InvalidIdentifier, rust/crates/vckss-core/src/error.rs:14


# ------------------------------------------------------------------ hierarchy

def _assert_same_hierarchy(oracle, n, u, v, w, max_levels=128):
    g = oracle.Graph(n, u, v, w)
    ref = oracle.Cmg(g, direct_threshold=700, max_levels=max_levels).levels()
    got, _ = F.build_hierarchy(n, u, v, w, direct_threshold=700, max_levels=max_levels)
    assert len(got) == len(ref)
    for i, (a, b) in enumerate(zip(got, ref)):
        for key in ("n", "m", "nnz", "repeat", "terminal", "coarse"):
            assert a[key] == b[key], (i, key, a[key], b[key])
        for key in ("u", "v"):
            assert np.array_equal(a[key], b[key]), (i, key)
        for key in ("w", "diag"):
            assert np.array_equal(a[key].view(np.uint64), b[key].view(np.uint64)), (i, key)
        if a["coarse"]:
            assert np.array_equal(a["labels"], b["labels"]), (i, "labels")
    return got


def test_hierarchy_chain(oracle):
    n = 5000
    u = np.arange(n - 1)
    _assert_same_hierarchy(oracle, n, u, u + 1, 1.0 + (u % 7) * 0.25)


def test_hierarchy_grid(oracle):
    s = 120
    idx = np.arange(s * s).reshape(s, s)
    u = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    v = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    w = 1.0 + ((u * 7 + v * 3) % 11) / 10.0
    levels = _assert_same_hierarchy(oracle, s * s, u, v, w)
    assert len(levels) >= 2


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_hierarchy_random(oracle, seed):
    rng = np.random.default_rng(seed)
    n = 4000
    u = rng.integers(0, n, 3 * n)
    v = rng.integers(0, n, 3 * n)
    keep = u != v
    u, v = u[keep], v[keep]
    # two components: a path over the first half, random edges stay as drawn
    path = np.arange(n // 2 - 1)
    u = np.concatenate([u, path])
    v = np.concatenate([v, path + 1])
    w = rng.uniform(0.1, 3.0, len(u))
    _assert_same_hierarchy(oracle, n, u, v, w)


def test_hierarchy_max_levels(oracle):
    n = 3000
    u = np.arange(n - 1)
    _assert_same_hierarchy(oracle, n, u, u + 1, np.ones(n - 1), max_levels=2)


@pytest.mark.parametrize("seed", [5, 6])
def test_hierarchy_hybrid_graph(oracle, seed):
    from fevc_amd import synth
    spec = synth.PanelSpec(workers=20000, firms=1500, periods=5, mobility=0.2, seed=seed,
                           name=f"h{seed}")
    data = synth.generate(spec)
    cols = [data[k] for k in ("worker", "firm", "deletion", "outcome", "frequency", "target")]
    p = oracle.Problem(*cols)
    g, aux = p.hybrid_graph()
    u, v, w, _ = g.export()
    levels = _assert_same_hierarchy(oracle, g.n, u, v, w)
    assert len(levels) >= 2


def test_full_reference_abi_is_exported():
    """Every function of the reference header links; the ones outside the
    path are typed refusals (status 11 with a message, schema queries 0)."""
    import ctypes as C
    lib = F.library()
    names = load_golden("abi_symbols.json")["functions"]
    assert len(names) == 147
    for f in names:
        assert hasattr(lib, f["name"]), f["name"]
    implemented = {name for name, _, _ in F.SIGNATURES}
    refused = [f for f in names if f["name"] not in implemented and f["returns"] == "int32_t"]
    # the Stata shim's own call sequence (vckss_plugin.c) is implemented
    for name in ("vckss_rust_backend_request_capability_v3", "vckss_rust_engine_admit_prepare_v4",
                 "vckss_rust_engine_prepare_interrupt_v4", "vckss_rust_engine_preparation_receipt_v4",
                 "vckss_rust_engine_solve_interrupt_v5", "vckss_rust_engine_memory_forecast_v1",
                 "vckss_rust_engine_result_v1", "vckss_rust_engine_performance_receipt_v1",
                 "vckss_rust_engine_detailed_receipt_v6", "vckss_rust_engine_detailed_receipt_v7",
                 "vckss_rust_engine_rhs_receipts_v1", "vckss_rust_engine_full_cmg_receipt_v1"):
        assert name in implemented, name
    assert len(refused) + len(implemented & {f["name"] for f in names}) >= 140
    fn = getattr(lib, refused[0]["name"])
    fn.restype = C.c_int32
    assert fn() == 11
    assert refused[0]["name"].encode() in lib.vckss_rust_last_error()
    schema = [f for f in names if f["returns"] == "uint32_t" and f["name"] != "vckss_rust_abi_version"]
    g = getattr(lib, schema[0]["name"])
    g.restype = C.c_uint32
    assert g() == 0


def test_numerical_preflight_validation():
    """numerical_api.rs:113-196, 237-279: schema and reserved fields are ABI
    errors (10); projection, component, control and non-full-CMG requests are
    unsupported (11); one probe is invalid input (20)."""
    import ctypes as C
    lib = F.library()
    assert lib.vckss_rust_numerical_schema_v1() == 1 and lib.vckss_rust_numerical_schema_v2() == 2

    def status(**changes):
        r = F.default_numerical_request(probes=16, tolerance=1e-8)
        for path, value in changes.items():
            obj = r
            parts = path.split(".")
            for part in parts[:-1]:
                obj = getattr(obj, part)
            setattr(obj, parts[-1], value)
        return lib.vckss_rust_engine_numerical_preflight_v2(C.byref(r))

    assert status() == 0
    assert status(**{"v1.schema_version": 1}) == 10
    assert status(**{"v1.reserved": 1}) == 10
    assert status(reserved=1) == 10
    assert status(full_cmg=2) == 10
    assert status(**{"v1.projection_requested": 1}) == 11
    assert status(**{"v1.controls_count": 1}) == 11
    assert status(full_cmg=0) == 11
    assert status(**{"v1.point.probes": 1}) == 20
    assert status(**{"v1.threads": 0}) == 20
    assert status(**{"v1.point.deletion_mode": 0}) == 11
