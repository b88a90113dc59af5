"""GPU parity of the hot path against the oracle, through the C ABI.

Bars (stated per test):
  * Counter-V1 atoms: bit-exact (integer work);
  * level products A x: bit-exact where a row is summed serially (K = 64,
    at most 64 entries; renumbered rows keep the reference's neighbour
    order), 1e-13 of the row scale where rows are split over threads;
  * one V-cycle: 1e-10 relative (FP64; the coarse means are tree sums on the
    device and the terminal solve is a dense grounded inverse instead of LDL);
  * PCG: certified residual below the requested tolerance, iteration counts
    within one of the oracle's, solution within 1e-6 relative;
  * end to end: the reference KAT at its own 3e-12 tolerance
    (rust/crates/vckss-core/tests/rust_mata_shared_atoms.rs:39-294), and a
    synthetic panel against the oracle engine at 1e-8 relative on components
    with tight solves, PCG-tolerance agreement at the default tolerance.
Only one engine context is live at a time (the reference registry keeps one),
so every test prepares its own.
"""
import functools

import numpy as np
import pytest

from conftest import load_golden

import fevc_amd as F
from fevc_amd import synth

pytestmark = pytest.mark.gpu

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")
TAG_LEVERAGE = 0x4C45564552414745
TAG_TARGET = 0x0054415247455401
SPEC = synth.PanelSpec(workers=4000, firms=300, periods=5, mobility=0.2, seed=11, name="gpu-parity")


@functools.lru_cache(maxsize=None)
def panel(heavy_frequency=False):
    d = synth.generate(SPEC)
    cols = [np.asarray(d[k]) for k in COLS]
    if heavy_frequency:
        rng = np.random.default_rng(7)
        cols[4] = rng.integers(1, 300, len(cols[0])).astype(np.uint64)
    return tuple(cols)


def relerr(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("heavy", [False, True])
def test_atoms_bit_exact(oracle, heavy):
    cols = panel(heavy)
    p = oracle.Problem(*cols)
    e = F.Engine(*cols)
    try:
        domains = [(0, TAG_LEVERAGE, p.array("unit_rank"), p.array("unit_count")),
                   (1, TAG_TARGET, p.array("stratum_rank"), p.array("stratum_count"))]
        if heavy:
            assert p.array("unit_count").max() > 64  # multi-word atoms
        for domain, tag, entity, trials in domains:
            for seed, first, columns in [(8675309, 0, 1), (8675309, 3, 7), (42, 64, 33), (1, 5, 64)]:
                got = e.atoms(seed, domain, first, columns)
                want = oracle.rademacher_fill(seed, tag, first, columns, entity, trials)
                assert np.array_equal(got, want), (domain, seed, first, columns)
    finally:
        e.release()


def _hybrid(oracle, cols):
    p = oracle.Problem(*cols)
    g, aux = p.hybrid_graph()
    return p, g


def test_level_products(oracle):
    # Rows of at most 64 entries at K = 64 are summed serially in the
    # reference's neighbour order: bit-exact.  Other rows are split over
    # slices/chunks: rounding-level agreement, 1e-13 of the row's scale.
    cols = panel(heavy_frequency=True)
    p, g = _hybrid(oracle, cols)
    e = F.Engine(*cols)
    try:
        levels = oracle.Cmg(g, direct_threshold=700, max_levels=128).levels()
        assert len(levels) == e.dims["levels"]
        rng = np.random.default_rng(3)
        for li, lv in enumerate(levels):
            gl = oracle.Graph(lv["n"], lv["u"], lv["v"], lv["w"])
            deg = np.bincount(lv["u"], minlength=lv["n"]) + np.bincount(lv["v"], minlength=lv["n"])
            for K in (1, 4, 32, 64):
                X = rng.standard_normal((lv["n"], K))
                Y = e.spmv(li, X.ravel(), K).reshape(lv["n"], K)
                for j in range(K):
                    want = gl.matvec(X[:, j])
                    scale = lv["diag"] * (np.abs(X[:, j]) + np.abs(X[:, j]).max())
                    assert np.all(np.abs(Y[:, j] - want) <= 1e-13 * scale + 1e-300), (li, K, j)
                    if K == 64:
                        short = deg <= 64
                        assert np.array_equal(Y[short, j].view(np.uint64), want[short].view(np.uint64)), (li, j)
    finally:
        e.release()


@pytest.mark.parametrize("K", [1, 8, 64])
def test_vcycle_matches_oracle(oracle, K):
    cols = panel()
    p, g = _hybrid(oracle, cols)
    e = F.Engine(*cols)
    try:
        cmg = oracle.Cmg(g, direct_threshold=700, max_levels=128)
        rng = np.random.default_rng(K)
        R = rng.standard_normal((g.n, K))
        R -= R.mean(axis=0)
        Z = e.vcycle(R.ravel(), K).reshape(g.n, K)
        Z32 = e.vcycle(R.ravel(), K, precision=32).reshape(g.n, K)
        for j in range(K):
            want = cmg.apply(R[:, j])
            assert relerr(Z[:, j], want) < 1e-10, j
            assert relerr(Z32[:, j], want) < 1e-4, j
    finally:
        e.release()


@pytest.mark.parametrize("precision", [64, 32])
def test_pcg_certified(oracle, precision):
    cols = panel()
    p, g = _hybrid(oracle, cols)
    e = F.Engine(*cols)
    try:
        cmg = oracle.Cmg(g, direct_threshold=700, max_levels=128)
        K, tol = 16, 1e-10
        rng = np.random.default_rng(5)
        B = rng.standard_normal((g.n, K))
        B -= B.mean(axis=0)
        X, iters = e.pcg(B.ravel(), K, tolerance=tol, precision=precision)
        X = X.reshape(g.n, K)
        _, _, _, diag = g.export()
        bound = 2.0 * diag.max()
        for j in range(K):
            r = B[:, j] - g.matvec(X[:, j])
            assert np.linalg.norm(r) <= 1.01 * tol * (np.linalg.norm(B[:, j]) + bound * np.linalg.norm(X[:, j]))
            xr, d = cmg.solve(B[:, j], rel=tol, max_iterations=1000, recompute=25)
            assert abs(int(iters[j]) - int(d.iterations)) <= (1 if precision == 64 else 3), j
            assert relerr(X[:, j] - X[:, j].mean(), xr - xr.mean()) < 1e-6
    finally:
        e.release()


# ------------------------------------------------------------------ end to end

def _kat():
    K = load_golden("shared_atoms_kat.json")
    inp = K["input"]
    cols = [np.array([inp[k][r] for r in inp["order"]], dtype=np.float64) for k in COLS]
    return K, cols


@pytest.mark.parametrize("widths", [(32, 16), (2, 2), (1, 1)])
def test_kat_end_to_end(widths):
    K, cols = _kat()
    o = K["options"]
    e = F.Engine(*cols)
    try:
        e.configure(64, *widths)
        res = e.solve(seed=o["seed"], probes=o["probes"], tolerance=o["pcg_tolerance"],
                      fit_tolerance=o["pcg_tolerance"], maximum_iterations=o["maximum_iterations"],
                      residual_replacement_interval=o["residual_replacement_interval"])
        tol = K["tolerance"]
        for key in ("plugin", "correction", "corrected"):
            assert np.max(np.abs(np.asarray(res[key]) - K[key])) < tol, key
        assert np.max(np.abs(np.asarray(res["numerical_mcse"]) - K["numerical_mcse"])) < tol
        arrays = e.arrays(o["probes"])
        for name, key in [("fitted_cell", "fitted_cell"), ("projection_share", "unit_projection_share"),
                          ("residual_share", "unit_residual_share"), ("finite_bias", "unit_finite_bias"),
                          ("finite_variance", "unit_finite_variance"), ("residual_mass", "unit_residual_mass"),
                          ("deleted_mass", "unit_deleted_mass"),
                          ("cell_correction_weight", "cell_correction_weight"),
                          ("target_draws", "target_draws")]:
            assert np.max(np.abs(arrays[name] - np.asarray(K[key]))) < tol, name
    finally:
        e.release()


def _compare_engine(oracle, cols, probes, tolerance, rel_components, abs_arrays, precision=64,
                    widths=(32, 16), chunk=None):
    p = oracle.Problem(*cols)
    want = p.run(probes=probes, fit_tolerance=1e-10,
                 probe_tolerance=tolerance if tolerance is not None else 1e-6)
    e = F.Engine(*cols)
    try:
        e.configure(precision, *widths)
        if chunk is not None:
            e.set_chunk(chunk)
        got = e.solve(probes=probes, tolerance=tolerance)
        arrays = e.arrays(probes)
        for key, okey in (("plugin", "plugin"), ("correction", "correction"), ("corrected", "corrected")):
            g, w = np.asarray(got[key]), np.asarray(want[okey])
            assert np.all(np.abs(g - w) <= rel_components * np.maximum(np.abs(w), 1e-3)), (key, g, w)
        for name in ("fitted_cell", "projection_share", "residual_share", "cell_correction_weight"):
            scale = max(np.max(np.abs(want[name])), 1.0)
            assert np.max(np.abs(arrays[name] - want[name])) <= abs_arrays * scale, name
        assert np.max(arrays["leverage_complete_residual"]) <= max(10 * (tolerance or 1e-6), 1e-11)
        return got, want, arrays
    finally:
        e.release()


def test_panel_tight_tolerance(oracle):
    _compare_engine(oracle, list(panel()), probes=24, tolerance=1e-11, rel_components=1e-8,
                    abs_arrays=1e-8)


def test_panel_default_tolerance(oracle):
    # default probe tolerance 1e-6 (full_cmg.rs:29): agreement at the solve tolerance
    _compare_engine(oracle, list(panel()), probes=24, tolerance=None, rel_components=1e-4,
                    abs_arrays=1e-4)


def test_panel_fp32_vcycle(oracle):
    # FP32 V-cycle, FP64 Krylov: same certified tolerance, same answers
    _compare_engine(oracle, list(panel()), probes=16, tolerance=1e-11, rel_components=1e-8,
                    abs_arrays=1e-8, precision=32)


def test_panel_width_invariance(oracle):
    a, _, _ = _compare_engine(oracle, list(panel()), probes=12, tolerance=1e-11, rel_components=1e-8,
                              abs_arrays=1e-8, widths=(5, 3))
    b, _, _ = _compare_engine(oracle, list(panel()), probes=12, tolerance=1e-11, rel_components=1e-8,
                              abs_arrays=1e-8, widths=(64, 32))
    for key in ("plugin", "correction", "corrected"):
        assert np.allclose(a[key], b[key], rtol=1e-8, atol=0)


def test_heavy_frequency_panel(oracle):
    _compare_engine(oracle, list(panel(True)), probes=8, tolerance=1e-11, rel_components=1e-8,
                    abs_arrays=1e-8)


@pytest.mark.parametrize("chunk", [16, 128])
def test_panel_chunk_sizes(oracle, chunk):
    # the segment chunk size only reorders sums: same answers at the same tolerance
    _compare_engine(oracle, list(panel(True)), probes=8, tolerance=1e-11, rel_components=1e-8,
                    abs_arrays=1e-8, precision=32, widths=(64, 32), chunk=chunk)


def test_autotune_then_solve(oracle):
    cols = list(panel())
    e = F.Engine(*cols)
    try:
        t = e.autotune(precision=32, tolerance=1e-8, batches=1)
        assert t["width"] in (16, 32, 64) and t["chunk"] in (32, 64, 128)
        assert len(t["trials"]) == 9 and all(r > 0 for _, _, r in t["trials"])
        rates = {(int(w), int(c)): r for w, c, r in t["trials"]}
        assert t["rate"] == rates[(t["width"], t["chunk"])]
        assert t["rate"] == max(rates.values()) or (t["width"], t["chunk"]) == (64, 64)
        got = e.solve(probes=16, tolerance=1e-11)
    finally:
        e.release()
    want = oracle.Problem(*cols).run(probes=16, fit_tolerance=1e-10, probe_tolerance=1e-11)
    for key in ("plugin", "correction", "corrected"):
        g, w = np.asarray(got[key]), np.asarray(want[key])
        assert np.all(np.abs(g - w) <= 1e-8 * np.maximum(np.abs(w), 1e-3)), (key, g, w)


def test_error_paths():
    K, cols = _kat()
    e = F.Engine(*cols)
    with pytest.raises(F.EngineError) as ei:
        e.configure(16, 32, 16)
    assert ei.value.code == 20
    with pytest.raises(F.EngineError) as ei:
        e.solve(probes=1)
    assert ei.value.code == 20
    g = e.generation
    e.release()
    r = F.ResultV1()
    import ctypes as C
    assert F.library().vckss_rust_engine_result_v1(g, C.byref(r), C.sizeof(r)) == 80
