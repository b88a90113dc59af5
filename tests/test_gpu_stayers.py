"""The stayer hybrid (stayers(both)) on the GPU, through the C ABI, against the
oracle (its pins are in tests/test_oracle_stayers.py): movers keep match
deletion, eligible stayers are appended by vckss_rust_engine_augment_stayers
and deleted one physical copy at a time (stayer_hybrid.rs:71-348,
generic_jla.rs:2170-2213, 5499-5540).

Bars: components within 1e-8 relative and per-unit arrays within 1e-8 of
their scale with tight solves (the JLA route), 1e-9 for the exact route; an
augmentation without stayer rows reproduces the mover result bit for bit.
"""
import functools

import numpy as np
import pytest

import fevc_amd as F
from fevc_amd import synth

pytestmark = pytest.mark.gpu

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")
SPEC = synth.PanelSpec(workers=4000, firms=300, periods=5, mobility=0.2, seed=17, name="gpu-stayers")


@functools.lru_cache(maxsize=None)
def panel():
    d = synth.generate(SPEC)
    cols = [np.asarray(d[k], float) for k in COLS]
    cols[4] = np.random.default_rng(17).integers(1, 4, cols[0].size).astype(float)
    return tuple(cols)


@functools.lru_cache(maxsize=None)
def stayers():
    from oracle import oracle as O
    cols = panel()
    keep = O.Problem(*cols).mask()
    return synth.eligible_stayers(cols[0], cols[1], cols[3], cols[4], cols[5], keep)


def _engine(precision=64, widths=(32, 16), with_stayers=True):
    e = F.Engine(*panel())
    e.configure(precision, *widths)
    if with_stayers:
        e.augment_stayers(*stayers())
    return e


@pytest.mark.parametrize("precision,widths", [(64, (32, 16)), (32, (64, 32)), (64, (5, 3))])
def test_hybrid_matches_oracle(oracle, precision, widths):
    p = oracle.Problem(*panel())
    h = p.with_stayers(*stayers())
    want = h.run(probes=16, fit_tolerance=1e-10, probe_tolerance=1e-11)
    e = _engine(precision, widths)
    try:
        print(f"hybrid: movers {p.workers} workers, stayers {int(stayers()[1].max())} workers, "
              f"{h.units} units; precision {precision}, widths {widths}", flush=True)
        got = e.solve(probes=16, tolerance=1e-11)
        arrays = e.arrays(16)
    finally:
        e.release()
    for key in ("plugin", "correction", "corrected"):
        g, w = np.asarray(got[key]), np.asarray(want[key])
        assert np.all(np.abs(g - w) <= 1e-8 * np.maximum(np.abs(w), 1e-3)), (key, g, w)
    for name in ("fitted_cell", "projection_share", "residual_share", "finite_bias", "finite_variance",
                 "residual_mass", "deleted_mass", "cell_correction_weight"):
        scale = max(np.max(np.abs(want[name])), 1.0)
        assert np.max(np.abs(arrays[name] - want[name])) <= 1e-8 * scale, name


def test_hybrid_exact_matches_oracle():
    from oracle import exact as X
    from oracle import oracle as O
    d = synth.generate(synth.PanelSpec(workers=600, firms=60, periods=5, mobility=0.3, seed=5, name="exact-stayers"))
    cols = [np.asarray(d[k], float) for k in COLS]
    keep = O.Problem(*cols).mask()
    st = synth.eligible_stayers(cols[0], cols[1], cols[3], cols[4], cols[5], keep)
    mw, mf, md, my, mq, mt = [c[keep] for c in cols]
    sf, sw, sy, sq, stt = st
    w = np.concatenate((mw, mw.max() + 1 + sw))
    f = np.concatenate((mf, np.unique(mf)[sf.astype(int) - 1]))
    g = np.concatenate((md, md.max() + 1 + np.arange(sf.size)))
    stay = np.r_[np.zeros(mw.size, bool), np.ones(sf.size, bool)]
    want = X.exact_estimator(w, f, g, np.concatenate((my, sy)), np.concatenate((mq, sq)),
                             np.concatenate((mt, stt)), mode="hybrid", stayer_rows=stay)
    e = F.Engine(*cols)
    try:
        e.augment_stayers(*st)
        res, rec = e.solve_exact(exact_limit=10000)
        split = e.stayer_hybrid_result()  # ffi_engine.rs:8772-8840
    finally:
        e.release()
    for k, wk in (("mover_correction", "correction_movers"), ("stayer_correction", "correction_stayers")):
        w3 = np.asarray(want[wk])[:3]
        assert np.all(np.abs(split[k][:3] - w3) <= 1e-9 * np.abs(w3) + 1e-12), (k, split[k], w3)
    assert np.allclose(split["mover_correction"] + split["stayer_correction"], split["correction"], rtol=0,
                       atol=1e-10)
    assert split["deletion_units"] == want["deletion_units"] and split["schema_version"] == 1
    for k in ("plugin", "correction", "corrected"):
        got, w_ = np.asarray(res[k], float), np.asarray(want[k])
        assert np.all(np.abs(got - w_) <= 1e-9 * np.abs(w_) + 1e-12), (k, got, w_)
    assert rec["deletion_units"] == want["deletion_units"]
    assert abs(rec["max_leverage"] - want["max_leverage"]) <= 1e-10


def test_zero_stayers_reproduce_movers():
    a = _engine(with_stayers=False)
    try:
        ra = a.solve(probes=16, tolerance=1e-10)
    finally:
        a.release()
    b = _engine(with_stayers=False)
    try:
        b.augment_stayers(*[np.zeros(0) for _ in range(5)])
        rb = b.solve(probes=16, tolerance=1e-10)
    finally:
        b.release()
    for k in ("plugin", "correction", "corrected", "numerical_mcse"):
        assert np.array_equal(np.asarray(ra[k]), np.asarray(rb[k])), k


def test_hybrid_numerical_matches_oracle(oracle):
    """Numerical MCSE of the hybrid (mover blocks and stayer copies) against the
    oracle, with the bars of tests/test_gpu_numerical.py."""
    from test_gpu_numerical import _compare
    R, tol = 24, 1e-11
    h = oracle.Problem(*panel()).with_stayers(*stayers())
    want = h.run(probes=R, leverage_width=4, target_width=4, fit_tolerance=1e-10, probe_tolerance=tol,
                 numerical=True)
    e = _engine(64, (64, 32))
    try:
        e.smoother(0)
        res, got = e.solve_numerical(probes=R, tolerance=tol)
    finally:
        e.release()
    _compare(want, got, 1e-6)
    for key in ("plugin", "correction", "corrected"):
        assert np.allclose(res[key], want[key], rtol=1e-8, atol=0)
