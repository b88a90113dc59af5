"""The stayer hybrid (stayers(both)) in the oracle: movers keep match deletion,
eligible stayers (workers of one retained firm with at least two physical
observations) are appended and deleted one physical copy at a time
(stayer_hybrid.rs:71-348; generic_jla.rs:2170-2213, 5499-5540, 6019-6053).

Pins:
- the Rust-route exact hybrid (oracle/exact.py, mode="hybrid") against the
  reference's dense Python-oracle form with every stayer copy its own
  deletion unit (fevc/tests/python/oracle.py:97-320), 1e-10;
- the JLA hybrid against the exact hybrid within 12 MCSE + 5e-7, the bar of
  rust/crates/vckss-core/tests/generic_jla.rs:1243-1287;
- an augmentation with no stayer rows reproduces the mover result bit for bit
  (generic_jla.rs tests: zero-stayer hybrid == mover result).
"""
import numpy as np
import pytest

from fevc_amd import synth

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")


def hybrid_panel(seed, workers=400, firms=40):
    d = synth.generate(synth.PanelSpec(workers=workers, firms=firms, periods=4, mobility=0.3, seed=seed,
                                       name="stayer-hybrid"))
    cols = [np.asarray(d[k], float) for k in COLS]
    rng = np.random.default_rng(seed)
    cols[4] = rng.integers(1, 4, cols[0].size).astype(float)
    return cols


def combined_rows(cols, keep, stayers):
    """Stored rows of the hybrid: retained movers, then the stayers."""
    mw, mf, md, my, mq, mt = [c[keep] for c in cols]
    sf, sw, sy, sq, st = stayers
    w = np.concatenate((mw, mw.max() + 1 + sw))
    f = np.concatenate((mf, np.unique(mf)[sf.astype(int) - 1]))
    g = np.concatenate((md, md.max() + 1 + np.arange(sf.size)))
    stay = np.r_[np.zeros(mw.size, bool), np.ones(sf.size, bool)]
    return w, f, g, np.concatenate((my, sy)), np.concatenate((mq, sq)), np.concatenate((mt, st)), stay


def test_exact_hybrid_restatements_agree():
    from oracle import exact as X
    from oracle import oracle as O
    cols = hybrid_panel(2)
    keep = O.Problem(*cols).mask()
    st = synth.eligible_stayers(cols[0], cols[1], cols[3], cols[4], cols[5], keep)
    assert st[0].size > 0
    w, f, g, y, q, t, stay = combined_rows(cols, keep, st)
    ex = X.exact_estimator(w, f, g, y, q, t, mode="hybrid", stayer_rows=stay)
    # every stayer copy its own deletion unit, movers by match
    idx = np.r_[np.flatnonzero(~stay), np.repeat(np.flatnonzero(stay), q[stay].astype(int))]
    q2 = np.r_[q[~stay], np.ones(int(q[stay].sum()))]
    t2 = np.r_[t[~stay], np.repeat(t[stay] / q[stay], q[stay].astype(int))]
    g2 = np.r_[g[~stay], g.max() + 1 + np.arange(int(q[stay].sum()))]
    dk = X.dense_kss(y[idx], w[idx], f[idx], frequency=q2, target=t2, deletion_id=g2)
    for k in ("plugin", "correction", "corrected"):
        assert np.allclose(ex[k], dk[k], rtol=1e-10, atol=1e-13), k
    assert ex["deletion_units"] == np.unique(g[~stay]).size + int(q[stay].sum())


@pytest.mark.parametrize("seed", [5, 8])
def test_hybrid_jla_matches_exact_within_mcse(oracle, seed):
    from oracle import exact as X
    cols = hybrid_panel(seed)
    p = oracle.Problem(*cols)
    keep = p.mask()
    st = synth.eligible_stayers(cols[0], cols[1], cols[3], cols[4], cols[5], keep)
    h = p.with_stayers(*st)
    assert h.units == p.units + st[0].size and h.workers == p.workers + int(st[1].max())
    r = h.run(probes=400, probe_tolerance=1e-11, fit_tolerance=1e-11, leverage_width=16, target_width=16)
    w, f, g, y, q, t, stay = combined_rows(cols, keep, st)
    ex = X.exact_estimator(w, f, g, y, q, t, mode="hybrid", stayer_rows=stay)
    assert np.allclose(r["plugin"], ex["plugin"], rtol=1e-9, atol=1e-12)
    dev = np.abs(np.array(r["correction"]) - ex["correction"])
    assert np.all(dev <= 12.0 * np.array(r["mcse"]) + 5e-7), (dev, r["mcse"])


def test_zero_stayers_reproduce_the_mover_result(oracle):
    cols = hybrid_panel(3)
    p = oracle.Problem(*cols)
    empty = [np.zeros(0) for _ in range(5)]
    h = p.with_stayers(*empty)
    a = p.run(probes=16, probe_tolerance=1e-10)
    b = h.run(probes=16, probe_tolerance=1e-10)
    for k in ("plugin", "correction", "corrected", "mcse"):
        assert np.array_equal(a[k], b[k]), k


def test_hybrid_numerical_scores_are_centred(oracle):
    """Numerical MCSE of the hybrid: mover blocks (whose generic linearisation,
    generic_jla/numerical.rs:300-347 with k = 1/(1 - h), equals the compressed
    unit linearisation) plus stayer copies; the replay reproduces the leverage
    phase, so every fold's scores average to zero, at any replay batch width."""
    cols = hybrid_panel(5)
    p = oracle.Problem(*cols)
    st = synth.eligible_stayers(cols[0], cols[1], cols[3], cols[4], cols[5], p.mask())
    h = p.with_stayers(*st)
    runs = [h.run(probes=48, leverage_width=w, target_width=8, probe_tolerance=1e-11, fit_tolerance=1e-11,
                  numerical=True)["numerical"] for w in (8, 1)]
    n = runs[0]
    assert n["status"] == 2 and n["certified_replay_rhs"] == 48
    scale = np.sqrt(np.abs(np.diag(np.asarray(n["leverage"]).reshape(3, 3))).max() * 48 * 47)
    assert np.max(np.abs(np.asarray(n["score_means"]))) <= 1e-9 * scale
    for k in ("conditional", "leverage", "usable", "mcse"):
        assert np.allclose(np.asarray(runs[1][k]), np.asarray(n[k]), rtol=1e-9, atol=1e-18), k
