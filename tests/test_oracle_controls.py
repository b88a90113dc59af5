"""Pins of the oracle's fixed-offset nuisance with controls
(oracle/controls.py; generic_jla.rs:1955-2060, fevc/tests/python/
oracle.py:242-275).

- The fitted control index of the FWL fit (CMG_FULL_V2 solves of the
  controls and the outcome, pseudo-inverse of the control Schur complement)
  against a dense weighted least-squares fit of the full joint design.
- A rank-deficient control set (a duplicated column, or a column that is a
  firm indicator, absorbed by the firm effects) is refused with
  SINGULAR_INFORMATION (82).
- The exact estimator on the working outcome against the dense oracle's
  fixed-offset estimate (the reference's own dense form) on the same rows.
- The JLA estimate on the working outcome within Monte Carlo error of it.
"""
import numpy as np
import pytest

from fevc_amd import synth

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")


def _panel(seed, workers=600, firms=60):
    spec = synth.PanelSpec(workers=workers, firms=firms, periods=5, mobility=0.3, seed=seed, name="controls")
    d = synth.generate(spec)
    z, y = synth.year_controls(spec, d, effect=0.3)
    cols = [np.asarray(d[k]) for k in COLS]
    cols[3] = y
    rng = np.random.default_rng(seed + 50)
    cols[4] = rng.integers(1, 4, len(y)).astype(np.uint64)
    return cols, z


@pytest.mark.parametrize("seed", [3, 8])
def test_fitted_index_matches_dense_joint_fit(oracle, seed):
    from oracle import controls as X
    cols, z = _panel(seed)
    fo = X.fixed_offset(cols, z, fit_tolerance=1e-12)
    want = X.dense_joint_index(cols, z)
    got = fo["index"][fo["mask"]]
    assert np.max(np.abs(got - want)) <= 1e-9 * max(np.abs(want).max(), 1.0)
    assert np.max(np.abs(fo["beta"])) > 0.05  # the year effects are found


def test_rank_deficient_controls_are_refused(oracle):
    """A duplicated control (singular raw Gram) and a firm indicator (absorbed
    by the firm effects: singular residualised information) are refused with
    SINGULAR_INFORMATION, as model_solver.rs:1799-1870 refuses them."""
    from oracle import controls as X
    cols, z = _panel(5)
    firm_one = (np.asarray(cols[1]) == np.asarray(cols[1])[0]).astype(float)
    for wide in (np.column_stack([z, z[:, 2]]), np.column_stack([z, firm_one])):
        with pytest.raises(X.ControlError) as err:
            X.fixed_offset(cols, wide, fit_tolerance=1e-12)
        assert err.value.code == X.SINGULAR_INFORMATION
    # a control that is not collinear stays accepted
    rng = np.random.default_rng(1)
    X.fixed_offset(cols, np.column_stack([z, rng.normal(size=len(firm_one))]), fit_tolerance=1e-12)


def test_exact_on_working_outcome_matches_dense_fixed_offset(oracle):
    from oracle import controls as X
    from oracle import exact as E
    cols, z = _panel(4, workers=300, firms=30)
    fo = X.fixed_offset(cols, z, fit_tolerance=1e-12)
    m = fo["mask"]
    w, f, g, _, fr, t = [np.asarray(c)[m] for c in cols]
    ours = E.exact_estimator(w, f, g, fo["outcome"][m], fr, t)
    dense = E.dense_kss(np.asarray(cols[3])[m], w, f, controls=z[m], frequency=fr, target=t,
                        nuisance="fixedoffset")
    for k in ("plugin", "correction", "corrected"):
        a, b = np.asarray(ours[k]), np.asarray(dense[k])
        assert np.all(np.abs(a - b) <= 1e-8 * np.maximum(np.abs(b), 1e-6)), (k, a, b)


def test_jla_on_working_outcome_within_mcse_of_dense(oracle):
    from oracle import controls as X
    from oracle import exact as E
    cols, z = _panel(6, workers=400, firms=40)
    fo = X.fixed_offset(cols, z, fit_tolerance=1e-12)
    m = fo["mask"]
    jcols = list(cols)
    jcols[3] = fo["outcome"]
    r = oracle.Problem(*jcols).run(probes=400, leverage_width=16, target_width=16, probe_tolerance=1e-11,
                                   fit_tolerance=1e-11)
    w, f, g, _, fr, t = [np.asarray(c)[m] for c in cols]
    dense = E.dense_kss(np.asarray(cols[3])[m], w, f, controls=z[m], frequency=fr, target=t,
                        nuisance="fixedoffset")
    assert np.allclose(r["plugin"], dense["plugin"], rtol=1e-8, atol=1e-12)
    dev = np.abs(np.array(r["correction"]) - dense["correction"])
    assert np.all(dev <= 12.0 * np.array(r["mcse"]) + 5e-7), (dev, r["mcse"])


def test_joint_exact_directions_match_dense_joint(oracle):
    """The joint route's algebra (control block makers, control coefficients
    in the target predictions) with exact block leverages and the full
    target probe law equals the dense oracle's joint estimate."""
    from oracle import controls as X
    from oracle import exact as E
    cols, z = _panel(4, workers=200, firms=25)
    ours = X.joint_exact_directions(cols, z)
    m = oracle.Problem(*cols).mask()
    w, f, g, y, fr, t = [np.asarray(c)[m] for c in cols]
    dense = E.dense_kss(y, w, f, controls=z[m], frequency=fr, target=t, nuisance="joint")
    for k in ("plugin", "correction", "corrected"):
        a, b = np.asarray(ours[k]), np.asarray(dense[k])
        assert np.all(np.abs(a - b) <= 1e-8 * np.maximum(np.abs(b), 1e-6)), (k, a, b)
    # the controls matter: the estimate without them differs
    plain = E.dense_kss(y, w, f, frequency=fr, target=t)
    assert np.max(np.abs(np.asarray(plain["correction"]) - dense["correction"])) > 1e-4


def test_joint_jla_within_mcse_of_dense(oracle):
    from oracle import controls as X
    from oracle import exact as E
    cols, z = _panel(6, workers=400, firms=40)
    r = X.joint_jla(cols, z, probes=400, leverage_width=16, target_width=16, probe_tolerance=1e-11,
                    fit_tolerance=1e-11)
    m = oracle.Problem(*cols).mask()
    w, f, g, y, fr, t = [np.asarray(c)[m] for c in cols]
    dense = E.dense_kss(y, w, f, controls=z[m], frequency=fr, target=t, nuisance="joint")
    assert np.allclose(r["plugin"], dense["plugin"], rtol=1e-8, atol=1e-12)
    dev = np.abs(r["correction"] - dense["correction"])
    assert np.all(dev <= 12.0 * r["mcse"] + 5e-7), (dev, r["mcse"])
