"""Dense exact estimator on the GPU, through the C ABI, against the oracle
(oracle/exact.py restates exact_estimator.rs:490-1244; its pins are in
tests/test_oracle_exact.py).

Floating point: plug-in, correction and corrected components within 1e-9
relative (a 1e-12 absolute floor), the maximum block leverage within 1e-10.
The device forms the inverse by a blocked Cholesky of the scaled, shifted
information and combines match blocks in closed form; the oracle factors the
same matrix with LAPACK and solves every block maker densely.
"""
import functools

import numpy as np
import pytest

import fevc_amd as F
from fevc_amd import synth

pytestmark = pytest.mark.gpu

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")
SMALL = synth.PanelSpec(workers=600, firms=60, periods=5, mobility=0.3, seed=5, name="exact-small")
# the small plain-method panel: 2000 workers x 5 periods, 200 firms, mobility 0.2
CONFIG0 = synth.PanelSpec(workers=2000, firms=200, periods=5, mobility=0.2, seed=1, name="exact-config0")


@functools.lru_cache(maxsize=None)
def columns(spec):
    d = synth.generate(spec)
    return tuple(np.asarray(d[k]) for k in COLS)


@functools.lru_cache(maxsize=None)
def oracle_exact(spec):
    from oracle import exact as X
    from oracle import oracle as O
    cols = columns(spec)
    keep = O.Problem(*cols).mask()
    return X.exact_estimator(*[c[keep] for c in cols])


def _arrays(res):
    return {k: np.asarray(v, float) for k, v in res.items()}


def _close(got, want, rel=1e-9, floor=1e-12):
    got, want = np.asarray(got), np.asarray(want)
    assert np.all(np.abs(got - want) <= rel * np.abs(want) + floor), (got, want)


@pytest.mark.parametrize("spec", [SMALL, CONFIG0], ids=["small", "config0"])
def test_exact_matches_oracle(oracle, spec):
    want = oracle_exact(spec)
    e = F.Engine(*columns(spec))
    try:
        print(f"exact {spec.name}: workers {want['workers']} firms {want['firms']}", flush=True)
        res, rec = e.solve_exact(exact_limit=10000)
    finally:
        e.release()
    res = _arrays(res)
    print(f"  factor + fit {rec['seconds_factor']:.4f}s, cells {rec['seconds_cells']:.4f}s, "
          f"rcond >= {rec['information_rcond']:.3g}, max leverage {rec['max_leverage']:.6f}", flush=True)
    for k in ("plugin", "correction", "corrected"):
        _close(res[k], want[k])
    assert rec["parameters"] == want["workers"] + want["firms"] - 1
    assert rec["deletion_units"] == want["deletion_units"]
    assert abs(rec["max_leverage"] - want["max_leverage"]) <= 1e-10
    assert abs(rec["deletion_rank_gap"] - want["deletion_rank_gap"]) <= 1e-10
    assert rec["inverse_relres"] < 1e-10 and rec["fit_relres"] <= rec["fit_tolerance"]
    assert 0 < rec["information_rcond"] <= want["information_rcond"] * (1 + 1e-9)
    assert rec["weighted_rss"] == pytest.approx(want["weighted_rss"], rel=1e-9)
    assert np.all(res["numerical_mcse"] == 0.0)


def test_auto_routes_by_exact_limit():
    e = F.Engine(*columns(SMALL))
    try:
        exact, rec = e.solve_exact(exact_limit=10000)
        exact = _arrays(exact)
        auto, _ = e.solve_exact(algorithm=0, exact_limit=rec["parameters"])
        auto = _arrays(auto)
        assert e.exact_receipt()["parameters"] == rec["parameters"]
        for k in ("plugin", "correction", "corrected"):
            assert np.array_equal(auto[k], exact[k])
        jla = _arrays(e.solve(probes=64, tolerance=1e-10))
        with pytest.raises(F.EngineError) as ei:  # the last solve was not exact
            e.exact_receipt()
        assert ei.value.code == 11
        assert np.all(jla["numerical_mcse"] > 0)
        # one below the dimension AUTO resolves to the legacy JLA route (full_cmg_v2 = 0),
        # which this engine does not run
        with pytest.raises(F.EngineError) as ei:
            e.solve_exact(algorithm=0, exact_limit=rec["parameters"] - 1)
        assert ei.value.code == 11
        # ffi_engine.rs:10366-10410: CMG_FULL_V2 takes JLA only
        for algorithm in (0, 1):
            with pytest.raises(F.EngineError) as ei:
                e.solve_exact(algorithm=algorithm, exact_limit=10000, full_cmg_v2=1)
            assert ei.value.code == 11
    finally:
        e.release()


def test_jla_agrees_with_exact_within_monte_carlo_error():
    e = F.Engine(*columns(SMALL))
    try:
        exact, _ = e.solve_exact(exact_limit=10000)
        jla = e.solve(probes=1024, tolerance=1e-10)
    finally:
        e.release()
    exact, jla = _arrays(exact), _arrays(jla)
    # plug-in is the same fit; the correction is an unbiased JLA estimate of
    # the exact one, with the reported Monte Carlo standard error
    _close(jla["plugin"], exact["plugin"], rel=1e-7)
    z = np.abs(jla["corrected"] - exact["corrected"]) / jla["numerical_mcse"]
    assert np.all(z < 5.0), (jla["corrected"], exact["corrected"], jla["numerical_mcse"])


def test_exact_limits_and_refusals():
    e = F.Engine(*columns(SMALL))
    try:
        with pytest.raises(F.EngineError) as ei:
            e.solve_exact(exact_limit=10)
        assert ei.value.code == 40
        with pytest.raises(F.EngineError) as ei:
            e.solve_exact(exact_limit=10000, blocksize_limit=1)
        assert ei.value.code == 40
        with pytest.raises(F.EngineError) as ei:
            r = e.numerical_request(probes=8)
            r.v1.point.algorithm = 1
            r.v1.point.exact_estimator_limit = 10000
            e.solve_numerical(r)
        assert ei.value.code == 11
    finally:
        e.release()


@functools.lru_cache(maxsize=None)
def oracle_observation(spec):
    from oracle import exact as X
    from oracle import oracle as O
    cols = columns(spec)
    keep = O.Problem(*cols, observation=True).mask()
    return keep, X.exact_estimator(*[c[keep] for c in cols], mode="observation")


@pytest.mark.parametrize("spec", [SMALL, CONFIG0], ids=["small", "config0"])
def test_exact_observation_deletion_matches_oracle(oracle, spec):
    keep, want = oracle_observation(spec)
    e = F.Engine(*columns(spec), deletion_mode="observation")
    try:
        assert np.array_equal(e.mask(), keep)
        print(f"observation {spec.name}: retained rows {int(keep.sum())}", flush=True)
        res, rec = e.solve_exact(exact_limit=10000)
    finally:
        e.release()
    res = _arrays(res)
    for k in ("plugin", "correction", "corrected"):
        _close(res[k], want[k])
    assert rec["deletion_units"] == want["deletion_units"]
    assert abs(rec["max_leverage"] - want["max_leverage"]) <= 1e-10
    assert abs(rec["deletion_rank_gap"] - want["deletion_rank_gap"]) <= 1e-10


def test_observation_deletion_refusals():
    e = F.Engine(*columns(SMALL), deletion_mode="observation")
    try:
        e.deletion_mode = 1  # a solve whose deletion mode differs from the prepared graph's
        with pytest.raises(F.EngineError) as ei:
            e.solve_exact(exact_limit=10000)
        assert ei.value.code == 20
        e.deletion_mode = 2
        e.solve_exact(exact_limit=10000)
    finally:
        e.release()
