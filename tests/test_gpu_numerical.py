"""Numerical MCSE attachment on the GPU, through the C ABI, against the oracle
(engine/numerical.rs, numerical_mc.rs restated in oracle/engine.c and
oracle/numerical.c, pinned in tests/test_oracle_numerical.py).

The device replays the leverage probes in batches of the leverage width
(the reference replays at most 4 at a time): the replayed solutions agree to
the solve tolerance, so the leverage covariance agrees to that tolerance and
the integer bookkeeping exactly.
"""
import functools

import numpy as np
import pytest

import fevc_amd as F
from fevc_amd import synth

pytestmark = pytest.mark.gpu

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")
SPEC = synth.PanelSpec(workers=4000, firms=300, periods=5, mobility=0.2, seed=11, name="gpu-numerical")
DEEP = synth.PanelSpec(workers=300000, firms=15000, periods=5, mobility=0.1, pareto_shape=1.1, seed=24,
                       name="gpu-numerical-deep")


@functools.lru_cache(maxsize=None)
def columns(spec):
    d = synth.generate(spec)
    return tuple(np.asarray(d[k]) for k in COLS)


@functools.lru_cache(maxsize=None)
def oracle_numerical(spec, probes, tolerance):
    from oracle import oracle as O
    p = O.Problem(*columns(spec))
    return p.run(probes=probes, leverage_width=4, target_width=4, fit_tolerance=1e-10,
                 probe_tolerance=tolerance, numerical=True)


def _device(spec, probes, tolerance, mode, precision=64):
    e = F.Engine(*columns(spec))
    try:
        e.configure(precision, 64, 32)
        e.smoother(mode)
        res, num = e.solve_numerical(probes=probes, tolerance=tolerance)
        rhs = e.numerical_rhs()
        work = e.numerical_cmg_work()
        return res, num, rhs, work
    finally:
        e.release()


def _compare(want, got, rel):
    wn = want["numerical"]
    assert got["status"] == wn["status"] == 2
    for k in ("leverage_probes", "target_probes"):
        assert got[k] == wn[k]
    assert (got["target_fold_a"], got["target_fold_b"]) == (wn["fold_a"], wn["fold_b"])
    assert got["certified_replay_rhs"] == got["executed_replay_rhs"] == wn["certified_replay_rhs"]
    assert got["attempted_replay_rhs"] == wn["attempted_replay_rhs"]
    assert got["replay_generator_words"] == wn["replay_generator_words"]
    assert got["point_rhs"] == wn["point_rhs"]
    assert got["failed_replay_probe"] == 0xFFFFFFFF and got["replay_error_code"] == 0
    for k in ("minimum_constrained", "minimum_residual_margin", "maximum_sensitivity_ratio"):
        assert abs(got[k] - wn[k]) <= rel * abs(wn[k]), k
    scale = {k: max(np.abs(wn[k]).max(), 1e-300) for k in ("conditional", "leverage", "usable")}
    for k in ("conditional", "leverage", "usable"):
        assert np.max(np.abs(got[k] - wn[k])) <= rel * scale[k], (k, got[k], wn[k])
    assert np.all(np.abs(got["mcse"] - wn["mcse"]) <= rel * np.abs(wn["mcse"])), (got["mcse"], wn["mcse"])
    # fold score means vanish up to rounding; compare on the scores' own scale
    R = wn["leverage_probes"]
    score_scale = np.sqrt(np.abs(np.diag(wn["leverage"])).max() * R * (R - 1))
    assert np.max(np.abs(got["score_means"] - wn["score_means"])) <= rel * score_scale


def test_numerical_matches_oracle_reference_cycle(oracle):
    R, tol = 24, 1e-11
    want = oracle_numerical(SPEC, R, tol)
    res, got, rhs, work = _device(SPEC, R, tol, mode=0)
    _compare(want, got, 1e-6)
    # the point estimate is the plain solve's
    for key in ("plugin", "correction", "corrected"):
        assert np.allclose(res[key], want[key], rtol=1e-8, atol=0)
    assert [p for _, p, _ in rhs] == list(range(R))
    assert all(phase == 1 for phase, _, _ in rhs)
    assert max(c for _, _, c in rhs) <= max(10 * tol, 1e-11)
    assert work["logical_rhs_count"] == 1 + 3 * R + R


def test_numerical_matches_oracle_tuned_preconditioner(oracle):
    R, tol = 24, 1e-11
    want = oracle_numerical(SPEC, R, tol)
    _, got, _, _ = _device(SPEC, R, tol, mode=1, precision=32)
    _compare(want, got, 1e-6)


def test_numerical_on_a_deep_hierarchy(oracle):
    R, tol = 16, 1e-10
    want = oracle_numerical(DEEP, R, tol)
    _, got, _, _ = _device(DEEP, R, tol, mode=1, precision=32)
    _compare(want, got, 1e-5)


def test_numerical_under_observation_deletion(oracle):
    """Observation deletion (generic_jla/numerical.rs:172-751): per-copy
    pullbacks, observation folds and replayed copy signs, against the oracle."""
    from oracle import oracle as O
    cols = list(columns(SPEC))
    cols[4] = np.random.default_rng(9).integers(1, 4, len(cols[0])).astype(np.float64)
    R, tol = 24, 1e-11
    want = O.Problem(*cols, observation=True).run(probes=R, leverage_width=4, target_width=4, fit_tolerance=1e-10,
                                                 probe_tolerance=tol, numerical=True)
    for precision in (64, 32):
        e = F.Engine(*cols, deletion_mode="observation")
        try:
            e.configure(precision, 64, 32)
            print(f"observation numerical, precision {precision}", flush=True)
            res, got = e.solve_numerical(probes=R, tolerance=tol)
        finally:
            e.release()
        _compare(want, got, 1e-6)
        for key in ("plugin", "correction", "corrected"):
            assert np.allclose(res[key], want[key], rtol=1e-8, atol=0)


def test_numerical_result_requires_the_attachment():
    e = F.Engine(*columns(SPEC))
    try:
        e.solve(probes=8, tolerance=1e-8)
        with pytest.raises(F.EngineError) as ei:
            e.numerical_result()
        assert ei.value.code == 11
    finally:
        e.release()
