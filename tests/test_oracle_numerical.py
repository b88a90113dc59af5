"""Pins of the oracle's numerical MCSE (oracle/numerical.c, the numerical
attachment of oracle/engine.c), CPU only.

* derivative fields against the reference's complex-step golden vectors
  (rust/crates/vckss-core/tests/numerical_mc.rs:118-777, extracted into
  tests/golden/numerical_mc_vectors.json) and against central finite
  differences of the functions they differentiate;
* the covariance policy cases of numerical_mc.rs tests 7-50 and 109-117;
* the fold estimator's exact unbiasedness by enumeration (tests 52-107);
* finalize against a numpy eigendecomposition;
* engine-level invariants of the replay.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def _golden():
    with open(os.path.join(HERE, "golden", "numerical_mc_vectors.json")) as fh:
        return json.load(fh)["cases"]


@pytest.mark.parametrize("case", range(16))
def test_derivatives_match_complex_step_vectors(case):
    c = _golden()[case]
    d = O.finite_derivative(c["u"], c["r"])
    assert d is not None
    for got, want in zip(list(d["db"]) + list(d["dv"]), c["db"] + c["dv"]):
        assert abs(got - want) <= 1e-11 * max(abs(want), 1.0)
    g = O.fd_observation(d, c["control"])
    for got, want in zip(g, c["observation"]):
        assert abs(got - want) <= 1e-11 * max(abs(want), 1.0)


def _fields(u, r):
    d = O.finite_derivative(u, r)
    return np.array([d["h"], d["m"], d["bias"], d["variance"]])


def test_derivatives_match_finite_differences():
    rng = np.random.default_rng(3)
    for _ in range(20):
        u = np.array([rng.uniform(0.05, 0.6), rng.uniform(0.2, 0.9), rng.uniform(0.0, 0.05),
                      rng.uniform(0.0, 0.9), rng.uniform(0.0, 0.05)])
        r = 32
        d = O.finite_derivative(u, r)
        jac = np.stack([d["dh"], -d["dh"], d["db"], d["dv"]])  # dm = -dh (m = 1 - h)
        for j in range(5):
            e = np.zeros(5)
            e[j] = 1e-6
            fd = (_fields(u + e, r) - _fields(u - e, r)) / 2e-6
            assert np.allclose(fd, jac[:, j], rtol=1e-6, atol=1e-9), (j, fd, jac[:, j])
        if d["nonsmooth"]:
            continue
        control = 0.04

        def F(v):
            x = O.finite_derivative(v, r)
            ell = x["m"] - control
            return 1.0 / ell + x["bias"] / ell ** 2 - x["variance"] / ell ** 3
        g = O.fd_observation(d, control)
        for j in range(5):
            e = np.zeros(5)
            e[j] = 1e-6
            fd = (F(u + e) - F(u - e)) / 2e-6
            assert abs(fd - g[j]) <= 1e-6 * max(abs(fd), 1e-3), (j, fd, g[j])


def test_scaled_covariance_policy_and_signed_stage():
    for scale in (1e-120, 1.0, 1e120):
        cond = np.diag([scale, 2 * scale, 3 * scale])
        signed = np.diag([-0.2 * scale, 0.4 * scale, -0.1 * scale])
        f = O.numerical_finalize(cond, signed)
        assert f["status"] == 2
        assert abs(f["mcse"][3] / np.sqrt(scale) - np.sqrt(14.8)) < 1e-12
        assert f["mcse"][0] < np.sqrt(scale)
    eye = np.eye(3)
    assert O.numerical_finalize(eye, np.diag([-2.0, 0, 0]))["status"] == 4
    assert O.numerical_finalize(eye, np.diag([-1.0 - 1e-13, 0, 0]))["status"] == 3
    large = O.numerical_finalize(eye * 1e308, np.zeros((3, 3)))
    assert np.all(np.isfinite(large["mcse"]))
    draws = np.zeros((4, 3))
    draws[:, 0] = [1e154, -1e154, 1e154, -1e154]
    assert abs(O.cross_covariance(draws, draws)[0, 0] / 1e307 - 10.0 / 3.0) < 1e-14


def test_clipped_adjustment_and_exact_zero():
    d = O.finite_derivative([0.2, 0.8, 0.01, 0.01, 1.0], 32)
    assert d["nonsmooth"]
    assert O.fd_observation(d, 0.0) is None
    z = np.zeros((3, 3))
    assert np.array_equal(O.cross_covariance(z, z), z)
    f = O.numerical_finalize(z, z)
    assert f["status"] == 2 and np.array_equal(f["mcse"], np.zeros(4))


def test_fold_estimator_is_unbiased_by_enumeration():
    for targets in (2, 4, 5):
        for mx, my in ((0.0, 0.0), (0.7, -0.4)):
            total, naive, cases = 0.0, 0.0, 0
            for xb in range(8):
                x = np.array([mx + (0.3 if (xb >> r) & 1 else -0.3) for r in range(3)])
                xm = x.mean()
                for yb in range(1 << targets):
                    y = np.array([my + (0.8 if (yb >> t) & 1 else -0.8) for t in range(targets)])
                    folds = [y[p::2].mean() for p in (0, 1)]
                    sc = [np.stack([(x - xm) * f, np.zeros(3), np.zeros(3)], 1) for f in folds]
                    draws = np.stack([xm * y, np.zeros(targets), np.zeros(targets)], 1)
                    alls = np.stack([(x - xm) * y.mean(), np.zeros(3), np.zeros(3)], 1)
                    cond = O.cross_covariance(draws, draws)[0, 0]
                    total += cond + O.cross_covariance(sc[0], sc[1])[0, 0]
                    naive += cond + O.cross_covariance(alls, alls)[0, 0]
                    cases += 1
            truth = my * my * 0.09 / 3 + mx * mx * 0.64 / targets + 0.09 * 0.64 / (3 * targets)
            assert abs(total / cases - truth) < 1e-13
            assert naive / cases > truth * 1.001


def test_finalize_against_eigendecomposition():
    rng = np.random.default_rng(5)
    for _ in range(30):
        a = rng.standard_normal((3, 3))
        cond = a @ a.T
        b = rng.standard_normal((3, 3)) * 0.3
        lev = (b + b.T) / 2
        f = O.numerical_finalize(cond, lev)
        raw = cond + lev
        w = np.linalg.eigvalsh(raw)
        scale = max(np.abs(cond).max(), np.abs(lev).max())
        tau = 1e-12 * (np.linalg.norm(cond / scale) + np.linalg.norm(lev / scale))
        if w.min() / scale < -tau * 10:
            assert f["status"] == 4
            continue
        if w.min() / scale > tau * 10:
            assert f["status"] == 2
            assert np.allclose(f["usable"], raw, rtol=0, atol=1e-14 * scale)
        c = np.array([1.0, 1.0, 2.0])
        assert np.allclose(f["mcse"][:3], np.sqrt(np.maximum(np.diag(f["usable"]), 0)), rtol=1e-12)
        assert np.isclose(f["mcse"][3], np.sqrt(max(c @ f["usable"] @ c, 0)), rtol=1e-10)
        # the PSD projection keeps eigenvectors and clips negative eigenvalues
        if f["status"] == 3:
            wu, vu = np.linalg.eigh(f["usable"])
            assert wu.min() >= -1e-12 * scale
            assert np.allclose(np.sort(np.maximum(w, 0)), np.sort(wu), atol=1e-10 * scale)


def _panel(workers=2000, seed=1):
    from fevc_amd import synth
    d = synth.generate(synth.PanelSpec(workers=workers, firms=workers // 10, periods=5, mobility=0.2, seed=seed,
                                       name="numerical"))
    return O.Problem(*[d[k] for k in ("worker", "firm", "deletion", "outcome", "frequency", "target")])


def test_engine_attachment_invariants():
    p = _panel()
    R = 20
    a = p.run(probes=R, leverage_width=4, target_width=4, probe_tolerance=1e-11, numerical=True)
    b = p.run(probes=R, leverage_width=1, target_width=2, probe_tolerance=1e-11, numerical=True)
    na, nb = a["numerical"], b["numerical"]
    assert na["status"] == 2
    assert na["certified_replay_rhs"] == na["attempted_replay_rhs"] == R
    assert (na["fold_a"], na["fold_b"], na["point_rhs"]) == (10, 10, 1 + 3 * R)
    assert na["failed_replay_probe"] == 0xFFFFFFFF
    # conditional covariance is the scaled cross product of the target draws
    draws = a["target_draws"][:, :3]
    dc = draws - draws.mean(0)
    assert np.allclose(na["conditional"], dc.T @ dc / (R * (R - 1)), rtol=1e-10, atol=0)
    assert np.all(na["usable"] == na["usable"].T)
    # replay batches of 4 or 1 give the same numbers to the solve tolerance
    for k in ("conditional", "leverage", "mcse"):
        assert np.allclose(na[k], nb[k], rtol=1e-6, atol=1e-18), k
    # the basic MCSE of the point result stays the draws' standard error
    assert np.allclose(a["mcse"], np.sqrt(np.diag(np.cov(a["target_draws"].T)) / R)[[0, 1, 2, 3]], rtol=1e-10)
