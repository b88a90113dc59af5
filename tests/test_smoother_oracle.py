"""Pins of the learned-smoother oracle (oracle/smoother.py), CPU only.

The learned smoother has no reference counterpart, so its oracle is pinned by
mathematics: the gradient against central finite differences of the loss,
the plain weights (omega = 0.5) against the oracle's restatement of the
reference CMG cycle (orc_cmg_apply, cycle.rs:136-311), and the features
against brute-force loops.
"""
import numpy as np
import pytest

from oracle import smoother as S


def _grid(s, seed=0):
    idx = np.arange(s * s).reshape(s, s)
    u = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    v = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    rng = np.random.default_rng(seed)
    return s * s, u, v, rng.uniform(0.2, 3.0, len(u))


def _random(n, seed=1):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, 3 * n)
    v = rng.integers(0, n, 3 * n)
    keep = u != v
    path = np.arange(n - 1)
    u = np.concatenate([u[keep], path])
    v = np.concatenate([v[keep], path + 1])
    return n, u, v, rng.uniform(0.1, 4.0, len(u))


def _levels(oracle, n, u, v, w, direct=40):
    g = oracle.Graph(n, u, v, w)
    cmg = oracle.Cmg(g, direct_threshold=direct, max_levels=32)
    return cmg, S.oracle_levels(cmg)


def _errors(n, k, seed):
    E = np.random.default_rng(seed).standard_normal((n, k))
    return E - E.mean(0)


@pytest.mark.parametrize("case", ["grid", "random"])
def test_gradient_matches_finite_differences(oracle, case):
    n, u, v, w = _grid(14) if case == "grid" else _random(300)
    _, levels = _levels(oracle, n, u, v, w)
    assert len(S.smoothed_levels(levels)) >= 2
    feats = S.prepare(levels)
    theta = S.initial_params(seed=3, scale=0.8)
    rng = np.random.default_rng(5)
    theta += 0.3 * rng.standard_normal(S.P) / np.sqrt(S.H)
    E = _errors(n, 4, 7)
    L, g = S.loss_and_grad(theta, levels, feats, E)
    assert 0.0 < L < 1.0
    h = 1e-6
    # every output-layer parameter, and a sample of the rest
    idx = list(range(S.OFF["a"][0], S.P)) + sorted(rng.choice(S.OFF["a"][0], 60, replace=False))
    for i in idx:
        d = np.zeros(S.P)
        d[i] = h
        fd = (S.loss(theta + d, levels, feats, E) - S.loss(theta - d, levels, feats, E)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(fd), 1e-4), (i, fd, g[i])
    # directional derivative along a random direction through all parameters
    d = rng.standard_normal(S.P)
    d *= 1e-6 / np.linalg.norm(d)
    fd = (S.loss(theta + d, levels, feats, E) - S.loss(theta - d, levels, feats, E)) / 2
    assert abs(fd - g @ d) <= 1e-6 * abs(fd)


def test_planted_gradient_mistake_is_caught(oracle):
    """A dropped adjoint term (the post-smoothing weight gradient) must fail
    the finite-difference check that pins the oracle."""
    n, u, v, w = _grid(12)
    _, levels = _levels(oracle, n, u, v, w)
    feats = S.prepare(levels)
    theta = S.initial_params(seed=2, scale=0.8)
    E = _errors(n, 3, 1)
    real = S.cycle_adjoint

    def broken(levels_, ws, store, lam):
        lr, gws = real(levels_, ws, store, lam)
        l0 = S.smoothed_levels(levels_)[0]
        gws[l0] = gws[l0] * 0.5
        return lr, gws
    S.cycle_adjoint = broken
    try:
        _, g = S.loss_and_grad(theta, levels, feats, E)
    finally:
        S.cycle_adjoint = real
    i = S.OFF["a0"][0]
    h = 1e-6
    d = np.zeros(S.P)
    d[i] = h
    fd = (S.loss(theta + d, levels, feats, E) - S.loss(theta - d, levels, feats, E)) / (2 * h)
    assert abs(fd - g[i]) > 1e-3 * abs(fd)


@pytest.mark.parametrize("case", ["grid", "random"])
def test_plain_weights_reproduce_cmg_cycle(oracle, case):
    n, u, v, w = _grid(16, seed=4) if case == "grid" else _random(400, seed=6)
    cmg, levels = _levels(oracle, n, u, v, w)
    # single visits: every non-terminal child runs one iteration
    for l, lv in enumerate(levels[:-1]):
        if not levels[l + 1]["terminal"]:
            assert lv["repeat"] == 1, "fixture must have single-visit levels"
    ws = [None if lv["terminal"] else 0.5 / lv["diag"] for lv in levels]
    rng = np.random.default_rng(0)
    r = rng.standard_normal((n, 3))
    r -= r.mean(0)
    y, _ = S.cycle_forward(levels, ws, r)
    for k in range(3):
        want = cmg.apply(r[:, k])
        assert np.linalg.norm(y[:, k] - want) <= 1e-12 * np.linalg.norm(want)


def test_features_brute_force(oracle):
    n, u, v, w = _random(60, seed=9)
    _, levels = _levels(oracle, n, u, v, w, direct=10)
    lv = levels[0]
    X, rs = S.features(lv, 0)
    nbr = [dict() for _ in range(lv["n"])]
    for a, b, c in zip(lv["u"], lv["v"], lv["w"]):
        nbr[a][b] = nbr[a].get(b, 0.0) + c
        nbr[b][a] = nbr[b].get(a, 0.0) + c
    lab = lv["labels"]
    logs = [np.log(d) for d in lv["diag"]]
    mu = np.mean(logs)
    for i in range(lv["n"]):
        s = sum(nbr[i].values())
        want = [1.0, logs[i] - mu, np.log(1 + len(nbr[i])) / 4, max(nbr[i].values()) / s,
                sum(c for j, c in nbr[i].items() if lab[j] == lab[i]) / s,
                np.log(sum(1 for k in range(lv["n"]) if lab[k] == lab[i])) / 4, 0.0,
                sum(c * (logs[j] - logs[i]) for j, c in nbr[i].items()) / s]
        assert np.allclose(X[i], want, rtol=1e-12, atol=1e-14), i
        assert rs[i] == pytest.approx(1.0 / s, rel=1e-14)


def test_initial_params_give_plain_weights(oracle):
    n, u, v, w = _grid(10)
    _, levels = _levels(oracle, n, u, v, w)
    feats = S.prepare(levels)
    theta = S.initial_params()
    theta[S.OFF["a"][0]:S.OFF["a"][0] + S.H] = 0.0
    for l, (om, _) in S.omegas(theta, levels, feats).items():
        assert np.allclose(om, 0.5, rtol=1e-14)
