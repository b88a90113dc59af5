"""Learned smoother weights on the GPU, through the C ABI.

The device network, training cycle and gradient are checked against the CPU
oracle (oracle/smoother.py, itself pinned by finite differences); training
must lower the loss; the learned weights must cut PCG iterations on panels
the network was not trained on; and the plain cycle must take over when the
learned weights do not help.
"""
import functools

import numpy as np
import pytest

import fevc_amd as F
from fevc_amd import synth
from oracle import smoother as S

pytestmark = pytest.mark.gpu

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")
ONE_LEVEL = synth.PanelSpec(workers=60000, firms=3000, periods=5, mobility=0.15, pareto_shape=1.1, seed=22,
                            name="smoother-one-level")
DEEP = synth.PanelSpec(workers=300000, firms=15000, periods=5, mobility=0.1, pareto_shape=1.1, seed=24,
                       name="smoother-deep")
DEEP_SKIP = synth.PanelSpec(workers=600000, firms=30000, periods=5, mobility=0.2, pareto_shape=1.2, seed=25,
                            name="smoother-deep-skip")
HELD_OUT = synth.PanelSpec(workers=400000, firms=20000, periods=5, mobility=0.12, pareto_shape=1.15, seed=9001,
                           name="smoother-held-out")


@functools.lru_cache(maxsize=None)
def columns(spec):
    d = synth.generate(spec)
    return tuple(np.asarray(d[k]) for k in COLS)


@functools.lru_cache(maxsize=None)
def oracle_side(spec):
    from oracle import oracle as O
    p = O.Problem(*columns(spec))
    g, _ = p.hybrid_graph()
    cmg = O.Cmg(g, direct_threshold=700, max_levels=128)
    levels = S.oracle_levels(cmg)
    return g.n, levels, S.prepare(levels)


def _theta(seed):
    theta = S.initial_params(seed=seed, scale=0.8)
    rng = np.random.default_rng(seed + 1)
    theta += 0.3 * rng.standard_normal(S.P) / np.sqrt(S.H)
    return theta


@pytest.mark.parametrize("spec", [ONE_LEVEL, DEEP], ids=["one-level", "deep"])
def test_omega_matches_oracle(oracle, spec):
    n, levels, feats = oracle_side(spec)
    theta = _theta(3)
    want = np.concatenate([om for _, (om, _) in sorted(S.omegas(theta, levels, feats).items())])
    e = F.Engine(*columns(spec))
    try:
        got = e.smoother_omega(theta)
    finally:
        e.release()
    assert got.shape == want.shape
    assert np.max(np.abs(got - want)) <= 1e-12
    assert np.all((got > 0) & (got < S.OMEGA_MAX))


@pytest.mark.parametrize("spec", [ONE_LEVEL, DEEP], ids=["one-level", "deep"])
def test_bf16_inference_matches_oracle(oracle, spec):
    """The solve path's BF16 inference (matrix-core layer products, FP32
    accumulation) against the oracle's BF16 emulation of the same network:
    activations round to BF16 at the same places, so the omegas agree up to
    the accumulation order's rare one-ulp BF16 flips; against the FP64
    network they agree to the BF16 model error."""
    n, levels, feats = oracle_side(spec)
    theta = _theta(3)
    emu = np.concatenate([forward_bf16(l, theta, levels, feats) for l in sorted(feats)])
    full = np.concatenate([om for _, (om, _) in sorted(S.omegas(theta, levels, feats).items())])
    e = F.Engine(*columns(spec))
    try:
        got = e.smoother_omega(theta, bf16=True)
    finally:
        e.release()
    assert got.shape == emu.shape
    d = np.abs(got - emu)
    print(f"{spec.name}: BF16 omega vs emulation max {d.max():.2e} mean {d.mean():.2e}; "
          f"vs FP64 max {np.abs(got - full).max():.2e}")
    assert d.max() <= 5e-3 and d.mean() <= 2e-4
    assert np.max(np.abs(got - full)) <= 3e-2
    assert np.all((got > 0) & (got < S.OMEGA_MAX))


def forward_bf16(l, theta, levels, feats):
    X, rs = feats[l]
    return S.forward_bf16(theta, X, levels[l], rs)


@pytest.mark.parametrize("spec", [ONE_LEVEL, DEEP], ids=["one-level", "deep"])
def test_gradient_matches_oracle(oracle, spec):
    n, levels, feats = oracle_side(spec)
    theta = _theta(5)
    E = np.random.default_rng(11).standard_normal((n, 32))
    E -= E.mean(0)
    L_want, g_want = S.loss_and_grad(theta, levels, feats, E)
    e = F.Engine(*columns(spec))
    try:
        L_got, g_got = e.smoother_gradient(theta, E)
    finally:
        e.release()
    assert abs(L_got - L_want) <= 1e-10 * abs(L_want)
    assert np.linalg.norm(g_got - g_want) <= 1e-9 * np.linalg.norm(g_want)


def test_training_lowers_loss():
    e = F.Engine(*columns(DEEP))
    try:
        state = np.zeros(3 * S.P)
        state[:S.P] = S.initial_params(seed=1)
        losses = e.smoother_train(state, 0, 40, 3e-3, seed=7, passes=2)
    finally:
        e.release()
    assert np.all(np.isfinite(losses))
    assert losses[-5:].mean() < 0.8 * losses[:5].mean(), losses


def _solve(cols, mode, params=None, probes=32, tolerance=1e-10):
    e = F.Engine(*cols)
    try:
        e.configure(32, 32, 16)
        e.smoother(mode, params)
        res = e.solve(probes=probes, tolerance=tolerance, fit_tolerance=tolerance)
        status = e.smoother_status()
        arr = e.arrays(probes)
        return res, status, arr
    finally:
        e.release()


def test_builtin_weights_cut_iterations_on_held_out_panel():
    """What the trained network gains: PCG iterations of the same leverage
    probes with the learned weights against the plain cycle, on a panel the
    network was not trained on, with every component within 1e-6."""
    params = F.builtin_smoother()
    assert params is not None, "the library must carry trained smoother weights"
    cols = columns(HELD_OUT)
    plain, sp, ap = _solve(cols, 0)
    learned, sl, al = _solve(cols, 1)
    assert sl["mode"] == 1 and sl["active"], sl
    assert sl["learned_iterations"] < sl["plain_iterations"]
    it_plain = np.mean(ap["leverage_iterations"])
    it_learned = np.mean(al["leverage_iterations"])
    print(f"held-out panel: {it_plain:.2f} -> {it_learned:.2f} PCG iterations per leverage solve "
          f"(calibration {sl['plain_iterations']} -> {sl['learned_iterations']})")
    assert it_learned <= 0.9 * it_plain
    for key in ("plugin", "correction", "corrected"):
        a, b = np.asarray(learned[key]), np.asarray(plain[key])
        assert np.all(np.abs(a - b) <= 1e-6 * np.maximum(np.abs(b), 1e-3)), key


def test_plain_cycle_takes_over_when_learned_weights_do_not_help():
    # a0 far negative: omega ~ 0, a useless smoother; the calibration keeps
    # the plain weights and, on a one-level hierarchy (no shape to tune),
    # the solve equals the plain solve bit for bit
    theta = S.initial_params(seed=2)
    theta[S.OFF["a0"][0]] = -12.0
    cols = columns(ONE_LEVEL)
    plain, _, _ = _solve(cols, 0, probes=8)
    got, st, _ = _solve(cols, 1, params=theta, probes=8)
    assert st["mode"] == 1 and not st["active"] and not st["shaped"]
    assert st["learned_iterations"] == st["plain_iterations"]
    for key in ("plugin", "correction", "corrected"):
        assert np.array_equal(np.asarray(got[key]), np.asarray(plain[key])), key


def test_tuned_shape_skips_levels_without_losing_iterations():
    """The tuned cycle shape on a deep hierarchy: fewer levels visited,
    within 10% of the reference cycle's PCG iterations, same components."""
    cols = columns(DEEP_SKIP)
    plain, sp, ap = _solve(cols, 0)
    tuned, st, at = _solve(cols, 1, params=_useless_weights())
    assert not st["active"]
    assert st["shaped"], st
    assert len(st["shape"]) < len(sp["shape"]) and st["shape"][0] == 0 and st["shape"][-1] == sp["shape"][-1]
    assert st["learned_iterations"] <= 1.1 * st["plain_iterations"]
    it_plain, it_tuned = np.mean(ap["leverage_iterations"]), np.mean(at["leverage_iterations"])
    print(f"tuned shape {st['shape']} (of {sp['shape']}): {it_plain:.2f} -> {it_tuned:.2f} iterations per solve")
    assert it_tuned <= 1.1 * it_plain
    for key in ("plugin", "correction", "corrected"):
        a, b = np.asarray(tuned[key]), np.asarray(plain[key])
        assert np.all(np.abs(a - b) <= 1e-6 * np.maximum(np.abs(b), 1e-3)), key


def _useless_weights():
    theta = S.initial_params(seed=2)
    theta[S.OFF["a0"][0]] = -12.0
    return theta


def test_learned_solve_matches_oracle(oracle):
    params = F.builtin_smoother()
    assert params is not None
    cols = columns(ONE_LEVEL)
    want = oracle.Problem(*cols).run(probes=16, fit_tolerance=1e-10, probe_tolerance=1e-11)
    e = F.Engine(*cols)
    try:
        e.configure(64, 16, 8)
        e.smoother(1)
        got = e.solve(probes=16, tolerance=1e-11, fit_tolerance=1e-10)
    finally:
        e.release()
    for key in ("plugin", "correction", "corrected"):
        g, w = np.asarray(got[key]), np.asarray(want[key])
        assert np.all(np.abs(g - w) <= 1e-8 * np.maximum(np.abs(w), 1e-3)), (key, g, w)


def test_smoother_argument_errors():
    e = F.Engine(*columns(ONE_LEVEL))
    try:
        with pytest.raises(F.EngineError) as ei:
            e.smoother(2)
        assert ei.value.code == 20
        with pytest.raises(F.EngineError) as ei:
            e.smoother(1, np.zeros(10))
        assert ei.value.code == 20
        bad = S.initial_params()
        bad[0] = np.nan
        with pytest.raises(F.EngineError) as ei:
            e.smoother_omega(bad)
        assert ei.value.code == 20
    finally:
        e.release()
