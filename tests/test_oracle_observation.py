"""Observation-deletion JLA route of the oracle (generic engine without
controls: generic_jla.rs:5470-6017, statistical_batches.rs:15-206, the
target phase of :8554-8906), pinned by the properties the reference's own
tests assert for it (rust/crates/vckss-core/tests/generic_jla.rs):

- against the dense exact estimator (pinned in tests/test_oracle_exact.py):
  |JLA - exact| <= 12 MCSE + 5e-7 for every correction component (:1243-1287);
- row-order reversal leaves the results unchanged (:3029-3075; bitwise there,
  here up to summation order, 1e-12);
- splitting a physical row into copies 1 and f - 1 of the same observation
  class leaves the results unchanged within 2e-9 (:3029-3075): this fixes
  the Counter-V1 addressing of copies (class rank, copy offset);
- the per-row leverage estimates are unbiased for the exact leverages with
  the Johnson-Lindenstrauss spread sqrt(2 / probes).
"""
import numpy as np
import pytest

from fevc_amd import synth

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")


def _panel(seed, workers=300, firms=30):
    d = synth.generate(synth.PanelSpec(workers=workers, firms=firms, periods=4, mobility=0.3, seed=seed,
                                       name="observation-jla"))
    cols = [np.asarray(d[k], float) for k in COLS]
    rng = np.random.default_rng(seed + 100)
    cols[4] = rng.integers(1, 4, cols[0].size).astype(float)
    cols[5] = cols[4] * rng.uniform(0.5, 1.5, cols[0].size)
    return cols


def _run(oracle, cols, probes=400):
    p = oracle.Problem(*cols, observation=True)
    return p, p.run(probes=probes, leverage_width=16, target_width=16, probe_tolerance=1e-11, fit_tolerance=1e-11)


@pytest.mark.parametrize("seed", [3, 7])
def test_observation_jla_matches_exact_within_mcse(oracle, seed):
    from oracle import exact as X
    cols = _panel(seed)
    p, r = _run(oracle, cols)
    keep = p.mask()
    ex = X.exact_estimator(*[c[keep] for c in cols], mode="observation")
    assert np.allclose(r["plugin"], ex["plugin"], rtol=1e-9, atol=1e-12)
    dev = np.abs(np.array(r["correction"]) - ex["correction"])
    assert np.all(dev <= 12.0 * np.array(r["mcse"]) + 5e-7), (dev, r["mcse"])
    assert np.all(np.array(r["mcse"]) > 0)


def test_observation_leverages_are_unbiased(oracle):
    from oracle.exact import expanded_design
    cols = _panel(11)
    p, r = _run(oracle, cols, probes=400)
    keep = p.mask()
    w, f, g, y, fr, t = [c[keep] for c in cols]
    des = expanded_design(y, w, f, frequency=fr)
    x = des["x"]
    h = np.einsum("ij,ij->i", x @ np.linalg.pinv(x.T @ x), x)
    exact = np.zeros(y.size)
    exact[des["stored"]] = h
    rel = r["projection_share"] / exact - 1.0
    assert abs(rel.mean()) < 4.0 * rel.std() / np.sqrt(rel.size)
    assert 0.7 * np.sqrt(2 / 400) < rel.std() < 1.3 * np.sqrt(2 / 400)


def test_observation_results_ignore_row_order(oracle):
    cols = _panel(5)
    _, a = _run(oracle, cols, probes=64)
    _, b = _run(oracle, [c[::-1].copy() for c in cols], probes=64)
    for k in ("plugin", "correction", "corrected", "mcse"):
        assert np.allclose(a[k], b[k], rtol=1e-12, atol=1e-15), k


def test_observation_physical_split_preserves_results(oracle):
    cols = _panel(9)
    p, base = _run(oracle, cols, probes=64)
    keep = p.mask()
    row = int(np.flatnonzero(keep & (cols[4] == 3))[0])
    # copy 1 stays in place, copies 2..f follow it with their own deletion identifier
    pieces = [c.copy() for c in cols]
    extra = [np.array([c[row]]) for c in cols]
    pieces[4][row], extra[4][0] = 1.0, cols[4][row] - 1.0
    pieces[5][row], extra[5][0] = cols[5][row] / cols[4][row], cols[5][row] * (cols[4][row] - 1) / cols[4][row]
    extra[2][0] = cols[2].max() + 1
    split = [np.insert(c, row + 1, e) for c, e in zip(pieces, extra)]
    q, res = _run(oracle, split, probes=64)
    assert q.rows == p.rows + 1 and q.physical == p.physical
    for k in ("plugin", "correction", "corrected", "mcse"):
        assert np.allclose(base[k], res[k], rtol=2e-9, atol=1e-12), (k, base[k], res[k])
    # a different copy numbering (the second piece restarting at copy 0) is detectable
    other = [c.copy() for c in split]
    other[3][row + 1] = np.nextafter(other[3][row + 1], np.inf)  # its own class
    _, moved = _run(oracle, other, probes=64)
    assert not np.allclose(base["correction"], moved["correction"], rtol=1e-6, atol=0)


# ---------------------------------------------------------------- numerical MCSE
# generic_jla/numerical.rs:172-751 for observation units.  Pins: the replayed
# leverage probes regenerate the leverage phase's copy signs and predictions,
# so every fold's scores average to zero (the offsets are the linearisation
# at the probe means); the result does not depend on the replay batch width;
# the reported MCSE is of the size of the estimator's spread across seeds.
# The copy pullback coefficients (numerical_mc.rs:137-145) are pinned by
# central differences of the copy's correction value below.
def _copy_points(seed, n):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        a2 = rng.uniform(0.05, 0.4)
        a4 = a2 * a2 * rng.uniform(1.0, 3.0)
        yield np.array([a2, a4, a2 * rng.uniform(0.6, 1.0), a4 * rng.uniform(0.6, 1.0)])


def _central(oracle, x, r, h=1e-5):
    out = np.zeros(4)
    for j in range(4):
        e = np.zeros(4)
        e[j] = h * max(abs(x[j]), 1e-3)
        hi = oracle.copy_pullback(x + e, r)
        lo = oracle.copy_pullback(x - e, r)
        assert hi[0] == 0 and lo[0] == 0
        out[j] = (hi[1] - lo[1]) / (2 * e[j])
    return out


def test_copy_pullback_central_differences(oracle):
    """The pullback of one copy's linearised correction to the row moments
    (a2, a4) and the copy correlations (c1, c3) equals central differences of
    the correction value; a transposed pair or a changed coefficient fails."""
    checked = 0
    caught = {"swap": 0, "sign": 0, "scale": 0}
    for x in _copy_points(11, 40):
        st, _, pb = oracle.copy_pullback(x, 64)
        if st != 0:
            continue
        fd = _central(oracle, x, 64)
        tol = 1e-6 * np.max(np.abs(fd))
        assert np.all(np.abs(pb - fd) <= tol), (x, pb, fd)
        checked += 1
        for name, bad in (("swap", pb[[0, 1, 3, 2]]), ("sign", pb * [1, 1, -1, 1]),
                          ("scale", pb * [1, 1, 1, 1.5])):
            caught[name] += bool(np.any(np.abs(bad - fd) > tol))
    assert checked >= 30
    assert all(v >= checked // 2 for v in caught.values()), caught
def test_observation_numerical_scores_are_centred(oracle):
    cols = _panel(3)
    p = oracle.Problem(*cols, observation=True)
    r = p.run(probes=64, leverage_width=8, target_width=8, probe_tolerance=1e-11, fit_tolerance=1e-11,
              numerical=True)
    n = r["numerical"]
    assert n["status"] == 2 and n["certified_replay_rhs"] == 64
    scale = np.sqrt(np.abs(np.diag(np.asarray(n["leverage"]).reshape(3, 3))).max() * 64 * 63)
    assert np.max(np.abs(np.asarray(n["score_means"]))) <= 1e-9 * scale
    # one replay Counter-V1 word per class word and column
    assert n["replay_generator_words"] % 64 == 0 and n["replay_generator_words"] > 0
    q = p.run(probes=64, leverage_width=1, target_width=8, probe_tolerance=1e-11, fit_tolerance=1e-11,
              numerical=True)["numerical"]
    for k in ("conditional", "leverage", "usable", "mcse"):
        assert np.allclose(np.asarray(q[k]), np.asarray(n[k]), rtol=1e-9, atol=1e-18), k
    assert q["replay_generator_words"] == n["replay_generator_words"]


def test_observation_numerical_mcse_has_the_spread_of_the_estimator(oracle):
    cols = _panel(3)
    p = oracle.Problem(*cols, observation=True)
    est, mcse = [], []
    for seed in range(24):
        r = p.run(seed=2000 + seed, probes=64, leverage_width=8, target_width=8, probe_tolerance=1e-11,
                  fit_tolerance=1e-11, numerical=True)
        est.append(r["corrected"][:3])
        mcse.append(r["numerical"]["mcse"][:3])
    spread, reported = np.std(est, axis=0, ddof=1), np.mean(mcse, axis=0)
    assert np.all(reported > 0.5 * spread) and np.all(reported < 2.0 * spread), (reported, spread)
