"""Pins of the dense exact oracle (oracle/exact.py), CPU only.

* the registered dense fixture of fevc/tests/python/test_dense_oracle.py:54-76
  (fixed-offset nuisance, two controls, match identifiers): plug-in and
  correction to 2e-13;
* the seven-edge closed-form leverages 11/15 and 3/5
  (test_dense_oracle.py:186-200), through the Rust-route restatement;
* the known effects of exact.rs:397-414;
* block Woodbury deleted residuals equal direct refits without the block
  (test_dense_oracle.py:79-90);
* frequency weights equal literal expansion (test_dense_oracle.py:93-119);
* the Rust-route restatement (zero-sum embedding, square-root frequency
  blocks) equals the expanded quotient restatement on leave-out connected
  panels, and both satisfy the accounting identities.
"""
import json
import os

import numpy as np
import pytest

from oracle import exact as X

HERE = os.path.dirname(os.path.abspath(__file__))
COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")


def _fixture():
    with open(os.path.join(HERE, "golden", "exact_dense_fixture.json")) as fh:
        return json.load(fh)


def test_registered_dense_fixture():
    d = _fixture()["dense_fixture"]
    r = X.dense_kss(np.array(d["y"]), d["worker"], d["firm"], controls=np.array(d["controls"]),
                    deletion="match", deletion_id=d["match"], nuisance="fixedoffset")
    tol = d["tolerance"]
    assert np.allclose(r["plugin"], d["fixedoffset_match_plugin"], atol=tol, rtol=tol)
    assert np.allclose(r["correction"], d["fixedoffset_match_correction"], atol=tol, rtol=tol)


def test_seven_edge_leverages():
    s = _fixture()["seven_edges"]
    e = np.array(s["edges"])
    n = len(e)
    r = X.exact_estimator(e[:, 0], e[:, 1], np.arange(n), np.linspace(-1, 1, n), np.ones(n), np.ones(n))
    assert np.allclose(r["unit_leverage"], s["leverage"], rtol=0, atol=2e-15)
    assert r["max_leverage"] == pytest.approx(11 / 15, abs=2e-15)


def test_known_effects():
    k = _fixture()["known_effects"]
    r = X.exact_estimator(k["worker"], k["firm"], np.arange(4), k["outcome"], np.ones(4), np.ones(4))
    assert np.allclose(r["beta"][:2], k["worker_effect"], atol=1e-12)
    assert np.allclose(r["beta"][2:], k["firm_effect"], atol=1e-12)
    assert np.allclose(r["residual"], 0.0, atol=1e-12)


def test_block_woodbury_equals_direct_refit():
    d = _fixture()["dense_fixture"]
    r = X.dense_kss(np.array(d["y"]), d["worker"], d["firm"], controls=np.array(d["controls"]),
                    deletion="match", deletion_id=d["match"])
    direct = X.refit_residuals(r["x"], r["y"], r["deletion"])
    for block, want in direct.items():
        use = r["deletion"] == block
        xg = r["x"][use]
        maker = np.eye(use.sum()) - xg @ r["inverse"] @ xg.T
        got = np.linalg.solve(maker, r["residual"][use])
        assert np.allclose(got, want, atol=2e-10, rtol=2e-10)


def _panel(seed, workers=80, firms=20, periods=4):
    from fevc_amd import synth
    from oracle import oracle as O
    d = synth.generate(synth.PanelSpec(workers=workers, firms=firms, periods=periods, mobility=0.5,
                                       seed=seed, name="exact"))
    cols = [np.asarray(d[k]) for k in COLS]
    keep = O.Problem(*cols).mask()
    return [c[keep] for c in cols]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_rust_route_equals_expanded_quotient(seed):
    w, f, g, y, fr, t = _panel(seed)
    a = X.exact_estimator(w, f, g, y, fr, t)
    b = X.dense_kss(y, w, f, frequency=fr, target=t, deletion="match", deletion_id=g)
    for k in ("plugin", "correction", "corrected"):
        assert np.allclose(a[k], b[k], rtol=1e-10, atol=1e-12), (k, a[k], b[k])
    assert a["deletion_units"] == b["deletion_units"]
    assert a["max_leverage"] == pytest.approx(b["leverage_max"], abs=1e-12)
    for r in (a, b):
        for k in ("plugin", "correction"):
            v = r[k]
            assert abs(v[3] - (v[0] + v[1] + 2 * v[2])) < 1e-12


def test_frequency_equals_literal_expansion():
    w, f, g, y, fr, t = _panel(4)
    rng = np.random.default_rng(4)
    fr = rng.integers(1, 4, fr.size)
    t = rng.uniform(0.5, 2.0, fr.size)
    a = X.exact_estimator(w, f, g, y, fr, t)
    idx = np.repeat(np.arange(fr.size), fr)
    b = X.exact_estimator(w[idx], f[idx], g[idx], y[idx], np.ones(idx.size), (t / fr)[idx])
    for k in ("plugin", "correction"):
        assert np.allclose(a[k], b[k], rtol=1e-10, atol=1e-12), k
    assert a["max_leverage"] == pytest.approx(b["max_leverage"], abs=1e-12)


def test_nonestimable_block_is_rejected():
    # worker 2 is observed at one firm only: deleting that match leaves it unidentified
    w = [0, 0, 1, 1, 2]
    f = [0, 1, 0, 1, 1]
    with pytest.raises(X.ExactError) as e:
        X.exact_estimator(w, f, np.arange(5), [1.0, 0.5, -0.2, 0.3, 0.9], np.ones(5), np.ones(5))
    assert e.value.code == X.NONESTIMABLE_DELETION


# ---------------------------------------------------------------- observation deletion
# graph.rs:1268-1298: (worker, firm, deletion, frequency) rows and the retained mask
OBSERVATION_CASES = [
    ([(1, 1, 1, 2), (2, 1, 2, 1), (2, 2, 3, 1), (3, 1, 4, 1), (3, 2, 5, 1)], [True] * 5),
    ([(1, 1, 1, 1), (2, 1, 2, 1), (2, 2, 3, 1), (3, 1, 4, 1), (3, 2, 5, 1)], [False, True, True, True, True]),
]


def _observation_columns(rows):
    a = np.array(rows, dtype=np.float64)
    n = len(rows)
    return [a[:, 0], a[:, 1], a[:, 2], np.linspace(-1.0, 1.0, n), a[:, 3], np.ones(n)]


@pytest.mark.parametrize("case", range(len(OBSERVATION_CASES)))
def test_observation_selection_reference_cases(case):
    from oracle import oracle as O
    rows, want = OBSERVATION_CASES[case]
    cols = _observation_columns(rows)
    assert list(O.Problem(*cols, observation=True).mask()) == want
    import fevc_amd as F
    e = F.Engine(*cols, upload=False, deletion_mode="observation")
    try:
        assert list(e.mask()) == want
    finally:
        e.release()


def _observation_panel(seed, workers=120, firms=20, periods=4):
    from fevc_amd import synth
    from oracle import oracle as O
    d = synth.generate(synth.PanelSpec(workers=workers, firms=firms, periods=periods, mobility=0.3,
                                       seed=seed, name="exact-observation"))
    cols = [np.asarray(d[k]) for k in COLS]
    keep = O.Problem(*cols, observation=True).mask()
    return cols, keep


def _worker_articulations(w, f):
    """Brute force: workers whose removal disconnects the worker-firm graph."""
    def components(rows_w, rows_f):
        nodes = {("w", a) for a in rows_w} | {("f", b) for b in rows_f}
        parent = {v: v for v in nodes}

        def find(v):
            while parent[v] != v:
                parent[v] = parent[parent[v]]
                v = parent[v]
            return v
        for a, b in zip(rows_w, rows_f):
            parent[find(("w", a))] = find(("f", b))
        return len({find(v) for v in nodes})
    base = components(w, f)
    out = []
    for a in np.unique(w):
        keep = w != a
        if keep.any() and components(w[keep], f[keep]) > base:
            out.append(a)
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_observation_selection_host_matches_oracle(seed):
    import fevc_amd as F
    cols, keep = _observation_panel(seed)
    e = F.Engine(*cols, upload=False, deletion_mode="observation")
    try:
        assert np.array_equal(e.mask(), keep)
    finally:
        e.release()
    w, f, fr = cols[0][keep], cols[1][keep], cols[4][keep]
    # the retained graph: connected, no physical-singleton worker, no worker articulation
    phys = {a: fr[w == a].sum() for a in np.unique(w)}
    assert min(phys.values()) >= 2
    assert not _worker_articulations(w, f)
    # stayers with several copies are kept
    firms_per_worker = {a: np.unique(f[w == a]).size for a in np.unique(w)}
    assert min(firms_per_worker.values()) == 1


def test_observation_restatements_agree():
    cols, keep = _observation_panel(3)
    w, f, g, y, fr, t = [c[keep] for c in cols]
    a = X.exact_estimator(w, f, g, y, fr, t, mode="observation")
    b = X.dense_kss(y, w, f, frequency=fr, target=t, deletion="observation")
    for k in ("plugin", "correction", "corrected"):
        assert np.allclose(a[k], b[k], rtol=1e-10, atol=1e-12), (k, a[k], b[k])
    assert a["deletion_units"] == b["deletion_units"] == int(fr.sum())
    assert a["max_leverage"] == pytest.approx(b["leverage_max"], abs=1e-12)


def test_observation_frequency_deletes_one_physical_copy():
    cols, keep = _observation_panel(4)
    w, f, g, y, fr, t = [c[keep] for c in cols]
    rng = np.random.default_rng(8)
    fr = rng.integers(2, 4, fr.size)
    t = rng.uniform(0.5, 2.0, fr.size)
    a = X.exact_estimator(w, f, g, y, fr, t, mode="observation")
    idx = np.repeat(np.arange(fr.size), fr)
    b = X.exact_estimator(w[idx], f[idx], np.arange(idx.size), y[idx], np.ones(idx.size), (t / fr)[idx],
                          mode="observation")
    for k in ("plugin", "correction"):
        assert np.allclose(a[k], b[k], rtol=1e-10, atol=1e-12), k
