"""Pins of the oracle's CMG_FULL_V2 refinement ladder and harmonic lift.

- refinement_pcg_options, failing_columns, euclidean_norm and the factor
  schedule against the reference's own known answers (full_cmg.rs:1985-2006,
  2120-2139);
- the harmonic lift (cmg_impl.rs.in:459-500) against brute-force weighted
  means, and the closed form that the hybrid Laplacian of the lifted vector
  vanishes on auxiliary rows and equals the dense firm Schur complement on
  firm rows (cmg_impl.rs.in:771-794 states the same identity);
- the reference's grid-and-hub fixture (full_cmg.rs:2008-2118): the batch
  solve refines, certifies the gate and returns the closed-form solution.
"""
import numpy as np
import pytest

from refinement_fixture import grid_and_hub, outcome_rhs, smooth_psi


def test_refinement_options_known_answers(oracle):
    # full_cmg.rs:1985-2004
    rel, ab = oracle.refinement_options(1e-10, 0.1, 1e-9, 4.0)
    assert rel == 0.0 and abs(ab - 4.0e-10) <= 1e-24
    rel, ab = oracle.refinement_options(1e-10, 0.01, 1e-9, 0.0)
    assert abs(rel - 1.0e-12) <= 1e-26 and ab == 0.0
    # a nonfinite smallest norm also keeps the relative stop
    rel, ab = oracle.refinement_options(1e-8, 0.001, 1e-7, float("inf"))
    assert abs(rel - 1e-11) <= 1e-25 and ab == 0.0


def test_refinement_schedule_and_gate_known_answers(oracle):
    # full_cmg.rs:2120-2139
    assert oracle.refinement_factors() == [0.1, 0.01, 0.001, 0.0]
    assert oracle.failing_columns([1e-7, 1e-5, 1e-7, float("nan")], [1e-7, 1e-7, 1e-5, 1e-7], 1e-6) == [1, 2, 3]
    assert oracle.failing_columns([1e-6, 1e-6], [1e-6, float("inf")], 1e-6) == [1]
    # full_cmg.rs:2005-2006
    assert abs(oracle.euclidean_norm([3.0e-200, 4.0e-200]) - 5.0e-200) <= 1e-214
    assert oracle.euclidean_norm([0.0, 0.0]) == 0.0


def _star_panel(seed):
    """Workers with up to 8 distinct firms, so the degree-4 hybrid graph has
    auxiliary star vertices (workers with more than four cells)."""
    rng = np.random.default_rng(seed)
    w, f, q = [], [], []
    for i in range(30):
        k = int(rng.integers(2, 9))
        for j in rng.choice(12, size=k, replace=False):
            w.append(i + 1)
            f.append(int(j) + 1)
            q.append(int(rng.integers(1, 5)))
    n = len(w)
    return [np.array(w, float), np.array(f, float), np.arange(1, n + 1, dtype=float), np.zeros(n),
            np.array(q, float), np.ones(n)]


def _dense_hybrid(g):
    u, v, wt, diag = g.export()
    L = np.zeros((g.n, g.n))
    np.add.at(L, (u, v), -wt)
    np.add.at(L, (v, u), -wt)
    L[np.arange(g.n), np.arange(g.n)] = diag
    return L


def _dense_firm_schur(p):
    cw, cf, wt = p.array("cell_worker"), p.array("cell_firm"), p.array("cell_weight")
    D = np.zeros((p.workers, p.firms))
    D[cw, cf] = wt
    return np.diag(D.sum(0)) - D.T @ np.diag(1 / D.sum(1)) @ D


@pytest.mark.parametrize("seed", range(4))
def test_harmonic_lift_brute_force_and_schur(oracle, seed):
    cols = _star_panel(seed)
    p = oracle.Problem(*cols, mask=np.ones(len(cols[0]), bool))
    g, aux = p.hybrid_graph()
    assert aux > 0 and g.n == p.firms + aux
    firm = np.random.default_rng(100 + seed).normal(size=p.firms)
    lifted = p.hybrid_lift(firm)
    assert np.array_equal(lifted[:p.firms], firm)
    # brute force: every auxiliary vertex is the weighted mean of its star
    L = _dense_hybrid(g)
    for a in range(p.firms, g.n):
        w = -L[a, :p.firms]
        assert np.all(L[a, p.firms:][np.arange(g.n - p.firms) != a - p.firms] == 0.0)
        assert abs(lifted[a] - (w @ firm) / w.sum()) <= 1e-13 * max(1.0, np.abs(firm).max())
    # closed form: L lift = (S_F firm, 0)
    y = L @ lifted
    assert np.max(np.abs(y[p.firms:])) <= 1e-12 * np.abs(L).max() * np.abs(firm).max()
    assert np.allclose(y[:p.firms], _dense_firm_schur(p) @ firm, atol=1e-12 * np.abs(L).max())


def test_lift_mistake_is_caught(oracle):
    """A lift that dropped the weights (plain neighbour mean) fails the
    harmonic identity, so the check above has teeth."""
    cols = _star_panel(0)
    p = oracle.Problem(*cols, mask=np.ones(len(cols[0]), bool))
    g, _ = p.hybrid_graph()
    L = _dense_hybrid(g)
    firm = np.random.default_rng(7).normal(size=p.firms)
    wrong = p.hybrid_lift(firm)
    for a in range(p.firms, g.n):
        nb = L[a, :p.firms] != 0
        wrong[a] = firm[nb].mean()
    assert np.max(np.abs((L @ wrong)[p.firms:])) > 1e-6


def probe_triple(p, wr, fr, seed=3):
    rng = np.random.default_rng(seed)
    a = rng.choice([-1.0, 1.0], p.cells)
    pw, pf = np.zeros(p.workers), np.zeros(p.firms)
    np.add.at(pw, p.array("cell_worker"), a)
    np.add.at(pf, p.array("cell_firm"), a)
    return np.stack([wr, 0 * wr, pw], 1), np.stack([fr, 0 * fr, pf], 1)


@pytest.fixture(scope="module")
def grid(oracle):
    cols, grid_firms = grid_and_hub()
    p = oracle.Problem(*cols)
    assert p.rows == len(cols[0])  # every row retained, as compress(all) in the reference
    psi = smooth_psi(p.firms, grid_firms)
    wr, fr = outcome_rhs(p.array("cell_worker"), p.array("cell_firm"), p.array("cell_weight"), p.workers,
                         p.firms, psi)
    return p, psi, wr, fr


def test_grid_fixture_refines_and_certifies(grid):
    p, psi, wr, fr = grid
    tol = 1e-10
    gate = max(10 * tol, 1e-11)  # complete_residual_tolerance (full_cmg.rs:399-412)
    wo, fo, rec, ref = p.solve_columns(wr[:, None], fr[:, None], tolerance=tol, threads=4)
    assert ref["rounds"] >= 1 and ref["refined"] >= 1  # the operator-scaled stop left the gate failing
    assert rec["complete"][0] <= gate and rec["reduced"][0] <= gate
    m = psi.sum() / p.firms
    assert np.max(np.abs(fo[:, 0] - (psi - m))) <= 1e-8 * np.abs(psi).max()
    assert np.max(np.abs(wo[:, 0] - m)) <= 1e-8 * np.abs(psi).max()


def test_grid_fixture_refines_failing_columns_only(grid):
    """Three columns: the smooth solution, a zero right-hand side and a
    Rademacher probe (one sign per cell added to its worker and firm, so the
    two-way system stays consistent).  The hub makes the smooth and the probe
    column fail the first pass; the zero column passes and is not refined."""
    p, psi, wr, fr = grid
    wr3, fr3 = probe_triple(p, wr, fr)
    wo, fo, rec, ref = p.solve_columns(wr3, fr3, tolerance=1e-10, threads=4)
    assert ref["rounds"] >= 1 and ref["refined"] == 2 * ref["rounds"]
    assert np.all(rec["complete"] <= 1e-9) and np.all(rec["reduced"] <= 1e-9)
    assert np.all(wo[:, 1] == 0.0) and np.all(fo[:, 1] == 0.0)
    m = psi.sum() / p.firms
    assert np.max(np.abs(fo[:, 0] - (psi - m))) <= 1e-8 * np.abs(psi).max()
