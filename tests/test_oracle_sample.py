"""Oracle sample selection against a brute-force restatement of the
leave-out rules (graph.rs:262-512): articulation workers by deletion + BFS,
bridge deletion units by deletion + connectivity, largest component by
(firm count, mass) with a tie refused.  Also the hybrid graph against the dense
firm Schur complement (cmg_impl.rs.in:771-820 closed form)."""
import numpy as np
import pytest


def _components(nodes, edges):
    parent = {x: x for x in nodes}

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x
    for a, b in edges:
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[ra] = rb
    comp = {}
    for x in nodes:
        comp.setdefault(find(x), []).append(x)
    return list(comp.values())


def brute_select(worker, firm, deletion, freq):
    n = len(worker)
    active = np.ones(n, bool)

    def largest():
        rows = np.flatnonzero(active)
        edges = {(("w", worker[r]), ("f", firm[r])) for r in rows}
        nodes = {x for e in edges for x in e}
        comps = _components(nodes, edges)
        ranks = []
        for c in comps:
            cs = set(c)
            firms = sum(1 for x in c if x[0] == "f")
            mass = sum(int(freq[r]) for r in rows if ("w", worker[r]) in cs)
            ranks.append((firms, mass, cs))
        best = max((r[0], r[1]) for r in ranks)
        winners = [r for r in ranks if (r[0], r[1]) == best]
        if len(winners) != 1:
            raise ValueError("tie")
        keep = winners[0][2]
        for r in rows:
            if ("w", worker[r]) not in keep:
                active[r] = False

    def units_per_worker():
        cnt = {}
        for r in np.flatnonzero(active):
            cnt.setdefault(worker[r], set()).add(deletion[r])
        return {w: len(s) for w, s in cnt.items()}

    largest()
    upw = units_per_worker()
    for r in np.flatnonzero(active):
        if upw[worker[r]] <= 1:
            active[r] = False
    largest()
    while True:
        largest()
        upw = units_per_worker()
        bad = {w for w, c in upw.items() if c == 1}
        if bad:
            for r in np.flatnonzero(active):
                if worker[r] in bad:
                    active[r] = False
            continue
        rows = np.flatnonzero(active)
        edges = {(("w", worker[r]), ("f", firm[r])) for r in rows}
        nodes = {x for e in edges for x in e}
        base = len(_components(nodes, edges))
        arts = set()
        for w in {worker[r] for r in rows}:
            x = ("w", w)
            sub = [e for e in edges if x not in e]
            subnodes = {y for e in sub for y in e} | (nodes - {x} - {y for e in edges for y in e if x in e})
            subnodes = (nodes - {x})
            if len(_components(subnodes, sub)) > base:
                arts.add(w)
        if arts:
            for r in rows:
                if worker[r] in arts:
                    active[r] = False
            continue
        units = {}
        for r in rows:
            units[deletion[r]] = (("w", worker[r]), ("f", firm[r]))
        bridges = set()
        ulist = list(units.items())
        for d, (a, b) in ulist:
            rest = [e for dd, e in ulist if dd != d]
            comps = _components(nodes, rest)
            ca = next(c for c in comps if a in c)
            if b not in ca:
                bridges.add(d)
        if not bridges:
            break
        for r in rows:
            if deletion[r] in bridges:
                active[r] = False
    return active


def random_panel(seed, workers=24, firms=6, periods=4, p_move=0.35):
    rng = np.random.default_rng(seed)
    w, f, d = [], [], []
    for i in range(workers):
        cur = rng.integers(firms)
        for t in range(periods):
            if t and rng.random() < p_move:
                cur = rng.integers(firms)
            w.append(i + 1)
            f.append(int(cur) + 1)
    w = np.array(w); f = np.array(f)
    # default deletion id: the worker-firm pair
    d = w * 1000 + f
    freq = rng.integers(1, 4, len(w))
    return w, f, d, freq


@pytest.mark.parametrize("seed", range(12))
def test_selection_matches_brute_force(oracle, seed):
    w, f, d, q = random_panel(seed)
    y = np.linspace(-1, 1, len(w))
    try:
        expected = brute_select(w, f, d, q)
    except ValueError:
        with pytest.raises(oracle.OracleError):
            oracle.Problem(w, f, d, y, q, np.ones(len(w)))
        return
    if expected.sum() == 0:
        with pytest.raises(oracle.OracleError):
            oracle.Problem(w, f, d, y, q, np.ones(len(w)))
        return
    try:
        p = oracle.Problem(w, f, d, y, q, np.ones(len(w)))
    except oracle.OracleError as e:
        assert e.code in (30, 31), str(e)
        return
    assert np.array_equal(p.mask(), expected)


def _dense_firm_schur(p):
    cw, cf, wt = p.array("cell_worker"), p.array("cell_firm"), p.array("cell_weight")
    D = np.zeros((p.workers, p.firms))
    D[cw, cf] = wt
    dw, df = D.sum(1), D.sum(0)
    return np.diag(df) - D.T @ np.diag(1 / dw) @ D


def test_degree_four_hybrid_equals_firm_schur(oracle):
    # cmg_impl.rs.in:797-820 fixture (degree-five worker becomes a star)
    w = [1, 1, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 4]
    f = [1, 2, 1, 2, 3, 1, 2, 3, 4, 1, 2, 3, 4, 5]
    q = [2, 3, 1, 2, 3, 1, 2, 3, 4, 1, 3, 5, 7, 9]
    n = len(w)
    from oracle import oracle as O
    # compress(&vec![true; rows]) as in the reference unit test
    p = O.Problem(w, f, list(range(1, n + 1)), [0.0] * n, q, [1.0] * n, mask=np.ones(n, bool))
    assert p.rows == n
    g, aux = p.hybrid_graph()
    assert aux == 1 and g.n == p.firms + 1
    u, v, wt, diag = g.export()
    L = np.zeros((g.n, g.n))
    np.add.at(L, (u, v), -wt); np.add.at(L, (v, u), -wt)
    L[np.arange(g.n), np.arange(g.n)] = diag
    F = p.firms
    S = L[:F, :F] - L[:F, F:] @ np.linalg.solve(L[F:, F:], L[F:, :F])
    assert np.allclose(S, _dense_firm_schur(p), atol=1e-12)
