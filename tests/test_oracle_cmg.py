"""Oracle CMG against the frozen FNV-1a digests of the vendored crate
(rust/vendor/cmg/tests/fixtures/frozen_bits.txt, [default] and
[default+release]) plus closed-form checks that a transposed operand, a dropped
term or a wrong sign would fail."""
import numpy as np
import pytest

from conftest import load_golden

D = load_golden("cmg_frozen_digests.json")["digests"]


@pytest.mark.parametrize("case", sorted(D))
def test_frozen_digest(oracle, case):
    family, kind = case.split("/")
    assert f"{oracle.frozen_digest(family, kind):016x}" == D[case]


def _random_graph(rng, n, m, components=1):
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = u != v
    u, v = u[keep], v[keep]
    # spanning path keeps it connected
    path = np.arange(n - 1)
    u = np.concatenate([u, path])
    v = np.concatenate([v, path + 1])
    w = rng.uniform(0.1, 3.0, len(u))
    return oracle_graph(n, u, v, w)


def oracle_graph(n, u, v, w):
    from oracle import oracle as O
    return O.Graph(n, u, v, w)


def _dense(g):
    u, v, w, d = g.export()
    L = np.zeros((g.n, g.n))
    np.add.at(L, (u, v), -w)
    np.add.at(L, (v, u), -w)
    L[np.arange(g.n), np.arange(g.n)] = d
    return L


def test_matvec_is_dense_laplacian(oracle):
    rng = np.random.default_rng(3)
    g = _random_graph(rng, 60, 200)
    L = _dense(g)
    x = rng.normal(size=60)
    assert np.allclose(g.matvec(x), L @ x, atol=1e-12)
    assert np.allclose(L.sum(axis=1), 0, atol=1e-12)


def test_cycle_is_symmetric_positive_on_quotient(oracle):
    # The stationary cycle is an SPSD operator on the range of L
    # (tests/preconditioner_properties.rs): x'My == y'Mx and x'Mx > 0.
    rng = np.random.default_rng(5)
    g = _random_graph(rng, 1500, 5000)
    c = oracle.Cmg(g, direct_threshold=64)
    assert len(c.levels()) >= 2
    a = rng.normal(size=1500); a -= a.mean()
    b = rng.normal(size=1500); b -= b.mean()
    Ma, Mb = c.apply(a), c.apply(b)
    Ma -= Ma.mean(); Mb -= Mb.mean()
    assert abs(a @ Mb - b @ Ma) <= 1e-10 * abs(a @ Ma)
    assert a @ Ma > 0 and b @ Mb > 0


def test_pcg_matches_dense_pseudoinverse(oracle):
    rng = np.random.default_rng(7)
    g = _random_graph(rng, 300, 900)
    c = oracle.Cmg(g, direct_threshold=40)
    L = _dense(g)
    xt = rng.normal(size=300)
    b = L @ xt
    x, d = c.solve(b, rel=1e-12, max_iterations=500)
    ref = np.linalg.pinv(L) @ b
    assert d.iterations > 0
    assert np.allclose(x - x.mean(), ref - ref.mean(), atol=1e-8 * np.abs(ref).max())


def test_hierarchy_labels_partition_and_contraction(oracle):
    # coarse graph == R L R^T (exact contraction, coarsen.rs:350-357)
    rng = np.random.default_rng(11)
    g = _random_graph(rng, 800, 2400)
    c = oracle.Cmg(g, direct_threshold=32)
    lv = c.levels()
    L0 = _dense(g)
    lab = lv[0]["labels"]
    R = np.zeros((lv[0]["coarse"], g.n))
    R[lab, np.arange(g.n)] = 1
    coarse = R @ L0 @ R.T
    Lc = np.zeros_like(coarse)
    n1 = lv[1]["n"]
    np.add.at(Lc, (lv[1]["u"], lv[1]["v"]), -lv[1]["w"])
    np.add.at(Lc, (lv[1]["v"], lv[1]["u"]), -lv[1]["w"])
    Lc[np.arange(n1), np.arange(n1)] = lv[1]["diag"]
    assert np.allclose(coarse, Lc, atol=1e-10)
