"""Device layout of the CMG hierarchy (CPU checks, no GPU).

The device renumbers every level so that aggregates are contiguous, and cuts
rows and aggregates into chunks.  These tests check the layout against the
host hierarchy it comes from, and run a numpy restatement of the device
V-cycle over the renumbered layout (fused scale + residual, contiguous
aggregation, prolong + smooth, dense grounded terminal inverse) against the
oracle's V-cycle (preconditioner/cycle.rs:136-478).
"""
import numpy as np
import pytest

F = pytest.importorskip("fevc_amd")


def _graphs():
    rng = np.random.default_rng(0)
    n = 4000
    u = rng.integers(0, n, 3 * n)
    v = rng.integers(0, n, 3 * n)
    keep = u != v
    u, v = u[keep], v[keep]
    path = np.arange(n - 1)
    yield "random", n, np.concatenate([u, path]), np.concatenate([v, path + 1]), rng.uniform(0.1, 3.0, len(u) + n - 1)
    s = 90
    idx = np.arange(s * s).reshape(s, s)
    gu = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    gv = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    yield "grid", s * s, gu, gv, 1.0 + ((gu * 7 + gv * 3) % 11) / 10.0
    # a star-heavy graph: hubs with thousands of neighbours (long rows, big aggregates)
    hubs = 6
    m = 6000
    leaf = np.arange(hubs, hubs + m)
    hu = np.concatenate([leaf % hubs, leaf[:-1]])
    hv = np.concatenate([leaf, leaf[1:]])
    yield "hubs", hubs + m, hu, hv, 0.5 + (hv % 5) * 0.3


def _panel_graph(oracle):
    from fevc_amd import synth
    d = synth.generate(synth.PanelSpec(workers=20000, firms=1500, periods=5, mobility=0.2,
                                       pareto_shape=1.1, size_cap=500.0, seed=9, name="layout"))
    p = oracle.Problem(*[d[k] for k in ("worker", "firm", "deletion", "outcome", "frequency", "target")])
    g, _ = p.hybrid_graph()
    u, v, w, _ = g.export()
    return g.n, u, v, w


def _dense_rows(lay, n):
    A = np.zeros((n, n))
    rp, col, val = lay["row_ptr"], lay["col"], lay["val"]
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            A[i, i] += val[k]
            A[i, col[k]] -= val[k]
    return A


def _check_chunks(chunks, multi, ptr, cap):
    segs = len(ptr) - 1
    covered = np.zeros(segs, np.int64)
    last_end = {}
    for q, (s, b, e, m) in enumerate(chunks):
        assert ptr[s] <= b <= e <= ptr[s + 1]
        assert e - b <= cap
        assert last_end.get(s, ptr[s]) == b  # contiguous, in order
        last_end[s] = e
        covered[s] += 1
        if m != 0xFFFFFFFF:
            first, end = multi[m]
            assert first <= q < end
            assert all(chunks[first:end, 0] == s)
    for s in range(segs):
        assert last_end.get(s, None) == ptr[s + 1] or (ptr[s] == ptr[s + 1] and covered[s] == 1)
    single = chunks[:, 3] == 0xFFFFFFFF
    assert np.all(np.bincount(chunks[single, 0], minlength=segs)[covered > 1] == 0)


@pytest.mark.parametrize("case", ["random", "grid", "hubs", "panel"])
def test_layout_invariants(oracle, case):
    if case == "panel":
        n, u, v, w = _panel_graph(oracle)
    else:
        n, u, v, w = next((g[1:] for g in _graphs() if g[0] == case))
    levels, _ = F.build_hierarchy(n, u, v, w)
    lays = F.hierarchy_layout(n, u, v, w, chunk=64)
    assert len(lays) == len(levels)
    for li, (lv, lay) in enumerate(zip(levels, lays)):
        nl = lv["n"]
        perm = lay["perm"].astype(np.int64)
        assert sorted(perm.tolist()) == list(range(nl))
        # renumbered CSR == original graph
        if nl <= 3000:
            A = _dense_rows(lay, nl)
            B = np.zeros((nl, nl))
            np.add.at(B, (lv["u"], lv["v"]), -lv["w"])
            np.add.at(B, (lv["v"], lv["u"]), -lv["w"])
            B[np.arange(nl), np.arange(nl)] = lv["diag"]
            assert np.allclose(A, B[np.ix_(perm, perm)], rtol=0, atol=1e-12)
        deg = np.diff(lay["row_ptr"].astype(np.int64))
        odeg = np.bincount(lv["u"], minlength=nl) + np.bincount(lv["v"], minlength=nl)
        assert np.array_equal(deg, odeg[perm])
        assert np.allclose(lay["inv"], np.where(lv["diag"][perm] > 0, 0.5 / lv["diag"][perm], 0.0))
        _check_chunks(lay["row_chunks"], lay["row_multi"], lay["row_ptr"].astype(np.int64), 64)
        if li + 1 < len(levels):
            lab = lay["lab"].astype(np.int64)
            assert np.all(np.diff(lab) >= 0)
            cperm = lays[li + 1]["perm"].astype(np.int64)
            cpos = np.empty_like(cperm)
            cpos[cperm] = np.arange(len(cperm))
            assert np.array_equal(lab, cpos[lv["labels"][perm]])
            agg = lay["agg_ptr"].astype(np.int64)
            assert np.array_equal(agg, np.concatenate([[0], np.cumsum(np.bincount(lab, minlength=lay["coarse"]))]))
            _check_chunks(lay["agg_chunks"], lay["agg_multi"], agg, 64)
    if case in ("hubs", "panel"):
        assert any(len(l["row_multi"]) for l in lays)


def _check_grouped(chunks, multi, ptr, cap, group):
    """Grouped chunk table: every segment is finished exactly once, either in
    a group of 2..group consecutive short segments (at most cap / 4 entries
    each) or by the plain single/multi chunks."""
    segs = len(ptr) - 1
    tag = chunks[:, 3]
    is_group = (tag > 0xFFFFFF00) & (tag != 0xFFFFFFFF)
    owner = np.zeros(segs, np.int64)
    for s, b, e, m in chunks[is_group]:
        c = int(m) - 0xFFFFFF00
        assert 2 <= c <= group
        assert b == ptr[s] and e == ptr[s + c]
        assert np.all(np.diff(ptr[s:s + c + 1]) <= cap // 4)
        owner[s:s + c] += 1
    plain = chunks[~is_group]
    grouped = owner > 0
    assert np.all(owner <= 1)
    assert np.all(np.isin(plain[:, 0], np.flatnonzero(grouped), invert=True))
    rest = np.flatnonzero(~grouped)
    assert np.array_equal(np.unique(plain[:, 0]), rest)
    # the plain chunks of the other segments are those of the ungrouped table
    full, _ = F.chunk_table(ptr, cap, 1)
    keep = np.isin(full[:, 0], rest)
    assert np.array_equal(plain[:, :3], full[keep, :3])
    for q in np.flatnonzero(plain[:, 3] != 0xFFFFFFFF):
        first, end = multi[plain[q, 3]]
        idx = np.flatnonzero(~is_group)[q]
        assert first <= idx < end and np.all(chunks[first:end, 0] == plain[q, 0])


@pytest.mark.parametrize("group", [1, 2, 4, 16])
def test_grouped_chunk_tables(oracle, group):
    rng = np.random.default_rng(group)
    deg = np.concatenate([rng.integers(0, 20, 3000), [0, 0, 0, 500, 5000, 1, 16, 17, 64, 65]])
    rng.shuffle(deg)
    ptrs = [np.concatenate([[0], np.cumsum(deg)]), np.zeros(1, np.int64), np.array([0, 0, 0, 3]),
            np.concatenate([[0], np.cumsum(np.full(37, 16))])]
    n, u, v, w = _panel_graph(oracle)
    lays = F.hierarchy_layout(n, u, v, w, chunk=64)
    ptrs += [lays[0]["row_ptr"].astype(np.int64), lays[0]["agg_ptr"].astype(np.int64)]
    for ptr in ptrs:
        for cap in (16, 64):
            chunks, multi = F.chunk_table(ptr, cap, group)
            if group == 1:
                _check_chunks(chunks, multi, ptr, cap)
            else:
                _check_grouped(chunks, multi, ptr, cap, group)
    # the finest level of a panel has mostly short rows: grouping cuts the chunk count
    full, _ = F.chunk_table(lays[0]["row_ptr"], 64, 1)
    grouped, _ = F.chunk_table(lays[0]["row_ptr"], 64, 4)
    if group == 4:
        assert len(grouped) < 0.7 * len(full)
    with pytest.raises(F.EngineError):
        F.chunk_table(np.array([0, 5, 3]), 64, group)
    with pytest.raises(F.EngineError):
        F.chunk_table(np.array([0, 5]), 64, 0)


def _vcycle(lays, levels, r0):
    """numpy restatement of the device V-cycle over the renumbered layout."""
    def matvec(lay, x):
        rp, col, val = lay["row_ptr"].astype(np.int64), lay["col"], lay["val"]
        rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
        return np.bincount(rows, weights=val * (x[rows] - x[col]), minlength=len(rp) - 1)

    def level(l, braw, mean, iterations):
        lay, lv = lays[l], levels[l]
        b = braw - mean
        if lv["terminal"]:
            if lv["terminal"] == 1:
                return lay["terminal_inverse"] @ b
            return lay["inv"] * b
        agg = lay["agg_ptr"].astype(np.int64)
        lab = lay["lab"].astype(np.int64)
        x = None
        for it in range(iterations):
            if it == 0:
                x = lay["inv"] * b
            else:
                x = x + lay["inv"] * (b - matvec(lay, x))
            r = b - matvec(lay, x)
            coarse = np.add.reduceat(r, agg[:-1]) if len(r) else r
            e = level(l + 1, coarse, coarse.mean(), lv["repeat"])
            # prolong + smooth from the kept residual: A x' = (b - r) + A P e
            pe = e[lab]
            x = x + pe + lay["inv"] * (r - matvec(lay, pe))
        return x

    return level(0, r0, 0.0, 1)


@pytest.mark.parametrize("case", ["random", "grid", "hubs", "panel"])
def test_renumbered_vcycle_matches_oracle(oracle, case):
    if case == "panel":
        n, u, v, w = _panel_graph(oracle)
    else:
        n, u, v, w = next((g[1:] for g in _graphs() if g[0] == case))
    levels, _ = F.build_hierarchy(n, u, v, w)
    lays = F.hierarchy_layout(n, u, v, w, chunk=64)
    cmg = oracle.Cmg(oracle.Graph(n, u, v, w), direct_threshold=700, max_levels=128)
    rng = np.random.default_rng(1)
    for _ in range(3):
        r = rng.standard_normal(n)
        r -= r.mean()
        want = cmg.apply(r)
        perm = lays[0]["perm"].astype(np.int64)
        z = np.empty(n)
        z[perm] = _vcycle(lays, levels, r[perm])
        assert np.linalg.norm(z - want) <= 1e-10 * np.linalg.norm(want)
