"""HIP-graph capture of the PCG iteration, through the C ABI.

A graph replays exactly the kernels the engine launches one by one, with the
same arguments and launch shapes, so every result must be bit-identical to
the solve without graphs: the point estimate, the numerical MCSE replay, both
V-cycle precisions, and the tuned and reference preconditioners.  Every solve
replays the two captured graphs once per PCG iteration, with the host's own
flag copies and polls between the replays; a changed configuration (width,
precision, preconditioner) is captured again, an unchanged one reuses its
graphs.
"""
import time

import numpy as np
import pytest

import fevc_amd as F
from fevc_amd import synth

pytestmark = pytest.mark.gpu

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")
SPEC = synth.PanelSpec(workers=4000, firms=300, periods=5, mobility=0.2, seed=11, name="gpu-graphs")


def _engine(spec=SPEC):
    d = synth.generate(spec)
    return F.Engine(*[np.asarray(d[k]) for k in COLS])


def _same(a, b):
    for k in ("plugin", "correction", "corrected", "numerical_mcse"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (k, a[k], b[k])


def test_graph_solves_are_bit_identical():
    e = _engine()
    try:
        for precision, mode in ((32, 1), (64, 0), (32, 0)):
            e.configure(precision, 32, 16)
            e.smoother(mode)
            e.graphs(False)
            print(f"precision {precision}, smoother mode {mode}", flush=True)
            plain = e.solve(probes=48, tolerance=1e-9)
            c0, l0 = e.graphs(True)
            first = e.solve(probes=48, tolerance=1e-9)
            c1, l1 = e.graphs()
            again = e.solve(probes=48, tolerance=1e-9)
            c2, l2 = e.graphs()
            _same(plain, first)
            _same(plain, again)
            assert c1 > c0 and l1 > l0  # captured and replayed
            assert c2 == c1 and l2 > l1  # unchanged configuration: graphs reused
    finally:
        e.release()


def test_graph_numerical_replay_is_bit_identical():
    e = _engine()
    try:
        e.configure(32, 64, 32)
        e.graphs(False)
        r0, n0 = e.solve_numerical(probes=16, tolerance=1e-10)
        e.graphs(True)
        r1, n1 = e.solve_numerical(probes=16, tolerance=1e-10)
        _same(r0, r1)
        for k in ("mcse", "conditional", "leverage", "usable"):
            assert np.array_equal(n0[k], n1[k]), k
        assert n0["certified_replay_rhs"] == n1["certified_replay_rhs"]
    finally:
        e.release()


def test_graph_solve_time():
    """Wall time of the same leverage solve with and without graphs (reported)."""
    e = _engine(synth.PanelSpec(workers=200000, firms=10000, periods=5, mobility=0.1, seed=3, name="graph-time"))
    try:
        out = {}
        for width in (16, 64):
            e.configure(32, width, width // 2)
            for on in (False, True, False, True):
                e.graphs(on)
                e.solve(probes=width, tolerance=1e-8)  # warm (capture on the first graph solve)
                t0 = time.perf_counter()
                e.solve(probes=4 * width, tolerance=1e-8)
                out.setdefault((width, on), []).append(time.perf_counter() - t0)
        for width in (16, 64):
            a, b = min(out[(width, False)]), min(out[(width, True)])
            print(f"width {width}: {a * 1e3:.1f} ms without graphs, {b * 1e3:.1f} ms with ({a / b:.2f}x)", flush=True)
    finally:
        e.release()


def test_graph_solves_with_observation_units_are_bit_identical():
    """Observation deletion and the stayer hybrid: the captured iterations
    replay the same kernels, whatever the deletion units."""
    d = synth.generate(synth.PanelSpec(workers=4000, firms=300, periods=5, mobility=0.2, seed=19, name="graphs-obs"))
    cols = [np.asarray(d[k], float) for k in COLS]
    cols[4] = np.random.default_rng(19).integers(1, 4, len(cols[0])).astype(float)
    for kind in ("observation", "stayers"):
        e = F.Engine(*cols, deletion_mode="observation" if kind == "observation" else "match")
        try:
            if kind == "stayers":
                e.augment_stayers(*synth.eligible_stayers(cols[0], cols[1], cols[3], cols[4], cols[5], e.mask()))
            e.configure(32, 32, 16)
            e.graphs(False)
            print(f"graphs with {kind} units", flush=True)
            plain = e.solve(probes=32, tolerance=1e-9)
            e.graphs(True)
            graphed = e.solve(probes=32, tolerance=1e-9)
            c, l = e.graphs()
            assert c > 0 and l > 0
            _same(plain, graphed)
        finally:
            e.release()
