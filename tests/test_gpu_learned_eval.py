"""The built-in learned smoother on its held-out evaluation graph.

Panel: 10,000,000 workers x T = 3 (30M rows), 500,000 firms with Pareto(1.05)
sizes truncated at 5,000, mobility 0.1, seed 7 - a graph of another size,
period count and firm-size law than the 50k-1M-worker training panels.  The
same 64 leverage probes (identical Counter-V1 draws) are solved to rel.res
1e-8 with the reference cycle and with the library's default preconditioner
(learned weights by BF16 inference, tuned cycle shape, chosen by the
calibration batch).  Bars: the learned configuration is active and needs
fewer PCG iterations per solve than the reference cycle; the 16-probe
estimate's components agree within 1e-6 relative between the two (both
certified to the same tolerance).
"""
import numpy as np
import pytest

import fevc_amd as F
from fevc_amd import synth

pytestmark = pytest.mark.gpu

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")
BIG = synth.PanelSpec(workers=10_000_000, firms=500_000, periods=3, mobility=0.1, pareto_shape=1.05,
                      size_cap=5000.0, seed=7, name="akm-10M-workers-500k-firms-T3")


def test_learned_smoother_on_held_out_10M_graph():
    assert F.builtin_smoother() is not None, "the library must carry trained smoother weights"
    print("learned eval: generate", BIG.name, flush=True)
    d = synth.generate(BIG)
    cols = [np.asarray(d[k]) for k in COLS]
    del d
    print("learned eval: prepare", flush=True)
    e = F.Engine(*cols)
    del cols
    try:
        print("learned eval: dims", e.dims, flush=True)
        e.configure(32, 64, 32)
        out = {}
        for mode in (0, 1):
            e.smoother(mode)
            t = e.time_leverage(64, 32, 1e-8, 1)
            res = e.solve(probes=16, tolerance=1e-8, fit_tolerance=1e-8)
            out[mode] = (t["iterations"] / t["solves"], res, e.smoother_status())
            print(f"learned eval: mode {mode}: {out[mode][0]:.2f} iterations per solve, status {out[mode][2]}",
                  flush=True)
    finally:
        e.release()
    (it_plain, plain, _), (it_learned, learned, st) = out[0], out[1]
    assert st["active"], st
    assert it_learned < it_plain, (it_learned, it_plain)
    for key in ("plugin", "correction", "corrected"):
        a, b = np.asarray(learned[key]), np.asarray(plain[key])
        assert np.all(np.abs(a - b) <= 1e-6 * np.abs(b)), (key, a, b)
