"""The headline workload at full size on the GPU against the oracle.

Panel: 2,000,000 workers x T=5 (10M rows), 100,000 firms with Pareto(1.1)
sizes truncated at 5,000, mobility 0.1, seed 1, match deletion; the
leave-out connected set keeps about 680k workers, 89k firms and 1.47M cells.
The device runs the production configuration (learned and tuned
preconditioner, batched lockstep PCG, FP32 V-cycle with FP64 Krylov) on
8 leverage and 8 target probes at rel.res 1e-8; the oracle runs the plain
FP64 restatement of the reference engine on identical Counter-V1 draws.

Bars: the four variance components (plug-in, correction, corrected) within
1e-6 relative; every unit's leverage shares within 1e-6 absolute (shares lie
in [0, 1]; each side solves to rel.res 1e-8 and the complete-residual gate
1e-7); every leverage solve's complete residual within the gate; the
estimate's refinement counters reported.
"""
import numpy as np
import pytest

import fevc_amd as F
from fevc_amd import synth

pytestmark = pytest.mark.gpu

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")


def test_headline_panel_matches_oracle(oracle):
    print("headline: generate", synth.SCALE.name, flush=True)
    d = synth.generate(synth.SCALE)
    cols = [np.asarray(d[k]) for k in COLS]
    del d
    print("headline: device prepare", flush=True)
    e = F.Engine(*cols)
    try:
        dims = e.dims
        print("headline: dims", dims, flush=True)
        e.configure(32, 64, 32)
        before = e.counters()
        got = e.solve(probes=8, tolerance=1e-8, fit_tolerance=1e-8)
        after = e.counters()
        arrays = e.arrays(8)
        status = e.smoother_status()
    finally:
        e.release()
    print("headline: device", got["corrected"], "smoother", status["active"], status["shape"],
          "refinement rounds", after["refine_rounds"] - before["refine_rounds"], flush=True)
    print("headline: oracle prepare and run", flush=True)
    p = oracle.Problem(*cols)
    assert (p.workers, p.firms, p.cells, p.units) == (dims["workers"], dims["firms"], dims["cells"], dims["units"])
    want = p.run(probes=8, leverage_width=8, target_width=4, threads=16, fit_tolerance=1e-8,
                 probe_tolerance=1e-8)
    print("headline: oracle", want["corrected"], flush=True)
    for key in ("plugin", "correction", "corrected"):
        g, w = np.asarray(got[key]), np.asarray(want[key])
        err = np.abs(g - w) / np.abs(w)
        print(f"headline: {key} max rel error {err.max():.2e}")
        assert np.all(err <= 1e-6), (key, g, w)
    for name in ("projection_share", "residual_share"):
        diff = np.max(np.abs(arrays[name] - want[name]))
        print(f"headline: {name} max abs difference {diff:.2e}")
        assert diff <= 1e-6, name
    assert np.max(arrays["leverage_complete_residual"]) <= 1e-7
