"""Uneven-size sample batches on the GPU (vckss_mi355x_batch_*_v1): several
independent subsamples of different sizes prepared together and estimated
concurrently, each on its own HIP stream.

Bars: every subsample's components within 1e-8 relative of the oracle's
estimate of that subsample alone (observation deletion, the workload's
mode, and match deletion); the concurrent run bit-identical to one sample at
a time (each sample's engine is deterministic whatever runs beside it); the
sample order follows the identifiers; a sample that fails preparation fails
the batch with its status.
"""
import functools

import numpy as np
import pytest

import fevc_amd as F
from fevc_amd import synth

pytestmark = pytest.mark.gpu

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")


@functools.lru_cache(maxsize=None)
def workload():
    cols, specs = synth.subsamples(count=6, median=2000, low=600, high=6000, sigma=0.8, seed=5)
    # identifiers out of order and not contiguous: samples follow the identifier
    ids = np.array([40, 7, 19, 3, 25, 11], np.uint64)
    cols["sample"] = ids[cols["sample"].astype(np.int64)]
    return cols, specs, ids


@pytest.mark.parametrize("mode", ["observation", "match"])
def test_batch_matches_oracle_per_sample(oracle, mode):
    cols, specs, ids = workload()
    probes = 32
    b = F.SampleBatch(cols["sample"], *[cols[k] for k in COLS], deletion_mode=mode, threads=4)
    try:
        print(f"batch: {b.samples} samples, {len(cols['worker'])} rows, mode {mode}", flush=True)
        assert b.samples == len(ids)
        together = b.solve(probes=probes, tolerance=1e-11, fit_tolerance=1e-10, precision=64, streams=3)
        alone = b.solve(probes=probes, tolerance=1e-11, fit_tolerance=1e-10, precision=64, streams=1)
    finally:
        b.release()
    for k, sid in enumerate(np.sort(ids)):
        got = together[k]
        assert got["sample"] == sid
        for key in ("plugin", "correction", "corrected", "mcse"):
            assert np.array_equal(got[key], alone[k][key]), (k, key)
        m = cols["sample"] == sid
        sub = [np.asarray(cols[c])[m] for c in COLS]
        want = oracle.Problem(*sub, observation=mode == "observation").run(
            probes=probes, leverage_width=32, target_width=32, fit_tolerance=1e-10, probe_tolerance=1e-11)
        for key in ("plugin", "correction", "corrected"):
            g, w = np.asarray(got[key]), np.asarray(want[key])
            assert np.all(np.abs(g - w) <= 1e-8 * np.maximum(np.abs(w), 1e-3)), (k, key, g, w)


def test_batch_sample_failure_reports_status():
    cols, specs, ids = workload()
    c = {k: np.asarray(v).copy() for k, v in cols.items()}
    # a sample of two rows cannot hold a leave-out connected set
    extra = {k: c[k][:2].copy() for k in c}
    extra["sample"][:] = 99
    c = {k: np.concatenate([c[k], extra[k]]) for k in c}
    with pytest.raises(F.EngineError) as err:
        F.SampleBatch(c["sample"], *[c[k] for k in COLS], deletion_mode="match")
    assert err.value.code != 0 and "sample 6" in str(err.value)
