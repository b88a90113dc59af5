"""Uneven-size sample batches on the GPU (vckss_mi355x_batch_*_v1/_v2):
several independent subsamples of different sizes prepared together and
estimated concurrently, each on its own HIP stream, or in lockstep (the PCG
calls of all samples run as one: each pass a single launch over every
sample, with per-sample block ranges, convergence gates and lane states).

Bars: every subsample's components within 1e-8 relative of the oracle's
estimate of that subsample alone (observation deletion, the workload's
mode, and match deletion); the concurrent and the lockstep runs
bit-identical to one sample at a time (each sample's passes are the ones it
runs alone, with its own grid and reduction order), with the FP64 and the
FP32 V-cycle, with every sample in one lockstep group and with fewer
lockstep threads than samples; the sample order follows the identifiers; a
sample that fails preparation fails the batch with its status.
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


@pytest.mark.parametrize("mode,precision,streams", [("observation", 64, 6), ("observation", 32, 6),
                                                    ("match", 32, 6), ("observation", 32, 4)])
def test_lockstep_matches_one_sample_at_a_time(mode, precision, streams):
    cols, specs, ids = workload()
    probes = 32
    b = F.SampleBatch(cols["sample"], *[cols[k] for k in COLS], deletion_mode=mode, threads=4)
    try:
        print(f"lockstep batch: {b.samples} samples, mode {mode}, precision {precision}, {streams} threads",
              flush=True)
        alone = b.solve(probes=probes, tolerance=1e-10, fit_tolerance=1e-10, precision=precision, streams=1)
        print("lockstep batch: one at a time done", flush=True)
        together = b.solve(probes=probes, tolerance=1e-10, fit_tolerance=1e-10, precision=precision,
                           streams=streams, lockstep=True)
        print("lockstep batch: lockstep done", flush=True)
        again = b.solve(probes=probes, tolerance=1e-10, fit_tolerance=1e-10, precision=precision,
                        streams=streams, lockstep=True)
    finally:
        b.release()
    for k, sid in enumerate(np.sort(ids)):
        assert together[k]["sample"] == sid
        for key in ("plugin", "correction", "corrected", "mcse"):
            assert np.array_equal(together[k][key], alone[k][key]), (k, key, together[k][key], alone[k][key])
            assert np.array_equal(again[k][key], alone[k][key]), (k, key)
        for key in ("iterations", "max_complete_residual", "units"):
            assert together[k][key] == alone[k][key], (k, key)


def test_lockstep_mode_argument():
    cols, specs, ids = workload()
    b = F.SampleBatch(cols["sample"], *[cols[k] for k in COLS], deletion_mode="match", threads=4)
    try:
        code = F.library().vckss_mi355x_batch_solve_v2(b.handle, 1, 8, 1e-8, 1e-8, 32, 4, 2)
        assert code == 20
    finally:
        b.release()
