"""Observation deletion on the JLA route (the generic engine without controls,
generic_jla.rs:5470-6017 and its target phase :8554-8906) on the GPU, through
the C ABI, against the oracle (its pins are in tests/test_oracle_observation.py).

Bars: Counter-V1 atoms bit-exact (integer work; one per stored row, the sum
of the signs of its copies [offset, offset + f) of its class); components
within 1e-8 relative with tight solves (1e-11), per-row arrays within 1e-8
of their scale, as for the match route.
"""
import functools

import numpy as np
import pytest

import fevc_amd as F
from fevc_amd import synth

pytestmark = pytest.mark.gpu

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")
SPEC = synth.PanelSpec(workers=4000, firms=300, periods=5, mobility=0.2, seed=13, name="gpu-observation")


@functools.lru_cache(maxsize=None)
def panel(heavy=False):
    d = synth.generate(SPEC)
    cols = [np.asarray(d[k], float) for k in COLS]
    rng = np.random.default_rng(21)
    cols[4] = rng.integers(1, 300 if heavy else 4, cols[0].size).astype(float)
    cols[5] = cols[4] * rng.uniform(0.5, 1.5, cols[0].size)
    # duplicated rows share an observation class and continue its copies
    dup = rng.choice(cols[0].size, 200, replace=False)
    cols = [np.concatenate((c, c[dup])) for c in cols]
    cols[2][-200:] = cols[2].max() + 1 + np.arange(200)
    return tuple(cols)


@pytest.mark.parametrize("heavy", [False, True])
def test_observation_atoms_bit_exact(oracle, heavy):
    cols = panel(heavy)
    p = oracle.Problem(*cols, observation=True)
    e = F.Engine(*cols, deletion_mode="observation")
    try:
        offsets = e.array(26, e.dims["units"])
        print(f"rows {e.dims['units']}, largest copy offset {offsets.max():.0f}", flush=True)
        if heavy:
            assert offsets.max() > 64  # classes spanning several Counter-V1 words
        for seed, first, columns in [(8675309, 0, 1), (8675309, 3, 7), (42, 64, 33), (1, 5, 64)]:
            got = e.atoms(seed, 0, first, columns)
            want = p.observation_atoms(seed, first, columns)
            assert np.array_equal(got, want), (seed, first, columns)
    finally:
        e.release()


def _compare(oracle, cols, probes, precision=64, widths=(32, 16)):
    p = oracle.Problem(*cols, observation=True)
    want = p.run(probes=probes, fit_tolerance=1e-10, probe_tolerance=1e-11)
    e = F.Engine(*cols, deletion_mode="observation")
    try:
        e.configure(precision, *widths)
        print(f"observation: rows {e.dims['rows']}, physical {e.dims['physical']}, precision {precision}, "
              f"widths {widths}", flush=True)
        got = e.solve(probes=probes, tolerance=1e-11)
        arrays = e.arrays(probes)
    finally:
        e.release()
    for key in ("plugin", "correction", "corrected"):
        g, w = np.asarray(got[key]), np.asarray(want[key])
        assert np.all(np.abs(g - w) <= 1e-8 * np.maximum(np.abs(w), 1e-3)), (key, g, w)
    for name in ("fitted_cell", "projection_share", "residual_share", "finite_bias", "finite_variance",
                 "residual_mass", "deleted_mass", "cell_correction_weight"):
        scale = max(np.max(np.abs(want[name])), 1.0)
        assert np.max(np.abs(arrays[name] - want[name])) <= 1e-8 * scale, name
    return got


def test_observation_engine_matches_oracle(oracle):
    _compare(oracle, list(panel()), probes=24)


def test_observation_heavy_frequencies(oracle):
    _compare(oracle, list(panel(True)), probes=8)


def test_observation_fp32_vcycle_and_widths(oracle):
    a = _compare(oracle, list(panel()), probes=12, precision=32, widths=(64, 32))
    b = _compare(oracle, list(panel()), probes=12, widths=(5, 3))  # batches straddle probe blocks
    for key in ("plugin", "correction", "corrected"):
        assert np.allclose(a[key], b[key], rtol=1e-8, atol=0)


def test_observation_refusals():
    e = F.Engine(*panel(), deletion_mode="observation")
    try:
        e.deletion_mode = 1  # solve mode differs from the prepared graph's (ffi_engine.rs:5705-5710)
        with pytest.raises(F.EngineError) as ei:
            e.solve(probes=8, tolerance=1e-8)
        assert ei.value.code == 20
    finally:
        e.release()
