"""Controls on the JLA route (the generic engine with Q > 0 controls under
match deletion, generic_jla.rs:1930-2140, 5313-5470, 6053-6273, 8554-8906) on
the GPU, through the C ABI, against the oracle (oracle/controls.py; its pins
are in tests/test_oracle_controls.py).

Bars: with tight solves (fit 1e-10, probes 1e-11) the components, the control
coefficients and every per-unit and per-probe array within 1e-8 of their
scale, for the joint nuisance (the match blocks' exact control part and the
joint solve's control coefficients in the target predictions) and the fixed
offset (the two-way route on the working outcome y - Z beta).  Block makers
of both forms are covered: year dummies (n x n makers, n = rows of the unit
<= 1 + Q, 5 and 10 rows) and two continuous controls (the reduced
(1 + Q) x (1 + Q) maker, units of 5 rows).  Refusals: a singular control
block (82), controls on the exact route, under observation deletion, with
the numerical attachment or with stayers (11).
"""
import functools

import numpy as np
import pytest

import fevc_amd as F
from fevc_amd import synth

pytestmark = pytest.mark.gpu

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")


@functools.lru_cache(maxsize=None)
def panel(kind="years", workers=3000, periods=5, seed=17):
    spec = synth.PanelSpec(workers=workers, firms=workers // 12, periods=periods, mobility=0.25, seed=seed,
                           name="gpu-controls")
    d = synth.generate(spec)
    z, y = synth.year_controls(spec, d, years=periods + 6 if kind == "years" else periods + 1, effect=0.3)
    cols = [np.asarray(d[k], float) for k in COLS]
    rng = np.random.default_rng(seed + 3)
    if kind == "continuous":
        z = rng.normal(size=(y.size, 2)) + 0.2 * z[:, :2]
        y = y + z @ np.array([0.4, -0.25])
    cols[3] = y
    cols[4] = rng.integers(1, 4, y.size).astype(float)
    cols[5] = cols[4] * rng.uniform(0.5, 1.5, y.size)
    return tuple(cols), z


def _close(got, want, name, tol=1e-8):
    got, want = np.asarray(got, float), np.asarray(want, float)
    scale = max(np.max(np.abs(want)), 1e-3)
    assert got.shape == want.shape, (name, got.shape, want.shape)
    err = np.max(np.abs(got - want))
    assert err <= tol * scale, (name, err, scale)


def _solve(cols, z, probes, nuisance, precision=64, widths=(32, 16)):
    e = F.Engine(*cols, controls=z)
    try:
        e.configure(precision, *widths)
        print(f"controls: q {z.shape[1]}, rows {e.dims['rows']}, units {e.dims['units']}, nuisance {nuisance}, "
              f"precision {precision}, widths {widths}", flush=True)
        got = e.solve(probes=probes, tolerance=1e-11, fit_tolerance=1e-10, nuisance=nuisance)
        got["arrays"] = e.arrays(probes)
        got["control_fit"] = e.control_fit()
        got["dims"] = dict(e.dims)
    finally:
        e.release()
    return got


@pytest.mark.parametrize("kind,periods,precision,widths", [
    ("years", 5, 64, (32, 16)),
    ("years", 10, 32, (64, 32)),
    ("continuous", 5, 64, (5, 3)),
])
def test_joint_controls_match_oracle(oracle, kind, periods, precision, widths):
    from oracle import controls as X
    cols, z = panel(kind, periods=periods)
    probes = 24
    want = X.joint_jla(list(cols), z, probes=probes, leverage_width=16, target_width=8, fit_tolerance=1e-10,
                       probe_tolerance=1e-11)
    got = _solve(cols, z, probes, "joint", precision, widths)
    cf = got["control_fit"]
    assert cf["controls"] == z.shape[1] and not cf["fixed_offset"]
    _close(cf["beta"], want["beta"], "beta")
    for key in ("plugin", "correction", "corrected"):
        _close(got[key], want[key], key)
    lay, fit = want["layout"], want["fit"]
    units = got["dims"]["units"]
    a = got["arrays"]
    _close(a["projection_share"], want["projection_share"], "projection_share")
    _close(a["finite_bias"], want["finite_bias"], "finite_bias")
    _close(a["residual_mass"], np.bincount(lay["unit"], lay["f"] * fit["e"], minlength=units), "residual_mass")
    B = np.bincount(lay["unit"], np.sqrt(lay["f"]) * want["values"], minlength=units)
    _close(a["deleted_mass"], B, "deleted_mass")
    _close(a["target_draws"], want["draws"], "target_draws")
    # the controls move the estimate: the plain two-way route differs
    plain = oracle.Problem(*cols).run(probes=probes, leverage_width=16, target_width=8, fit_tolerance=1e-10,
                                      probe_tolerance=1e-11)
    assert np.max(np.abs(np.asarray(plain["correction"]) - want["correction"])) > 1e-6


@pytest.mark.parametrize("precision", [64, 32])
def test_fixed_offset_controls_match_oracle(oracle, precision):
    from oracle import controls as X
    cols, z = panel("years")
    probes = 24
    fo = X.fixed_offset(list(cols), z, fit_tolerance=1e-10)
    work = list(cols)
    work[3] = fo["outcome"]
    want = oracle.Problem(*work).run(probes=probes, leverage_width=16, target_width=8, fit_tolerance=1e-10,
                                     probe_tolerance=1e-11)
    got = _solve(cols, z, probes, "fixed_offset", precision)
    cf = got["control_fit"]
    assert cf["fixed_offset"]
    _close(cf["beta"], fo["beta"], "beta")
    for key in ("plugin", "correction", "corrected"):
        _close(got[key], want[key], key)
    a = got["arrays"]
    for name in ("fitted_cell", "projection_share", "residual_mass", "deleted_mass", "cell_correction_weight"):
        _close(a[name], want[name], name)
    _close(a["target_draws"], want["target_draws"], "target_draws")


def test_control_refusals(oracle):
    cols, z = panel("years", workers=800)
    firm_one = (cols[1] == cols[1][0]).astype(float)
    for bad in (np.column_stack([z, z[:, 1]]), np.column_stack([z, firm_one])):
        e = F.Engine(*cols, controls=bad)
        try:
            with pytest.raises(F.EngineError) as err:
                e.solve(probes=4, tolerance=1e-10)
            assert err.value.code == 82, err.value
        finally:
            e.release()
    e = F.Engine(*cols, controls=z)
    try:
        with pytest.raises(F.EngineError) as err:
            e.solve_exact()
        assert err.value.code == 11
        with pytest.raises(F.EngineError) as err:
            e.solve_numerical(probes=4)
        assert err.value.code == 11
        # the generation still solves after the refusals
        r = e.solve(probes=4, tolerance=1e-10)
        assert np.all(np.isfinite(r["corrected"]))
    finally:
        e.release()
    e = F.Engine(*cols, controls=z, deletion_mode="observation")
    try:
        with pytest.raises(F.EngineError) as err:
            e.solve(probes=4, tolerance=1e-10)
        assert err.value.code == 11
    finally:
        e.release()
