"""End-to-end known-answer test of the oracle engine: the reference's
shared-atoms KAT (rust/crates/vckss-core/tests/rust_mata_shared_atoms.rs:39-294)."""
import numpy as np
import pytest

from conftest import load_golden

K = load_golden("shared_atoms_kat.json")


def kat_columns():
    inp = K["input"]
    order = inp["order"]
    return [np.array([inp[k][r] for r in order]) for k in
            ("worker", "firm", "deletion", "outcome", "frequency", "target")]


@pytest.fixture(scope="module")
def kat(oracle):
    p = oracle.Problem(*kat_columns())
    o = K["options"]
    res = p.run(seed=o["seed"], probes=o["probes"], leverage_width=o["leverage_width"],
                target_width=o["target_width"], fit_tolerance=o["pcg_tolerance"],
                probe_tolerance=o["pcg_tolerance"], max_iterations=o["maximum_iterations"],
                recompute_interval=o["residual_replacement_interval"])
    return p, res


def test_selection_and_dimensions(kat):
    p, _ = kat
    d = K["dimensions"]
    order = K["input"]["order"]
    assert list(p.mask()) == [r < 8 for r in order]
    assert (p.input_rows, p.rows, p.physical, p.workers, p.firms, p.cells, p.units,
            p.exact_strata, p.strata) == (
        d["input_rows"], d["retained_rows"], d["retained_physical_mass"], d["workers"],
        d["firms"], d["cells"], d["deletion_units"], d["exact_target_strata"],
        d["plan_target_strata"])
    assert p.target_total == d["target_total"]


@pytest.mark.parametrize("name,key", [
    ("fitted_cell", "fitted_cell"), ("projection_share", "unit_projection_share"),
    ("residual_share", "unit_residual_share"), ("finite_bias", "unit_finite_bias"),
    ("finite_variance", "unit_finite_variance"), ("residual_mass", "unit_residual_mass"),
    ("deleted_mass", "unit_deleted_mass"), ("cell_correction_weight", "cell_correction_weight"),
    ("target_draws", "target_draws")])
def test_arrays(kat, name, key):
    _, res = kat
    assert np.max(np.abs(np.asarray(res[name]) - np.asarray(K[key]))) < K["tolerance"]


@pytest.mark.parametrize("key", ["plugin", "correction", "corrected"])
def test_components(kat, key):
    _, res = kat
    assert np.max(np.abs(np.asarray(res[key]) - np.asarray(K[key]))) < K["tolerance"]


def test_mcse_and_residuals(kat):
    _, res = kat
    assert np.max(np.abs(np.asarray(res["mcse"]) - np.asarray(K["numerical_mcse"]))) < K["tolerance"]
    assert res["max_complete_residual"] <= 1e-11


def test_width_invariance(oracle):
    # engine.rs:4188: batch width does not change any result
    p = oracle.Problem(*kat_columns())
    a = p.run(probes=6, leverage_width=1, target_width=1, fit_tolerance=1e-13, probe_tolerance=1e-13)
    b = p.run(probes=6, leverage_width=4, target_width=3, fit_tolerance=1e-13, probe_tolerance=1e-13)
    assert np.array_equal(a["target_draws"], b["target_draws"])
    assert np.array_equal(a["projection_share"], b["projection_share"])
