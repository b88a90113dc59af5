"""Failure handling of the solve path on the GPU, forced from the host.

  * Interrupts (ffi_engine.rs:1412-1503, engine.rs:1153, 1376): the caller's
    poll runs at every leverage and target batch; a poll that returns 1 at
    its n-th call stops the solve with status 1 after exactly n polls, and
    the generation is Failed afterwards (context.rs:150-205: a later solve
    returns 81).
  * Deterministic gate failures: a block tolerance no unit can meet makes the
    leverage adjustment fail on every unit; the code reported is that of the
    first unit in unit order (engine.rs:2269-2410 stops there), the same on
    every run and the oracle's code.
  * The PCG iteration limit (status 63) on the device and in the oracle.
  * Breakdown fallback: a constant smoother weight omega = 3 makes the
    V-cycle indefinite; PCG reports the loss of positive definiteness and the
    solve is redone with the reference cycle, whose result it returns.
No kernel faults or hangs: every condition ends in a normal status.
"""
import functools

import numpy as np
import pytest

import fevc_amd as F
from fevc_amd import synth

pytestmark = pytest.mark.gpu

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")
# 1500 firms: a hybrid Laplacian of about 1500 vertices, so the hierarchy has
# smoothed levels above its direct terminal and PCG needs several iterations
SPEC = synth.PanelSpec(workers=20000, firms=1500, periods=5, mobility=0.2, seed=11, name="gpu-robustness")


@functools.lru_cache(maxsize=None)
def panel():
    d = synth.generate(SPEC)
    return tuple(np.asarray(d[k]) for k in COLS)


def _solve_counting(poll_result, interval=1):
    """A fresh generation solved with a counting poll; returns (polls, error)."""
    calls = []
    e = F.Engine(*panel())
    try:
        e.configure(64, 8, 4)
        e.smoother(0)
        try:
            e.solve(probes=16, tolerance=1e-8, checkpoint_interval=interval,
                    poll=lambda _: calls.append(1) or poll_result(len(calls)))
        except F.EngineError as ex:
            with pytest.raises(F.EngineError) as again:
                e.solve(probes=16, tolerance=1e-8)
            return len(calls), ex.code, again.value.code
        return len(calls), 0, None
    finally:
        e.release()


def test_interrupt_at_nth_poll_then_failed_state():
    polls, code, _ = _solve_counting(lambda n: 0)
    # solve entry, setup, fit, 2 leverage batches, adjustment, 4 target
    # batches and finalize at least (refinement rounds add checkpoints)
    assert code == 0 and polls >= 11, polls
    for n in (1, 4, polls):
        seen, code, later = _solve_counting(lambda k, n=n: int(k == n))
        assert (seen, code, later) == (n, 1, 81), (n, seen, code, later)
    # with interval 2 every second checkpoint polls, the first one included
    seen, code, _ = _solve_counting(lambda n: 0, interval=2)
    assert code == 0 and seen == (polls + 1) // 2


def test_gate_failure_is_first_unit_and_matches_oracle(oracle):
    cols = panel()
    # block tolerance 0.999999: 1 - P or the moment total fails at every unit
    with pytest.raises(oracle.OracleError) as want:
        oracle.Problem(*cols).run(probes=8, probe_tolerance=1e-8, block_tolerance=0.999999)
    codes = []
    for _ in range(2):
        e = F.Engine(*cols)
        try:
            e.configure(64, 8, 4)
            with pytest.raises(F.EngineError) as ex:
                e.solve(probes=8, tolerance=1e-8, block_tolerance=0.999999)
            codes.append((ex.value.code, str(ex.value)))
            with pytest.raises(F.EngineError) as again:
                e.solve(probes=8, tolerance=1e-8)
            assert again.value.code == 81
        finally:
            e.release()
    print("gate failure:", codes, "oracle:", want.value.code)
    assert codes[0] == codes[1]
    assert "zero-based unit 0" in codes[0][1]
    assert codes[0][0] == want.value.code


def test_iteration_limit(oracle):
    cols = panel()
    with pytest.raises(oracle.OracleError) as want:
        oracle.Problem(*cols).run(probes=4, probe_tolerance=1e-10, fit_tolerance=1e-10, max_iterations=2)
    e = F.Engine(*cols)
    try:
        e.configure(64, 4, 2)
        e.smoother(0)
        with pytest.raises(F.EngineError) as ex:
            e.solve(probes=4, tolerance=1e-10, maximum_iterations=2)
    finally:
        e.release()
    assert ex.value.code == 63 == want.value.code


def test_breakdown_falls_back_to_reference_cycle():
    cols = panel()
    e = F.Engine(*cols)
    try:
        e.configure(64, 16, 8)
        e.smoother(0)
        plain = e.solve(probes=16, tolerance=1e-10)
        plain_arrays = e.arrays(16)
        e.smoother_force(3.0)
        status = e.smoother_status()
        assert status["active"] and not status["breakdown"]
        forced = e.solve(probes=16, tolerance=1e-10)
        forced_arrays = e.arrays(16)
        status = e.smoother_status()
    finally:
        e.release()
    print("breakdown fallback status:", status)
    assert status["breakdown"] and not status["active"]
    for key in ("plugin", "correction", "corrected"):
        assert forced[key] == plain[key], key
    assert np.array_equal(forced_arrays["projection_share"], plain_arrays["projection_share"])
