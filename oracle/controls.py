"""Fixed-offset nuisance with controls on the JLA route (test
infrastructure only: tests/ import it as the checker, never the product).

The reference's generic engine, with nuisance FIXED_OFFSET and Q > 0
controls (rust/crates/vckss-core/src/generic_jla.rs:1955-2060): the joint fit
of the worker and firm effects and the controls gives the control
coefficients; the working outcome is the outcome net of the fitted control
index, y - Z beta_Z; the leave-out estimator then runs on the two-way
effects alone with that outcome (fevc/tests/python/oracle.py:242-275 for the
dense form).  The joint fit is restated by Frisch-Waugh-Lovell: the Q + 1
columns (controls and outcome) are fitted on the two-way effects with the
oracle's CMG_FULL_V2 batch solve, their frequency-weighted residuals give
S = Zr' F Zr and b = Zr' F yr, and beta_Z = S^+ b with the pseudo-inverse
over eigenvalues above rank_tolerance times the largest (the fitted index is
the reference's for any basis of the control span).

Pins (tests/test_oracle_controls.py): the fitted control index against a
dense weighted least-squares fit of the full joint design; the exact
estimator on the working outcome against the dense oracle's fixed-offset
estimate; the JLA estimate within Monte Carlo error of it.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O


def fixed_offset(cols, controls, *, fit_tolerance=1e-10, rank_tolerance=1e-10, threads=1, observation=False):
    """Working outcome (input row order) and the control fit of a panel.

    cols: worker, firm, deletion, outcome, frequency, target (input rows);
    controls: rows x Q.  Returns dict(outcome, beta, index, rank, S, b,
    mask, fit_receipt)."""
    cols = [np.asarray(c) for c in cols]
    z = np.asarray(controls, np.float64).reshape(len(cols[0]), -1)
    q = z.shape[1]
    p = O.Problem(*cols, observation=observation)
    m = p.mask()
    _, wi = np.unique(cols[0][m], return_inverse=True)
    _, fi = np.unique(cols[1][m], return_inverse=True)
    if wi.max() + 1 != p.workers or fi.max() + 1 != p.firms:
        raise AssertionError("dense identifiers differ from the prepared problem's")
    f = cols[4][m].astype(np.float64)
    y = np.column_stack([z[m], cols[3][m].astype(np.float64)])
    wr = np.zeros((p.workers, q + 1))
    fr = np.zeros((p.firms, q + 1))
    np.add.at(wr, wi, f[:, None] * y)
    np.add.at(fr, fi, f[:, None] * y)
    ws, fs, rec, _ = p.solve_columns(wr, fr, tolerance=fit_tolerance, threads=threads)
    res = y - (ws[wi] + fs[fi])
    zr, yr = res[:, :q], res[:, q]
    S = zr.T @ (f[:, None] * zr)
    b = zr.T @ (f * yr)
    lam, vec = np.linalg.eigh(S)
    keep = lam > rank_tolerance * max(lam.max(), 0.0)
    beta = vec[:, keep] @ ((vec[:, keep].T @ b) / lam[keep])
    out = np.asarray(cols[3], np.float64).copy()
    index = z @ beta
    out[m] = out[m] - index[m]
    return dict(outcome=out, beta=beta, index=index, rank=int(keep.sum()), S=S, b=b, mask=m, fit_receipt=rec)


def dense_joint_index(cols, controls):
    """The fitted control index of the full joint design by dense weighted
    least squares (worker dummies, firm dummies but one, controls; literal
    frequency weights), on the retained rows of the oracle's selection."""
    cols = [np.asarray(c) for c in cols]
    z = np.asarray(controls, np.float64).reshape(len(cols[0]), -1)
    m = O.Problem(*cols).mask()
    _, wi = np.unique(cols[0][m], return_inverse=True)
    _, fi = np.unique(cols[1][m], return_inverse=True)
    W, F = wi.max() + 1, fi.max() + 1
    x = np.column_stack([np.eye(W)[wi], np.eye(F)[fi][:, 1:], z[m]])
    sw = np.sqrt(cols[4][m].astype(np.float64))
    coef, *_ = np.linalg.lstsq(x * sw[:, None], cols[3][m] * sw, rcond=None)
    return z[m] @ coef[W + F - 1:]
