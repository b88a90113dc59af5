"""Dense exact KSS estimator, restated in numpy (test infrastructure only).

Two independent restatements of the exact leave-out estimator:

* ``exact_estimator`` follows the Rust engine's exact route
  (rust/crates/vckss-core/src/exact_estimator.rs:490-1244, match deletion,
  joint nuisance, no controls, no centering): the permutation-equivariant
  W + F embedding whose firm block is constrained to zero sum
  (``build_design``, :1817-1848), the zero-sum quotient inverse with the
  unit null projector (dense.rs:89-160), the fit gates (:1871-1957), and the
  match-block maker with square-root frequency scaling (:1448-1696) against
  the target moments (:1698-1787).
* ``dense_kss`` follows the reference's own dense falsification oracle
  (fevc/tests/python/oracle.py:97-320): literal frequency expansion, the
  worker / all-but-one-firm / control quotient, joint or fixed-offset
  nuisance, and the general block leave-out identity.

Only tests/ import this module; it is the checker, never the product.

Pins (tests/test_oracle_exact.py): the registered dense fixture of
fevc/tests/python/test_dense_oracle.py:54-76 (plug-in and correction to
2e-13), the seven-edge leverages 11/15 and 3/5
(test_dense_oracle.py:186-200), the known effects of exact.rs:397-414, block
Woodbury residuals against direct refits (test_dense_oracle.py:79-90),
frequency weights against literal expansion (:93-119), and the two
restatements against each other on random panels.
"""
from __future__ import annotations

import numpy as np


class ExactError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


SINGULAR_INFORMATION = 82
INVERSE_RESIDUAL_FAILED = 83
NONESTIMABLE_DELETION = 75
FULL_RESIDUAL_FAILED = 64
BLOCK_INVERSE_FAILED = 76


def _codes(values):
    _, code = np.unique(np.asarray(values), return_inverse=True)
    return code.astype(np.int64)


def _bilinear(left, right, workers, worker_share, firm_share, cell_w, cell_f, cell_mass):
    """TargetMoments::bilinear (exact_estimator.rs:1736-1786)."""
    lw, rw = left[:workers], right[:workers]
    lf, rf = left[workers:], right[workers:]
    lwm, rwm = worker_share @ lw, worker_share @ rw
    lfm, rfm = firm_share @ lf, firm_share @ rf
    worker = np.sum(worker_share * lw * rw) - lwm * rwm
    firm = np.sum(firm_share * lf * rf) - lfm * rfm
    a = np.sum(cell_mass * lw[cell_w] * rf[cell_f])
    b = np.sum(cell_mass * rw[cell_w] * lf[cell_f])
    cov = 0.5 * (a + b - lwm * rfm - rwm * lfm)
    return np.array([worker, firm, cov, worker + firm + 2.0 * cov])


def _zero_sum_inverse(info, workers, firms, tolerance):
    """invert_scaled_zero_sum_quotient (dense.rs:89-160): certify on the
    diagonally scaled unshifted matrix, invert the matrix plus the unit
    projector 11'/F on the firm block."""
    p = info.shape[0]
    d = np.diag(info)
    if np.any(~np.isfinite(d)) or np.any(d <= 0):
        raise ExactError(SINGULAR_INFORMATION, "symmetric information has a nonpositive diagonal")
    q = 1.0 / np.sqrt(d)
    scaled = 0.5 * (info + info.T) * np.outer(q, q)
    ev = np.linalg.eigvalsh(scaled)
    err = 64.0 * np.finfo(float).eps * p * max(np.linalg.norm(scaled), 1.0)
    if not (ev[-1] > 0) or abs(ev[0]) > err or ev[1] - err <= tolerance * (ev[-1] + err):
        raise ExactError(SINGULAR_INFORMATION, "full-firm zero-sum quotient is singular")
    rcond = (ev[1] - err) / (ev[-1] + err)
    aug = info.copy()
    aug[workers:, workers:] += 1.0 / firms
    s = 1.0 / np.sqrt(np.diag(aug))
    sa = aug * np.outer(s, s)
    chol = np.linalg.cholesky(sa)
    inv_s = np.linalg.inv(chol)
    inv_s = inv_s.T @ inv_s
    inverse = inv_s * np.outer(s, s)
    inverse = 0.5 * (inverse + inverse.T)
    # compatible scaled residual (dense.rs:581-682): columns orthogonal to
    # the scaled null direction
    null = np.zeros(p)
    null[workers:] = 1.0 / q[workers:]
    null /= np.linalg.norm(null)
    proj = np.eye(p) - np.outer(null, null)
    inv_q = inverse / np.outer(q, q)
    res = (scaled @ inv_q - np.eye(p)) @ proj
    relres = np.max(np.linalg.norm(proj @ res, axis=0))
    if relres > max(100.0 * tolerance, 1e-10):
        raise ExactError(INVERSE_RESIDUAL_FAILED, "zero-sum quotient inverse failed its residual gate")
    return inverse, rcond, relres


def exact_estimator(worker, firm, deletion, outcome, frequency, target, *, rank_tolerance=1e-10,
                    block_tolerance=1e-10, solver_tolerance=1e-10, mode="match", stayer_rows=None):
    """run_exact_estimator (exact_estimator.rs:490-1244) for match deletion
    (or, with mode="observation", physical-observation deletion,
    exact_estimator.rs:914-987; with mode="hybrid", the stayer hybrid of
    stayer_hybrid.rs:71-348: match blocks for the rows outside the boolean
    `stayer_rows`, physical-observation deletion for the stayer rows), joint
    nuisance, no controls and no centering, on retained stored rows."""
    w = _codes(worker)
    f = _codes(firm)
    g = _codes(deletion)
    y = np.asarray(outcome, float)
    fr = np.asarray(frequency, float)
    tw = np.asarray(target, float)
    rows = y.size
    W, F = int(w.max()) + 1, int(f.max()) + 1
    p = W + F
    observed = np.ones(rows, bool) if mode == "observation" else (
        np.asarray(stayer_rows, bool) if mode == "hybrid" else np.zeros(rows, bool))
    if mode != "observation":
        for code in np.unique(g[~observed]):
            use = (g == code) & ~observed
            if np.unique(w[use]).size != 1 or np.unique(f[use]).size != 1:
                raise ExactError(20, "a deletion unit spans two cells")
    # build_design: worker dummy, firm coordinates e_f - 1/F
    x = np.zeros((rows, p))
    x[np.arange(rows), w] = 1.0
    x[:, W:] = -1.0 / F
    x[np.arange(rows), W + f] += 1.0
    info = x.T @ (fr[:, None] * x)
    inverse, rcond, relres = _zero_sum_inverse(info, W, F, rank_tolerance)
    rhs = x.T @ (fr * y)
    beta = inverse @ rhs
    beta[W:] -= beta[W:].mean()  # center_firm_coordinates
    # complete_fit_residual: plain worker / firm normal equations
    fit = beta[w] + beta[W + f]
    nrhs = np.zeros(p)
    nres = np.zeros(p)
    np.add.at(nrhs, w, fr * y)
    np.add.at(nrhs, W + f, fr * y)
    np.add.at(nres, w, fr * (y - fit))
    np.add.at(nres, W + f, fr * (y - fit))
    fit_rel = np.linalg.norm(nres) / np.linalg.norm(nrhs)
    if fit_rel > max(1e-11, 10.0 * solver_tolerance):
        raise ExactError(FULL_RESIDUAL_FAILED, "complete original fit residual exceeds tolerance")
    residual = y - x @ beta
    # plug-in (jla.rs:286-364)
    total = tw.sum()
    wm = np.sum(tw * beta[w]) / total
    fm = np.sum(tw * beta[W + f]) / total
    cw, cf = beta[w] - wm, beta[W + f] - fm
    plugin = np.array([np.sum(tw * cw * cw), np.sum(tw * cf * cf), np.sum(tw * cw * cf),
                       np.sum(tw * (cw + cf) ** 2)]) / total
    # target moments
    worker_share = np.bincount(w, tw, W) / total
    firm_share = np.bincount(f, tw, F) / total
    cell_key = w * F + f
    cells, cell_index = np.unique(cell_key, return_inverse=True)
    cell_w, cell_f = cells // F, cells % F
    cell_mass = np.bincount(cell_index, tw, cells.size) / total
    factor = np.linalg.cholesky(inverse)
    correction = np.zeros(4)
    max_leverage = 0.0
    rank_gap = np.inf
    h_unit = np.zeros(int(g.max()) + 1)
    correction_copies = np.zeros(4)
    if observed.any():
        design_inverse = x @ inverse
        for row in np.flatnonzero(observed):
            z = design_inverse[row]
            lev = float(x[row] @ z)
            if not np.isfinite(lev) or lev < -100.0 * rank_tolerance:
                raise ExactError(NONESTIMABLE_DELETION, "physical observation leverage is invalid")
            maker = 1.0 - lev
            if maker <= block_tolerance:
                raise ExactError(NONESTIMABLE_DELETION, "physical observation deletion has leverage at or above one")
            max_leverage = max(max_leverage, lev)
            rank_gap = min(rank_gap, maker)
            scale = fr[row] * y[row] * residual[row] / maker
            correction += scale * _bilinear(z, z, W, worker_share, firm_share, cell_w, cell_f, cell_mass)
    correction_copies[:3] = correction[:3]
    match_codes = np.unique(g[~observed])
    for code in match_codes:
        idx = np.flatnonzero((g == code) & ~observed)
        root = np.sqrt(fr[idx])
        bd = root[:, None] * x[idx]
        ty = root * y[idx]
        te = root * residual[idx]
        block_inverse = bd @ inverse
        low = bd @ factor
        maker = np.eye(idx.size) - low @ low.T
        ev = np.linalg.eigvalsh(0.5 * (maker + maker.T))
        if ev[0] <= block_tolerance:
            raise ExactError(NONESTIMABLE_DELETION, "a match residual block is singular")
        deleted = np.linalg.solve(maker, te)
        rel = np.linalg.norm(maker @ deleted - te) / max(np.linalg.norm(te), 1e-300)
        if rel > max(100.0 * rank_tolerance, 1e-10):
            raise ExactError(BLOCK_INVERSE_FAILED, "a match residual solve failed")
        max_leverage = max(max_leverage, 1.0 - ev[0])
        rank_gap = min(rank_gap, ev[0])
        h_unit[code] = (1.0 - ev[0]) / fr[idx].sum()
        left = block_inverse.T @ ty
        right = block_inverse.T @ deleted
        correction += _bilinear(left, right, W, worker_share, firm_share, cell_w, cell_f, cell_mass)
    correction[3] = correction[0] + correction[1] + 2.0 * correction[2]
    return dict(plugin=plugin, correction=correction, corrected=plugin - correction, beta=beta,
                residual=residual, max_leverage=max_leverage, deletion_rank_gap=rank_gap,
                deletion_units=int(match_codes.size + fr[observed].sum()),
                correction_stayers=correction_copies, correction_movers=correction - correction_copies,
                weighted_rss=float(np.sum(fr * residual ** 2)),
                information_rcond=rcond, inverse_relres=relres, fit_relres=fit_rel,
                unit_leverage=h_unit, workers=W, firms=F)


def expanded_design(y, worker, firm, controls=None, frequency=None, target=None, deletion="match",
                    deletion_id=None):
    """build_expanded_design (oracle.py:97-200)."""
    y = np.asarray(y, float)
    n = y.size
    wc, fc = _codes(worker), _codes(firm)
    z = np.empty((n, 0)) if controls is None else np.asarray(controls, float).reshape(n, -1)
    freq = np.ones(n, np.int64) if frequency is None else np.asarray(frequency).astype(np.int64)
    tgt = freq.astype(float) if target is None else np.asarray(target, float)
    if deletion == "observation":
        dstored = np.arange(n)
    elif deletion_id is None:
        _, dstored = np.unique(np.column_stack((wc, fc)), axis=0, return_inverse=True)
        dstored = dstored.ravel()
    else:
        dstored = _codes(deletion_id)
    idx = np.repeat(np.arange(n), freq)
    W, F = int(wc.max()) + 1, int(fc.max()) + 1
    d = np.eye(W)[wc[idx]]
    fq = np.eye(F)[fc[idx]][:, :F - 1]
    x = np.column_stack((d, fq, z[idx]))
    wrows = np.column_stack((d, np.zeros((idx.size, F - 1 + z.shape[1]))))
    frows = np.column_stack((np.zeros((idx.size, W)), fq, np.zeros((idx.size, z.shape[1]))))
    dexp = np.arange(idx.size) if deletion == "observation" else dstored[idx]
    return dict(y=y[idx], x=x, wrows=wrows, frows=frows, stored=idx, deletion=dexp, target=tgt, W=W)


def _targets(des):
    """target_matrices (oracle.py:203-219)."""
    n = des["target"].size
    rep = np.array([int(np.flatnonzero(des["stored"] == r)[0]) for r in range(n)])
    aw, af = des["wrows"][rep], des["frows"][rep]
    share = des["target"] / des["target"].sum()
    center = np.diag(share) - np.outer(share, share)
    qw = aw.T @ center @ aw
    qf = af.T @ center @ af
    qc = 0.5 * (aw.T @ center @ af + af.T @ center @ aw)
    return qw, qf, qc, qw + qf + 2.0 * qc


def dense_kss(y, worker, firm, *, controls=None, frequency=None, target=None, deletion="match",
              deletion_id=None, nuisance="joint", block_tolerance=1e-10):
    """exact_kss (oracle.py:222-320): dense block leave-out on literal copies."""
    des = expanded_design(y, worker, firm, controls, frequency, target, deletion, deletion_id)
    full_x = des["x"]
    full_inverse = np.linalg.inv(full_x.T @ full_x)
    full_beta = full_inverse @ (full_x.T @ des["y"])
    qs_full = _targets(des)
    if nuisance == "fixedoffset":
        fe = np.any(des["wrows"] != 0, axis=0) | np.any(des["frows"] != 0, axis=0)
        wy = des["y"] - full_x[:, ~fe] @ full_beta[~fe]
        x = full_x[:, fe]
        qs = [q[np.ix_(fe, fe)] for q in qs_full]
    else:
        wy, x, qs = des["y"], full_x, qs_full
    inverse = np.linalg.inv(x.T @ x)
    beta = inverse @ (x.T @ wy)
    residual = wy - x @ beta
    plugin = np.array([beta @ q @ beta for q in qs])
    correction = np.zeros(4)
    lev = 0.0
    for block in np.unique(des["deletion"]):
        use = des["deletion"] == block
        xg = x[use]
        pg = xg @ inverse @ xg.T
        lev = max(lev, float(np.linalg.eigvalsh(pg).max()))
        maker = np.eye(pg.shape[0]) - pg
        if np.linalg.eigvalsh(maker).min() <= block_tolerance:
            raise ExactError(NONESTIMABLE_DELETION, "deletion block is not leave-out estimable")
        deleted = np.linalg.solve(maker, residual[use])
        for t, q in enumerate(qs):
            correction[t] += wy[use] @ (xg @ inverse @ q @ inverse @ xg.T) @ deleted
    return dict(plugin=plugin, correction=correction, corrected=plugin - correction, beta=beta,
                residual=residual, leverage_max=lev, x=x, y=wy, inverse=inverse,
                deletion=des["deletion"], deletion_units=int(np.unique(des["deletion"]).size))


def refit_residuals(des_x, des_y, deletion):
    """deleted_residuals_by_refit (oracle.py:420-432): direct refits without
    each block."""
    out = {}
    for block in np.unique(deletion):
        use = deletion == block
        beta, _, rank, _ = np.linalg.lstsq(des_x[~use], des_y[~use], rcond=None)
        if rank != des_x.shape[1]:
            raise ExactError(NONESTIMABLE_DELETION, "direct deleted fit lost rank")
        out[int(block)] = des_y[use] - des_x[use] @ beta
    return out
