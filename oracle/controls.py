"""Controls on the JLA route: the fixed-offset and joint nuisance (test
infrastructure only: tests/ import it as the checker, never the product).

The reference's generic engine with Q > 0 controls (rust/crates/vckss-core/
src/generic_jla.rs:1930-2140): the joint fit of the worker and firm effects
and the controls is restated by Frisch-Waugh-Lovell.  The Q + 1 columns
(controls and outcome) are fitted on the two-way effects with the oracle's
CMG_FULL_V2 batch solve; their frequency-weighted residuals give the control
Schur complement S = Zr' F Zr and b = Zr' F yr, and beta = S^-1 b.  A
singular or numerically unresolved control block is refused with
SINGULAR_INFORMATION (82), as model_solver.rs:1799-1870 and
dense.rs:57-87 refuse it: the raw control Gram Z' F Z must be nonsingular
after diagonal scaling, and the generalised eigenvalues of S against it
must keep a reciprocal condition above max(rank_tolerance, 1e-12).

- Fixed offset (generic_jla.rs:2012-2045): the working outcome is
  y - Z beta; the leave-out estimator then runs on the two-way effects alone
  (the compressed engine on the working outcome).
- Joint (generic_jla.rs:5313-5470, 6053-6273, 8554-8906,
  statistical_batches.rs:259-346; fevc/docs/BLOCK_CONTROL_DERIVATION.md):
  the leverage probes are the two-way ones (the unit finite values do not
  depend on the outcome or the controls); each match block's residual maker
  is I - C_g - h v v' with the exact control part C_g = U_g S^-1 U_g'
  (U = frequency-transformed residualised controls); the target probes'
  row predictions carry the control coefficients of the joint solve,
  p_r = (D u)_r + zr_r' beta_u with beta_u = -S^-1 Z' F D u.

Pins (tests/test_oracle_controls.py): the fitted control index against a
dense weighted least-squares fit of the full joint design; the exact
estimator on the fixed-offset working outcome against the dense oracle's
fixed-offset estimate; the joint route with exact block leverages and every
target direction of the probe law (no sampling) against the dense oracle's
joint estimate (fevc/tests/python/oracle.py:222-320); the JLA estimates
within Monte Carlo error of the dense ones.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O

SINGULAR_INFORMATION = 82
NONESTIMABLE_DELETION = 75
BLOCK_INVERSE_FAILED = 76


class ControlError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


def _scaled_spectrum(m):
    d = np.sqrt(np.diag(m))
    if not np.all(np.isfinite(d)) or np.any(d <= 0.0):
        return np.array([0.0, 1.0])
    return np.linalg.eigvalsh(m / np.outer(d, d))


def layout(p, cols, controls):
    """Stored-row layout of a prepared problem: input row, cell, dense worker
    and firm, frequency, outcome and control rows, and the unit of each
    stored row."""
    z = np.asarray(controls, np.float64).reshape(len(cols[0]), -1)
    ret = p.array("retained").astype(np.int64)
    rc = p.array("row_cell").astype(np.int64)
    wi = p.array("cell_worker").astype(np.int64)[rc]
    fi = p.array("cell_firm").astype(np.int64)[rc]
    ptr = p.array("unit_ptr").astype(np.int64)
    rows = p.array("unit_rows").astype(np.int64)
    unit = np.empty(p.rows, np.int64)
    for g in range(p.units):
        unit[rows[ptr[g]:ptr[g + 1]]] = g
    return dict(ret=ret, cell=rc, wi=wi, fi=fi, f=np.asarray(cols[4], np.float64)[ret],
                y=np.asarray(cols[3], np.float64)[ret], z=z[ret], unit=unit, ptr=ptr, rows=rows,
                target=np.asarray(cols[5], np.float64)[ret])


def control_fit(p, lay, *, fit_tolerance=1e-10, rank_tolerance=1e-10, threads=1):
    """The joint fit by FWL.  Returns beta, the Schur complement S and its
    inverse, the residualised controls zr and outcome yr, the joint residual
    e = yr - zr beta and the joint two-way coefficients (alpha, psi)."""
    z, y, f, wi, fi = lay["z"], lay["y"], lay["f"], lay["wi"], lay["fi"]
    q = z.shape[1]
    x = np.column_stack([z, y])
    wr = np.zeros((p.workers, q + 1))
    fr = np.zeros((p.firms, q + 1))
    np.add.at(wr, wi, f[:, None] * x)
    np.add.at(fr, fi, f[:, None] * x)
    ws, fs, rec, _ = p.solve_columns(wr, fr, tolerance=fit_tolerance, inner_ratio=0.01, threads=threads)
    res = x - (ws[wi] + fs[fi])
    zr, yr = res[:, :q], res[:, q]
    g0 = z.T @ (f[:, None] * z)
    lam0 = _scaled_spectrum(g0)
    if not lam0[-1] > 0.0 or lam0[0] <= 1e-12 * lam0[-1]:
        raise ControlError(SINGULAR_INFORMATION, "the control Gram is singular")
    S = zr.T @ (f[:, None] * zr)
    S = 0.5 * (S + S.T)
    whitener = np.linalg.cholesky(np.linalg.inv(g0))
    gen = np.linalg.eigvalsh(whitener.T @ S @ whitener)
    if not gen[-1] > 0.0 or gen[0] / gen[-1] <= max(rank_tolerance, 1e-12):
        raise ControlError(SINGULAR_INFORMATION, "the residualised control information is singular")
    lam = _scaled_spectrum(S)
    if not lam[-1] > 0.0 or lam[0] <= rank_tolerance * lam[-1]:
        raise ControlError(SINGULAR_INFORMATION, "the control Schur complement is singular")
    b = zr.T @ (f * yr)
    s_inv = np.linalg.inv(S)
    s_inv = 0.5 * (s_inv + s_inv.T)
    beta = np.linalg.solve(S, b)
    return dict(beta=beta, S=S, b=b, s_inv=s_inv, zr=zr, yr=yr, e=yr - zr @ beta,
                alpha=ws[:, q] - ws[:, :q] @ beta, psi=fs[:, q] - fs[:, :q] @ beta,
                fit_receipt=rec, generalized_rcond=float(gen[0] / gen[-1]))


def fixed_offset(cols, controls, *, fit_tolerance=1e-10, rank_tolerance=1e-10, threads=1, observation=False):
    """Working outcome (input row order) and the control fit of a panel.

    cols: worker, firm, deletion, outcome, frequency, target (input rows);
    controls: rows x Q.  Returns dict(outcome, beta, index, S, b, mask,
    fit_receipt)."""
    cols = [np.asarray(c) for c in cols]
    z = np.asarray(controls, np.float64).reshape(len(cols[0]), -1)
    p = O.Problem(*cols, observation=observation)
    lay = layout(p, cols, z)
    fit = control_fit(p, lay, fit_tolerance=fit_tolerance, rank_tolerance=rank_tolerance, threads=threads)
    out = np.asarray(cols[3], np.float64).copy()
    index = z @ fit["beta"]
    m = p.mask()
    out[m] = out[m] - index[m]
    return dict(outcome=out, beta=fit["beta"], index=index, S=fit["S"], b=fit["b"], mask=m,
                fit_receipt=fit["fit_receipt"])


def plugin(lay, alpha, psi):
    """jla.rs:286-364: target-weighted variances of the fitted effects."""
    t = lay["target"]
    a, s = alpha[lay["wi"]], psi[lay["fi"]]
    total = t.sum()
    a = a - (t @ a) / total
    s = s - (t @ s) / total
    out = np.array([t @ (a * a), t @ (s * s), t @ (a * s), t @ ((a + s) * (a + s))]) / total
    return out


def match_blocks(lay, fit, projection, bias, variance, *, block_tolerance=1e-10):
    """generic_jla.rs:6053-6273 with the controls: per match unit the block
    residual maker M = I - L L', L = [sqrt(h) v, sqrt(f) zr Phi] (Phi Phi' =
    S^-1), and the adjusted deleted residual M^-1 sqrt(f) e + B M^-1 v
    (v' M^-1 sqrt(f) e) - V M^-1 v (v' M^-1 v)(v' M^-1 sqrt(f) e), per stored
    row (frequency-transformed)."""
    phi = np.linalg.cholesky(fit["s_inv"])
    values = np.zeros(len(lay["y"]))
    ptr, rows = lay["ptr"], lay["rows"]
    for g in range(len(ptr) - 1):
        r = rows[ptr[g]:ptr[g + 1]]
        root = np.sqrt(lay["f"][r])
        common = root / np.sqrt(lay["f"][r].sum())
        low = np.column_stack([np.sqrt(projection[g]) * common, root[:, None] * (fit["zr"][r] @ phi)])
        maker = np.eye(len(r)) - low @ low.T
        if np.linalg.eigvalsh(maker).min() <= block_tolerance:
            raise ControlError(NONESTIMABLE_DELETION, "a match residual block is singular")
        sol = np.linalg.solve(maker, np.column_stack([root * fit["e"][r], common]))
        transformed, inverse_common = sol[:, 0], sol[:, 1]
        ct, ci = common @ transformed, common @ inverse_common
        values[r] = transformed + bias[g] * inverse_common * ct - variance[g] * inverse_common * ci * ct
    return values


def joint_draw(lay, fit, values, sides):
    """statistical_batches.rs:259-346 for match deletion: one target draw
    from the two FE solutions of a direction, sides = ((u_w, u_f) of the
    worker-side solve, (u_w, u_f) of the firm-side solve).  The joint
    solve adds the control coefficients beta_u = -S^-1 Z' F D u to every row
    prediction (D u)_r + z_r' beta_u - (D V beta_u)_r = (D u)_r + zr_r' beta_u."""
    f, y, z, zr = lay["f"], lay["y"], lay["z"], fit["zr"]
    pred = []
    for uw, uf in sides:
        base = uw[lay["wi"]] + uf[lay["fi"]]
        beta = -fit["s_inv"] @ (z.T @ (f * base))
        pred.append(base + zr @ beta)
    pred.append(pred[0] + pred[1])
    units = len(lay["ptr"]) - 1
    out = []
    for pr in pred:
        first = np.bincount(lay["unit"], weights=f * y * pr, minlength=units)
        second = np.bincount(lay["unit"], weights=np.sqrt(f) * values * pr, minlength=units)
        out.append(float(first @ second))
    w, fm, t = out
    return np.array([w, fm, 0.5 * (t - w - fm), t])


def joint_jla(cols, controls, *, seed=8675309, probes=200, leverage_width=8, target_width=8, threads=1,
              fit_tolerance=1e-10, probe_tolerance=1e-10, rank_tolerance=1e-10, block_tolerance=1e-10):
    """The joint-nuisance JLA estimate with controls under match deletion:
    the compressed engine's leverage finite values and target solutions
    (same Counter-V1 draws), the joint fit and the control block algebra."""
    cols = [np.asarray(c) for c in cols]
    p = O.Problem(*cols)
    lay = layout(p, cols, controls)
    fit = control_fit(p, lay, fit_tolerance=fit_tolerance, rank_tolerance=rank_tolerance, threads=threads)
    run = p.run(seed=seed, probes=probes, leverage_width=leverage_width, target_width=target_width,
                threads=threads, fit_tolerance=fit_tolerance, probe_tolerance=probe_tolerance,
                rank_tolerance=rank_tolerance, block_tolerance=block_tolerance, targets=True)
    values = match_blocks(lay, fit, run["projection_share"], run["finite_bias"], run["finite_variance"],
                          block_tolerance=block_tolerance)
    tw, tf = run["target_worker"], run["target_firm"]
    draws = np.array([joint_draw(lay, fit, values, ((tw[2 * t], tf[2 * t]), (tw[2 * t + 1], tf[2 * t + 1])))
                      for t in range(probes)])
    correction = draws.mean(axis=0)
    mcse = draws.std(axis=0, ddof=1) / np.sqrt(probes)
    pl = plugin(lay, fit["alpha"], fit["psi"])
    return dict(plugin=pl, correction=correction, corrected=pl - correction, mcse=mcse, draws=draws,
                beta=fit["beta"], values=values, projection_share=run["projection_share"],
                finite_bias=run["finite_bias"], finite_variance=run["finite_variance"], fit=fit, layout=lay)


def joint_exact_directions(cols, controls, *, fit_tolerance=1e-12, rank_tolerance=1e-10):
    """The joint route with exact block leverages h_g = F_g x_g' (A'FA)^-1 x_g,
    no finite-probe adjustment, and the expectation over the target probe
    law in place of sampled probes: every stratum's direction
    sqrt(mass_s / T) (e_cell(s) - t / T) with mass_s its physical count times
    its per-copy mass, solved by a dense two-way fit.  Equals the exact
    leave-out correction when the route's algebra is right."""
    cols = [np.asarray(c) for c in cols]
    p = O.Problem(*cols)
    lay = layout(p, cols, controls)
    fit = control_fit(p, lay, fit_tolerance=fit_tolerance, rank_tolerance=rank_tolerance)
    W, F = p.workers, p.firms
    design = np.zeros((len(lay["y"]), W + F - 1))
    design[np.arange(len(lay["y"])), lay["wi"]] = 1.0
    has = lay["fi"] > 0
    design[np.flatnonzero(has), W + lay["fi"][has] - 1] = 1.0
    gram_inv = np.linalg.inv(design.T @ (lay["f"][:, None] * design))
    ptr, rows = lay["ptr"], lay["rows"]
    h = np.array([lay["f"][rows[ptr[g]:ptr[g + 1]]].sum()
                  * design[rows[ptr[g]]] @ gram_inv @ design[rows[ptr[g]]] for g in range(p.units)])
    zero = np.zeros(p.units)
    values = match_blocks(lay, fit, h, zero, zero)
    cw, cf = p.array("cell_worker").astype(np.int64), p.array("cell_firm").astype(np.int64)
    s_cell = p.array("stratum_cell").astype(np.int64)
    mass = p.array("stratum_count") * p.array("stratum_per_copy")
    centre = p.array("cell_target_sum") / p.target_total

    def solve(wrhs, frhs):
        u = gram_inv @ np.concatenate([wrhs, frhs[1:]])
        return u[:W], np.concatenate([[0.0], u[W:]])

    total = np.zeros(4)
    for s in range(len(s_cell)):
        d = -np.sqrt(mass[s] / p.target_total) * centre
        d[s_cell[s]] += np.sqrt(mass[s] / p.target_total)
        wrhs = np.bincount(cw, weights=d, minlength=W)
        frhs = np.bincount(cf, weights=d, minlength=F)
        worker_side = solve(wrhs, np.zeros(F))
        firm_side = solve(np.zeros(W), frhs)
        total += joint_draw(lay, fit, values, (worker_side, firm_side))
    pl = plugin(lay, fit["alpha"], fit["psi"])
    return dict(plugin=pl, correction=total, corrected=pl - total, beta=fit["beta"], leverage=h)


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
