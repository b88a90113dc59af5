"""Seeded synthetic AKM panels for the leave-out KSS workloads.

The generator restates the data-generating process the workloads are quoted
on: worker and firm effects with sorting, assortative moves between firm
psi-deciles, heteroskedastic noise, optional Pareto firm sizes.  All draws come
from numpy's PCG64 with a recorded seed, so every panel is reproducible.

Conventions chosen where the description leaves freedom (stated here once):
  * psi ~ N(0, 0.5^2); firms are ranked by psi into 10 deciles.
  * alpha = rho * z(psi_d) + sqrt(1 - rho^2) * N(0,1), where z(psi_d) is the
    standardized psi-decile centre of the worker's home decile, rho = 0.2.
  * A worker with alpha-decile q draws a destination decile d with
    probability proportional to exp(0.5 * c(q, d)), c(q, d) = 1 - 2|q - d|/9
    (a rank similarity in [-1, 1]); a firm inside the decile is drawn with
    probability proportional to its size weight.
  * Each period after the first, a worker moves with probability `mobility`
    to a firm drawn the same way (a draw of the current firm is a stay).
  * y = alpha + psi + eps, eps ~ N(0, 0.3^2 (1 + u)), u ~ U(0, 1).
  * Pareto firm sizes: size weight = (1 + Pareto(shape)) truncated at `cap`.
  * The deletion unit is the worker-firm match (default), frequency 1,
    target weight 1.
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass(frozen=True)
class PanelSpec:
    workers: int
    firms: int
    periods: int
    mobility: float
    pareto_shape: float | None = None
    size_cap: float = 5000.0
    seed: int = 1
    name: str = "panel"


SMALL = PanelSpec(workers=2_000, firms=200, periods=5, mobility=0.2, seed=1,
                  name="akm-2k-workers-200-firms-T5")
SCALE = PanelSpec(workers=2_000_000, firms=100_000, periods=5, mobility=0.1,
                  pareto_shape=1.1, size_cap=5000.0, seed=1,
                  name="akm-2M-workers-100k-firms-T5")


def scaled(spec: PanelSpec, factor: float, name: str | None = None) -> PanelSpec:
    return dataclasses.replace(
        spec, workers=max(50, int(spec.workers * factor)),
        firms=max(10, int(spec.firms * factor)),
        name=name or f"{spec.name}-x{factor:g}")


def generate(spec: PanelSpec):
    """Return dict of columns: worker, firm, deletion (u64), outcome (f64),
    frequency (u64), target (f64) plus the true effects."""
    rng = np.random.default_rng(spec.seed)
    W, F, T = spec.workers, spec.firms, spec.periods
    psi = rng.normal(0.0, 0.5, F)
    order = np.argsort(psi, kind="stable")
    decile_of_firm = np.empty(F, np.int64)
    decile_of_firm[order] = (np.arange(F) * 10) // F
    if spec.pareto_shape is None:
        size = np.ones(F)
    else:
        size = np.minimum(1.0 + rng.pareto(spec.pareto_shape, F), spec.size_cap)
    # firms of each decile with cumulative size weights for inverse-CDF draws
    dec_firms = [np.flatnonzero(decile_of_firm == d) for d in range(10)]
    dec_cdf = []
    for d in range(10):
        wts = size[dec_firms[d]]
        c = np.cumsum(wts)
        dec_cdf.append(c / c[-1])
    q = np.arange(10)
    sim = 1.0 - 2.0 * np.abs(q[:, None] - q[None, :]) / 9.0
    trans = np.exp(0.5 * sim)
    trans /= trans.sum(axis=1, keepdims=True)
    trans_cdf = np.cumsum(trans, axis=1)
    # worker home decile and effect
    home = rng.integers(0, 10, W)
    centre = np.array([psi[dec_firms[d]].mean() for d in range(10)])
    z = (centre - centre.mean()) / centre.std()
    rho = 0.2
    alpha = rho * z[home] + np.sqrt(1 - rho * rho) * rng.normal(0.0, 1.0, W)
    arank = np.argsort(np.argsort(alpha, kind="stable"), kind="stable")
    wdec = (arank * 10) // W

    def draw_firm(dec_source):
        u = rng.random(len(dec_source))
        dest = np.minimum((trans_cdf[dec_source] < u[:, None]).sum(axis=1), 9)
        out = np.empty(len(dec_source), np.int64)
        v = rng.random(len(dec_source))
        for d in range(10):
            sel = np.flatnonzero(dest == d)
            if sel.size:
                idx = np.searchsorted(dec_cdf[d], v[sel], side="right")
                out[sel] = dec_firms[d][np.minimum(idx, len(dec_firms[d]) - 1)]
        return out

    firm = np.empty((T, W), np.int64)
    firm[0] = draw_firm(wdec)
    for t in range(1, T):
        move = rng.random(W) < spec.mobility
        firm[t] = firm[t - 1]
        idx = np.flatnonzero(move)
        if idx.size:
            firm[t, idx] = draw_firm(wdec[idx])
    worker = np.broadcast_to(np.arange(W)[None, :], (T, W))
    worker = worker.T.reshape(-1)
    firm = firm.T.reshape(-1)
    eps = rng.normal(0.0, 1.0, W * T) * 0.3 * np.sqrt(1.0 + rng.random(W * T))
    y = alpha[worker] + psi[firm] + eps
    w64 = worker.astype(np.uint64) + 1
    f64 = firm.astype(np.uint64) + 1
    return dict(
        worker=w64, firm=f64, deletion=w64 * np.uint64(F + 1) + f64,
        outcome=y.astype(np.float64), frequency=np.ones(W * T, np.uint64),
        target=np.ones(W * T, np.float64), alpha=alpha, psi=psi)


def year_controls(spec: PanelSpec, data, years: int = 11, effect: float = 0.1):
    """Year-dummy controls for a panel of `generate` (rows are worker-major,
    row = worker * T + period): worker w enters in calendar year s_w, uniform
    in 0 .. years - T, and is observed in years s_w .. s_w + T - 1.  Returns
    the dummies of years 1 .. years - 1 (year 0 the baseline; 10 controls with
    the default 11 years), rows x (years - 1), and the outcome plus a year
    effect of standard deviation `effect`."""
    W, T = spec.workers, spec.periods
    if years < T + 1:
        raise ValueError("need at least T + 1 calendar years")
    rng = np.random.default_rng(spec.seed + 0x5EED)
    start = rng.integers(0, years - T + 1, W)
    year = (start[:, None] + np.arange(T)[None, :]).reshape(-1)
    z = np.zeros((W * T, years - 1))
    has = year > 0
    z[np.flatnonzero(has), year[has] - 1] = 1.0
    gamma = np.concatenate([[0.0], rng.normal(0.0, effect, years - 1)])
    return z, np.asarray(data["outcome"], np.float64) + gamma[year]


def eligible_stayers(worker, firm, outcome, frequency, target, mover_mask):
    """Eligible stayer rows of a panel for the stayer hybrid, in the form the
    augmentation entry point takes (stayer_hybrid.rs:23-34): workers outside
    the retained mover sample whose rows all lie in one retained mover firm
    and who represent at least two physical observations.  Returns
    (firm, worker, outcome, frequency, target) with 1-based dense
    retained-mover firm indices (retained firm identifiers in ascending order,
    the prepared problem's firm order) and 1-based dense stayer-worker
    indices (ascending identifiers)."""
    worker, firm = np.asarray(worker), np.asarray(firm)
    mask = np.asarray(mover_mask, bool)
    retained_firms = np.unique(firm[mask])
    retained_workers = set(np.unique(worker[mask]).tolist())
    frequency = np.asarray(frequency, float)
    rows = []
    order = np.argsort(worker, kind="stable")
    bounds = np.flatnonzero(np.diff(worker[order])) + 1
    for idx in np.split(order, bounds):
        w = worker[idx[0]]
        if w in retained_workers or np.unique(firm[idx]).size != 1:
            continue
        if frequency[idx].sum() < 2 or not np.isin(firm[idx[0]], retained_firms):
            continue
        rows.append(np.sort(idx))
    if not rows:
        return tuple(np.zeros(0) for _ in range(5))
    rows_all = np.concatenate(rows)
    firm_index = np.searchsorted(retained_firms, firm[rows_all]) + 1.0
    worker_index = np.concatenate([np.full(r.size, k + 1.0) for k, r in enumerate(rows)])
    return (firm_index, worker_index, np.asarray(outcome, float)[rows_all], frequency[rows_all],
            np.asarray(target, float)[rows_all])


def subsamples(count: int = 64, median: int = 40_000, low: int = 5_000, high: int = 500_000, periods: int = 4,
               sigma: float = 1.0, seed: int = 11, scale: float = 1.0):
    """Many independent region-year subsamples of uneven size: worker counts
    log-normal with the given median and log-sd `sigma`, clipped to [low,
    high] (all three times `scale`), firms = workers / 20, T = `periods`,
    mobility U(0.05, 0.25) per subsample, the outcome DGP of `generate` with
    uniform firm sizes.  Returns (columns dict with a `sample` column, list
    of the subsamples' PanelSpecs); subsample k's rows carry sample id k."""
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.exp(np.log(median) + sigma * rng.standard_normal(count)), low, high) * scale
    mobility = rng.uniform(0.05, 0.25, count)
    specs, parts = [], []
    for k in range(count):
        w = max(50, int(sizes[k]))
        spec = PanelSpec(workers=w, firms=max(10, w // 20), periods=periods, mobility=float(mobility[k]),
                         seed=seed * 1000 + k, name=f"subsample-{k}")
        d = generate(spec)
        d = {c: np.asarray(d[c]) for c in ("worker", "firm", "deletion", "outcome", "frequency", "target")}
        d["sample"] = np.full(d["worker"].size, k, np.uint64)
        specs.append(spec)
        parts.append(d)
    cols = {c: np.concatenate([p[c] for p in parts]) for c in parts[0]}
    return cols, specs
