"""The reference's refinement fixture (full_cmg.rs:2008-2098,
refinement_certifies_columns_with_small_rhs_relative_to_solution).

A side x side grid of firms; one worker per grid edge (one copy at each end)
and, per grid firm, `hub_workers` workers with one copy there and
`hub_copies` copies at a hub firm.  The hub inflates the operator bound
2 max degree, so CMG's operator-scaled PCG stop
||r|| <= rel (||b|| + 2 maxdeg ||x||) accepts a smooth, large solution whose
residual relative to ||b|| is above the complete-residual gate: the batch
solve must refine.  With y = F psi the exact solution is alpha = 0,
psi_f = psi(f) (psi = 0 at the hub); after the firm-mean normalisation
(cmg_impl.rs.in:550-570) that is alpha = m, psi - m with m the plain mean of
psi over all firms.
"""
import numpy as np


def grid_and_hub(side=40, hub_copies=100_000, hub_workers=20):
    """Columns (worker, firm, deletion, outcome, frequency, target) as
    float64 arrays, and the number of grid firms."""
    firms = side * side
    hub = firms + 1
    worker, firm, freq = [], [], []
    nw = 0
    for r in range(side):
        for c in range(side):
            here = r * side + c
            neighbours = ([here + 1] if c + 1 < side else []) + ([here + side] if r + 1 < side else [])
            for there in neighbours:
                nw += 1
                for j in (here, there):
                    worker.append(nw)
                    firm.append(j + 1)
                    freq.append(1)
            for _ in range(hub_workers):
                nw += 1
                for j, copies in ((here + 1, 1), (hub, hub_copies)):
                    worker.append(nw)
                    firm.append(j)
                    freq.append(copies)
    rows = len(worker)
    cols = (np.array(worker, np.float64), np.array(firm, np.float64), np.arange(1, rows + 1, dtype=np.float64),
            np.zeros(rows), np.array(freq, np.float64), np.ones(rows))
    return cols, firms


def smooth_psi(firms_total, grid_firms, side=40):
    """psi of every dense firm index (grid firms in id order, hub last)."""
    j = np.arange(firms_total)
    r, c = (j // side).astype(float), (j % side).astype(float)
    psi = 1.0e3 * np.cos(np.pi * r / side) * np.cos(np.pi * c / side)
    psi[j >= grid_firms] = 0.0
    return psi


def outcome_rhs(cell_worker, cell_firm, cell_weight, workers, firms, psi):
    """Normal-equation right-hand sides X'W y of y = F psi, from the cells."""
    wr = np.zeros(workers)
    fr = np.zeros(firms)
    y = cell_weight * psi[cell_firm]
    np.add.at(wr, cell_worker, y)
    np.add.at(fr, cell_firm, y)
    return wr, fr
