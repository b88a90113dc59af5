"""ctypes binding of the CPU oracle (test infrastructure only).

The oracle is a plain-C restatement of the reference engine; see oracle/orc.h
for the reference file:line each function follows.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_LIB = None
LEVERAGE_TAG = 0x4C45564552414745
TARGET_TAG = 0x0054415247455401


class OracleError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"oracle error {code}: {message}")
        self.code = code


def _u64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def _f64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class EngineOptions(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("probes", C.c_uint32),
                ("leverage_width", C.c_uint32), ("target_width", C.c_uint32),
                ("threads", C.c_uint32), ("fit_tolerance", C.c_double),
                ("probe_tolerance", C.c_double), ("max_iterations", C.c_uint32),
                ("recompute_interval", C.c_uint32), ("rank_tolerance", C.c_double),
                ("block_tolerance", C.c_double)]


class EngineSummary(C.Structure):
    _fields_ = [("plugin", C.c_double * 4), ("correction", C.c_double * 4),
                ("corrected", C.c_double * 4), ("mcse", C.c_double * 4),
                ("weighted_rss", C.c_double), ("leverage_iterations", C.c_uint64),
                ("target_iterations", C.c_uint64), ("fit_iterations", C.c_uint64),
                ("max_complete_residual", C.c_double), ("solve_seconds", C.c_double)]


class PcgOptions(C.Structure):
    _fields_ = [("relative_tolerance", C.c_double), ("absolute_tolerance", C.c_double),
                ("max_iterations", C.c_uint64), ("recompute_interval", C.c_uint64)]


class PcgDiag(C.Structure):
    _fields_ = [("iterations", C.c_uint64), ("restarts", C.c_uint64),
                ("initial_residual_norm", C.c_double), ("residual_norm", C.c_double),
                ("relative_residual", C.c_double), ("backward_error", C.c_double),
                ("tolerance", C.c_double), ("rhs_projection_norm", C.c_double)]


class CmgOptions(C.Structure):
    _fields_ = [("direct_threshold", C.c_uint64), ("max_levels", C.c_uint64),
                ("max_hierarchy_nnz_factor", C.c_double), ("low_degree_threshold", C.c_double)]


class NumericalResult(C.Structure):
    _fields_ = [("status", C.c_uint32), ("leverage_probes", C.c_uint32), ("target_probes", C.c_uint32),
                ("fold_a", C.c_uint32), ("fold_b", C.c_uint32), ("failed_replay_probe", C.c_uint32),
                ("replay_error_code", C.c_uint32), ("reserved", C.c_uint32),
                ("certified_replay_rhs", C.c_uint64), ("attempted_replay_rhs", C.c_uint64),
                ("replay_generator_words", C.c_uint64), ("point_rhs", C.c_uint64),
                ("psd_adjustment", C.c_double), ("minimum_constrained", C.c_double),
                ("minimum_residual_margin", C.c_double), ("maximum_sensitivity_ratio", C.c_double),
                ("score_means", C.c_double * 6), ("conditional", C.c_double * 9),
                ("leverage", C.c_double * 9), ("raw", C.c_double * 9), ("usable", C.c_double * 9),
                ("mcse", C.c_double * 4), ("max_point_complete_residual", C.c_double),
                ("max_replay_complete_residual", C.c_double)]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = np.array(list(v)) if hasattr(v, "__len__") else v
        for k in ("conditional", "leverage", "raw", "usable"):
            out[k] = out[k].reshape(3, 3)
        out["score_means"] = out["score_means"].reshape(2, 3)
        return out


def finite_derivative(u, r):
    """numerical_mc.rs FiniteDerivative::new: dict or None."""
    out = np.zeros(20)
    if not lib().orc_finite_derivative(np.ascontiguousarray(u, np.float64).ctypes.data, r, out.ctypes.data):
        return None
    return dict(h=out[0], m=out[1], bias=out[2], variance=out[3], dh=out[4:9], db=out[9:14], dv=out[14:19],
                nonsmooth=bool(out[19]), raw=out)


def fd_observation(fd, control=0.0):
    g = np.zeros(5)
    if not lib().orc_fd_observation(fd["raw"].ctypes.data, control, g.ctypes.data):
        return None
    return g


def copy_pullback(x, r, control=0.0):
    """One observation copy (generic_jla/numerical.rs:213-256): (status,
    value, pullback) for x = (a2, a4, c1, c3)."""
    u, fd, pb, v = np.zeros(5), np.zeros(20), np.zeros(4), C.c_double()
    st = lib().orc_copy_pullback(np.ascontiguousarray(x, np.float64).ctypes.data, control, r, u.ctypes.data,
                                 fd.ctypes.data, C.byref(v), pb.ctypes.data)
    return st, v.value, pb


def cross_covariance(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    out = np.zeros(9)
    _check(lib().orc_cross_covariance(a.ctypes.data, b.ctypes.data, len(a), out.ctypes.data))
    return out.reshape(3, 3)


def numerical_finalize(conditional, leverage):
    raw, usable, mcse = np.zeros(9), np.zeros(9), np.zeros(4)
    status = C.c_uint32()
    psd = C.c_double()
    c = np.ascontiguousarray(conditional, np.float64).ravel()
    lv = np.ascontiguousarray(leverage, np.float64).ravel()
    lib().orc_numerical_finalize(c.ctypes.data, lv.ctypes.data, raw.ctypes.data, usable.ctypes.data,
                                 mcse.ctypes.data, C.byref(status), C.byref(psd))
    return dict(raw=raw.reshape(3, 3), usable=usable.reshape(3, 3), mcse=mcse, status=int(status.value),
                psd_adjustment=psd.value)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "liborc.so")
        if not os.path.exists(path):
            from . import build as _b  # type: ignore
            _b.build()
        L = C.CDLL(path)
        L.orc_last_error.restype = C.c_char_p
        L.orc_rng_word.restype = C.c_uint64
        L.orc_observation_atoms.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p]
        L.orc_prepare_stayers.argtypes = [C.c_void_p, C.c_uint64] + [C.c_void_p] * 6
        L.orc_graph_from_edges.restype = C.c_void_p
        L.orc_graph_from_edges.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_graph_vertices.restype = C.c_uint64
        L.orc_graph_vertices.argtypes = [C.c_void_p]
        L.orc_graph_edges.restype = C.c_uint64
        L.orc_graph_edges.argtypes = [C.c_void_p]
        L.orc_graph_export.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        L.orc_graph_free.argtypes = [C.c_void_p]
        L.orc_graph_matvec.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_cmg_build.argtypes = [C.c_void_p, CmgOptions, C.POINTER(C.c_void_p)]
        L.orc_cmg_free.argtypes = [C.c_void_p]
        L.orc_cmg_levels.restype = C.c_uint64
        L.orc_cmg_levels.argtypes = [C.c_void_p]
        L.orc_cmg_level_info.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
        L.orc_cmg_level_labels.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
        L.orc_cmg_level_graph.restype = C.c_void_p
        L.orc_cmg_level_graph.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_cmg_apply.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_pcg_solve.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, PcgOptions, C.POINTER(PcgDiag)]
        L.orc_prepare.argtypes = [C.c_uint64] + [C.c_void_p] * 6 + [C.POINTER(C.c_void_p)]
        L.orc_prepare_masked.argtypes = [C.c_uint64] + [C.c_void_p] * 7 + [C.POINTER(C.c_void_p)]
        L.orc_prepare_observation.argtypes = [C.c_uint64] + [C.c_void_p] * 6 + [C.POINTER(C.c_void_p)]
        L.orc_problem_free.argtypes = [C.c_void_p]
        L.orc_problem_dims.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_problem_target_total.restype = C.c_double
        L.orc_problem_target_total.argtypes = [C.c_void_p]
        L.orc_problem_topology_checksum.restype = C.c_uint64
        L.orc_problem_topology_checksum.argtypes = [C.c_void_p]
        L.orc_problem_mask.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_problem_array.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_hybrid_graph.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
        L.orc_engine_run.argtypes = [C.c_void_p, EngineOptions, C.POINTER(C.c_void_p)]
        L.orc_engine_run_targets.argtypes = [C.c_void_p, EngineOptions, C.POINTER(C.c_void_p)]
        L.orc_engine_result_free.argtypes = [C.c_void_p]
        L.orc_engine_result_summary.argtypes = [C.c_void_p, C.POINTER(EngineSummary)]
        L.orc_engine_result_array.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_uint64]
        L.orc_engine_run_numerical.argtypes = [C.c_void_p, EngineOptions, C.POINTER(C.c_void_p),
                                               C.POINTER(NumericalResult)]
        L.orc_finite_derivative.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
        L.orc_fd_observation.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
        L.orc_copy_pullback.argtypes = [C.c_void_p, C.c_double, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p]
        L.orc_cross_covariance.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        L.orc_numerical_finalize.argtypes = [C.c_void_p] * 5 + [C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
        L.orc_time_probe_solves.argtypes = [C.c_void_p, EngineOptions, C.c_uint32, C.c_uint32,
                                            C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
        L.orc_frozen_digest.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_uint64)]
        L.orc_philox_block.argtypes = [C.c_uint64] * 5 + [C.c_void_p]
        L.orc_rng_word.argtypes = [C.c_uint64] * 5
        L.orc_rademacher_sum.argtypes = [C.c_uint64] * 5 + [C.POINTER(C.c_int64)]
        L.orc_rademacher_fill.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                          C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        L.orc_set_empty_sum_negative.argtypes = [C.c_int]
        L.orc_refinement_options.argtypes = [C.c_double] * 4 + [C.POINTER(C.c_double)] * 2
        L.orc_refinement_factor.restype = C.c_double
        L.orc_refinement_factor.argtypes = [C.c_uint32]
        L.orc_failing_columns.restype = C.c_uint64
        L.orc_failing_columns.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_double, C.c_void_p]
        L.orc_euclidean_norm.restype = C.c_double
        L.orc_euclidean_norm.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_hybrid_lift.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
        L.orc_solve_columns.argtypes = ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_double, C.c_double,
                                         C.c_uint32, C.c_uint32, C.c_uint32] + [C.c_void_p] * 7)
        _LIB = L
    return _LIB


def _check(rc: int):
    if rc != 0:
        raise OracleError(rc, lib().orc_last_error().decode())


# ---------------- RNG ----------------

def philox_block(seed, tag, block, entity, word):
    out = np.zeros(4, dtype=np.uint64)
    lib().orc_philox_block(seed, tag, block, entity, word, out.ctypes.data)
    return [int(x) for x in out]


def rng_word(seed, tag, probe, entity, word):
    return int(lib().orc_rng_word(seed, tag, probe, entity, word))


def rademacher_sum(seed, tag, probe, entity, trials):
    out = C.c_int64()
    _check(lib().orc_rademacher_sum(seed, tag, probe, entity, trials, C.byref(out)))
    return out.value


def rademacher_fill(seed, tag, first_probe, columns, entity, trials):
    entity = np.ascontiguousarray(entity, dtype=np.uint64)
    trials = np.ascontiguousarray(trials, dtype=np.uint64)
    out = np.zeros(columns * len(entity), dtype=np.int64)
    _check(lib().orc_rademacher_fill(seed, tag, first_probe, columns, entity.ctypes.data,
                                     trials.ctypes.data, len(entity), out.ctypes.data))
    return out.reshape(columns, len(entity))


# ---------------- CMG_FULL_V2 batch-solve pieces ----------------

def refinement_options(pcg_tolerance, factor, gate, smallest_norm):
    """(relative, absolute) PCG stop of one refinement round (full_cmg.rs:1375-1390)."""
    rel, ab = C.c_double(), C.c_double()
    lib().orc_refinement_options(pcg_tolerance, factor, gate, smallest_norm, C.byref(rel), C.byref(ab))
    return rel.value, ab.value


def refinement_factors():
    return [float(lib().orc_refinement_factor(r)) for r in range(4)]


def failing_columns(complete, reduced, tolerance):
    """Columns the complete-residual gate sends to refinement (full_cmg.rs:1445-1470)."""
    c = np.ascontiguousarray(complete, np.float64)
    r = np.ascontiguousarray(reduced, np.float64)
    out = np.zeros(len(c), np.uint64)
    n = lib().orc_failing_columns(c.ctypes.data, r.ctypes.data, len(c), tolerance, out.ctypes.data)
    return [int(x) for x in out[:n]]


def euclidean_norm(v):
    v = np.ascontiguousarray(v, np.float64)
    return float(lib().orc_euclidean_norm(v.ctypes.data, len(v)))


# ---------------- CMG ----------------

def frozen_digest(family: str, kind: str) -> int:
    out = C.c_uint64()
    _check(lib().orc_frozen_digest(family.encode(), kind.encode(), C.byref(out)))
    return out.value


class Graph:
    def __init__(self, n, u, v, w):
        u = np.ascontiguousarray(u, dtype=np.uint32)
        v = np.ascontiguousarray(v, dtype=np.uint32)
        w = np.ascontiguousarray(w, dtype=np.float64)
        self._h = lib().orc_graph_from_edges(n, len(u), u.ctypes.data, v.ctypes.data, w.ctypes.data)
        if not self._h:
            raise OracleError(50, lib().orc_last_error().decode())
        self._owned = True

    @classmethod
    def _borrow(cls, handle):
        g = cls.__new__(cls)
        g._h = handle
        g._owned = False
        return g

    @property
    def n(self):
        return int(lib().orc_graph_vertices(self._h))

    @property
    def m(self):
        return int(lib().orc_graph_edges(self._h))

    def export(self):
        m, n = self.m, self.n
        u = np.zeros(m, np.uint32); v = np.zeros(m, np.uint32)
        w = np.zeros(m, np.float64); d = np.zeros(n, np.float64)
        lib().orc_graph_export(self._h, u.ctypes.data, v.ctypes.data, w.ctypes.data, d.ctypes.data)
        return u, v, w, d

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        lib().orc_graph_matvec(self._h, x.ctypes.data, y.ctypes.data)
        return y

    def __del__(self):
        if getattr(self, "_owned", False) and self._h:
            lib().orc_graph_free(self._h)
            self._h = None


class Cmg:
    def __init__(self, graph: Graph, direct_threshold=700, max_levels=32):
        self.graph = graph
        h = C.c_void_p()
        _check(lib().orc_cmg_build(graph._h, CmgOptions(direct_threshold, max_levels, 5.0, 0.125),
                                   C.byref(h)))
        self._h = h.value

    def levels(self):
        out = []
        for i in range(int(lib().orc_cmg_levels(self._h))):
            info = np.zeros(6, np.uint64)
            lib().orc_cmg_level_info(self._h, i, info.ctypes.data)
            g = Graph._borrow(lib().orc_cmg_level_graph(self._h, i))
            labels = None
            if info[5] or i + 1 < int(lib().orc_cmg_levels(self._h)):
                labels = np.zeros(int(info[0]), np.uint32)
                if lib().orc_cmg_level_labels(self._h, i, labels.ctypes.data) != 0:
                    labels = None
            u, v, w, d = g.export()
            out.append(dict(n=int(info[0]), m=int(info[1]), nnz=int(info[2]),
                            repeat=int(info[3]), terminal=int(info[4]), coarse=int(info[5]),
                            u=u, v=v, w=w, diag=d, labels=labels))
        return out

    def apply(self, rhs, project=False):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        out = np.zeros(self.graph.n)
        _check(lib().orc_cmg_apply(self._h, rhs.ctypes.data, out.ctypes.data, int(project)))
        return out

    def solve(self, rhs, rel=1e-8, abs_tol=0.0, max_iterations=1000, recompute=25, guess=None):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros(self.graph.n)
        d = PcgDiag()
        g = None if guess is None else np.ascontiguousarray(guess, dtype=np.float64)
        _check(lib().orc_pcg_solve(self.graph._h, self._h, rhs.ctypes.data,
                                   None if g is None else g.ctypes.data, x.ctypes.data,
                                   PcgOptions(rel, abs_tol, max_iterations, recompute), C.byref(d)))
        return x, d

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_cmg_free(self._h)
            self._h = None


# ---------------- sample + engine ----------------

class Problem:
    ARRAYS = {"cell_worker": (0, np.uint32, 5), "cell_firm": (1, np.uint32, 5),
              "cell_weight": (2, np.float64, 5), "unit_cell": (3, np.uint32, 6),
              "unit_count": (4, np.uint64, 6), "unit_rank": (5, np.uint64, 6),
              "stratum_cell": (6, np.uint32, 8), "stratum_count": (7, np.uint64, 8),
              "stratum_rank": (8, np.uint64, 8), "stratum_per_copy": (9, np.float64, 8),
              "cell_outcome_sum": (10, np.float64, 5), "cell_target_sum": (11, np.float64, 5),
              "unit_outcome_sum": (12, np.float64, 6), "unit_offset": (13, np.uint64, 6),
              "retained": (14, np.uint64, 1), "row_cell": (15, np.uint32, 1),
              "unit_ptr": (16, np.uint64, 9), "unit_rows": (17, np.uint64, 1)}

    def __init__(self, worker, firm, deletion, outcome, frequency, target, mask=None, observation=False):
        cols = [np.ascontiguousarray(worker, np.uint64), np.ascontiguousarray(firm, np.uint64),
                np.ascontiguousarray(deletion, np.uint64), np.ascontiguousarray(outcome, np.float64),
                np.ascontiguousarray(frequency, np.uint64), np.ascontiguousarray(target, np.float64)]
        h = C.c_void_p()
        if observation:
            _check(lib().orc_prepare_observation(len(cols[0]), *[c.ctypes.data for c in cols], C.byref(h)))
        elif mask is None:
            _check(lib().orc_prepare(len(cols[0]), *[c.ctypes.data for c in cols], C.byref(h)))
        else:
            m = np.ascontiguousarray(mask, dtype=np.uint8)
            _check(lib().orc_prepare_masked(len(cols[0]), *[c.ctypes.data for c in cols],
                                            m.ctypes.data, C.byref(h)))
        self._adopt(h.value)

    def _adopt(self, handle):
        self._h = handle
        d = np.zeros(9, np.uint64)
        lib().orc_problem_dims(self._h, d.ctypes.data)
        (self.input_rows, self.rows, self.physical, self.workers, self.firms, self.cells,
         self.units, self.exact_strata, self.strata) = [int(x) for x in d]
        self.target_total = float(lib().orc_problem_target_total(self._h))
        self.topology_checksum = int(lib().orc_problem_topology_checksum(self._h))

    def with_stayers(self, firm, worker, outcome, frequency, target):
        """The stayer hybrid of this (match-deletion) problem: eligible stayer
        rows with 1-based dense retained-mover firm and stayer-worker indices."""
        cols = [np.ascontiguousarray(firm, np.uint64), np.ascontiguousarray(worker, np.uint64),
                np.ascontiguousarray(outcome, np.float64), np.ascontiguousarray(frequency, np.uint64),
                np.ascontiguousarray(target, np.float64)]
        h = C.c_void_p()
        _check(lib().orc_prepare_stayers(self._h, len(cols[0]), *[c.ctypes.data for c in cols], C.byref(h)))
        out = Problem.__new__(Problem)
        out._adopt(h.value)
        out.mover = self
        return out

    def mask(self):
        m = np.zeros(self.input_rows, np.uint8)
        lib().orc_problem_mask(self._h, m.ctypes.data)
        return m.astype(bool)

    def array(self, name):
        which, dtype, dim = self.ARRAYS[name]
        n = int([self.input_rows, self.rows, self.physical, self.workers, self.firms, self.cells,
                 self.units, self.exact_strata, self.strata, self.units + 1][dim])
        out = np.zeros(n, dtype)
        _check(lib().orc_problem_array(self._h, which, out.ctypes.data))
        return out

    def observation_atoms(self, seed, first, columns):
        out = np.zeros(columns * self.units, np.int64)
        _check(lib().orc_observation_atoms(self._h, seed, first, columns, out.ctypes.data))
        return out.reshape(columns, self.units)

    def hybrid_graph(self):
        h = C.c_void_p()
        aux = C.c_uint64()
        _check(lib().orc_hybrid_graph(self._h, C.byref(h), C.byref(aux)))
        g = Graph._borrow(h.value)
        g._owned = True
        return g, int(aux.value)

    def hybrid_lift(self, firm):
        """Harmonic lift of a firm vector onto the hybrid vertices (cmg_impl.rs.in:459-500)."""
        g, aux = self.hybrid_graph()
        n = g.n
        f = np.ascontiguousarray(firm, np.float64)
        out = np.zeros(n)
        _check(lib().orc_hybrid_lift(self._h, f.ctypes.data, out.ctypes.data, n))
        return out

    def solve_columns(self, worker_rhs, firm_rhs, tolerance=1e-10, inner_ratio=1.0, threads=1,
                      max_iterations=10000, recompute_interval=100):
        """One CMG_FULL_V2 batch solve (full_cmg.rs:809-1111) of the columns of
        worker_rhs (workers x k) and firm_rhs (firms x k)."""
        wr = np.ascontiguousarray(np.asarray(worker_rhs, np.float64).T)
        fr = np.ascontiguousarray(np.asarray(firm_rhs, np.float64).T)
        k = wr.shape[0]
        wo, fo = np.zeros_like(wr), np.zeros_like(fr)
        comp, red = np.zeros(k), np.zeros(k)
        it = np.zeros(k, np.uint64)
        rounds, refined = C.c_uint64(), C.c_uint64()
        _check(lib().orc_solve_columns(self._h, wr.ctypes.data, fr.ctypes.data, k, tolerance, inner_ratio,
                                       max_iterations, recompute_interval, threads, wo.ctypes.data,
                                       fo.ctypes.data, comp.ctypes.data, red.ctypes.data, it.ctypes.data,
                                       C.byref(rounds), C.byref(refined)))
        return wo.T, fo.T, {"complete": comp, "reduced": red, "iterations": it.astype(int)}, \
            {"rounds": int(rounds.value), "refined": int(refined.value)}

    def run(self, seed=8675309, probes=200, leverage_width=8, target_width=8, threads=1,
            fit_tolerance=1e-10, probe_tolerance=1e-6, max_iterations=10000,
            recompute_interval=100, rank_tolerance=1e-10, block_tolerance=1e-10, numerical=False,
            targets=False):
        """The compressed JLA engine.  With `targets`, the result also holds
        every target probe's two FE solutions: target_worker (2R x workers) and
        target_firm (2R x firms), row 2t the worker-side solve of probe t."""
        opt = EngineOptions(seed, probes, leverage_width, target_width, threads, fit_tolerance,
                            probe_tolerance, max_iterations, recompute_interval, rank_tolerance,
                            block_tolerance)
        h = C.c_void_p()
        num = NumericalResult() if numerical else None
        if numerical:
            _check(lib().orc_engine_run_numerical(self._h, opt, C.byref(h), C.byref(num)))
        elif targets:
            _check(lib().orc_engine_run_targets(self._h, opt, C.byref(h)))
        else:
            _check(lib().orc_engine_run(self._h, opt, C.byref(h)))
        try:
            s = EngineSummary()
            lib().orc_engine_result_summary(h.value, C.byref(s))
            res = dict(plugin=list(s.plugin), correction=list(s.correction),
                       corrected=list(s.corrected), mcse=list(s.mcse),
                       weighted_rss=s.weighted_rss, leverage_iterations=s.leverage_iterations,
                       target_iterations=s.target_iterations, fit_iterations=s.fit_iterations,
                       max_complete_residual=s.max_complete_residual,
                       solve_seconds=s.solve_seconds)
            names = ["fitted_cell", "projection_share", "residual_share", "finite_bias",
                     "finite_variance", "residual_mass", "deleted_mass", "cell_correction_weight",
                     "target_draws", "leverage_iterations_per_probe"]
            sizes = [self.cells] + [self.units] * 6 + [self.cells, 4 * probes, probes]
            for i, (name, n) in enumerate(zip(names, sizes)):
                a = np.zeros(n)
                _check(lib().orc_engine_result_array(h.value, i, a.ctypes.data, n))
                res[name] = a.reshape(probes, 4) if name == "target_draws" else a
            if numerical:
                res["numerical"] = num.as_dict()
            if targets:
                for i, (name, n) in enumerate((("target_worker", self.workers), ("target_firm", self.firms))):
                    a = np.zeros(2 * probes * n)
                    _check(lib().orc_engine_result_array(h.value, 10 + i, a.ctypes.data, a.size))
                    res[name] = a.reshape(2 * probes, n)
            return res
        finally:
            lib().orc_engine_result_free(h.value)

    def time_probe_solves(self, count, first=0, seed=8675309, threads=1, probe_tolerance=1e-8,
                          max_iterations=10000, recompute_interval=100):
        opt = EngineOptions(seed, 2, 1, 1, threads, probe_tolerance, probe_tolerance,
                            max_iterations, recompute_interval, 1e-10, 1e-10)
        sec = C.c_double()
        it = C.c_uint64()
        _check(lib().orc_time_probe_solves(self._h, opt, first, count, C.byref(sec), C.byref(it)))
        return sec.value, int(it.value)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_problem_free(self._h)
            self._h = None
