"""MI355X leave-out (KSS) engine: Python host mirror of the C ABI.

The product is libvckss_mi355x.so (fevc_amd/csrc, built by fevc_amd/build.py).
This module binds its vckss_rust_* entry points the way the reference Stata
shim binds the Rust plugin (rust/stata_backend/cshim/stata_entry.c:1836-1852
prepare, 2530-2588 solvefull, result, release) and exposes the extension
entry points used by the benchmark and the parity tests.  It never computes
anything itself: a missing library or device raises instead of falling back.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = ["EngineError", "Engine", "library", "build_hierarchy", "hierarchy_layout"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libvckss_mi355x.so")
_LIB = None


class EngineError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"vckss status {code}: {message}")
        self.code = code


Poll = C.CFUNCTYPE(C.c_int32, C.c_void_p)


class PrepareRequestV4(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("struct_size", C.c_uint32), ("rows", C.c_uint64),
                ("cleanup_abandoned", C.c_uint32), ("reserved", C.c_uint32),
                ("memory_limit_bytes", C.c_uint64), ("caller_copy_bytes", C.c_uint64),
                ("deletion_mode", C.c_uint32), ("controls_count", C.c_uint32),
                ("reserved_2", C.c_uint64), ("implicit_match", C.c_uint32), ("reserved_3", C.c_uint32)]


class PrepareRequestInterruptV3(C.Structure):
    _fields_ = [("options", PrepareRequestV4), ("interrupt_poll", C.c_void_p),
                ("interrupt_context", C.c_void_p), ("checkpoint_interval", C.c_uint32),
                ("reserved", C.c_uint32)]


class ColumnsV3(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("reserved", C.c_uint32), ("rows", C.c_uint64),
                ("worker", C.c_void_p), ("firm", C.c_void_p), ("deletion", C.c_void_p),
                ("outcome", C.c_void_p), ("frequency", C.c_void_p), ("target_weight", C.c_void_p),
                ("controls", C.c_void_p), ("controls_count", C.c_uint32), ("reserved_2", C.c_uint32),
                ("probe_order", C.c_void_p), ("probeorder_supplied", C.c_uint32),
                ("reserved_3", C.c_uint32)]


class SolveRequestV5(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32), ("struct_size", C.c_uint32), ("seed", C.c_uint64),
        ("probes", C.c_uint32), ("leverage_batch_width", C.c_uint32),
        ("target_batch_width", C.c_uint32), ("deletion_mode", C.c_uint32),
        ("rng_contract", C.c_uint32), ("solver_route", C.c_uint32),
        ("allow_automatic_cmg_setup_fallback", C.c_uint32), ("exact_dimension_limit", C.c_uint64),
        ("cmg_minimum_dimension", C.c_uint64), ("pcg_tolerance", C.c_double),
        ("maximum_iterations", C.c_uint32), ("residual_replacement_interval", C.c_uint32),
        ("rank_tolerance", C.c_double), ("block_tolerance", C.c_double),
        ("cmg_terminal_vertices", C.c_uint64), ("cmg_dense_vertex_cap", C.c_uint64),
        ("cmg_maximum_levels", C.c_uint64), ("cmg_aggregate_cap", C.c_uint64),
        ("cmg_minimum_reduction", C.c_double), ("cmg_jacobi_weight", C.c_double),
        ("cmg_pre_sweeps", C.c_uint32), ("cmg_post_sweeps", C.c_uint32),
        ("cmg_maximum_edge_complexity", C.c_double), ("cmg_maximum_vertex_complexity", C.c_double),
        ("cmg_memory_limit_bytes", C.c_uint64),
        ("algorithm", C.c_uint32), ("nuisance_mode", C.c_uint32),
        ("exact_estimator_limit", C.c_uint64), ("blocksize_limit", C.c_uint64),
        ("engine", C.c_uint32), ("batch_mode", C.c_uint32), ("stayers_mode", C.c_uint32),
        ("target_weight_mode", C.c_uint32), ("deletion_unit_source", C.c_uint32),
        ("probeorder_supplied", C.c_uint32), ("wallseconds_supplied", C.c_uint32),
        ("capability_schema", C.c_uint32), ("capability_profile", C.c_uint32),
        ("frequency_use", C.c_uint32), ("physical_limit", C.c_uint64),
        ("request_signature", C.c_uint64), ("reserved_3", C.c_uint64),
        ("leverage_batch_mode", C.c_uint32), ("target_batch_mode", C.c_uint32),
        ("wallseconds", C.c_double), ("reserved_4", C.c_uint64),
        ("threads", C.c_uint32), ("tolerance_supplied", C.c_uint32), ("full_cmg_v2", C.c_uint32),
        ("reserved_5", C.c_uint32)]


class SolveRequestV4(C.Structure):
    """The V4 point prefix of the solve request (the numerical requests embed it)."""
    _fields_ = SolveRequestV5._fields_[:-4]


class NumericalRequestV1(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("struct_size", C.c_uint32), ("schema_version", C.c_uint32),
                ("threads", C.c_uint32), ("tolerance_supplied", C.c_uint32),
                ("projection_requested", C.c_uint32), ("component_requested", C.c_uint32),
                ("reserved", C.c_uint32), ("controls_count", C.c_uint32), ("reserved_2", C.c_uint32),
                ("point", SolveRequestV4)]


class NumericalRequestV2(C.Structure):
    _fields_ = [("v1", NumericalRequestV1), ("execution_mode", C.c_uint32),
                ("component_batch_mode", C.c_uint32), ("full_cmg", C.c_uint32), ("reserved", C.c_uint32)]


class NumericalResultV1(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("schema_version", C.c_uint32), ("generation", C.c_uint64),
                ("status", C.c_uint32), ("engine", C.c_uint32), ("leverage_probes", C.c_uint32),
                ("target_probes", C.c_uint32), ("target_fold_a", C.c_uint32), ("target_fold_b", C.c_uint32),
                ("certified_replay_rhs", C.c_uint64), ("executed_replay_rhs", C.c_uint64),
                ("attempted_replay_rhs", C.c_uint64), ("replay_generator_words", C.c_uint64),
                ("allocation_bound_bytes", C.c_uint64), ("failed_replay_probe", C.c_uint32),
                ("replay_error_code", C.c_uint32), ("psd_adjustment", C.c_double),
                ("minimum_constrained", C.c_double), ("minimum_residual_margin", C.c_double),
                ("maximum_sensitivity_ratio", C.c_double), ("score_means", C.c_double * 6),
                ("conditional", C.c_double * 9), ("leverage", C.c_double * 9), ("raw", C.c_double * 9),
                ("usable", C.c_double * 9), ("mcse", C.c_double * 4), ("point_rhs", C.c_uint64),
                ("maximum_point_complete_residual", C.c_double),
                ("maximum_replay_complete_residual", C.c_double)]


class NumericalRhsV1(C.Structure):
    _fields_ = [("schema_version", C.c_uint32), ("phase", C.c_uint32), ("probe", C.c_uint32),
                ("reserved", C.c_uint32), ("complete_residual", C.c_double)]


class FullCmgModelReceiptV1(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("schema_version", C.c_uint32), ("generation", C.c_uint64),
                ("controls_count", C.c_uint32), ("nuisance_mode", C.c_uint32),
                ("logical_rhs_count", C.c_uint64), ("explicit_options_rhs_count", C.c_uint64),
                ("controlled_rhs_count", C.c_uint64), ("control_refinement_rhs_count", C.c_uint64)]


assert C.sizeof(NumericalRequestV1) == 328 and C.sizeof(NumericalRequestV2) == 344
assert C.sizeof(NumericalResultV1) == 512 and C.sizeof(NumericalRhsV1) == 24


class SolveRequestInterruptV5(C.Structure):
    _fields_ = [("options", SolveRequestV5), ("interrupt_poll", C.c_void_p),
                ("interrupt_context", C.c_void_p), ("checkpoint_interval", C.c_uint32),
                ("reserved", C.c_uint32)]


class ComponentVector(C.Structure):
    _fields_ = [("worker", C.c_double), ("firm", C.c_double), ("covariance", C.c_double),
                ("total", C.c_double)]

    def as_list(self):
        return [self.worker, self.firm, self.covariance, self.total]


class ResultV1(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("reserved", C.c_uint32), ("generation", C.c_uint64),
                ("plugin", ComponentVector), ("correction", ComponentVector),
                ("corrected", ComponentVector), ("numerical_mcse", ComponentVector)]


class SnapshotV1(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("state", C.c_uint32), ("generation", C.c_uint64),
                ("last_released_generation", C.c_uint64)]


# (name, restype, argtypes) of every exported symbol declared in include/*.h
class StayerRequestV1(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("struct_size", C.c_uint32), ("rows", C.c_uint64),
                ("controls_count", C.c_uint32), ("reserved", C.c_uint32), ("caller_copy_bytes", C.c_uint64)]


class StayerRequestInterruptV1(C.Structure):
    _fields_ = [("options", StayerRequestV1), ("interrupt_poll", C.c_void_p), ("interrupt_context", C.c_void_p),
                ("checkpoint_interval", C.c_uint32), ("reserved", C.c_uint32)]


class StayerColumnsV1(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("reserved", C.c_uint32), ("rows", C.c_uint64),
                ("firm", C.c_void_p), ("worker", C.c_void_p), ("outcome", C.c_void_p), ("frequency", C.c_void_p),
                ("target_weight", C.c_void_p), ("controls", C.c_void_p), ("controls_count", C.c_uint32),
                ("reserved_2", C.c_uint32)]


class StayerReceiptV1(C.Structure):
    _fields_ = ([("struct_size", C.c_uint32), ("schema_version", C.c_uint32), ("generation", C.c_uint64)] +
                [(k, C.c_uint64) for k in ("mover_stored_rows", "stayer_stored_rows", "combined_stored_rows",
                                           "mover_physical_mass", "stayer_physical_mass", "combined_physical_mass",
                                           "mover_workers", "stayer_workers", "combined_workers", "firms",
                                           "mover_deletion_units", "stayer_deletion_units",
                                           "combined_deletion_units")] +
                [(k, C.c_double) for k in ("mover_target_mass", "stayer_target_mass", "combined_target_mass")] +
                [(k, C.c_uint64) for k in ("topology_checksum", "memory_limit_bytes", "caller_copy_bytes",
                                           "augmentation_peak_forecast_bytes", "augmented_resident_bytes",
                                           "total_prepared_resident_bytes")])


class _Vec4(C.Structure):
    _fields_ = [("worker", C.c_double), ("firm", C.c_double), ("covariance", C.c_double), ("total", C.c_double)]


class StayerHybridResultV1(C.Structure):
    _fields_ = ([("struct_size", C.c_uint32), ("schema_version", C.c_uint32), ("generation", C.c_uint64)] +
                [(k, _Vec4) for k in ("plugin", "correction", "corrected", "mover_correction",
                                      "stayer_correction")] +
                [("weighted_rss", C.c_double)] +
                [(k, C.c_uint64) for k in ("parameters", "full_parameters", "correction_parameters",
                                           "deletion_units")] +
                [(k, C.c_double) for k in ("max_leverage", "information_rcond", "inverse_relres",
                                           "inverse_original_relres", "inverse_sqrt_relres", "maker_relres",
                                           "full_fit_relres", "working_fit_relres", "fit_residual_tolerance",
                                           "control_basis_relres", "control_basis_forward_error",
                                           "deletion_rank_gap", "firm_zero_sum_residual")] +
                [(k, C.c_uint64) for k in ("peak_forecast_bytes", "fit_peak_forecast_bytes",
                                           "correction_peak_forecast_bytes", "topology_checksum")] +
                [("accounting_residual", C.c_double)])


SIGNATURES = [
    ("vckss_rust_abi_version", C.c_uint32, []),
    ("vckss_rust_backend_version", C.c_char_p, []),
    ("vckss_rust_capabilities_json", C.c_char_p, []),
    ("vckss_rust_last_error", C.c_char_p, []),
    ("vckss_rust_selftest", C.c_int32, []),
    ("vckss_rust_engine_default_prepare_request_interrupt_v3", C.c_int32,
     [C.POINTER(PrepareRequestInterruptV3), C.c_uint32]),
    ("vckss_rust_engine_default_solve_request_interrupt_v5", C.c_int32,
     [C.POINTER(SolveRequestInterruptV5), C.c_uint32]),
    ("vckss_rust_engine_prepare_interrupt_v4", C.c_int32,
     [C.POINTER(PrepareRequestInterruptV3), C.POINTER(ColumnsV3), C.POINTER(C.c_uint64), C.c_uint32]),
    ("vckss_rust_engine_retained_mask_v1", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint64]),
    ("vckss_rust_backend_capabilities_v1", C.c_int32, [C.c_void_p, C.c_uint32]),
    ("vckss_rust_backend_request_capability_v1", C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32]),
    ("vckss_rust_backend_request_capability_v2", C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32]),
    ("vckss_rust_backend_request_capability_v3", C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32]),
    ("vckss_rust_engine_admit_prepare_v4", C.c_int32, [C.c_void_p, C.c_uint32]),
    ("vckss_rust_engine_preparation_receipt_v4", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint32]),
    ("vckss_rust_engine_memory_forecast_v1", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint32]),
    ("vckss_rust_engine_performance_receipt_v1", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint32]),
    ("vckss_rust_engine_detailed_receipt_v6", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint32]),
    ("vckss_rust_engine_detailed_receipt_v7", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint32]),
    ("vckss_rust_engine_full_cmg_receipt_v1", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint32]),
    ("vckss_rust_engine_rhs_receipts_v1", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint64]),
    ("vckss_rust_engine_default_solve_request_v6", C.c_int32, [C.c_void_p, C.c_uint32]),
    ("vckss_rust_engine_default_solve_request_interrupt_v6", C.c_int32, [C.c_void_p, C.c_uint32]),
    ("vckss_rust_engine_solve_v6", C.c_int32, [C.c_uint64, C.c_void_p]),
    ("vckss_rust_engine_solve_interrupt_v6", C.c_int32, [C.c_uint64, C.c_void_p]),
    ("vckss_rust_engine_default_solve_request_v7", C.c_int32, [C.c_void_p, C.c_uint32]),
    ("vckss_rust_engine_default_solve_request_interrupt_v7", C.c_int32, [C.c_void_p, C.c_uint32]),
    ("vckss_rust_engine_solve_v7", C.c_int32, [C.c_uint64, C.c_void_p]),
    ("vckss_rust_engine_solve_interrupt_v7", C.c_int32, [C.c_uint64, C.c_void_p]),
    ("vckss_rust_engine_default_stayer_augmentation_request_interrupt_v1", C.c_int32,
     [C.POINTER(StayerRequestInterruptV1), C.c_uint32]),
    ("vckss_rust_engine_augment_stayers_v1", C.c_int32,
     [C.c_uint64, C.POINTER(StayerRequestV1), C.POINTER(StayerColumnsV1)]),
    ("vckss_rust_engine_augment_stayers_interrupt_v1", C.c_int32,
     [C.c_uint64, C.POINTER(StayerRequestInterruptV1), C.POINTER(StayerColumnsV1)]),
    ("vckss_rust_engine_stayer_augmentation_receipt_v1", C.c_int32,
     [C.c_uint64, C.POINTER(StayerReceiptV1), C.c_uint32]),
    ("vckss_rust_engine_stayer_hybrid_result_v1", C.c_int32,
     [C.c_uint64, C.POINTER(StayerHybridResultV1), C.c_uint32]),
    ("vckss_rust_engine_solve_v5", C.c_int32, [C.c_uint64, C.POINTER(SolveRequestV5)]),
    ("vckss_rust_engine_solve_interrupt_v5", C.c_int32,
     [C.c_uint64, C.POINTER(SolveRequestInterruptV5)]),
    ("vckss_rust_engine_result_v1", C.c_int32, [C.c_uint64, C.POINTER(ResultV1), C.c_uint32]),
    ("vckss_rust_engine_release_v1", C.c_int32, [C.c_uint64]),
    ("vckss_rust_engine_clear_abandoned_v1", C.c_int32, []),
    ("vckss_rust_engine_snapshot_v1", C.c_int32, [C.POINTER(SnapshotV1), C.c_uint32]),
    ("vckss_rust_engine_last_error", C.c_char_p, []),
    ("vckss_mi355x_configure_v1", C.c_int32, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]),
    ("vckss_mi355x_dims_v1", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint32]),
    ("vckss_mi355x_array_v1", C.c_int32, [C.c_uint64, C.c_uint32, C.c_void_p, C.c_uint64]),
    ("vckss_mi355x_atoms_v1", C.c_int32,
     [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, C.c_void_p, C.c_uint64]),
    ("vckss_mi355x_spmv_v1", C.c_int32, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    ("vckss_mi355x_spmv_noalloc_v1", C.c_int32, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    ("vckss_mi355x_vcycle_v1", C.c_int32, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    ("vckss_mi355x_pcg_v1", C.c_int32,
     [C.c_uint64, C.c_uint32, C.c_double, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("vckss_mi355x_time_leverage_v1", C.c_int32,
     [C.c_uint64, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.POINTER(C.c_double),
      C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_double), C.POINTER(C.c_double),
      C.POINTER(C.c_uint64)]),
    ("vckss_mi355x_first_pass_v1", C.c_int32, [C.c_uint64, C.c_uint32, C.POINTER(C.c_uint32)]),
    ("vckss_mi355x_kernel_timing_v1", C.c_int32,
     [C.c_uint64, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    ("vckss_mi355x_solve_columns_v1", C.c_int32,
     [C.c_uint64, C.c_uint32, C.c_double, C.c_double, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("vckss_mi355x_stream_copy_v1", C.c_int32, [C.c_uint64, C.c_uint64, C.POINTER(C.c_double)]),
    ("vckss_mi355x_counters_v1", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint32]),
    ("vckss_mi355x_set_chunk_v1", C.c_int32, [C.c_uint64, C.c_uint32]),
    ("vckss_rust_numerical_schema_v1", C.c_uint32, []),
    ("vckss_rust_numerical_schema_v2", C.c_uint32, []),
    ("vckss_rust_engine_numerical_preflight_v1", C.c_int32, [C.c_void_p]),
    ("vckss_rust_engine_numerical_preflight_v2", C.c_int32, [C.c_void_p]),
    ("vckss_rust_engine_solve_numerical_v1", C.c_int32, [C.c_uint64, C.c_void_p]),
    ("vckss_rust_engine_solve_numerical_interrupt_v1", C.c_int32,
     [C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    ("vckss_rust_engine_solve_numerical_interrupt_v2", C.c_int32,
     [C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    ("vckss_rust_engine_numerical_result_v1", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint32]),
    ("vckss_rust_engine_numerical_rhs_v1", C.c_int32,
     [C.c_uint64, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("vckss_rust_engine_numerical_cmg_work_v1", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint32]),
    ("vckss_mi355x_smoother_v1", C.c_int32, [C.c_uint64, C.c_uint32, C.c_void_p, C.c_uint32]),
    ("vckss_mi355x_smoother_status_v1", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint32]),
    ("vckss_mi355x_smoother_force_v1", C.c_int32, [C.c_uint64, C.c_double]),
    ("vckss_mi355x_exact_receipt_v1", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint32]),
    ("vckss_mi355x_controls_v1", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint32]),
    ("vckss_mi355x_batch_prepare_v1", C.c_int32, [C.c_uint64] + [C.c_void_p] * 7 +
     [C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    ("vckss_mi355x_batch_solve_v1", C.c_int32, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_double, C.c_double,
                                                C.c_uint32, C.c_uint32]),
    ("vckss_mi355x_batch_solve_v2", C.c_int32, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_double, C.c_double,
                                                 C.c_uint32, C.c_uint32, C.c_uint32]),
    ("vckss_mi355x_batch_result_v1", C.c_int32, [C.c_uint64, C.c_uint32, C.c_void_p, C.c_uint32]),
    ("vckss_mi355x_batch_release_v1", C.c_int32, [C.c_uint64]),
    ("vckss_mi355x_graphs_v1", C.c_int32,
     [C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("vckss_mi355x_cycle_shape_v1", C.c_int32, [C.c_uint64, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    ("vckss_mi355x_smoother_params_v1", C.c_int32, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    ("vckss_mi355x_smoother_omega_v1", C.c_int32,
     [C.c_uint64, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("vckss_mi355x_smoother_infer_v1", C.c_int32,
     [C.c_uint64, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("vckss_mi355x_smoother_gradient_v1", C.c_int32,
     [C.c_uint64, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(C.c_double), C.c_void_p]),
    ("vckss_mi355x_smoother_train_v1", C.c_int32,
     [C.c_uint64, C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint32, C.c_double, C.c_uint64, C.c_uint32,
      C.c_void_p]),
    ("vckss_mi355x_autotune_v1", C.c_int32,
     [C.c_uint64, C.c_uint32, C.c_double, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
      C.POINTER(C.c_double), C.c_void_p, C.c_uint32]),
    ("vckss_mi355x_hierarchy_build_v1", C.c_int32,
     [C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64,
      C.POINTER(C.c_uint64)]),
    ("vckss_mi355x_hierarchy_level_v1", C.c_int32,
     [C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("vckss_mi355x_hierarchy_levels_v1", C.c_int32,
     [C.c_uint64, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]),
    ("vckss_mi355x_hierarchy_free_v1", C.c_int32, [C.c_uint64]),
    ("vckss_mi355x_layout_v1", C.c_int32,
     [C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p] + [C.c_void_p] * 12),
    ("vckss_mi355x_chunk_table_v1", C.c_int32,
     [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("vckss_mi355x_prepare_host_v1", C.c_int32,
     [C.POINTER(PrepareRequestInterruptV3), C.POINTER(ColumnsV3), C.POINTER(C.c_uint64)]),
]


def library():
    """Load the built engine library; raises if it was never built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(_LIB_PATH):
            raise EngineError(11, f"{_LIB_PATH} is missing: run __graft_entry__.build() first")
        lib = C.CDLL(_LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def _check(code: int):
    if code != 0:
        raise EngineError(code, library().vckss_rust_engine_last_error().decode())


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def host_memory_bytes() -> int:
    """Physical memory of this host in bytes."""
    return os.sysconf("SC_PAGE_SIZE") * os.sysconf("SC_PHYS_PAGES")


def default_prepare_request(rows: int) -> PrepareRequestInterruptV3:
    req = PrepareRequestInterruptV3()
    _check(library().vckss_rust_engine_default_prepare_request_interrupt_v3(C.byref(req), C.sizeof(req)))
    req.options.rows = rows
    return req


def default_numerical_request(seed=8675309, probes=200, tolerance=None, fit_tolerance=1e-10,
                              maximum_iterations=10000, residual_replacement_interval=100):
    """A V2 numerical MCSE request on the CMG_FULL_V2 path (numerical_api.rs:199-235): the
    default solve request's V4 prefix as the point, JLA, full CMG."""
    base = default_solve_request().options
    r = NumericalRequestV2()
    C.memmove(C.byref(r.v1.point), C.byref(base), C.sizeof(SolveRequestV4))
    p = r.v1.point
    p.struct_size = C.sizeof(SolveRequestV4)
    p.algorithm = 2
    p.seed, p.probes, p.pcg_tolerance = seed, probes, fit_tolerance
    p.maximum_iterations, p.residual_replacement_interval = maximum_iterations, residual_replacement_interval
    v1 = r.v1
    v1.abi_version, v1.struct_size, v1.schema_version, v1.threads = 1, C.sizeof(r), 2, 1
    if tolerance is not None:
        p.pcg_tolerance = tolerance
        v1.tolerance_supplied = 1
    r.full_cmg = 1
    return r


def default_solve_request() -> SolveRequestInterruptV5:
    req = SolveRequestInterruptV5()
    _check(library().vckss_rust_engine_default_solve_request_interrupt_v5(C.byref(req), C.sizeof(req)))
    return req


class Engine:
    """One prepared context (generation) of the engine.

    Mirrors the reference call sequence prepare -> solve -> result -> release.
    """

    def __init__(self, worker, firm, deletion, outcome, frequency, target, *, upload=True,
                 poll=None, deletion_mode="match", memory_limit=None, controls=None):
        """deletion_mode "match" (the default) or "observation" (the dense
        exact route, and the JLA route without numerical MCSE).
        memory_limit is the whole-command host byte limit of the admission
        (default: the host's physical memory).  controls: rows x Q (Q <= 32)
        control columns; solves then take nuisance "joint" (the default) or
        "fixed_offset" on the JLA route under match deletion."""
        lib = library()
        self.deletion_mode = {"match": 1, "observation": 2}[deletion_mode]
        self.stayers_mode = 1  # movers; 2 (all) after augment_stayers
        cols = [_f64(worker), _f64(firm), _f64(deletion), _f64(outcome), _f64(frequency), _f64(target)]
        self._cols = cols
        rows = len(cols[0])
        self.controls = 0
        self._control_cols = []
        if controls is not None:
            z = np.asarray(controls, np.float64).reshape(rows, -1)
            self._control_cols = [np.ascontiguousarray(z[:, j]) for j in range(z.shape[1])]
            self.controls = len(self._control_cols)
        req = default_prepare_request(rows)
        req.options.cleanup_abandoned = 1
        req.options.caller_copy_bytes = rows * (6 + self.controls) * 8
        req.options.controls_count = self.controls
        req.options.memory_limit_bytes = memory_limit or host_memory_bytes()
        req.options.deletion_mode = self.deletion_mode
        self._poll = Poll(poll) if poll else None
        if self._poll:
            req.interrupt_poll = C.cast(self._poll, C.c_void_p)
            req.checkpoint_interval = 1
        c = ColumnsV3()
        c.struct_size = C.sizeof(c)
        c.rows = rows
        c.worker, c.firm, c.deletion, c.outcome, c.frequency, c.target_weight = [a.ctypes.data for a in cols]
        if self.controls:
            self._control_ptrs = (C.c_void_p * self.controls)(*[a.ctypes.data for a in self._control_cols])
            c.controls = C.cast(self._control_ptrs, C.c_void_p)
            c.controls_count = self.controls
        h = C.c_uint64()
        if upload:
            _check(lib.vckss_rust_engine_prepare_interrupt_v4(C.byref(req), C.byref(c), C.byref(h), 8))
        else:
            _check(lib.vckss_mi355x_prepare_host_v1(C.byref(req), C.byref(c), C.byref(h)))
        self.generation = h.value
        d = np.zeros(15, np.uint64)
        _check(lib.vckss_mi355x_dims_v1(self.generation, d.ctypes.data, 15))
        keys = ["input_rows", "rows", "physical", "workers", "firms", "cells", "units",
                "exact_strata", "strata", "hybrid_vertices", "hybrid_edges", "auxiliary", "levels",
                "terminal_dimension", "topology_checksum"]
        self.dims = {k: int(v) for k, v in zip(keys, d)}

    def augment_stayers(self, firm, worker, outcome, frequency, target):
        """Append eligible stayer rows (1-based dense retained-mover firm and
        stayer-worker indices); later solves take stayers_mode ALL."""
        cols = [_f64(firm), _f64(worker), _f64(outcome), _f64(frequency), _f64(target)]
        self._stayer_cols = cols
        rows = len(cols[0])
        r = StayerRequestInterruptV1()
        _check(library().vckss_rust_engine_default_stayer_augmentation_request_interrupt_v1(C.byref(r),
                                                                                             C.sizeof(r)))
        r.options.rows = rows
        r.options.caller_copy_bytes = rows * 5 * 8
        c = StayerColumnsV1()
        c.struct_size, c.rows = C.sizeof(c), rows
        if rows:
            c.firm, c.worker, c.outcome, c.frequency, c.target_weight = [a.ctypes.data for a in cols]
        _check(library().vckss_rust_engine_augment_stayers_interrupt_v1(self.generation, C.byref(r), C.byref(c)))
        self.stayers_mode = 2
        d = np.zeros(15, np.uint64)
        _check(library().vckss_mi355x_dims_v1(self.generation, d.ctypes.data, 15))
        self.dims.update({k: int(v) for k, v in zip(list(self.dims), d)})
        return self.stayer_receipt()

    def stayer_receipt(self):
        o = StayerReceiptV1()
        _check(library().vckss_rust_engine_stayer_augmentation_receipt_v1(self.generation, C.byref(o), C.sizeof(o)))
        return {name: getattr(o, name) for name, _ in o._fields_}

    def stayer_hybrid_result(self):
        """The exact stayer hybrid's result with its mover and stayer corrections."""
        o = StayerHybridResultV1()
        _check(library().vckss_rust_engine_stayer_hybrid_result_v1(self.generation, C.byref(o), C.sizeof(o)))
        out = {}
        for name, _ in o._fields_:
            v = getattr(o, name)
            out[name] = np.array([v.worker, v.firm, v.covariance, v.total]) if isinstance(v, _Vec4) else v
        return out

    def mask(self):
        m = np.zeros(self.dims["input_rows"], np.uint8)
        _check(library().vckss_rust_engine_retained_mask_v1(self.generation, m.ctypes.data, m.size))
        return m.astype(bool)

    def configure(self, precision=64, leverage_width=32, target_width=16):
        _check(library().vckss_mi355x_configure_v1(self.generation, precision, leverage_width,
                                                   target_width))

    def solve(self, seed=8675309, probes=200, tolerance=None, fit_tolerance=1e-10,
              maximum_iterations=10000, residual_replacement_interval=100, poll=None, checkpoint_interval=1,
              nuisance="joint", **overrides):
        """poll: optional callable returning 0 (continue) or 1 (user break),
        called at every `checkpoint_interval`-th checkpoint of the solve.
        nuisance: "joint" or "fixed_offset" (how prepared controls enter)."""
        req = default_solve_request()
        o = req.options
        o.algorithm = 2  # JLA
        o.nuisance_mode = {"joint": 1, "fixed_offset": 2}[nuisance]  # VCKSS_NUISANCE_*
        o.full_cmg_v2 = 1
        o.threads = 1
        o.deletion_mode = self.deletion_mode
        o.stayers_mode = self.stayers_mode
        o.seed = seed
        o.probes = probes
        o.pcg_tolerance = fit_tolerance
        o.maximum_iterations = maximum_iterations
        o.residual_replacement_interval = residual_replacement_interval
        if tolerance is not None:
            o.pcg_tolerance = tolerance
            o.tolerance_supplied = 1
        for k, v in overrides.items():
            setattr(o, k, v)
        cb = Poll(poll) if poll else None
        if cb:
            req.interrupt_poll = C.cast(cb, C.c_void_p)
            req.checkpoint_interval = checkpoint_interval
        _check(library().vckss_rust_engine_solve_interrupt_v5(self.generation, C.byref(req)))
        return self.result()

    def solve_exact(self, fit_tolerance=1e-10, exact_limit=None, algorithm=1, poll=None, **overrides):
        """Dense exact estimator (algorithm 1 = EXACT; 0 = AUTO routes here when
        W + F - 1 <= exact_limit), a V5 solve with full_cmg_v2 = 0: (result, receipt)."""
        kw = dict(full_cmg_v2=0)
        kw.update(overrides)
        if exact_limit is not None:
            kw["exact_estimator_limit"] = exact_limit
        res = self.solve(fit_tolerance=fit_tolerance, poll=poll, algorithm=algorithm, **kw)
        return res, self.exact_receipt()

    def graphs(self, enable=None):
        """Enable (True) or disable (False) HIP-graph iterations, or query
        (None); returns (captures, graph launches) so far."""
        c, l = C.c_uint64(), C.c_uint64()
        mode = 2 if enable is None else int(bool(enable))
        _check(library().vckss_mi355x_graphs_v1(self.generation, mode, C.byref(c), C.byref(l)))
        return c.value, l.value

    def exact_receipt(self):
        out = np.zeros(13)
        _check(library().vckss_mi355x_exact_receipt_v1(self.generation, out.ctypes.data, 13))
        keys = ("parameters", "deletion_units", "cells", "max_leverage", "deletion_rank_gap",
                "information_rcond", "inverse_relres", "fit_relres", "fit_tolerance",
                "firm_zero_sum_residual", "weighted_rss", "seconds_factor", "seconds_cells")
        d = dict(zip(keys, out.tolist()))
        for k in ("parameters", "deletion_units", "cells"):
            d[k] = int(d[k])
        return d

    def numerical_request(self, **kw):
        r = default_numerical_request(**kw)
        r.v1.point.deletion_mode = self.deletion_mode
        r.v1.point.stayers_mode = self.stayers_mode
        return r

    def solve_numerical(self, request=None, **kw):
        """Point estimate plus the numerical MCSE attachment; returns (result, numerical)."""
        r = request if request is not None else self.numerical_request(**kw)
        _check(library().vckss_rust_engine_solve_numerical_interrupt_v2(self.generation, C.byref(r), None, None, 0))
        return self.result(), self.numerical_result()

    def numerical_result(self):
        o = NumericalResultV1()
        _check(library().vckss_rust_engine_numerical_result_v1(self.generation, C.byref(o), C.sizeof(o)))
        out = {}
        for name, _ in o._fields_:
            v = getattr(o, name)
            out[name] = np.array(list(v)) if hasattr(v, "__len__") else v
        for k in ("conditional", "leverage", "raw", "usable"):
            out[k] = out[k].reshape(3, 3)
        out["score_means"] = out["score_means"].reshape(2, 3)
        return out

    def numerical_rhs(self):
        """(phase, probe, complete residual) of every certified replay RHS; the
        destination must hold them all (numerical_api.rs:541-608)."""
        cap = self.numerical_result()["certified_replay_rhs"]
        n = C.c_uint64()
        out = (NumericalRhsV1 * max(cap, 1))()
        _check(library().vckss_rust_engine_numerical_rhs_v1(self.generation, out, cap, C.byref(n)))
        return [(r.phase, r.probe, r.complete_residual) for r in out[:n.value]]

    def numerical_cmg_work(self):
        o = FullCmgModelReceiptV1()
        _check(library().vckss_rust_engine_numerical_cmg_work_v1(self.generation, C.byref(o), C.sizeof(o)))
        return {name: getattr(o, name) for name, _ in o._fields_}

    def result(self):
        r = ResultV1()
        _check(library().vckss_rust_engine_result_v1(self.generation, C.byref(r), C.sizeof(r)))
        return dict(plugin=r.plugin.as_list(), correction=r.correction.as_list(),
                    corrected=r.corrected.as_list(), numerical_mcse=r.numerical_mcse.as_list())

    def control_fit(self):
        """The last solve's control fit: q, fixed offset, the generalised
        reciprocal condition of the control information, the coefficients."""
        out = np.zeros(3 + max(self.controls, 1))
        _check(library().vckss_mi355x_controls_v1(self.generation, out.ctypes.data, out.size))
        q = int(out[0])
        return dict(controls=q, fixed_offset=bool(out[1]), rcond=float(out[2]), beta=out[3:3 + q].copy())

    def array(self, which: int, size: int):
        out = np.zeros(size, np.float64)
        _check(library().vckss_mi355x_array_v1(self.generation, which, out.ctypes.data, size))
        return out

    def arrays(self, probes):
        d = self.dims
        names = [("fitted_cell", 0, d["cells"]), ("projection_share", 1, d["units"]),
                 ("residual_share", 2, d["units"]), ("finite_bias", 3, d["units"]),
                 ("finite_variance", 4, d["units"]), ("residual_mass", 5, d["units"]),
                 ("deleted_mass", 6, d["units"]), ("cell_correction_weight", 7, d["cells"]),
                 ("target_draws", 8, 4 * probes), ("leverage_iterations", 9, probes),
                 ("target_iterations", 10, 2 * probes), ("leverage_complete_residual", 11, probes),
                 ("timing", 12, 4)]
        out = {n: self.array(w, s) for n, w, s in names}
        out["target_draws"] = out["target_draws"].reshape(probes, 4)
        return out

    def atoms(self, seed, domain, first_probe, columns):
        n = self.dims["units"] if domain == 0 else self.dims["strata"]
        out = np.zeros(n * columns, np.int64)
        _check(library().vckss_mi355x_atoms_v1(self.generation, seed, domain, first_probe, columns,
                                               out.ctypes.data, out.size))
        return out.reshape(columns, n)

    def level_vertices(self):
        """Vertex count of every CMG level of the prepared hierarchy."""
        return [int(v) for v in self.array(28, self.dims["levels"])]

    def spmv(self, level, x, columns):
        x = _f64(x)
        y = np.zeros_like(x)
        _check(library().vckss_mi355x_spmv_v1(self.generation, level, columns, x.ctypes.data,
                                              y.ctypes.data))
        return y

    def vcycle(self, r, columns, precision=64):
        r = _f64(r)
        z = np.zeros_like(r)
        _check(library().vckss_mi355x_vcycle_v1(self.generation, columns, precision, r.ctypes.data,
                                                z.ctypes.data))
        return z

    def pcg(self, rhs, columns, tolerance=1e-8, precision=64):
        rhs = _f64(rhs)
        x = np.zeros_like(rhs)
        it = np.zeros(columns, np.uint32)
        _check(library().vckss_mi355x_pcg_v1(self.generation, columns, tolerance, precision,
                                             rhs.ctypes.data, x.ctypes.data, it.ctypes.data))
        return x, it

    def kernel_timing(self, which):
        """HIP-event timing of a kernel in the last time_leverage run: 0 the
        finest A p (dominant), 1 the finest prolong + smooth."""
        ms, nb, n = C.c_double(), C.c_double(), C.c_uint64()
        _check(library().vckss_mi355x_kernel_timing_v1(self.generation, which, C.byref(ms), C.byref(nb),
                                                       C.byref(n)))
        return {"ms": ms.value, "bytes": nb.value, "launches": int(n.value)}

    def first_pass(self, mode=None):
        """Set the first-pass stop (0: the reference's; 1: gate-aware, the
        default) and return the previous mode; with None only query it."""
        prev = C.c_uint32()
        _check(library().vckss_mi355x_first_pass_v1(self.generation, 2 if mode is None else int(mode),
                                                    C.byref(prev)))
        return int(prev.value)

    def solve_columns(self, worker_rhs, firm_rhs, tolerance=1e-10, inner_ratio=1.0, precision=64,
                      maximum_iterations=10000, replacement_interval=100):
        """One CMG_FULL_V2 batch solve of host columns (workers x k, firms x k):
        returns worker and firm solutions, receipts per column and the
        refinement rounds / refined columns of the call."""
        wr = np.ascontiguousarray(np.asarray(worker_rhs, np.float64).T)
        fr = np.ascontiguousarray(np.asarray(firm_rhs, np.float64).T)
        k = wr.shape[0]
        wo = np.zeros_like(wr)
        fo = np.zeros_like(fr)
        rec = np.zeros((k, 5))
        ref = np.zeros(2, np.uint64)
        _check(library().vckss_mi355x_solve_columns_v1(self.generation, k, tolerance, inner_ratio, precision,
                                                       maximum_iterations, replacement_interval, wr.ctypes.data,
                                                       fr.ctypes.data, wo.ctypes.data, fo.ctypes.data,
                                                       rec.ctypes.data, ref.ctypes.data))
        receipts = {"iterations": rec[:, 0].astype(int), "restarts": rec[:, 1].astype(int),
                    "reduced": rec[:, 2], "complete": rec[:, 3], "rhs_norm": rec[:, 4]}
        return wo.T, fo.T, receipts, {"rounds": int(ref[0]), "refined": int(ref[1])}

    def time_leverage(self, width, precision, tolerance, batches):
        sec, ms, by = C.c_double(), C.c_double(), C.c_double()
        solves, iters, launches = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(library().vckss_mi355x_time_leverage_v1(
            self.generation, width, precision, tolerance, batches, C.byref(sec), C.byref(solves),
            C.byref(iters), C.byref(ms), C.byref(by), C.byref(launches)))
        return dict(seconds=sec.value, solves=solves.value, iterations=iters.value,
                    kernel_ms=ms.value, kernel_bytes=by.value, kernel_launches=launches.value)

    # ---- learned smoother weights (no counterpart in the reference)
    def smoother(self, mode, params=None):
        """mode 0: the reference cycle; 1: learned weights (built-in when params is None)."""
        if params is None:
            _check(library().vckss_mi355x_smoother_v1(self.generation, mode, None, 0))
        else:
            p = np.ascontiguousarray(params, np.float64)
            _check(library().vckss_mi355x_smoother_v1(self.generation, mode, p.ctypes.data, len(p)))

    def smoother_force(self, omega):
        """Constant smoother weight omega on every smoothed level, uncalibrated
        (host-forced fallback check)."""
        _check(library().vckss_mi355x_smoother_force_v1(self.generation, float(omega)))

    def smoother_status(self):
        """mode, learned weights in use (active), breakdown fallback, calibration
        iterations of the reference cycle and of the chosen configuration,
        tuned shape in use, and the visited levels."""
        out = np.zeros(10, np.uint32)
        _check(library().vckss_mi355x_smoother_status_v1(self.generation, out.ctypes.data, 10))
        n = C.c_uint32()
        _check(library().vckss_mi355x_cycle_shape_v1(self.generation, None, 0, C.byref(n)))
        shape = np.zeros(n.value, np.uint32)
        _check(library().vckss_mi355x_cycle_shape_v1(self.generation, shape.ctypes.data, n.value, C.byref(n)))
        return dict(mode=int(out[0]), active=bool(out[1]), breakdown=bool(out[2]),
                    plain_iterations=int(out[3]), learned_iterations=int(out[4]), shaped=bool(out[5]),
                    shape=[int(x) for x in shape],
                    calibration=[None if x == 0xFFFFFFFF else int(x) for x in out[6:10]])

    def smoother_omega(self, params, bf16=False):
        """omega of every smoothed level: the FP64 training forward, or with
        bf16 the solve path's BF16 inference."""
        fn = library().vckss_mi355x_smoother_infer_v1 if bf16 else library().vckss_mi355x_smoother_omega_v1
        p = np.ascontiguousarray(params, np.float64)
        written = C.c_uint64()
        _check(fn(self.generation, p.ctypes.data, len(p), None, 0, C.byref(written)))
        out = np.zeros(written.value)
        _check(fn(self.generation, p.ctypes.data, len(p), out.ctypes.data, len(out), C.byref(written)))
        return out

    def smoother_gradient(self, params, errors):
        p = np.ascontiguousarray(params, np.float64)
        e = np.ascontiguousarray(errors, np.float64)
        loss = C.c_double()
        grad = np.zeros(len(p))
        _check(library().vckss_mi355x_smoother_gradient_v1(self.generation, p.ctypes.data, len(p), e.ctypes.data,
                                                           e.shape[0], C.byref(loss), grad.ctypes.data))
        return loss.value, grad

    def smoother_train(self, state, step0, steps, lr, seed, passes=2):
        """Adam steps; `state` (3 x params) is updated in place; returns the losses."""
        assert state.dtype == np.float64 and state.flags.c_contiguous
        losses = np.zeros(max(steps, 1))
        _check(library().vckss_mi355x_smoother_train_v1(self.generation, state.ctypes.data, len(state), step0,
                                                        steps, lr, seed, passes, losses.ctypes.data))
        return losses[:steps]

    def set_chunk(self, chunk):
        _check(library().vckss_mi355x_set_chunk_v1(self.generation, chunk))

    def autotune(self, precision=32, tolerance=1e-8, batches=1):
        w, c, r = C.c_uint32(), C.c_uint32(), C.c_double()
        trials = np.zeros(3 * 9)
        _check(library().vckss_mi355x_autotune_v1(self.generation, precision, tolerance, batches,
                                                  C.byref(w), C.byref(c), C.byref(r),
                                                  trials.ctypes.data, 9))
        return dict(width=w.value, chunk=c.value, rate=r.value,
                    trials=[tuple(x) for x in trials.reshape(9, 3)])

    def counters(self):
        out = np.zeros(6, np.uint64)
        _check(library().vckss_mi355x_counters_v1(self.generation, out.ctypes.data, 6))
        keys = ["vcycles", "batches", "refine_rounds", "refined_columns", "iterations_launched"]
        return {k: int(v) for k, v in zip(keys, out)}

    def stream_copy(self, nbytes):
        ms = C.c_double()
        _check(library().vckss_mi355x_stream_copy_v1(self.generation, nbytes, C.byref(ms)))
        return ms.value

    # ---- receipts of the reference plugin ABI
    def preparation_receipt(self):
        return _out(library().vckss_rust_engine_preparation_receipt_v4, "VckssEnginePreparationReceiptV4",
                    self.generation)

    def memory_forecast(self):
        return _out(library().vckss_rust_engine_memory_forecast_v1, "VckssMemoryForecastV1", self.generation)

    def performance_receipt(self):
        return _out(library().vckss_rust_engine_performance_receipt_v1, "VckssEnginePerformanceReceiptV1",
                    self.generation)

    def detailed_receipt(self, version=7):
        return _out(getattr(library(), f"vckss_rust_engine_detailed_receipt_v{version}"),
                    f"VckssEngineDetailedReceiptV{version}", self.generation)

    def full_cmg_receipt(self):
        return _out(library().vckss_rust_engine_full_cmg_receipt_v1, "VckssFullCmgReceiptV1", self.generation)

    def rhs_receipts(self, rows):
        from . import abi_types

        out = (abi_types.STRUCTS["VckssEngineRhsReceiptV1"] * rows)()
        _check(library().vckss_rust_engine_rhs_receipts_v1(self.generation, out, rows))
        return list(out)

    def release(self):
        if self.generation:
            _check(library().vckss_rust_engine_release_v1(self.generation))
            self.generation = 0

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def _out(fn, struct, *lead):
    """Call an entry point that writes one header struct; return it."""
    from . import abi_types

    o = abi_types.new(struct)
    _check(fn(*lead, C.byref(o), C.sizeof(o)))
    return o


def backend_capabilities():
    return _out(library().vckss_rust_backend_capabilities_v1, "VckssBackendCapabilitiesV1")


def request_capability(request):
    """Request-capability receipt for a V1, V2 or V3 request structure."""
    version = {"VckssBackendRequestCapabilityRequestV1": 1, "VckssBackendRequestCapabilityRequestV2": 2,
               "VckssBackendRequestCapabilityRequestV3": 3}[type(request).__name__]
    fn = getattr(library(), f"vckss_rust_backend_request_capability_v{version}")
    from . import abi_types

    o = abi_types.new(f"VckssBackendRequestCapabilityReceiptV{version}")
    _check(fn(C.byref(request), C.byref(o), C.sizeof(o)))
    return o


def admit_prepare(request, probeorder_supplied=0):
    """Admission of a VckssEnginePrepareRequestV4 before its data (0 when admitted)."""
    return library().vckss_rust_engine_admit_prepare_v4(C.byref(request), probeorder_supplied)


def chunk_table(ptr, chunk=64, group=1):
    """Chunk table of segments with pointers `ptr` (host only): (desc as an
    (chunks, 4) array, multi ranges as (multi, 2)).  With group > 1, runs of
    short segments share a chunk tagged 0xffffff00 + count."""
    lib = library()
    ptr = np.ascontiguousarray(ptr, np.uint32)
    segments = max(len(ptr) - 1, 0)
    s = np.zeros(2, np.uint64)
    _check(lib.vckss_mi355x_chunk_table_v1(ptr.ctypes.data, segments, chunk, group, s.ctypes.data, None, None))
    desc = np.zeros(4 * int(s[0]), np.uint32)
    multi = np.zeros(2 * int(s[1]), np.uint32)
    _check(lib.vckss_mi355x_chunk_table_v1(ptr.ctypes.data, segments, chunk, group, s.ctypes.data,
                                           desc.ctypes.data if desc.size else None,
                                           multi.ctypes.data if multi.size else None))
    return desc.reshape(-1, 4), multi.reshape(-1, 2)


def hierarchy_layout(n, u, v, w, chunk=64, direct_threshold=700, max_levels=128):
    """Device layout of every level of the host CMG hierarchy of a graph."""
    lib = library()
    u = np.ascontiguousarray(u, np.uint32)
    v = np.ascontiguousarray(v, np.uint32)
    w = np.ascontiguousarray(w, np.float64)
    h = C.c_uint64()
    _check(lib.vckss_mi355x_hierarchy_build_v1(n, len(u), u.ctypes.data, v.ctypes.data,
                                               w.ctypes.data, direct_threshold, max_levels, C.byref(h)))
    try:
        nl, fnnz = C.c_uint32(), C.c_uint64()
        _check(lib.vckss_mi355x_hierarchy_levels_v1(h.value, C.byref(nl), C.byref(fnnz)))
        out = []
        for lv in range(nl.value):
            s = np.zeros(8, np.uint64)
            _check(lib.vckss_mi355x_layout_v1(h.value, lv, chunk, s.ctypes.data, *([None] * 12)))
            n_, nnz, nc, rch, rmu, ach, amu, nt = [int(x) for x in s]
            a = dict(perm=np.zeros(n_, np.uint32), row_ptr=np.zeros(n_ + 1, np.uint32),
                     col=np.zeros(nnz, np.uint32), val=np.zeros(nnz), inv=np.zeros(n_),
                     lab=np.zeros(n_ if nc else 0, np.uint32),
                     agg_ptr=np.zeros(nc + 1 if nc else 0, np.uint32),
                     row_chunks=np.zeros(4 * rch, np.uint32), row_multi=np.zeros(2 * rmu, np.uint32),
                     agg_chunks=np.zeros(4 * ach, np.uint32), agg_multi=np.zeros(2 * amu, np.uint32),
                     terminal_inverse=np.zeros(nt * nt))
            order = ["perm", "row_ptr", "col", "val", "inv", "lab", "agg_ptr", "row_chunks",
                     "row_multi", "agg_chunks", "agg_multi", "terminal_inverse"]
            _check(lib.vckss_mi355x_layout_v1(h.value, lv, chunk, s.ctypes.data,
                                              *[a[k].ctypes.data if a[k].size else None for k in order]))
            a["row_chunks"] = a["row_chunks"].reshape(-1, 4)
            a["agg_chunks"] = a["agg_chunks"].reshape(-1, 4)
            a["row_multi"] = a["row_multi"].reshape(-1, 2)
            a["agg_multi"] = a["agg_multi"].reshape(-1, 2)
            a["terminal_inverse"] = a["terminal_inverse"].reshape(nt, nt)
            a["n"], a["coarse"] = n_, nc
            out.append(a)
        return out
    finally:
        lib.vckss_mi355x_hierarchy_free_v1(h.value)


def builtin_smoother():
    """The library's built-in trained smoother parameters (None if it carries none)."""
    n = C.c_uint32()
    _check(library().vckss_mi355x_smoother_params_v1(None, 0, C.byref(n)))
    if not n.value:
        return None
    out = np.zeros(n.value)
    _check(library().vckss_mi355x_smoother_params_v1(out.ctypes.data, n.value, C.byref(n)))
    return out


def build_hierarchy(n, u, v, w, direct_threshold=700, max_levels=128):
    """Host CMG hierarchy of an arbitrary graph as a list of level dicts."""
    lib = library()
    u = np.ascontiguousarray(u, np.uint32)
    v = np.ascontiguousarray(v, np.uint32)
    w = np.ascontiguousarray(w, np.float64)
    h = C.c_uint64()
    _check(lib.vckss_mi355x_hierarchy_build_v1(n, len(u), u.ctypes.data, v.ctypes.data,
                                               w.ctypes.data, direct_threshold, max_levels, C.byref(h)))
    try:
        nl, fnnz = C.c_uint32(), C.c_uint64()
        _check(lib.vckss_mi355x_hierarchy_levels_v1(h.value, C.byref(nl), C.byref(fnnz)))
        levels = []
        for i in range(nl.value):
            info = np.zeros(6, np.uint64)
            _check(lib.vckss_mi355x_hierarchy_level_v1(h.value, i, info.ctypes.data, None, None, None,
                                                       None, None))
            n_i, m_i = int(info[0]), int(info[1])
            lab = np.zeros(n_i, np.uint32)
            eu = np.zeros(m_i, np.uint32)
            ev = np.zeros(m_i, np.uint32)
            ew = np.zeros(m_i, np.float64)
            dg = np.zeros(n_i, np.float64)
            _check(lib.vckss_mi355x_hierarchy_level_v1(h.value, i, info.ctypes.data, lab.ctypes.data,
                                                       eu.ctypes.data, ev.ctypes.data, ew.ctypes.data,
                                                       dg.ctypes.data))
            levels.append(dict(n=n_i, m=m_i, nnz=int(info[2]), repeat=int(info[3]),
                               terminal=int(info[4]), coarse=int(info[5]),
                               labels=lab if int(info[5]) else None, u=eu, v=ev, w=ew, diag=dg))
        return levels, int(fnnz.value)
    finally:
        lib.vckss_mi355x_hierarchy_free_v1(h.value)


class SampleBatch:
    """Independent samples (rows tagged by `sample`) prepared together and
    estimated concurrently on the GPU (vckss_mi355x_batch_*_v1)."""

    def __init__(self, sample, worker, firm, deletion, outcome, frequency, target, *,
                 deletion_mode="match", threads=8):
        cols = [_f64(sample), _f64(worker), _f64(firm), _f64(deletion), _f64(outcome), _f64(frequency),
                _f64(target)]
        self._cols = cols
        h, n = C.c_uint64(), C.c_uint32()
        _check(library().vckss_mi355x_batch_prepare_v1(
            len(cols[0]), *[c.ctypes.data for c in cols], {"match": 1, "observation": 2}[deletion_mode],
            threads, C.byref(h), C.byref(n)))
        self.handle, self.samples = h.value, n.value

    def solve(self, seed=8675309, probes=128, tolerance=1e-8, fit_tolerance=1e-8, precision=32, streams=4,
              lockstep=False):
        """Every sample's JLA estimate: up to `streams` samples at once, each on
        its own stream, or with `lockstep` up to `streams` samples whose PCG
        passes run as single launches over all of them."""
        _check(library().vckss_mi355x_batch_solve_v2(self.handle, seed, probes, tolerance, fit_tolerance,
                                                     precision, streams, 1 if lockstep else 0))
        return [self.result(i) for i in range(self.samples)]

    def result(self, i):
        out = np.zeros(24)
        _check(library().vckss_mi355x_batch_result_v1(self.handle, i, out.ctypes.data, out.size))
        keys = ("sample", "rows", "workers", "firms", "units", "seconds", "iterations", "max_complete_residual")
        r = dict(plugin=out[0:4].copy(), correction=out[4:8].copy(), corrected=out[8:12].copy(),
                 mcse=out[12:16].copy())
        r.update({k: float(v) for k, v in zip(keys, out[16:24])})
        return r

    def release(self):
        if getattr(self, "handle", 0):
            library().vckss_mi355x_batch_release_v1(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
