/* SPDX-License-Identifier: MIT
 *
 * C ABI of the MI355X leave-out engine (libvckss_mi355x.so).
 *
 * The engine is a drop-in for the compressed no-control match-deletion
 * improved-JLA path of the fevc Rust backend (route CMG_FULL_V2).  Every
 * declaration below keeps the name, argument meaning, struct layout and status
 * codes of the reference interface it replaces, cited as
 * rust/stata_backend/include/vckss_rust.h:<line>; the implementation lives in
 * fevc_amd/csrc/vckss_abi.cpp.  Requests outside this path are refused with
 * VCKSS_STATUS_UNSUPPORTED (11), as the reference refuses unsupported features.
 */
#ifndef VCKSS_MI355X_H
#define VCKSS_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Constants: vckss_rust.h:16-146 */
#define VCKSS_RUST_ABI_VERSION_V1 1u
#define VCKSS_DELETION_MATCH 1u
#define VCKSS_DELETION_OBSERVATION 2u
#define VCKSS_RNG_NONE 0u
#define VCKSS_RNG_COUNTER_V1 1u
#define VCKSS_ROUTE_AUTO 0u
#define VCKSS_ROUTE_EXACT 1u
#define VCKSS_ROUTE_DIAGONAL_PCG 2u
#define VCKSS_ROUTE_CMG_PCG 3u
#define VCKSS_INTERRUPT_CONTINUE 0
#define VCKSS_INTERRUPT_USER_BREAK 1
#define VCKSS_ALGORITHM_AUTO 0u
#define VCKSS_ALGORITHM_EXACT 1u
#define VCKSS_ALGORITHM_JLA 2u
#define VCKSS_NUISANCE_JOINT 1u
#define VCKSS_NUISANCE_FIXED_OFFSET 2u
#define VCKSS_ENGINE_AUTO_OR_UNSPECIFIED 0u
#define VCKSS_ENGINE_COMPRESSED 1u
#define VCKSS_ENGINE_GENERIC 2u
#define VCKSS_BATCH_MODE_AUTO 0u
#define VCKSS_BATCH_MODE_EXPLICIT 1u
#define VCKSS_STAYERS_MOVERS 1u
#define VCKSS_STAYERS_ALL 2u
#define VCKSS_TARGET_WEIGHT_FREQUENCY_DEFAULT 0u
#define VCKSS_TARGET_WEIGHT_STORED_ROW_EXPLICIT 1u
#define VCKSS_DELETION_SOURCE_CELL_DEFAULT 1u
#define VCKSS_DELETION_SOURCE_MATCH_ID_EXPLICIT 2u
#define VCKSS_DELETION_SOURCE_OBSERVATION_ROW 3u
#define VCKSS_REQUEST_CAPABILITY_SCHEMA_V1 1u
#define VCKSS_REQUEST_CAPABILITY_SCHEMA_V2 2u
#define VCKSS_REQUEST_CAPABILITY_SCHEMA_V3 3u
#define VCKSS_REQUEST_PROFILE_NONE 0u
#define VCKSS_REQUEST_PROFILE_EXACT_V1 1u
#define VCKSS_REQUEST_PROFILE_JLA_COUNTER_V1 2u
#define VCKSS_REQUEST_PROFILE_JLA_GENERIC_COUNTER_V1 3u
#define VCKSS_REQUEST_PROFILE_PLANNED_V1 4u
#define VCKSS_REQUEST_FREQUENCY_UNIT 0u
#define VCKSS_REQUEST_FREQUENCY_LITERAL 1u
#define VCKSS_ROUTE_NOT_APPLICABLE 4u
#define VCKSS_ENGINE_NOT_APPLICABLE 3u
#define VCKSS_BATCH_MODE_INDEPENDENT 2u
#define VCKSS_BATCH_MODE_NOT_APPLICABLE 3u
#define VCKSS_CORE_MATCH_GRAPH_READY (UINT64_C(1) << 0)
#define VCKSS_CORE_EXACT_READY (UINT64_C(1) << 1)
#define VCKSS_CORE_DIAGONAL_PCG_READY (UINT64_C(1) << 2)
#define VCKSS_CORE_BATCHED_PCG_READY (UINT64_C(1) << 3)
#define VCKSS_CORE_CMG_GRAPH_READY (UINT64_C(1) << 4)
#define VCKSS_CORE_SOLVER_ROUTER_READY (UINT64_C(1) << 5)
#define VCKSS_CORE_COUNTER_RNG_READY (UINT64_C(1) << 6)
#define VCKSS_CORE_JLA_PLAN_READY (UINT64_C(1) << 7)
#define VCKSS_CORE_FULL_CMG_V2_READY (UINT64_C(1) << 8)
#define VCKSS_CORE_FULL_CMG_POINT_ROUTING_V1_READY (UINT64_C(1) << 10)
#define VCKSS_CORE_FULL_CMG_MODEL_V1_READY (UINT64_C(1) << 11)
#define VCKSS_CORE_DELETION_UNIT_MOVERS_V1_READY (UINT64_C(1) << 15)
#define VCKSS_SUPPORT_EXACT (UINT64_C(1) << 0)
#define VCKSS_SUPPORT_JLA (UINT64_C(1) << 1)
#define VCKSS_SUPPORT_MATCH_DELETION (UINT64_C(1) << 2)
#define VCKSS_SUPPORT_OBSERVATION_DELETION (UINT64_C(1) << 3)
#define VCKSS_SUPPORT_CONTROLS (UINT64_C(1) << 4)
#define VCKSS_SUPPORT_DIAGONAL (UINT64_C(1) << 5)
#define VCKSS_SUPPORT_CMG (UINT64_C(1) << 6)
/* request-capability reason codes: vckss_rust.h:89-124 */
#define VCKSS_REQUEST_REASON_SUPPORTED 0u
#define VCKSS_REQUEST_REASON_UNKNOWN_SCHEMA 1u
#define VCKSS_REQUEST_REASON_UNKNOWN_ALGORITHM 2u
#define VCKSS_REQUEST_REASON_ALGORITHM_AUTO_UNRESOLVED 3u
#define VCKSS_REQUEST_REASON_UNKNOWN_DELETION 4u
#define VCKSS_REQUEST_REASON_UNKNOWN_NUISANCE 5u
#define VCKSS_REQUEST_REASON_UNKNOWN_SOLVER_ROUTE 6u
#define VCKSS_REQUEST_REASON_UNKNOWN_RNG_CONTRACT 7u
#define VCKSS_REQUEST_REASON_UNKNOWN_FREQUENCY_USE 8u
#define VCKSS_REQUEST_REASON_CONTROLS_LIMIT 9u
#define VCKSS_REQUEST_REASON_EXACT_RNG 10u
#define VCKSS_REQUEST_REASON_EXACT_SOLVER_ROUTE 11u
#define VCKSS_REQUEST_REASON_JLA_DELETION 12u
#define VCKSS_REQUEST_REASON_JLA_NUISANCE 13u
#define VCKSS_REQUEST_REASON_JLA_CONTROLS 14u
#define VCKSS_REQUEST_REASON_JLA_RNG 15u
#define VCKSS_REQUEST_REASON_UNKNOWN_ENGINE 16u
#define VCKSS_REQUEST_REASON_EXACT_ENGINE 17u
#define VCKSS_REQUEST_REASON_JLA_ENGINE_AUTO_UNRESOLVED 18u
#define VCKSS_REQUEST_REASON_JLA_GENERIC_SOLVER_ROUTE 19u
#define VCKSS_REQUEST_REASON_UNKNOWN_BATCH_MODE 20u
#define VCKSS_REQUEST_REASON_UNKNOWN_STAYERS_MODE 21u
#define VCKSS_REQUEST_REASON_UNKNOWN_TARGET_WEIGHT_MODE 22u
#define VCKSS_REQUEST_REASON_UNKNOWN_DELETION_UNIT_SOURCE 23u
#define VCKSS_REQUEST_REASON_PROBEORDER_UNSUPPORTED 24u
#define VCKSS_REQUEST_REASON_WALLSECONDS_UNSUPPORTED 25u
#define VCKSS_REQUEST_REASON_PHYSICAL_LIMIT 26u
#define VCKSS_REQUEST_REASON_DELETION_UNIT_SOURCE_MISMATCH 27u
#define VCKSS_REQUEST_REASON_BATCH_MODE_UNSUPPORTED 28u
#define VCKSS_REQUEST_REASON_STAYERS_MODE_UNSUPPORTED 29u
#define VCKSS_REQUEST_REASON_BATCH_SUMMARY_MISMATCH 30u
#define VCKSS_REQUEST_REASON_WALLSECONDS_VALUE 31u
#define VCKSS_REQUEST_REASON_FALLBACK_ROUTE_MISMATCH 32u
#define VCKSS_REQUEST_REASON_AUTO_RNG 33u
#define VCKSS_REQUEST_REASON_AUTO_ENGINE_COMPRESSED 34u
#define VCKSS_REQUEST_REASON_COMPRESSED_SCIENTIFIC_INELIGIBILITY 35u
#define VCKSS_PLAN_APPLICABILITY_NONE 0u
#define VCKSS_PLAN_APPLICABILITY_EXACT 1u
#define VCKSS_PLAN_APPLICABILITY_COMPRESSED 2u
#define VCKSS_PLAN_APPLICABILITY_GENERIC 3u
#define VCKSS_BATCH_SELECTION_NOT_APPLICABLE 0u
#define VCKSS_BATCH_SELECTION_AUTO 1u
#define VCKSS_BATCH_SELECTION_EXPLICIT 2u
/* generic execution modes of the V6/V7 solve requests
 * (rust/crates/vckss-plugin/src/generic_execution_api.rs:8-10) */
#define VCKSS_GENERIC_EXECUTION_DIAGONAL_QUEUE 1u
#define VCKSS_GENERIC_EXECUTION_DIRECT_ATTACHMENTS 2u
#define VCKSS_GENERIC_EXECUTION_RESOLVED_AUTO 3u

/* Status codes: rust/crates/vckss-core/src/error.rs:8-47 */
#define VCKSS_STATUS_OK 0
#define VCKSS_STATUS_USER_BREAK 1
#define VCKSS_STATUS_ABI_MISMATCH 10
#define VCKSS_STATUS_UNSUPPORTED 11
#define VCKSS_STATUS_INVALID_INPUT 20
#define VCKSS_STATUS_STALE_CONTEXT 80
#define VCKSS_STATUS_PANIC 99

typedef int32_t (*VckssInterruptPollV1)(void *context); /* vckss_rust.h:292 */

/* vckss_rust.h:269-296 */
typedef struct VckssEnginePrepareRequestV2 {
    uint32_t abi_version;
    uint32_t struct_size;
    uint64_t rows;
    uint32_t cleanup_abandoned;
    uint32_t reserved;
    uint64_t memory_limit_bytes;
    uint64_t caller_copy_bytes;
} VckssEnginePrepareRequestV2;

typedef struct VckssEnginePrepareRequestV3 {
    VckssEnginePrepareRequestV2 v2;
    uint32_t deletion_mode;
    uint32_t controls_count;
    uint64_t reserved_2;
} VckssEnginePrepareRequestV3;

typedef struct VckssEnginePrepareRequestV4 {
    VckssEnginePrepareRequestV3 v3;
    uint32_t implicit_match;
    uint32_t reserved_3;
} VckssEnginePrepareRequestV4;

/* vckss_rust.h:335-341 */
typedef struct VckssEnginePrepareRequestInterruptV3 {
    VckssEnginePrepareRequestV4 options;
    VckssInterruptPollV1 interrupt_poll;
    void *interrupt_context;
    uint32_t checkpoint_interval;
    uint32_t reserved;
} VckssEnginePrepareRequestInterruptV3;

/* vckss_rust.h:343-367: identifiers are exact binary64 integers */
typedef struct VckssEngineColumnsV1 {
    uint32_t struct_size;
    uint32_t reserved;
    uint64_t rows;
    const double *worker;
    const double *firm;
    const double *deletion;
    const double *outcome;
    const double *frequency;
    const double *target_weight;
} VckssEngineColumnsV1;

typedef struct VckssEngineColumnsV2 {
    VckssEngineColumnsV1 v1;
    const double *const *controls;
    uint32_t controls_count;
    uint32_t reserved_2;
} VckssEngineColumnsV2;

typedef struct VckssEngineColumnsV3 {
    VckssEngineColumnsV2 v2;
    const double *probe_order;
    uint32_t probeorder_supplied;
    uint32_t reserved_3;
} VckssEngineColumnsV3;

/* vckss_rust.h:501-575 */
typedef struct VckssEngineSolveRequestV1 {
    uint32_t abi_version;
    uint32_t struct_size;
    uint64_t seed;
    uint32_t probes;
    uint32_t leverage_batch_width;
    uint32_t target_batch_width;
    uint32_t deletion_mode;
    uint32_t rng_contract;
    uint32_t solver_route;
    uint32_t allow_automatic_cmg_setup_fallback;
    uint64_t exact_dimension_limit;
    uint64_t cmg_minimum_dimension;
    double pcg_tolerance;
    uint32_t maximum_iterations;
    uint32_t residual_replacement_interval;
    double rank_tolerance;
    double block_tolerance;
    uint64_t cmg_terminal_vertices;
    uint64_t cmg_dense_vertex_cap;
    uint64_t cmg_maximum_levels;
    uint64_t cmg_aggregate_cap;
    double cmg_minimum_reduction;
    double cmg_jacobi_weight;
    uint32_t cmg_pre_sweeps;
    uint32_t cmg_post_sweeps;
    double cmg_maximum_edge_complexity;
    double cmg_maximum_vertex_complexity;
    uint64_t cmg_memory_limit_bytes;
} VckssEngineSolveRequestV1;

typedef struct VckssEngineSolveRequestV2 {
    VckssEngineSolveRequestV1 v1;
    uint32_t algorithm;
    uint32_t nuisance_mode;
    uint64_t exact_estimator_limit;
    uint64_t blocksize_limit;
} VckssEngineSolveRequestV2;

typedef struct VckssEngineSolveRequestV3 {
    VckssEngineSolveRequestV2 v2;
    uint32_t engine;
    uint32_t batch_mode;
    uint32_t stayers_mode;
    uint32_t target_weight_mode;
    uint32_t deletion_unit_source;
    uint32_t probeorder_supplied;
    uint32_t wallseconds_supplied;
    uint32_t capability_schema;
    uint32_t capability_profile;
    uint32_t frequency_use;
    uint64_t physical_limit;
    uint64_t request_signature;
    uint64_t reserved_3;
} VckssEngineSolveRequestV3;

typedef struct VckssEngineSolveRequestV4 {
    VckssEngineSolveRequestV3 v3;
    uint32_t leverage_batch_mode;
    uint32_t target_batch_mode;
    double wallseconds;
    uint64_t reserved_4;
} VckssEngineSolveRequestV4;

typedef struct VckssEngineSolveRequestV5 {
    VckssEngineSolveRequestV4 v4;
    uint32_t threads;
    uint32_t tolerance_supplied;
    uint32_t full_cmg_v2;
    uint32_t reserved_5;
} VckssEngineSolveRequestV5;

/* vckss_rust.h:630-636 */
typedef struct VckssEngineSolveRequestInterruptV5 {
    VckssEngineSolveRequestV5 options;
    VckssInterruptPollV1 interrupt_poll;
    void *interrupt_context;
    uint32_t checkpoint_interval;
    uint32_t reserved;
} VckssEngineSolveRequestInterruptV5;

/* vckss_rust.h:577-588: V6 selects an executor for the generic engine and a
 * positive permitted thread count; V7 marks automatic component batches. */
typedef struct VckssEngineSolveRequestV6 {
    VckssEngineSolveRequestV4 v4;
    uint32_t execution_mode;
    uint32_t threads;
    uint64_t reserved_6;
} VckssEngineSolveRequestV6;

typedef struct VckssEngineSolveRequestV7 {
    VckssEngineSolveRequestV6 v6;
    uint32_t component_batch_mode;
    uint32_t reserved_7;
} VckssEngineSolveRequestV7;

/* vckss_rust.h:638-652 */
typedef struct VckssEngineSolveRequestInterruptV6 {
    VckssEngineSolveRequestV6 options;
    VckssInterruptPollV1 interrupt_poll;
    void *interrupt_context;
    uint32_t checkpoint_interval;
    uint32_t reserved;
} VckssEngineSolveRequestInterruptV6;

typedef struct VckssEngineSolveRequestInterruptV7 {
    VckssEngineSolveRequestV7 options;
    VckssInterruptPollV1 interrupt_poll;
    void *interrupt_context;
    uint32_t checkpoint_interval;
    uint32_t reserved;
} VckssEngineSolveRequestInterruptV7;

/* vckss_rust.h:168-260: backend capabilities and request-capability receipts */
typedef struct VckssBackendCapabilitiesV1 {
    uint32_t struct_size;
    uint32_t abi_version;
    uint64_t core_ready_flags;
    uint64_t support_flags;
    uint32_t deterministic_parallelism;
    uint32_t reserved;
} VckssBackendCapabilitiesV1;

typedef struct VckssBackendRequestCapabilityRequestV1 {
    uint32_t abi_version;
    uint32_t struct_size;
    uint32_t request_schema;
    uint32_t algorithm;
    uint32_t deletion_mode;
    uint32_t nuisance_mode;
    uint32_t solver_route;
    uint32_t rng_contract;
    uint32_t controls_count;
    uint32_t frequency_use;
    uint64_t reserved;
} VckssBackendRequestCapabilityRequestV1;

typedef struct VckssBackendRequestCapabilityReceiptV1 {
    uint32_t struct_size;
    uint32_t abi_version;
    uint32_t request_schema;
    uint32_t supported;
    uint32_t reason_code;
    uint32_t profile_code;
    uint32_t algorithm;
    uint32_t deletion_mode;
    uint32_t nuisance_mode;
    uint32_t solver_route;
    uint32_t rng_contract;
    uint32_t controls_count;
    uint32_t frequency_use;
    uint32_t reserved;
    uint64_t request_signature;
} VckssBackendRequestCapabilityReceiptV1;

typedef struct VckssBackendRequestCapabilityRequestV2 {
    VckssBackendRequestCapabilityRequestV1 v1;
    uint32_t engine;
    uint32_t batch_mode;
    uint32_t stayers_mode;
    uint32_t target_weight_mode;
    uint32_t deletion_unit_source;
    uint32_t probeorder_supplied;
    uint32_t wallseconds_supplied;
    uint32_t reserved_2;
    uint64_t physical_limit;
} VckssBackendRequestCapabilityRequestV2;

typedef struct VckssBackendRequestCapabilityReceiptV2 {
    VckssBackendRequestCapabilityReceiptV1 v1;
    uint32_t engine;
    uint32_t batch_mode;
    uint32_t stayers_mode;
    uint32_t target_weight_mode;
    uint32_t deletion_unit_source;
    uint32_t probeorder_supplied;
    uint32_t wallseconds_supplied;
    uint32_t reserved_2;
    uint64_t physical_limit;
} VckssBackendRequestCapabilityReceiptV2;

typedef struct VckssBackendRequestCapabilityRequestV3 {
    VckssBackendRequestCapabilityRequestV2 v2;
    uint32_t leverage_batch_mode;
    uint32_t target_batch_mode;
    uint32_t allow_automatic_cmg_setup_fallback;
    uint32_t reserved_3;
    double wallseconds;
    uint64_t reserved_4;
} VckssBackendRequestCapabilityRequestV3;

typedef struct VckssBackendRequestCapabilityReceiptV3 {
    VckssBackendRequestCapabilityReceiptV2 v2;
    uint32_t leverage_batch_mode;
    uint32_t target_batch_mode;
    uint32_t allow_automatic_cmg_setup_fallback;
    uint32_t reserved_3;
    double wallseconds;
    uint32_t algorithm_resolution_deferred;
    uint32_t engine_resolution_deferred;
    uint32_t route_resolution_deferred;
    uint32_t leverage_batch_resolution_deferred;
    uint32_t target_batch_resolution_deferred;
    uint32_t wall_advisory_only;
    uint64_t reserved_4;
} VckssBackendRequestCapabilityReceiptV3;

/* vckss_rust.h:724-768: preparation receipt (V2 fields with graph-selection
 * counts, V3 target mass, V4 controls and deletion mode) */
typedef struct VckssEnginePreparationReceiptV2 {
    uint32_t struct_size;
    uint32_t reserved;
    uint64_t generation;
    uint64_t input_rows;
    uint64_t retained_rows;
    uint64_t workers;
    uint64_t firms;
    uint64_t cells;
    uint64_t deletion_units;
    uint64_t target_strata;
    uint64_t memory_limit_bytes;
    uint64_t caller_copy_bytes;
    uint64_t preparation_peak_forecast_bytes;
    uint64_t prepared_resident_bytes;
    uint64_t graph_input_rows;
    uint64_t graph_retained_rows;
    uint64_t graph_input_physical_mass;
    uint64_t graph_retained_physical_mass;
    uint64_t graph_initial_components;
    uint64_t graph_maximum_components;
    uint64_t graph_initial_component_rows;
    uint64_t graph_mover_input_rows;
    uint64_t graph_initial_deletion_edges;
    uint64_t graph_retained_deletion_edges;
    uint64_t graph_insufficient_workers_removed;
    uint64_t graph_articulation_workers_removed;
    uint64_t graph_bridge_units_removed;
    uint64_t graph_bridge_rows_removed;
    uint64_t graph_degree_iterations;
    uint64_t graph_articulation_iterations;
    uint64_t graph_bridge_iterations;
    uint64_t graph_fixed_point_iterations;
} VckssEnginePreparationReceiptV2;

typedef struct VckssEnginePreparationReceiptV3 {
    VckssEnginePreparationReceiptV2 v2;
    double target_weight_sum;
} VckssEnginePreparationReceiptV3;

typedef struct VckssEnginePreparationReceiptV4 {
    VckssEnginePreparationReceiptV3 v3;
    uint32_t controls_count;
    uint32_t deletion_mode;
} VckssEnginePreparationReceiptV4;

/* vckss_rust.h:2322-2328 */
typedef struct VckssMemoryForecastV1 {
    uint32_t struct_size;
    uint32_t schema_version;
    uint64_t expected_peak_bytes;
    uint64_t admission_peak_bytes;
    uint64_t conditional_reserve_bytes;
} VckssMemoryForecastV1;

/* vckss_rust.h:1131-1271: detailed receipts V2-V6, each extending the last */
typedef struct VckssEngineDetailedReceiptV2 {
    uint32_t struct_size;
    uint32_t reserved;
    uint64_t generation;
    uint64_t seed;
    uint32_t probes_requested;
    uint32_t leverage_probes_accepted;
    uint32_t target_probes_accepted;
    uint32_t solver_requested;
    uint32_t solver_selected;
    uint32_t solver_fallback;
    int32_t solver_fallback_error;
    uint64_t solver_dimension;
    uint64_t leverage_batch_width;
    uint64_t target_batch_width;
    double rank_tolerance;
    double block_tolerance;
    double full_residual_tolerance;
    uint32_t full_fit_route;
    uint32_t full_fit_iterations;
    double full_fit_reduced_residual;
    double full_fit_complete_residual;
    uint32_t full_fit_zero_rhs;
    uint32_t reserved_1;
    uint64_t leverage_rhs_count;
    uint64_t target_rhs_count;
    double max_reduced_residual;
    double max_complete_residual;
    double max_leverage;
    double max_reciprocal_residual;
    double accounting_residual;
    uint64_t topology_checksum;
    uint64_t cmg_levels;
    uint64_t cmg_fine_vertices;
    uint64_t cmg_fine_edges;
    uint64_t cmg_terminal_vertices;
    double cmg_edge_complexity;
    double cmg_vertex_complexity;
    uint64_t cmg_structural_bytes;
    uint64_t cmg_workspace_bytes;
    uint64_t cmg_dense_factor_bytes;
    double full_fit_weighted_rss;
    uint64_t memory_limit_bytes;
    uint64_t caller_copy_bytes;
    uint64_t preparation_peak_forecast_bytes;
    uint64_t prepared_resident_bytes;
    uint64_t solver_setup_forecast_bytes;
    uint64_t leverage_phase_forecast_bytes;
    uint64_t target_phase_forecast_bytes;
    uint64_t result_forecast_bytes;
    uint64_t solve_peak_forecast_bytes;
    uint64_t command_peak_forecast_bytes;
} VckssEngineDetailedReceiptV2;

typedef struct VckssEngineDetailedReceiptV3 {
    VckssEngineDetailedReceiptV2 v2;
    uint32_t rng_contract;
    uint32_t reserved_2;
    uint64_t rhs_receipt_rows;
    uint64_t caller_result_copy_bytes;
} VckssEngineDetailedReceiptV3;

typedef struct VckssEngineDetailedReceiptV4 {
    VckssEngineDetailedReceiptV3 v3;
    uint32_t algorithm_requested;
    uint32_t algorithm_selected;
    uint32_t deletion_mode;
    uint32_t nuisance_mode;
    uint64_t parameters;
    uint64_t full_parameters;
    uint64_t correction_parameters;
    double information_rcond;
    double inverse_relres;
    uint64_t exact_peak_forecast_bytes;
} VckssEngineDetailedReceiptV4;

typedef struct VckssEngineDetailedReceiptV5 {
    VckssEngineDetailedReceiptV4 v4;
    uint64_t applicability_flags;
    double working_fit_complete_residual;
    double inverse_sqrt_relres;
    double maker_relres;
    double control_basis_relres;
    double control_basis_forward_error;
    double deletion_rank_gap;
    double firm_zero_sum_residual;
    uint64_t fit_peak_forecast_bytes;
    uint64_t correction_peak_forecast_bytes;
    double actual_accounting_residual;
} VckssEngineDetailedReceiptV5;

typedef struct VckssEngineDetailedReceiptV6 {
    VckssEngineDetailedReceiptV5 v5;
    uint32_t engine_requested;
    uint32_t engine_selected;
    uint64_t generic_applicability_flags;
    uint32_t controls_count;
    uint32_t rhs_receipt_schema;
    uint64_t control_projection_rhs_count;
    double control_rank_rcond;
    double control_rank_smallest_generalized_eigenvalue_lower;
    double control_rank_largest_generalized_eigenvalue_upper;
    double control_rank_projection_error_bound;
    double control_rank_normalization_error_bound;
    double control_rank_fe_information_eigenvalue_lower_bound;
    double control_rank_maximum_projection_residual;
    double control_rank_effective_tolerance;
    double control_rank_projection_pcg_tolerance;
    double control_rank_projection_residual_gate;
    double generic_control_basis_relres;
    double generic_control_basis_forward_error;
    double generic_control_schur_rcond;
    double generic_control_schur_relres;
    double generic_deletion_rank_gap;
    double full_joint_fit_complete_residual;
    double generic_working_fit_complete_residual;
    double generic_maker_relres;
    uint64_t canonicalization_peak_forecast_bytes;
    uint64_t generic_fit_peak_forecast_bytes;
    uint64_t geometry_peak_forecast_bytes;
    uint64_t generic_leverage_peak_forecast_bytes;
    uint64_t generic_target_peak_forecast_bytes;
    uint64_t maker_peak_forecast_bytes;
    uint64_t generic_result_forecast_bytes;
    uint64_t generic_peak_forecast_bytes;
    uint64_t rhs_v2_caller_copy_bytes;
    uint32_t capability_schema;
    uint32_t capability_profile;
    uint32_t batch_mode;
    uint32_t stayers_mode;
    uint32_t target_weight_mode;
    uint32_t deletion_unit_source;
    uint32_t probeorder_supplied;
    uint32_t wallseconds_supplied;
    uint32_t frequency_use;
    uint32_t reserved_6;
    uint64_t physical_limit;
    uint64_t request_signature;
} VckssEngineDetailedReceiptV6;

/* vckss_rust.h:1273-1439: the execution plan receipt the detailed receipt V7
 * appends (estimator resolution, solver, batches, wall model, counter RNG
 * work, memory) */
typedef struct VckssEstimatorResolutionReceiptV1 {
    uint32_t algorithm_schema;
    uint32_t algorithm_requested;
    uint32_t algorithm_selected;
    uint32_t algorithm_reason;
    uint32_t engine_schema;
    uint32_t engine_requested;
    uint32_t engine_selected;
    uint32_t engine_reason;
    uint32_t compressed_eligibility;
    uint32_t resolved_before_rng;
    uint32_t opportunistic_engine_fallback_allowed;
    uint32_t reserved;
    uint64_t identified_complexity;
    uint64_t exact_limit;
    uint64_t rng_draws_before_resolution;
    uint64_t counter_atoms_before_resolution;
} VckssEstimatorResolutionReceiptV1;

typedef struct VckssSolverExecutionReceiptV1 {
    uint32_t schema_version;
    uint32_t requested_route;
    uint32_t selected_route;
    uint32_t fallback_used;
    int32_t fallback_error;
    uint32_t full_setup_complete;
    uint32_t fe_setup_complete;
    uint32_t fe_hierarchy_reused;
    uint32_t plan_frozen_before_rng;
    uint32_t auto_route_contract;
    uint32_t threads_requested;
    uint32_t threads_used;
    uint32_t parallel_regions;
    uint32_t applicability;
    uint32_t reserved_1;
    uint32_t reserved_2;
    uint64_t planned_rhs;
    uint64_t auto_firm_threshold;
    uint64_t auto_rhs_threshold;
    uint64_t full_solver_dimension;
    uint64_t fe_solver_dimension;
    uint64_t logical_atoms_before_plan_freeze;
    uint64_t unique_words_before_plan_freeze;
    uint64_t physical_trials_before_plan_freeze;
} VckssSolverExecutionReceiptV1;

typedef struct VckssBatchPhasePlanReceiptV1 {
    uint32_t request_mode;
    uint32_t selection_reason;
    uint32_t applicability;
    uint32_t reserved;
    uint64_t requested_width;
    uint64_t selected_width;
    uint64_t probe_width_cap;
    uint64_t declared_threads;
    uint64_t thread_width_cap;
    uint64_t route_width_cap;
    uint64_t effective_width_cap;
    uint64_t hard_memory_bytes;
    uint64_t width_one_forecast_bytes;
    uint64_t selected_forecast_bytes;
} VckssBatchPhasePlanReceiptV1;

typedef struct VckssBatchExecutionReceiptV1 {
    uint32_t schema_version;
    uint32_t deterministic;
    uint32_t width_invariance_required;
    uint32_t arithmetic_contract;
    uint32_t whole_command_admitted;
    uint32_t applicability;
    uint64_t non_batched_peak_bytes;
    uint64_t selected_command_peak_bytes;
    VckssBatchPhasePlanReceiptV1 leverage;
    VckssBatchPhasePlanReceiptV1 target;
} VckssBatchExecutionReceiptV1;

typedef struct VckssWallExecutionReceiptV1 {
    uint32_t schema_version;
    uint32_t model_code;
    uint32_t status;
    uint32_t routing_effect;
    uint32_t requested_applicable;
    uint32_t forecast_applicable;
    uint32_t advisory_applicable;
    uint32_t margin_applicable;
    double requested_seconds;
    double forecast_seconds;
    double advisory_seconds;
    double advisory_margin_fraction;
    uint64_t preparation_work;
    uint64_t engine_setup_work;
    uint64_t full_fit_work;
    uint64_t leverage_work;
    uint64_t target_work;
    uint64_t result_export_work;
    uint64_t total_work;
} VckssWallExecutionReceiptV1;

typedef struct VckssCounterPhaseExecutionReceiptV1 {
    uint64_t planned_logical_atoms;
    uint64_t actual_logical_atoms;
    uint64_t planned_unique_packed_words;
    uint64_t actual_unique_packed_words;
    uint64_t planned_physical_trials;
    uint64_t actual_physical_trials;
    uint64_t planned_generator_work;
    uint64_t actual_generator_work;
} VckssCounterPhaseExecutionReceiptV1;

typedef struct VckssCounterExecutionReceiptV1 {
    uint32_t schema_version;
    uint32_t rng_contract;
    uint32_t generator_work_applicable;
    uint32_t completed;
    VckssCounterPhaseExecutionReceiptV1 leverage;
    VckssCounterPhaseExecutionReceiptV1 target;
    VckssCounterPhaseExecutionReceiptV1 total;
    uint64_t logical_atoms_before_plan_freeze;
    uint64_t unique_words_before_plan_freeze;
    uint64_t physical_trials_before_plan_freeze;
} VckssCounterExecutionReceiptV1;

typedef struct VckssExecutionMemoryReceiptV1 {
    uint32_t schema_version;
    uint32_t applicability;
    uint32_t peak_phase;
    uint32_t reserved;
    uint64_t hard_limit_bytes;
    uint64_t prepared_persistent_bytes;
    uint64_t setup_peak_bytes;
    uint64_t fit_peak_bytes;
    uint64_t correction_peak_bytes;
    uint64_t leverage_peak_bytes;
    uint64_t target_peak_bytes;
    uint64_t result_peak_bytes;
    uint64_t non_batched_peak_bytes;
    uint64_t command_peak_bytes;
    uint64_t shared_cmg_persistent_bytes;
    uint64_t full_control_block_persistent_bytes;
    uint64_t setup_transient_bytes;
    uint64_t cmg_preconditioner_workspace_bytes;
    uint64_t cmg_aggregated_cell_capacity_bytes;
    uint64_t cmg_group_index_bytes;
    uint64_t cmg_hybrid_graph_bytes;
    uint64_t retained_nq_bytes;
    uint64_t retained_q2_bytes;
} VckssExecutionMemoryReceiptV1;

typedef struct VckssExecutionPlanReceiptV1 {
    uint32_t struct_size;
    uint32_t schema_version;
    uint64_t generation;
    uint64_t applicability_flags;
    uint64_t contract_flags;
    uint64_t request_signature;
    VckssEstimatorResolutionReceiptV1 resolution;
    VckssSolverExecutionReceiptV1 solver;
    VckssBatchExecutionReceiptV1 batch;
    VckssWallExecutionReceiptV1 wall;
    VckssCounterExecutionReceiptV1 counter;
    VckssExecutionMemoryReceiptV1 memory;
} VckssExecutionPlanReceiptV1;

typedef struct VckssEngineDetailedReceiptV7 {
    VckssEngineDetailedReceiptV6 v6;
    VckssExecutionPlanReceiptV1 execution;
} VckssEngineDetailedReceiptV7;

/* vckss_rust.h:1441-1456 */
typedef struct VckssEnginePerformanceReceiptV1 {
    uint32_t struct_size;
    uint32_t schema_version;
    uint64_t generation;
    uint64_t applicability_flags;
    uint32_t algorithm_selected;
    uint32_t engine_selected;
    uint64_t ingest_ns;
    uint64_t canonicalize_ns;
    uint64_t graph_ns;
    uint64_t compress_ns;
    uint64_t plan_ns;
    uint64_t stayer_augmentation_ns;
    uint64_t solve_ns;
    uint64_t native_total_ns;
} VckssEnginePerformanceReceiptV1;

/* vckss_rust.h:1458-1509 */
typedef struct VckssFullCmgReceiptV1 {
    uint32_t struct_size;
    uint32_t schema_version;
    uint64_t generation;
    uint32_t backend_identity;
    uint32_t platform_os;
    uint32_t platform_arch;
    uint32_t batch_strategy_mask;
    uint8_t cmg_source_commit[40];
    uint32_t threads_requested;
    uint32_t threads_used;
    uint64_t maximum_concurrency;
    uint64_t vertices;
    uint64_t edges;
    uint64_t hierarchy_levels;
    uint64_t terminal_vertices;
    uint64_t graph_copy_bytes;
    uint64_t hierarchy_bytes;
    uint64_t plan_bytes;
    uint64_t workspace_bytes_each;
    uint64_t workspace_pool_bytes;
    uint64_t admitted_peak_bytes;
    double fit_effective_tolerance;
    double probe_effective_tolerance;
    double fit_initial_inner_tolerance;
    double probe_initial_inner_tolerance;
    uint64_t refinement_attempts;
    uint64_t refined_columns;
    uint64_t batch_calls;
    uint64_t rhs_count;
    uint64_t serial_batches;
    uint64_t planned_batches;
    uint64_t across_rhs_batches;
    uint64_t total_iterations;
    uint64_t total_operator_applications;
    uint64_t total_preconditioner_applications;
    double maximum_reduced_residual;
    double maximum_complete_residual;
    uint64_t graph_ns;
    uint64_t hierarchy_plan_ns;
    uint64_t rhs_ns;
    uint64_t solve_ns;
    uint64_t extraction_ns;
    uint64_t preparation_peak_bytes;
    uint64_t prepared_persistent_bytes;
    uint64_t non_cmg_command_peak_bytes;
    uint64_t pre_rng_forecast_bytes;
    uint64_t actual_retained_bytes;
    uint64_t allocator_allowance_bytes;
    uint64_t maximum_batch_rhs;
    uint64_t workspace_count;
} VckssFullCmgReceiptV1;

/* vckss_rust.h:1611-1621: one row per solved right-hand side */
typedef struct VckssEngineRhsReceiptV1 {
    uint32_t phase;
    uint32_t side;
    int64_t probe;
    uint32_t route;
    uint32_t iterations;
    uint32_t zero_rhs;
    uint32_t reserved;
    double reduced_residual;
    double complete_residual;
} VckssEngineRhsReceiptV1;

/* vckss_rust.h:369-398: eligible stayer rows, 1-based dense retained-mover
 * firm and stayer-worker indices as exact binary64 integers */
typedef struct VckssStayerAugmentationRequestV1 {
    uint32_t abi_version;
    uint32_t struct_size;
    uint64_t rows;
    uint32_t controls_count;
    uint32_t reserved;
    uint64_t caller_copy_bytes;
} VckssStayerAugmentationRequestV1;

typedef struct VckssStayerAugmentationRequestInterruptV1 {
    VckssStayerAugmentationRequestV1 options;
    VckssInterruptPollV1 interrupt_poll;
    void *interrupt_context;
    uint32_t checkpoint_interval;
    uint32_t reserved;
} VckssStayerAugmentationRequestInterruptV1;

typedef struct VckssStayerAugmentationColumnsV1 {
    uint32_t struct_size;
    uint32_t reserved;
    uint64_t rows;
    const double *firm;
    const double *worker;
    const double *outcome;
    const double *frequency;
    const double *target_weight;
    const double *const *controls;
    uint32_t controls_count;
    uint32_t reserved_2;
} VckssStayerAugmentationColumnsV1;

/* vckss_rust.h:770-796 */
typedef struct VckssStayerAugmentationReceiptV1 {
    uint32_t struct_size;
    uint32_t schema_version;
    uint64_t generation;
    uint64_t mover_stored_rows;
    uint64_t stayer_stored_rows;
    uint64_t combined_stored_rows;
    uint64_t mover_physical_mass;
    uint64_t stayer_physical_mass;
    uint64_t combined_physical_mass;
    uint64_t mover_workers;
    uint64_t stayer_workers;
    uint64_t combined_workers;
    uint64_t firms;
    uint64_t mover_deletion_units;
    uint64_t stayer_deletion_units;
    uint64_t combined_deletion_units;
    double mover_target_mass;
    double stayer_target_mass;
    double combined_target_mass;
    uint64_t topology_checksum;
    uint64_t memory_limit_bytes;
    uint64_t caller_copy_bytes;
    uint64_t augmentation_peak_forecast_bytes;
    uint64_t augmented_resident_bytes;
    uint64_t total_prepared_resident_bytes;
} VckssStayerAugmentationReceiptV1;

/* vckss_rust.h:1037-1052 */
typedef struct VckssComponentVectorV1 {
    double worker;
    double firm;
    double covariance;
    double total;
} VckssComponentVectorV1;

typedef struct VckssEngineResultV1 {
    uint32_t struct_size;
    uint32_t reserved;
    uint64_t generation;
    VckssComponentVectorV1 plugin;
    VckssComponentVectorV1 correction;
    VckssComponentVectorV1 corrected;
    VckssComponentVectorV1 numerical_mcse;
} VckssEngineResultV1;

/* vckss_rust.h:2337-2346: numerical MCSE request (opt-in attachment) */
typedef struct VckssNumericalRequestV1 {
    uint32_t abi_version, struct_size, schema_version, threads;
    uint32_t tolerance_supplied, projection_requested, component_requested, reserved;
    uint32_t controls_count, reserved_2;
    VckssEngineSolveRequestV4 point;
} VckssNumericalRequestV1;

/* vckss_rust.h:2348-2351 */
typedef struct VckssNumericalRequestV2 {
    VckssNumericalRequestV1 v1;
    uint32_t execution_mode, component_batch_mode, full_cmg, reserved;
} VckssNumericalRequestV2;

/* vckss_rust.h:2359-2370: status 1 exact_zero, 2 ok_local, 3 ok_local_psd_adjusted,
 * 4 unstable_nonpsd, 5 nonsmooth_adjustment, 6 nonfinite_derivative, 7 replay_failed */
typedef struct VckssNumericalResultV1 {
    uint32_t struct_size, schema_version;
    uint64_t generation;
    uint32_t status, engine, leverage_probes, target_probes, target_fold_a, target_fold_b;
    uint64_t certified_replay_rhs, executed_replay_rhs, attempted_replay_rhs, replay_generator_words,
        allocation_bound_bytes;
    uint32_t failed_replay_probe, replay_error_code;
    double psd_adjustment, minimum_constrained, minimum_residual_margin, maximum_sensitivity_ratio;
    double score_means[2][3];
    double conditional[3][3], leverage[3][3], raw[3][3], usable[3][3], mcse[4];
    uint64_t point_rhs;
    double maximum_point_complete_residual, maximum_replay_complete_residual;
} VckssNumericalResultV1;

/* vckss_rust.h:2371-2374 */
typedef struct VckssNumericalRhsV1 {
    uint32_t schema_version, phase, probe, reserved;
    double complete_residual;
} VckssNumericalRhsV1;

/* vckss_rust.h:1512-1522 */
typedef struct VckssFullCmgModelReceiptV1 {
    uint32_t struct_size;
    uint32_t schema_version;
    uint64_t generation;
    uint32_t controls_count;
    uint32_t nuisance_mode;
    uint64_t logical_rhs_count;
    uint64_t explicit_options_rhs_count;
    uint64_t controlled_rhs_count;
    uint64_t control_refinement_rhs_count;
} VckssFullCmgModelReceiptV1;

#if !defined(__cplusplus)
_Static_assert(sizeof(VckssNumericalRequestV1) == 328, "numerical request layout");
_Static_assert(sizeof(VckssNumericalRequestV2) == 344, "numerical V2 request layout");
_Static_assert(sizeof(VckssNumericalResultV1) == 512, "numerical result layout");
_Static_assert(sizeof(VckssNumericalRhsV1) == 24, "numerical RHS layout");
#endif

/* vckss_rust.h:1635-1640 */
typedef struct VckssEngineSnapshotV1 {
    uint32_t struct_size;
    uint32_t state;
    uint64_t generation;
    uint64_t last_released_generation;
} VckssEngineSnapshotV1;

/* vckss_rust.h:1642-1646 */
uint32_t vckss_rust_abi_version(void);
const char *vckss_rust_backend_version(void);
const char *vckss_rust_capabilities_json(void);
const char *vckss_rust_last_error(void);
int32_t vckss_rust_selftest(void);

/* vckss_rust.h:1713-1716 */
int32_t vckss_rust_engine_default_prepare_request_interrupt_v3(
    VckssEnginePrepareRequestInterruptV3 *output, uint32_t output_capacity_bytes);
/* vckss_rust.h:1745-1748 */
int32_t vckss_rust_engine_default_solve_request_interrupt_v5(
    VckssEngineSolveRequestInterruptV5 *output, uint32_t output_capacity_bytes);
/* vckss_rust.h:1799-1804: canonicalise, select the leave-out connected set,
 * compress, build the JLA plan, the hybrid Laplacian and its CMG hierarchy,
 * upload to the GPU.  *output_handle receives a nonzero generation. */
int32_t vckss_rust_engine_prepare_interrupt_v4(
    const VckssEnginePrepareRequestInterruptV3 *request, const VckssEngineColumnsV3 *columns,
    uint64_t *output_handle, uint32_t output_capacity_bytes);
/* vckss_rust.h:1717-1720, 1805-1819: the stayer hybrid (stayers(both)):
 * eligible stayer rows appended to a prepared match-deletion generation;
 * solves then take stayers_mode ALL (ffi_engine.rs:2867-2879, 3667-3676,
 * 4171-4262, 8737-8770) */
int32_t vckss_rust_engine_default_stayer_augmentation_request_interrupt_v1(
    VckssStayerAugmentationRequestInterruptV1 *output, uint32_t output_capacity_bytes);
int32_t vckss_rust_engine_augment_stayers_v1(uint64_t generation, const VckssStayerAugmentationRequestV1 *request,
                                             const VckssStayerAugmentationColumnsV1 *columns);
int32_t vckss_rust_engine_augment_stayers_interrupt_v1(uint64_t generation,
                                                       const VckssStayerAugmentationRequestInterruptV1 *request,
                                                       const VckssStayerAugmentationColumnsV1 *columns);
int32_t vckss_rust_engine_stayer_augmentation_receipt_v1(uint64_t generation, VckssStayerAugmentationReceiptV1 *output,
                                                         uint32_t output_capacity_bytes);

/* vckss_rust.h:1054-1086 */
typedef struct VckssStayerHybridResultV1 {
    uint32_t struct_size;
    uint32_t schema_version;
    uint64_t generation;
    VckssComponentVectorV1 plugin;
    VckssComponentVectorV1 correction;
    VckssComponentVectorV1 corrected;
    VckssComponentVectorV1 mover_correction;
    VckssComponentVectorV1 stayer_correction;
    double weighted_rss;
    uint64_t parameters;
    uint64_t full_parameters;
    uint64_t correction_parameters;
    uint64_t deletion_units;
    double max_leverage;
    double information_rcond;
    double inverse_relres;
    double inverse_original_relres;
    double inverse_sqrt_relres;
    double maker_relres;
    double full_fit_relres;
    double working_fit_relres;
    double fit_residual_tolerance;
    double control_basis_relres;
    double control_basis_forward_error;
    double deletion_rank_gap;
    double firm_zero_sum_residual;
    uint64_t peak_forecast_bytes;
    uint64_t fit_peak_forecast_bytes;
    uint64_t correction_peak_forecast_bytes;
    uint64_t topology_checksum;
    double accounting_residual;
} VckssStayerHybridResultV1;

/* vckss_rust.h:1978-1982: the exact stayer hybrid's receipt with its mover
 * and stayer corrections (ffi_engine.rs:6122-6145, 8772-8840) */
int32_t vckss_rust_engine_stayer_hybrid_result_v1(uint64_t generation, VckssStayerHybridResultV1 *output,
                                                  uint32_t output_capacity_bytes);

/* vckss_rust.h:1647-1665: backend capabilities and the request-capability
 * receipts the shim reconciles before preparing (stata_entry.c:1164-1300).
 * The receipt echoes the request, classifies it with the reference's reason
 * codes (then this engine's: no controls, joint nuisance, CMG route for JLA,
 * no probe-order keys, automatic batches) and carries the reference's FNV-1a
 * request signature. */
int32_t vckss_rust_backend_capabilities_v1(VckssBackendCapabilitiesV1 *output, uint32_t output_capacity_bytes);
int32_t vckss_rust_backend_request_capability_v1(const VckssBackendRequestCapabilityRequestV1 *request,
                                                 VckssBackendRequestCapabilityReceiptV1 *output,
                                                 uint32_t output_capacity_bytes);
int32_t vckss_rust_backend_request_capability_v2(const VckssBackendRequestCapabilityRequestV2 *request,
                                                 VckssBackendRequestCapabilityReceiptV2 *output,
                                                 uint32_t output_capacity_bytes);
int32_t vckss_rust_backend_request_capability_v3(const VckssBackendRequestCapabilityRequestV3 *request,
                                                 VckssBackendRequestCapabilityReceiptV3 *output,
                                                 uint32_t output_capacity_bytes);
/* vckss_rust.h:1765-1768: admission of a prepare request before its data
 * (stata_entry.c:1793-1798) */
int32_t vckss_rust_engine_admit_prepare_v4(const VckssEnginePrepareRequestV4 *request, uint32_t probeorder_supplied);
/* vckss_rust.h:1923-1927: counts, graph-selection receipt, target mass */
int32_t vckss_rust_engine_preparation_receipt_v4(uint64_t generation, VckssEnginePreparationReceiptV4 *output,
                                                 uint32_t output_capacity_bytes);
/* vckss_rust.h:2332: expected and admission peak bytes of the generation
 * (host arrays and hierarchy plus the device buffers' peak; 20% allowance) */
int32_t vckss_rust_engine_memory_forecast_v1(uint64_t generation, VckssMemoryForecastV1 *output,
                                             uint32_t output_capacity_bytes);
/* vckss_rust.h:2116-2120: phase wall times of the prepare and the solve */
int32_t vckss_rust_engine_performance_receipt_v1(uint64_t generation, VckssEnginePerformanceReceiptV1 *output,
                                                 uint32_t output_capacity_bytes);
/* vckss_rust.h:2101-2115: the solve's detailed receipt; V7 appends the
 * execution plan.  The capability echo (schema, profile, signature) is the
 * solve request's. */
int32_t vckss_rust_engine_detailed_receipt_v6(uint64_t generation, VckssEngineDetailedReceiptV6 *output,
                                              uint32_t output_capacity_bytes);
int32_t vckss_rust_engine_detailed_receipt_v7(uint64_t generation, VckssEngineDetailedReceiptV7 *output,
                                              uint32_t output_capacity_bytes);
/* vckss_rust.h:2121-2130: the CMG_FULL_V2 receipt and one row per solved
 * right-hand side (fit, leverage probes, worker and firm side of target
 * probes) */
int32_t vckss_rust_engine_full_cmg_receipt_v1(uint64_t generation, VckssFullCmgReceiptV1 *output,
                                              uint32_t output_capacity_bytes);
int32_t vckss_rust_engine_rhs_receipts_v1(uint64_t generation, VckssEngineRhsReceiptV1 *output,
                                          uint64_t output_capacity_rows);
/* vckss_rust.h:1584-1598 (and the V7 request default and solve of
 * generic_execution_api.rs:500-550): the generic engine's executor requests.
 * The direct executor on the CMG route with automatic batches runs this
 * engine's CMG_FULL_V2 path; the diagonal queue returns 11. */
int32_t vckss_rust_engine_default_solve_request_v6(VckssEngineSolveRequestV6 *output, uint32_t output_capacity_bytes);
int32_t vckss_rust_engine_default_solve_request_interrupt_v6(VckssEngineSolveRequestInterruptV6 *output,
                                                             uint32_t output_capacity_bytes);
int32_t vckss_rust_engine_solve_v6(uint64_t generation, const VckssEngineSolveRequestV6 *request);
int32_t vckss_rust_engine_solve_interrupt_v6(uint64_t generation, const VckssEngineSolveRequestInterruptV6 *request);
int32_t vckss_rust_engine_default_solve_request_v7(VckssEngineSolveRequestV7 *output, uint32_t output_capacity_bytes);
int32_t vckss_rust_engine_default_solve_request_interrupt_v7(VckssEngineSolveRequestInterruptV7 *output,
                                                             uint32_t output_capacity_bytes);
int32_t vckss_rust_engine_solve_v7(uint64_t generation, const VckssEngineSolveRequestV7 *request);
int32_t vckss_rust_engine_solve_interrupt_v7(uint64_t generation, const VckssEngineSolveRequestInterruptV7 *request);

/* vckss_rust.h:1928-1932 */
int32_t vckss_rust_engine_retained_mask_v1(uint64_t generation, uint8_t *output,
                                           uint64_t output_capacity_bytes);
/* vckss_rust.h:1949-1952 and 1969-1972: the per-probe path on the GPU */
int32_t vckss_rust_engine_solve_v5(uint64_t generation, const VckssEngineSolveRequestV5 *request);
int32_t vckss_rust_engine_solve_interrupt_v5(uint64_t generation,
                                             const VckssEngineSolveRequestInterruptV5 *request);
/* vckss_rust.h:1973-1977 */
int32_t vckss_rust_engine_result_v1(uint64_t generation, VckssEngineResultV1 *output,
                                    uint32_t output_capacity_bytes);
/* vckss_rust.h:2136-2142 */
int32_t vckss_rust_engine_release_v1(uint64_t generation);
int32_t vckss_rust_engine_clear_abandoned_v1(void);
int32_t vckss_rust_engine_snapshot_v1(VckssEngineSnapshotV1 *output, uint32_t output_capacity_bytes);
const char *vckss_rust_engine_last_error(void);

/* vckss_rust.h:2348-2388: the numerical MCSE attachment on the same path.
 * The point estimate equals the plain solve's; the leverage probes are then
 * replayed and scored against the unit linearisation and target folds. */
uint32_t vckss_rust_numerical_schema_v1(void);
uint32_t vckss_rust_numerical_schema_v2(void);
int32_t vckss_rust_engine_numerical_preflight_v1(const VckssNumericalRequestV1 *request);
int32_t vckss_rust_engine_numerical_preflight_v2(const VckssNumericalRequestV2 *request);
int32_t vckss_rust_engine_solve_numerical_v1(uint64_t generation, const VckssNumericalRequestV1 *request);
int32_t vckss_rust_engine_solve_numerical_interrupt_v1(uint64_t generation, const VckssNumericalRequestV1 *request,
                                                       VckssInterruptPollV1 callback, void *context,
                                                       uint32_t interval);
int32_t vckss_rust_engine_solve_numerical_interrupt_v2(uint64_t generation, const VckssNumericalRequestV2 *request,
                                                       VckssInterruptPollV1 callback, void *context,
                                                       uint32_t interval);
int32_t vckss_rust_engine_numerical_result_v1(uint64_t generation, VckssNumericalResultV1 *output,
                                              uint32_t output_capacity_bytes);
int32_t vckss_rust_engine_numerical_rhs_v1(uint64_t generation, VckssNumericalRhsV1 *output, uint64_t capacity,
                                           uint64_t *written);
int32_t vckss_rust_engine_numerical_cmg_work_v1(uint64_t generation, VckssFullCmgModelReceiptV1 *output,
                                                uint32_t output_capacity_bytes);

#ifdef __cplusplus
}
#endif
#endif
