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
#define VCKSS_REQUEST_CAPABILITY_SCHEMA_V3 3u
#define VCKSS_REQUEST_PROFILE_PLANNED_V1 4u
#define VCKSS_REQUEST_FREQUENCY_UNIT 0u
#define VCKSS_REQUEST_FREQUENCY_LITERAL 1u
#define VCKSS_CORE_MATCH_GRAPH_READY (UINT64_C(1) << 0)
#define VCKSS_CORE_CMG_GRAPH_READY (UINT64_C(1) << 4)
#define VCKSS_CORE_COUNTER_RNG_READY (UINT64_C(1) << 6)
#define VCKSS_CORE_JLA_PLAN_READY (UINT64_C(1) << 7)
#define VCKSS_CORE_FULL_CMG_V2_READY (UINT64_C(1) << 8)

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
