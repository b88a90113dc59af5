// Typed refusals for the rest of the reference plugin ABI.
//
// The reference header (rust/stata_backend/include/vckss_rust.h) declares 147
// functions; the Stata shim links 98 of them.  This library implements the
// CMG_FULL_V2 JLA path (vckss_abi.cpp).  Every other entry point is exported
// here so that the shim links unchanged: a status-returning call records
// "<name>: not supported by the MI355X engine" in the thread-local error and
// returns VCKSS_STATUS_UNSUPPORTED (11), as the reference does for features a
// build leaves out (rust/crates/vckss-core/src/error.rs:12); a schema query
// returns 0 (no schema offered).  C linkage does not encode parameter types,
// so the refusals take no parameters and leave the caller's arguments alone.
#include <cstdint>

namespace vk {
void record_unsupported(const char *name);
const char *thread_last_error();
}  // namespace vk

#define VK_REFUSE(name)                   \
  int32_t name(void) {                    \
    ::vk::record_unsupported(#name);      \
    return 11;                            \
  }
#define VK_NO_SCHEMA(name) \
  uint32_t name(void) { return 0u; }

extern "C" {

VK_REFUSE(vckss_rust_engine_admit_prepare_memory_v1)
VK_REFUSE(vckss_rust_engine_admit_prepare_probe_order_v1)
VK_REFUSE(vckss_rust_engine_admit_prepare_v2)
VK_REFUSE(vckss_rust_engine_admit_prepare_v3)
VK_REFUSE(vckss_rust_engine_augment_component_inference_interrupt_v1)
VK_REFUSE(vckss_rust_engine_augment_component_inference_interrupt_v2)
VK_REFUSE(vckss_rust_engine_augment_component_inference_interrupt_v3)
VK_REFUSE(vckss_rust_engine_augment_component_inference_interrupt_v4)
VK_REFUSE(vckss_rust_engine_augment_component_inference_interrupt_v5)
VK_REFUSE(vckss_rust_engine_augment_match_component_inference_interrupt_v1)
VK_REFUSE(vckss_rust_engine_augment_match_component_inference_interrupt_v2)
VK_REFUSE(vckss_rust_engine_augment_match_component_inference_interrupt_v3)
VK_REFUSE(vckss_rust_engine_augment_match_component_inference_interrupt_v4)
VK_REFUSE(vckss_rust_engine_augment_match_component_inference_interrupt_v5)
VK_REFUSE(vckss_rust_engine_augment_projection_interrupt_v1)
VK_REFUSE(vckss_rust_engine_centering_v1)
VK_REFUSE(vckss_rust_engine_component_batch_receipt_v1)
VK_REFUSE(vckss_rust_engine_component_inference_augmentation_receipt_v1)
VK_REFUSE(vckss_rust_engine_component_inference_result_v2)
VK_REFUSE(vckss_rust_engine_component_inference_result_v3)
VK_REFUSE(vckss_rust_engine_component_inference_result_v4)
VK_REFUSE(vckss_rust_engine_component_inference_result_v5)
VK_REFUSE(vckss_rust_engine_component_inference_unit_receipt_v1)
VK_REFUSE(vckss_rust_engine_component_inference_unit_receipt_v2)
VK_REFUSE(vckss_rust_engine_default_component_inference_augmentation_request_interrupt_v1)
VK_REFUSE(vckss_rust_engine_default_prepare_request_interrupt_v1)
VK_REFUSE(vckss_rust_engine_default_prepare_request_interrupt_v2)
VK_REFUSE(vckss_rust_engine_default_prepare_request_v3)
VK_REFUSE(vckss_rust_engine_default_projection_augmentation_request_interrupt_v1)
VK_REFUSE(vckss_rust_engine_default_solve_request_interrupt_v1)
VK_REFUSE(vckss_rust_engine_default_solve_request_interrupt_v2)
VK_REFUSE(vckss_rust_engine_default_solve_request_interrupt_v3)
VK_REFUSE(vckss_rust_engine_default_solve_request_interrupt_v4)
VK_REFUSE(vckss_rust_engine_default_solve_request_interrupt_v8)
VK_REFUSE(vckss_rust_engine_default_solve_request_v1)
VK_REFUSE(vckss_rust_engine_default_solve_request_v2)
VK_REFUSE(vckss_rust_engine_default_solve_request_v3)
VK_REFUSE(vckss_rust_engine_default_solve_request_v4)
VK_REFUSE(vckss_rust_engine_default_solve_request_v5)
VK_REFUSE(vckss_rust_engine_default_solve_request_v8)
VK_REFUSE(vckss_rust_engine_detailed_receipt_v1)
VK_REFUSE(vckss_rust_engine_detailed_receipt_v2)
VK_REFUSE(vckss_rust_engine_detailed_receipt_v3)
VK_REFUSE(vckss_rust_engine_detailed_receipt_v4)
VK_REFUSE(vckss_rust_engine_detailed_receipt_v5)
VK_REFUSE(vckss_rust_engine_exact_execution_receipt_v1)
VK_REFUSE(vckss_rust_engine_execution_plan_receipt_v1)
VK_REFUSE(vckss_rust_engine_full_cmg_model_receipt_v1)
VK_REFUSE(vckss_rust_engine_generic_execution_receipt_v1)
VK_REFUSE(vckss_rust_engine_generic_execution_receipt_v2)
VK_REFUSE(vckss_rust_engine_leaveout_residual_diagnostic_v1)
VK_REFUSE(vckss_rust_engine_memory_policy_v1)
VK_REFUSE(vckss_rust_engine_preparation_receipt_v1)
VK_REFUSE(vckss_rust_engine_preparation_receipt_v2)
VK_REFUSE(vckss_rust_engine_preparation_receipt_v3)
VK_REFUSE(vckss_rust_engine_prepare_interrupt_v1)
VK_REFUSE(vckss_rust_engine_prepare_interrupt_v2)
VK_REFUSE(vckss_rust_engine_prepare_interrupt_v3)
VK_REFUSE(vckss_rust_engine_prepare_memory_interrupt_v1)
VK_REFUSE(vckss_rust_engine_prepare_v1)
VK_REFUSE(vckss_rust_engine_prepare_v2)
VK_REFUSE(vckss_rust_engine_prepare_v3)
VK_REFUSE(vckss_rust_engine_projection_augmentation_receipt_v1)
VK_REFUSE(vckss_rust_engine_projection_result_receipt_v1)
VK_REFUSE(vckss_rust_engine_projection_result_v1)
VK_REFUSE(vckss_rust_engine_repeat_measurement_diagnostic_v1)
VK_REFUSE(vckss_rust_engine_repeat_measurement_leading_v1)
VK_REFUSE(vckss_rust_engine_rhs_receipts_v2)
VK_REFUSE(vckss_rust_engine_solve_exact_execution_interrupt_v1)
VK_REFUSE(vckss_rust_engine_solve_exact_execution_v1)
VK_REFUSE(vckss_rust_engine_solve_exact_legacy_execution_interrupt_v1)
VK_REFUSE(vckss_rust_engine_solve_exact_resolved_execution_interrupt_v2)
VK_REFUSE(vckss_rust_engine_solve_exact_resolved_execution_v2)
VK_REFUSE(vckss_rust_engine_solve_interrupt_v1)
VK_REFUSE(vckss_rust_engine_solve_interrupt_v2)
VK_REFUSE(vckss_rust_engine_solve_interrupt_v3)
VK_REFUSE(vckss_rust_engine_solve_interrupt_v4)
VK_REFUSE(vckss_rust_engine_solve_interrupt_v8)
VK_REFUSE(vckss_rust_engine_solve_numerical_generic_legacy_v2)
VK_REFUSE(vckss_rust_engine_solve_numerical_legacy_v2)
VK_REFUSE(vckss_rust_engine_solve_v1)
VK_REFUSE(vckss_rust_engine_solve_v2)
VK_REFUSE(vckss_rust_engine_solve_v3)
VK_REFUSE(vckss_rust_engine_solve_v4)
VK_REFUSE(vckss_rust_engine_solve_v8)
VK_REFUSE(vckss_rust_report_call_v1)
VK_REFUSE(vckss_rust_session_clear_abandoned_v1)
VK_REFUSE(vckss_rust_session_preparation_receipt_v1)
VK_REFUSE(vckss_rust_session_prepare_v1)
VK_REFUSE(vckss_rust_session_release_v1)
VK_REFUSE(vckss_rust_session_snapshot_v1)

VK_NO_SCHEMA(vckss_rust_centering_schema_v1)
VK_NO_SCHEMA(vckss_rust_component_inference_interface_version)
VK_NO_SCHEMA(vckss_rust_exact_execution_schema_v1)
VK_NO_SCHEMA(vckss_rust_exact_legacy_execution_schema_v1)
VK_NO_SCHEMA(vckss_rust_exact_resolved_execution_schema_v2)
VK_NO_SCHEMA(vckss_rust_projection_centering_schema_v1)

const char *vckss_rust_session_last_error(void) { return ::vk::thread_last_error(); }

}  // extern "C"
