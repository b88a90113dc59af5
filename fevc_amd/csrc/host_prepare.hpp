// Host-side preparation: canonical identifiers, leave-out connected set,
// compressed problem and the no-control JLA plan.
#pragma once

#include <cstdint>
#include <vector>

namespace vk {

struct Columns {
  uint64_t rows = 0;
  const uint64_t *worker = nullptr, *firm = nullptr, *deletion = nullptr;
  const double *outcome = nullptr;
  const uint64_t *frequency = nullptr;
  const double *target = nullptr;
};

struct SelectionReceipt {
  uint64_t input_rows = 0, retained_rows = 0, input_physical_mass = 0, retained_physical_mass = 0;
  uint64_t initial_components = 0, maximum_components = 0, initial_component_rows = 0;
  uint64_t mover_input_rows = 0, initial_deletion_edges = 0, retained_deletion_edges = 0;
  uint64_t insufficient_workers_removed = 0, articulation_workers_removed = 0;
  uint64_t bridge_units_removed = 0, bridge_rows_removed = 0;
  uint64_t degree_iterations = 0, articulation_iterations = 0, bridge_iterations = 0;
  uint64_t fixed_point_iterations = 0;
};

// The compressed problem (problem.rs:647-674) with the JLA plan (jla.rs:20-46)
// and the gather indices the device kernels read.
struct Prepared {
  uint64_t input_rows = 0, input_physical = 0;
  double input_target_total = 0.0;
  uint64_t input_topology_checksum = 0;
  std::vector<uint8_t> mask;
  SelectionReceipt selection;
  // wall time of the preparation phases (nanoseconds): validation and
  // canonical ids, graph selection, compression, JLA plan
  uint64_t canonicalize_ns = 0, graph_ns = 0, compress_ns = 0, plan_ns = 0;
  bool observation = false;  // observation deletion (else match deletion)
  bool stayers = false;      // a mover problem augmented with eligible stayers
  // units [0, match_units) are match deletion units; the rest are observation
  // units, one per stored row, in row order, ending with the last row
  uint64_t match_units = 0;

  uint64_t rows = 0, workers = 0, firms = 0, cells = 0, units = 0, exact_strata = 0, strata = 0;
  std::vector<uint32_t> row_worker, row_firm, row_deletion, row_cell;
  std::vector<double> outcome, target;
  std::vector<uint64_t> frequency;
  std::vector<uint32_t> cell_worker, cell_firm;
  std::vector<double> cell_weight, cell_outcome, cell_target;
  std::vector<uint64_t> worker_ptr;  // cells of worker w are [ptr[w], ptr[w+1])
  std::vector<uint64_t> firm_ptr;    // firm_cells[ptr[f]..] in (worker, cell) order
  std::vector<uint32_t> firm_cells;
  std::vector<uint64_t> unit_ptr;    // rows of unit g, ascending
  std::vector<uint32_t> unit_rows;
  uint64_t physical_total = 0;
  double target_total = 0.0;
  uint64_t topology_checksum = 0;

  std::vector<double> per_copy;
  std::vector<uint64_t> row_rank;
  std::vector<uint32_t> unit_cell;
  std::vector<uint64_t> unit_count, unit_rank;
  std::vector<uint64_t> unit_offset;  // observation deletion: first copy of the row in its class
  std::vector<double> unit_outcome, unit_target;
  std::vector<uint32_t> stratum_cell;
  std::vector<double> stratum_per_copy, stratum_target;
  std::vector<uint64_t> stratum_count, stratum_rank;
  std::vector<uint64_t> cell_unit_ptr;  // units of a cell, ascending
  std::vector<uint32_t> cell_units;
  std::vector<uint64_t> cell_stratum_ptr;  // strata of a cell are contiguous

  std::vector<uint64_t> row_input;  // input row of each stored row (ascending)
  // Controls (problem.rs:647-674 keeps them column-major per stored row):
  // stored row x controls, row-major, and per cell the frequency-weighted
  // sums over its rows (Kahan, rows ascending, as the cell outcome sums).
  uint32_t controls = 0;
  std::vector<double> control, cell_control;
};

// Validate, canonicalize, select (unless `mask` is given), compress, plan.
// Throws vk::Error with the reference status code.
// With `observation`, the observation-deletion graph selection (graph.rs:82-242)
// instead of the match-deletion one; the exact estimator then deletes every
// physical observation on its own.
Prepared prepare(const Columns &cols, const uint8_t *mask, bool observation = false);

// Attach `q` (<= 32) control columns of the input rows to a prepared problem:
// null pointers and nonfinite values are refused (status 20).
void attach_controls(Prepared &P, const double *const *columns, uint32_t q);

// Replace the stored-row outcome by `y`: the outcome, cell and unit sums are
// formed as prepare forms them; graph, cells and plan ranks are unchanged.
void set_outcome(Prepared &P, const std::vector<double> &y);

// Eligible stayer rows: 1-based dense retained-mover firm and stayer worker
// indices (stayer_hybrid.rs:23-34).
struct StayerColumns {
  uint64_t rows = 0;
  const uint64_t *firm = nullptr, *worker = nullptr, *frequency = nullptr;
  const double *outcome = nullptr, *target = nullptr;
};
struct StayerReceipt {
  uint64_t mover_stored_rows = 0, stayer_stored_rows = 0, combined_stored_rows = 0;
  uint64_t mover_physical_mass = 0, stayer_physical_mass = 0, combined_physical_mass = 0;
  uint64_t mover_workers = 0, stayer_workers = 0, combined_workers = 0, firms = 0;
  uint64_t mover_deletion_units = 0, stayer_deletion_units = 0, combined_deletion_units = 0;
  double mover_target_mass = 0.0, stayer_target_mass = 0.0, combined_target_mass = 0.0;
  uint64_t topology_checksum = 0;
};
// stayer_hybrid.rs:71-348 (prepare_exact_stayer_hybrid): validate the stayer
// rows against the prepared mover problem M, recompress movers and stayers
// together (every row retained), and address the stayer rows as observation
// units.  Throws vk::Error with the reference status code.
Prepared augment_stayers(const Prepared &M, const StayerColumns &s, StayerReceipt &rec);

}  // namespace vk
