/* Compressed problem + JLA plan layout of the oracle (test infrastructure).
 * Mirrors CompressedProblem (problem.rs:647-674) and JlaPlan (jla.rs:20-46). */
#ifndef ORC_SAMPLE_H
#define ORC_SAMPLE_H

#include <stdint.h>

typedef struct {
    uint64_t input_rows, retained_rows, retained_physical_mass;
    uint64_t insufficient_workers_removed, articulation_workers_removed;
    uint64_t bridge_units_removed, bridge_rows_removed;
    uint64_t degree_iterations, articulation_iterations, bridge_iterations;
    uint64_t fixed_point_iterations;
} orc_selection_receipt;

struct orc_problem {
    uint64_t input_rows;
    uint64_t input_topology_checksum;
    uint8_t *mask;
    orc_selection_receipt selection;
    /* compressed */
    uint64_t rows, workers, firms, cells, units, exact_strata, strata;
    uint64_t *retained;
    uint32_t *row_worker, *row_firm, *row_deletion, *row_cell;
    double *outcome, *target;
    uint64_t *frequency;
    uint32_t *cell_worker, *cell_firm;
    double *cell_weight, *cell_outcome_sum, *cell_target_sum;
    uint64_t *worker_ptr;             /* cells of worker w: [ptr[w], ptr[w+1]) */
    uint64_t *firm_ptr, *firm_cells;  /* cells of firm f in (worker, cell) order */
    uint64_t *unit_ptr, *unit_rows;   /* rows of deletion unit g, ascending */
    uint64_t physical_total;
    double target_total;
    uint64_t topology_checksum;
    /* JLA plan */
    double *per_copy;
    uint64_t *row_rank;
    uint32_t *unit_cell;
    uint64_t *unit_count, *unit_rank;
    double *unit_outcome, *unit_target;
    uint32_t *stratum_cell;
    double *stratum_per_copy, *stratum_target;
    uint64_t *stratum_count, *stratum_rank;
    /* observation deletion: one unit per stored row; its first copy in the
     * row's observation class (generic_jla.rs:4158-4215) */
    int observation;
    uint64_t *unit_offset;
    /* units [0, match_units) are match units, the rest observation units
     * (the last stored rows); stayers: a mover problem augmented with
     * eligible stayers (stayer_hybrid.rs:71-348) */
    uint64_t match_units;
    int stayers;
};

#endif
