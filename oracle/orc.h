/*
 * CPU oracle for the leave-out (KSS) CMG-PCG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Rust engine (johannes-schmieder/fevc, rust/crates/vckss-core and the
 * vendored CMG crate rust/vendor/cmg).  It is used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker
 * and as the reported CPU baseline.  The product library never links it.
 *
 * Every function names the reference file:line it restates.  Parity pins:
 *   - RNG: permanent vectors of rng.rs:366-419 (bit-exact).
 *   - CMG graph, hierarchy, LDL, cycle, PCG: FNV-1a digests of
 *     rust/vendor/cmg/tests/fixtures/frozen_bits.txt [default] (bit-exact).
 *   - Engine end to end: rust_mata_shared_atoms.rs known answers (3e-12).
 *   - Hybrid graph: cmg_impl.rs.in:749-835 weights and the dense Schur
 *     closed form.
 *   - Leave-out selection: brute force (articulation and bridge deletion
 *     plus BFS) on small panels.
 *   - Numerical MCSE: finite differences of the linearised functions, the
 *     covariance-policy cases and fold enumeration of
 *     rust/crates/vckss-core/tests/numerical_mc.rs, eigendecomposition.
 *   - CMG_FULL_V2 refinement ladder (full_cmg.rs:970-1088, 1375-1470): the
 *     reference's own known answers for refinement_pcg_options,
 *     failing_columns, euclidean_norm and the frozen factor schedule
 *     (full_cmg.rs:1985-2006, 2120-2139); its grid-and-hub fixture
 *     (full_cmg.rs:2008-2118), which must refine and then certify, against
 *     the closed-form solution (0, psi) of y = F psi.
 *   - Harmonic lift (cmg_impl.rs.in:459-500): brute-force weighted means of
 *     the auxiliary stars, and the closed form that the hybrid Laplacian on
 *     the lifted vector vanishes on auxiliary rows and equals the dense firm
 *     Schur complement on firm rows.
 */
#ifndef ORC_H
#define ORC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error transport (error.rs:8-47 codes) ---------------------------- */
enum {
    ORC_OK = 0,
    ORC_USER_BREAK = 1,
    ORC_UNSUPPORTED = 11,
    ORC_INVALID_INPUT = 20,
    ORC_INVALID_IDENTIFIER = 21,
    ORC_INVALID_WEIGHT = 22,
    ORC_INVALID_TARGET_WEIGHT = 23,
    ORC_GRAPH_EMPTY = 30,
    ORC_GRAPH_UNIDENTIFIED = 31,
    ORC_GRAPH_CERTIFICATE = 32,
    ORC_RESOURCE_LIMIT = 40,
    ORC_ALLOCATION = 41,
    ORC_CMG_SETUP = 50,
    ORC_CMG_APPLY = 51,
    ORC_PCG_CURVATURE = 60,
    ORC_PCG_PRECONDITIONER = 61,
    ORC_PCG_MAX_ITERATIONS = 63,
    ORC_FULL_RESIDUAL = 64,
    ORC_RNG_CONTRACT = 70,
    ORC_CORRECTION_NONFINITE = 71,
    ORC_ACCOUNTING = 72,
    ORC_JLA_MOMENT = 73,
    ORC_JLA_CONSTRAINT = 74,
    ORC_NONESTIMABLE = 75,
    ORC_BLOCK_INVERSE = 76,
    ORC_TARGET_CENTERING = 77,
    ORC_TARGET_IDENTITY = 78,
    ORC_NONFINITE_CORRECTED = 79,
    ORC_INVARIANT = 90
};

const char *orc_last_error(void);

/* ---- Counter-V1 RNG (rng.rs) ------------------------------------------ */
#define ORC_TAG_LEVERAGE 0x4c45564552414745ULL
#define ORC_TAG_TARGET 0x0054415247455401ULL

void orc_philox_block(uint64_t seed, uint64_t tag, uint64_t probe_block,
                      uint64_t entity, uint64_t word, uint64_t out[4]);
uint64_t orc_rng_word(uint64_t seed, uint64_t tag, uint64_t probe,
                      uint64_t entity, uint64_t word);
int orc_rademacher_sum(uint64_t seed, uint64_t tag, uint64_t probe,
                       uint64_t entity, uint64_t trials, int64_t *out);
int orc_rademacher_fill(uint64_t seed, uint64_t tag, uint64_t first_probe,
                        uint64_t columns, const uint64_t *entity,
                        const uint64_t *trials, uint64_t n, int64_t *out);

/* ---- sample pipeline (types.rs, problem.rs, graph.rs, jla.rs) ---------- */
typedef struct orc_problem orc_problem;

/* Canonicalize, select the match-deletion leave-out graph, compress and
 * build the no-control JLA plan.  Returns a code; *out on success. */
int orc_prepare(uint64_t rows, const uint64_t *worker, const uint64_t *firm,
                const uint64_t *deletion, const double *outcome,
                const uint64_t *frequency, const double *target,
                orc_problem **out);
int orc_prepare_masked(uint64_t rows, const uint64_t *worker, const uint64_t *firm,
                       const uint64_t *deletion, const double *outcome,
                       const uint64_t *frequency, const double *target, const uint8_t *mask,
                       orc_problem **out);
/* Observation-deletion graph selection (graph.rs:82-242) instead of the
 * match-deletion one; compression and plan as orc_prepare. */
int orc_prepare_observation(uint64_t rows, const uint64_t *worker, const uint64_t *firm,
                            const uint64_t *deletion, const double *outcome,
                            const uint64_t *frequency, const double *target, orc_problem **out);
/* The stayer hybrid (stayer_hybrid.rs:71-348): eligible stayer rows (1-based
 * dense retained-mover firm and stayer-worker indices) appended to the match
 * problem m; stayer rows become observation units. */
int orc_prepare_stayers(const orc_problem *m, uint64_t rows, const uint64_t *firm, const uint64_t *worker,
                        const double *outcome, const uint64_t *frequency, const double *target,
                        orc_problem **out);
void orc_problem_free(orc_problem *p);
/* Observation deletion: leverage atoms of `columns` probes from `first`, one
 * per stored row, out[col * rows + row] (generic_jla.rs:5660-5706). */
int orc_observation_atoms(const orc_problem *p, uint64_t seed, uint64_t first, uint64_t columns, int64_t *out);
/* dims[0..]: input_rows, retained_rows, retained_physical, workers, firms,
 * cells, deletion_units, target_strata_exact, plan_target_strata */
void orc_problem_dims(const orc_problem *p, uint64_t dims[9]);
double orc_problem_target_total(const orc_problem *p);
uint64_t orc_problem_topology_checksum(const orc_problem *p);
void orc_problem_mask(const orc_problem *p, uint8_t *mask);
/* which: 0 cell_worker(u32) 1 cell_firm(u32) 2 cell_weight(f64)
 *        3 unit_cell(u32) 4 unit_count(u64) 5 unit_rank(u64)
 *        6 stratum_cell(u32) 7 stratum_count(u64) 8 stratum_rank(u64)
 *        9 stratum_per_copy(f64) 10 cell_outcome_sum(f64)
 *        11 cell_target_sum(f64) 12 unit_outcome_sum(f64)
 *        13 unit_offset(u64; observation deletion, else zero) */
int orc_problem_array(const orc_problem *p, int which, void *dst);

/* ---- CMG (rust/vendor/cmg) --------------------------------------------- */
typedef struct orc_graph orc_graph;
typedef struct orc_cmg orc_cmg;

orc_graph *orc_graph_from_edges(uint64_t n, uint64_t m, const uint32_t *u,
                                const uint32_t *v, const double *w);
void orc_graph_free(orc_graph *g);
uint64_t orc_graph_vertices(const orc_graph *g);
uint64_t orc_graph_edges(const orc_graph *g);
void orc_graph_export(const orc_graph *g, uint32_t *u, uint32_t *v, double *w,
                      double *diag);
void orc_graph_matvec(const orc_graph *g, const double *x, double *y);

typedef struct {
    uint64_t direct_threshold;      /* 700 */
    uint64_t max_levels;            /* 32 (128 on the full-CMG route) */
    double max_hierarchy_nnz_factor; /* 5.0 */
    double low_degree_threshold;    /* 0.125 */
} orc_cmg_options;

int orc_cmg_build(const orc_graph *g, orc_cmg_options opt, orc_cmg **out);
void orc_cmg_free(orc_cmg *c);
uint64_t orc_cmg_levels(const orc_cmg *c);
/* info: n, edges, nnz, repeat(effective), terminal_reason(+1, 0 none), coarse */
void orc_cmg_level_info(const orc_cmg *c, uint64_t level, uint64_t info[6]);
int orc_cmg_level_labels(const orc_cmg *c, uint64_t level, uint32_t *labels);
const orc_graph *orc_cmg_level_graph(const orc_cmg *c, uint64_t level);
int orc_cmg_apply(const orc_cmg *c, const double *rhs, double *out, int project);

typedef struct {
    double relative_tolerance;
    double absolute_tolerance;
    uint64_t max_iterations;
    uint64_t recompute_interval;
} orc_pcg_options;

typedef struct {
    uint64_t iterations, restarts;
    double initial_residual_norm, residual_norm, relative_residual,
        backward_error, tolerance, rhs_projection_norm;
} orc_pcg_diag;

int orc_pcg_solve(const orc_graph *g, const orc_cmg *c, const double *rhs,
                  const double *guess, double *solution, orc_pcg_options opt,
                  orc_pcg_diag *diag);

/* frozen_bits.rs digests: family "chain", "chain_direct", "random", "grid";
 * kind "graph_from_edges", "build", "pcg", "pcg_guess". */
int orc_frozen_digest(const char *family, const char *kind, uint64_t *out);
/* sign of Rust's empty f64 sum (-0.0 since Rust 1.83); 1 = negative zero */
void orc_set_empty_sum_negative(int negative);

/* ---- hybrid graph + engine (cmg_impl.rs.in, operator.rs, engine.rs) ---- */
typedef struct {
    uint64_t seed;
    uint32_t probes;
    uint32_t leverage_width;
    uint32_t target_width;
    uint32_t threads;
    double fit_tolerance;      /* full-fit effective tolerance */
    double probe_tolerance;    /* probe effective tolerance */
    uint32_t max_iterations;
    uint32_t recompute_interval;
    double rank_tolerance;
    double block_tolerance;
} orc_engine_options;

typedef struct {
    double plugin[4], correction[4], corrected[4], mcse[4];
    double weighted_rss;
    uint64_t leverage_iterations, target_iterations, fit_iterations;
    double max_complete_residual;
    double solve_seconds;
} orc_engine_summary;

typedef struct orc_engine_result orc_engine_result;

int orc_hybrid_graph(const orc_problem *p, orc_graph **out,
                     uint64_t *auxiliary_count);
int orc_engine_run(const orc_problem *p, orc_engine_options opt,
                   orc_engine_result **out);
void orc_engine_result_free(orc_engine_result *r);
void orc_engine_result_summary(const orc_engine_result *r,
                               orc_engine_summary *s);
/* which: 0 fitted_cell 1 projection_share 2 residual_share 3 finite_bias
 *        4 finite_variance 5 residual_mass 6 deleted_mass
 *        7 cell_correction_weight 8 target_draws (4 per probe)
 *        9 leverage_iterations(per probe, as double) */
int orc_engine_result_array(const orc_engine_result *r, int which, double *dst,
                            uint64_t capacity);
/* As orc_engine_run, and it keeps every target probe's two FE solutions:
 * arrays 10 (2 R x workers) and 11 (2 R x firms), probe-major, worker-side
 * solve first. */
int orc_engine_run_targets(const orc_problem *p, orc_engine_options opt,
                           orc_engine_result **out);

/* ---- numerical MCSE attachment (engine/numerical.rs, numerical_mc.rs) --- */
typedef struct {
    /* 1 exact_zero, 2 ok_local, 3 ok_local_psd_adjusted, 4 unstable_nonpsd,
     * 5 nonsmooth_adjustment, 6 nonfinite_derivative, 7 replay_failed */
    uint32_t status, leverage_probes, target_probes, fold_a, fold_b;
    uint32_t failed_replay_probe, replay_error_code, reserved;
    uint64_t certified_replay_rhs, attempted_replay_rhs, replay_generator_words, point_rhs;
    double psd_adjustment, minimum_constrained, minimum_residual_margin, maximum_sensitivity_ratio;
    double score_means[6]; /* [fold][worker, firm, covariance] */
    double conditional[9], leverage[9], raw[9], usable[9], mcse[4];
    double max_point_complete_residual, max_replay_complete_residual;
} orc_numerical_result;

/* The engine run with the numerical attachment (engine.rs:1262-1545). */
int orc_engine_run_numerical(const orc_problem *p, orc_engine_options opt, orc_engine_result **out,
                             orc_numerical_result *num);
/* numerical_mc.rs pieces, exposed for the pins. */
int orc_finite_derivative(const double u[5], uint32_t r, double out[20]);
int orc_fd_observation(const double fd[20], double control, double grad[5]);
int orc_copy_pullback(const double x[4], double control, uint32_t r, double u[5], double fd[20], double *value,
                      double pb[4]);
int orc_cross_covariance(const double *a, const double *b, uint64_t n, double out[9]);
void orc_numerical_finalize(const double cond[9], const double lev[9], double raw[9], double usable[9],
                            double mcse[4], uint32_t *status, double *psd_adjustment);

/* ---- CMG_FULL_V2 batch solve pieces (full_cmg.rs, cmg_impl.rs.in) ------ */
/* refinement_pcg_options (full_cmg.rs:1375-1390) */
void orc_refinement_options(double pcg_tolerance, double factor, double gate, double smallest_norm,
                            double *relative, double *absolute);
/* REFINEMENT_FACTORS[round] (full_cmg.rs:37), 0 past the third round */
double orc_refinement_factor(uint32_t round);
/* failing_columns (full_cmg.rs:1445-1470); returns the count, indices in which */
uint64_t orc_failing_columns(const double *complete, const double *reduced, uint64_t k, double tolerance,
                             uint64_t *which);
/* euclidean_norm, scaled (full_cmg.rs:1392-1402) */
double orc_euclidean_norm(const double *v, uint64_t n);
/* lift_firm_into (cmg_impl.rs.in:459-500): firm values, then every auxiliary
 * vertex the weighted mean of its firm neighbours */
int orc_hybrid_lift(const orc_problem *p, const double *firm, double *lifted, uint64_t capacity);
/* solve_batch_with_options_and_interrupt (full_cmg.rs:809-1111) on k columns */
int orc_solve_columns(const orc_problem *p, const double *wr, const double *fr, uint64_t k, double tolerance,
                      double inner_ratio, uint32_t max_iterations, uint32_t recompute_interval, uint32_t threads,
                      double *worker_out, double *firm_out, double *complete, double *reduced,
                      uint64_t *iterations, uint64_t *refine_rounds, uint64_t *refined_columns);
/* Timed probe solves on an existing problem: solves `count` leverage probes
 * (first probe `first`) with `threads` host threads; returns seconds. */
int orc_time_probe_solves(const orc_problem *p, orc_engine_options opt,
                          uint32_t first, uint32_t count, double *seconds,
                          uint64_t *iterations);

#ifdef __cplusplus
}
#endif
#endif
