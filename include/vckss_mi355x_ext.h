/* SPDX-License-Identifier: MIT
 *
 * Extension and diagnostic entry points of libvckss_mi355x.so.
 *
 * These have no counterpart in the reference ABI.  They exist for callers
 * that tune the GPU batch shape and precision, and for the parity tests that
 * compare individual device stages (Counter-V1 atoms, level products, V-cycle,
 * PCG) and the host CMG hierarchy with the CPU oracle.  Status codes and the
 * thread-local error string follow the vckss_rust_* conventions.
 */
#ifndef VCKSS_MI355X_EXT_H
#define VCKSS_MI355X_EXT_H

#include <stdint.h>

#include "vckss_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* GPU execution shape for the next solve of `generation`:
 * precision 64 (FP64 V-cycle) or 32 (FP32 V-cycle, FP64 Krylov);
 * probes per leverage batch (1..64) and per target batch (1..32). */
int32_t vckss_mi355x_configure_v1(uint64_t generation, uint32_t precision,
                                  uint32_t leverage_width, uint32_t target_width);

/* dims: input_rows, retained_rows, physical, workers, firms, cells, units,
 * exact_strata, plan_strata, hybrid_vertices, hybrid_edges, auxiliary,
 * levels, terminal_dimension, topology_checksum */
int32_t vckss_mi355x_dims_v1(uint64_t generation, uint64_t *dims, uint32_t capacity);

/* Solved arrays of the last solve: 0 fitted_cell 1 projection_share
 * 2 residual_share 3 finite_bias 4 finite_variance 5 residual_mass
 * 6 deleted_mass 7 cell_correction_weight 8 target_draws (4 per probe)
 * 9 leverage iterations per probe 10 target iterations per column
 * 11 leverage complete residual per probe; 12 solve seconds (total,
 * leverage, target) and V-cycles applied. */
int32_t vckss_mi355x_array_v1(uint64_t generation, uint32_t which, double *out,
                              uint64_t capacity);

/* Counter-V1 atoms on the device: domain 0 leverage (deletion units),
 * 1 target (strata); out[col * entities + row]. */
int32_t vckss_mi355x_atoms_v1(uint64_t generation, uint64_t seed, uint32_t domain,
                              uint64_t first_probe, uint32_t columns, int64_t *out,
                              uint64_t capacity);

/* y = A_level x for `columns` interleaved lanes (FP64). */
int32_t vckss_mi355x_spmv_v1(uint64_t generation, uint32_t level, uint32_t columns,
                             const double *x, double *y);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP-graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The level product y = A x of vckss_mi355x_spmv_v1 on
 * caller-owned DEVICE memory: x and y hold n x columns doubles (original
 * vertex order, lanes interleaved, columns a power of two up to 64).  The
 * call allocates nothing: its working blocks (renumbered x and y and the
 * chunk partials of split rows, 2 n columns + chunks x columns doubles) live
 * in a fixed 32 MiB device scratch (4,194,304 doubles) of the library, one
 * per GPU for the life of the process; larger levels return 40.  The
 * generation must have been prepared with its device engine
 * (vckss_rust_engine_prepare_interrupt_v4).  x must be ready before the call
 * (synchronize the stream that wrote it); y is ready when it returns.  All
 * calls on one GPU share the scratch: calls that overlap on one GPU race. */
int32_t vckss_mi355x_spmv_noalloc_v1(uint64_t generation, uint32_t level, uint32_t columns, const double *x,
                                     double *y);

/* One V-cycle z = M r on the finest hybrid level. */
int32_t vckss_mi355x_vcycle_v1(uint64_t generation, uint32_t columns, uint32_t precision,
                               const double *r, double *z);

/* Certified lockstep PCG on the finest hybrid level; iterations[columns]. */
int32_t vckss_mi355x_pcg_v1(uint64_t generation, uint32_t columns, double tolerance,
                            uint32_t precision, const double *rhs, double *x,
                            uint32_t *iterations);

/* Average duration, algorithmic bytes per launch and working launches of a
 * timed kernel of the last vckss_mi355x_time_leverage_v1 run, from HIP events
 * on the engine stream (launches returning at the convergence gate excluded):
 * which 0 the finest A p of the Krylov iteration, 1 the finest prolong +
 * smooth pass of the V-cycle. */
int32_t vckss_mi355x_kernel_timing_v1(uint64_t generation, uint32_t which, double *milliseconds, double *bytes,
                                      uint64_t *launches);

/* Stop of the first PCG pass of every batch (no reference counterpart).
 * mode 1 (the default): the reference's certified stop
 * abs + rel (||b|| + 2 maxdeg ||x||) capped per column at
 * 0.1 x gate x ||b||, gate = max(10 tolerance, 1e-11): a heavy firm that
 * dominates the degree bound no longer lets the first pass stop above the
 * complete-residual gate, so batches meet the gate without a refinement
 * round and second extraction.  mode 0: the reference's stop alone (gate
 * failures are refined through the ladder of full_cmg.rs:970-1088).  mode 2
 * only queries.  previous (optional) receives the mode before the call.  The
 * gate and the refinement ladder apply in both modes. */
int32_t vckss_mi355x_first_pass_v1(uint64_t generation, uint32_t mode, uint32_t *previous);

/* One CMG_FULL_V2 batch solve (full_cmg.rs:809-1111,
 * solve_batch_with_options_and_interrupt) of 1..64 host columns: worker_rhs
 * (workers x columns) and firm_rhs (firms x columns), column-major, in the
 * prepared problem's dense worker and firm order.  The columns are solved in
 * lockstep to tolerance * inner_ratio, extracted, gated at
 * max(10 tolerance, 1e-11) on the reduced and complete residuals, and
 * refined through the 0.1 / 0.01 / 0.001 ladder.  Outputs: worker_out and
 * firm_out (same layout; firm part mean-normalised), receipts[5 x columns]
 * (iterations, restarts, reduced, complete residual, complete rhs norm) and
 * refine[2] (refinement rounds, refined columns). */
int32_t vckss_mi355x_solve_columns_v1(uint64_t generation, uint32_t columns, double tolerance, double inner_ratio,
                                      uint32_t precision, uint32_t maximum_iterations, uint32_t replacement_interval,
                                      const double *worker_rhs, const double *firm_rhs, double *worker_out,
                                      double *firm_out, double *receipts, uint64_t *refine);

/* Timed leverage batches (atoms, rhs, solve, extraction, moments) with the
 * dominant kernel's (the finest A p of the Krylov iteration) average duration
 * measured by HIP events on the engine stream. */
int32_t vckss_mi355x_time_leverage_v1(uint64_t generation, uint32_t width, uint32_t precision,
                                      double tolerance, uint32_t batches, double *seconds,
                                      uint64_t *solves, uint64_t *iterations,
                                      double *kernel_ms, double *kernel_bytes,
                                      uint64_t *kernel_launches);

/* Autotuner: times `batches` leverage batches (after one warm batch) for every
 * combination of probes per batch {16, 32, 64} and segment chunk size
 * {32, 64, 128} on this problem at `precision`, keeps the fastest device
 * layout (the upload shape, 64 and 64, unless another is 3% faster), and
 * configures the generation's batch widths (leverage = width,
 * target = width / 2).  trials (optional): (width, chunk, solves/s) triples. */
int32_t vckss_mi355x_autotune_v1(uint64_t generation, uint32_t precision, double tolerance, uint32_t batches,
                                 uint32_t *width, uint32_t *chunk, double *rate, double *trials,
                                 uint32_t trial_capacity);

/* Re-lays out the device hierarchy with segment chunks of at most `chunk`
 * entries (16..1024; 64 at upload).  Results change only by summation order. */
int32_t vckss_mi355x_set_chunk_v1(uint64_t generation, uint32_t chunk);

/* ---- Tuned preconditioner (no counterpart in the reference) -------------
 * Learned smoother weights: a message-passing network (8 node features, 2
 * layers of 32 channels) predicts a Jacobi weight omega_i in (0, 0.95) for
 * every vertex of every smoothed CMG level; the V-cycle then smooths with
 * omega_i / d_i where the reference uses 0.5 / d_i.  Parameters: 4481
 * doubles.  Cycle shape: the cycle may visit only level 0, the levels that at
 * least halve the last visited one, and the terminal, restricting through the
 * composite aggregation of the same hierarchy.
 *
 * mode 0: the reference cycle.  mode 1: a calibration batch of 16 leverage
 * probes to rel.res 1e-8 is solved with the reference cycle, the learned
 * weights (params, or the library's built-in trained weights when params is
 * NULL), the tuned shape and both; the configuration of least work
 * (iterations x level products per iteration) is kept, where learned weights
 * count only if they cut iterations against the plain weights on the same
 * shape and a shape only within 10% of the reference cycle's iterations.  A solve that breaks down with
 * it (status 60 or 61) is redone with the reference cycle, which then stays.
 * A new generation starts in mode 1 with the built-in weights when the
 * library carries them (calibrated on its first solve), else in mode 0. */
int32_t vckss_mi355x_smoother_v1(uint64_t generation, uint32_t mode, const double *params, uint32_t count);
/* out[10]: mode, learned weights in use, breakdown fallback taken, calibration
 * iterations with the reference cycle, with the chosen configuration, tuned
 * cycle shape in use, then the calibration iterations of the reference cycle,
 * the learned weights, the tuned shape and both (UINT32_MAX if not tried). */
int32_t vckss_mi355x_smoother_status_v1(uint64_t generation, uint32_t *out, uint32_t capacity);
/* Host-forced fallback check (no reference counterpart): every smoothed level
 * uses the constant smoother weight omega > 0 (inverse diagonal omega / d)
 * with no calibration, so the next solve runs with it; omega >= 1 makes the
 * V-cycle indefinite, PCG reports status 60 or 61 and the solve is redone
 * with the reference cycle (the fallback a solve with learned weights takes). */
int32_t vckss_mi355x_smoother_force_v1(uint64_t generation, double omega);
/* HIP-graph capture of the PCG iteration (no reference counterpart): enable
 * 1 captures the iteration body and the preconditioning sequence once per
 * configuration (cached by a fingerprint of every pointer, count and flag
 * their kernels read) and launches the two graphs per iteration; 0 launches
 * kernels one by one and frees the graphs; 2 only queries.  Timed runs
 * (`vckss_mi355x_time_leverage_v1`) always launch one by one.  captures and
 * launches (optional) count graph captures and graph launches so far. */
int32_t vckss_mi355x_graphs_v1(uint64_t generation, uint32_t enable, uint64_t *captures, uint64_t *launches);
/* Receipt of an exact-route solve (algorithm EXACT, or AUTO within
 * exact_estimator_limit; exact_estimator.rs:127-153): out[13] = identified
 * parameters, deletion units, cells, maximum block leverage, smallest block
 * maker eigenvalue, certified lower bound of the scaled quotient reciprocal
 * condition number, scaled inverse residual, complete fit residual and its
 * tolerance, firm zero-sum residual, weighted RSS, seconds for the
 * factorisation and fit, seconds for the cell pass.  11 if the result is a
 * JLA estimate. */
int32_t vckss_mi355x_exact_receipt_v1(uint64_t generation, double *out, uint32_t capacity);
/* Sample batches (no counterpart in the reference, whose caller runs one
 * sample per command; for callers with many independent samples, such as
 * region-year subsamples).  Rows carry a sample identifier (any exact binary64
 * integer); each sample is prepared as its own problem (the graph selection,
 * compression, plan and CMG hierarchy of vckss_rust_engine_prepare_v1, match
 * or observation deletion), on `threads` host threads.  *samples receives the
 * sample count; samples are numbered by ascending identifier.  A sample that
 * fails preparation fails the batch with its status. */
int32_t vckss_mi355x_batch_prepare_v1(uint64_t rows, const double *sample, const double *worker,
                                      const double *firm, const double *deletion, const double *outcome,
                                      const double *frequency, const double *target, uint32_t deletion_mode,
                                      uint32_t threads, uint64_t *handle, uint32_t *samples);
/* The JLA estimate of every sample (Counter-V1 seed, probes, probe and fit
 * tolerances, V-cycle precision 32 or 64; batches of up to 64 leverage and
 * 32 target probes), up to `streams` samples at once, each on its own HIP
 * stream; largest samples first.  Returns the status of the lowest-numbered
 * failing sample, after all have run. */
int32_t vckss_mi355x_batch_solve_v1(uint64_t handle, uint64_t seed, uint32_t probes, double tolerance,
                                    double fit_tolerance, uint32_t precision, uint32_t streams);
/* As v1, with `mode`: 0 runs up to `streams` samples at once, each alone on
 * its own stream (v1); 1 runs up to `streams` samples in lockstep: the PCG
 * calls of the samples waiting together run as one, each pass of the
 * lockstep certified PCG and its V-cycle a single launch over all of those
 * samples with per-sample block ranges, convergence gates and lane states.
 * Each sample's estimate is the same in both modes, bit for bit. */
int32_t vckss_mi355x_batch_solve_v2(uint64_t handle, uint64_t seed, uint32_t probes, double tolerance,
                                    double fit_tolerance, uint32_t precision, uint32_t streams, uint32_t mode);
/* out[0..16): plug-in, correction, corrected, MCSE (worker, firm, covariance,
 * total each); out[16..24): sample identifier, stored rows, workers, firms,
 * deletion units, solve seconds, PCG iterations, largest complete residual.
 * capacity >= 24. */
int32_t vckss_mi355x_batch_result_v1(uint64_t handle, uint32_t sample, double *out, uint32_t capacity);
int32_t vckss_mi355x_batch_release_v1(uint64_t handle);
/* The control fit of the last solve with controls (no counterpart in the
 * reference's C ABI; its generic-engine receipt holds the same quantities):
 * out[0] the control count q, out[1] 1 under the fixed-offset nuisance else
 * 0, out[2] the generalised reciprocal condition of the residualised control
 * information, out[3 .. 3 + q) the control coefficients.  capacity >= 3 + q;
 * 11 if the result has no controls. */
int32_t vckss_mi355x_controls_v1(uint64_t generation, double *out, uint32_t capacity);
/* Levels the cycle visits (all levels unless the tuned shape is in use). */
int32_t vckss_mi355x_cycle_shape_v1(uint64_t generation, uint32_t *levels, uint32_t capacity, uint32_t *count);
/* Built-in trained parameters (count = 0 when the library carries none). */
int32_t vckss_mi355x_smoother_params_v1(double *out, uint32_t capacity, uint32_t *count);
/* omega of every vertex of every smoothed level of the generation's hierarchy,
 * levels in order, original vertex order within a level. */
int32_t vckss_mi355x_smoother_omega_v1(uint64_t generation, const double *params, uint32_t count, double *omega,
                                       uint64_t capacity, uint64_t *written);
/* The same omega from the solve path's inference: BF16 weights and
 * activations with FP32 accumulation, layer products on the matrix cores. */
int32_t vckss_mi355x_smoother_infer_v1(uint64_t generation, const double *params, uint32_t count, double *omega,
                                       uint64_t capacity, uint64_t *written);
/* Training loss L = mean_j |e_j - M A e_j|_A^2 / |e_j|_A^2 over 32 error
 * columns (rows x 32, vertex-major, original hybrid-vertex order) and dL/dparams. */
int32_t vckss_mi355x_smoother_gradient_v1(uint64_t generation, const double *params, uint32_t count,
                                          const double *errors, uint64_t rows, double *loss, double *grad);
/* `steps` Adam steps (lr, beta 0.9 / 0.999) on this generation's hierarchy.
 * state: parameters, first moments, second moments (count = 3 x 4481),
 * updated in place; step0 = steps already taken.  Error columns come from
 * `seed` and `passes` plain-cycle error steps.  losses[steps]. */
int32_t vckss_mi355x_smoother_train_v1(uint64_t generation, double *state, uint32_t count, uint64_t step0,
                                       uint32_t steps, double learning_rate, uint64_t seed, uint32_t passes,
                                       double *losses);

/* Device counters of the generation since upload: V-cycles applied, solve
 * batches, refinement rounds, refined columns, PCG iterations launched
 * (including iterations queued after convergence, which return at once),
 * reserved. */
int32_t vckss_mi355x_counters_v1(uint64_t generation, uint64_t *out, uint32_t capacity);

/* Plain device-to-device copy of `bytes` through a streaming kernel
 * (k_stream_copy), for calibrating hardware traffic counters; returns the
 * kernel time in milliseconds. */
int32_t vckss_mi355x_stream_copy_v1(uint64_t generation, uint64_t bytes, double *milliseconds);

/* Host CMG hierarchy of an arbitrary graph (no GPU needed). */
int32_t vckss_mi355x_hierarchy_build_v1(uint64_t vertices, uint64_t edges, const uint32_t *u,
                                        const uint32_t *v, const double *w,
                                        uint64_t direct_threshold, uint64_t max_levels,
                                        uint64_t *handle);
/* info: n, edges, nnz, repeat, terminal (0 none, 1 direct, 2.. iterative), coarse */
int32_t vckss_mi355x_hierarchy_level_v1(uint64_t handle, uint32_t level, uint64_t *info,
                                        uint32_t *labels, uint32_t *eu, uint32_t *ev,
                                        double *ew, double *diag);
int32_t vckss_mi355x_hierarchy_levels_v1(uint64_t handle, uint32_t *levels, uint64_t *factor_nnz);
int32_t vckss_mi355x_hierarchy_free_v1(uint64_t handle);

/* Device layout of one hierarchy level, as uploaded: the renumbering
 * (perm: renumbered -> original), the renumbered CSR, inverse diagonal,
 * monotone coarse labels with contiguous aggregate ranges, and the chunk
 * tables (4 words per chunk: segment, begin, end, multi index or
 * 0xffffffff; 2 words per multi-chunk segment: first and end chunk).
 * sizes: n, nnz, coarse, row chunks, row multi, aggregate chunks, aggregate
 * multi, terminal dimension (direct terminal only, else 0).  Call with null
 * arrays to read the sizes first. */
int32_t vckss_mi355x_layout_v1(uint64_t handle, uint32_t level, uint32_t chunk, uint64_t *sizes,
                               uint32_t *perm, uint32_t *row_ptr, uint32_t *col, double *val,
                               double *inv, uint32_t *lab, uint32_t *agg_ptr, uint32_t *row_chunks,
                               uint32_t *row_multi, uint32_t *agg_chunks, uint32_t *agg_multi,
                               double *terminal_inverse);

/* Host only (no GPU): the chunk table of `segments` segments with pointers
 * ptr[0..segments] for chunks of at most `chunk` entries; with group > 1,
 * runs of up to `group` consecutive segments of at most chunk / 4 entries
 * share one chunk tagged 0xffffff00 + count.  sizes: chunks, multi-chunk
 * segments.  Call with null desc / mrange to read the sizes first. */
int32_t vckss_mi355x_chunk_table_v1(const uint32_t *ptr, uint64_t segments, uint32_t chunk, uint32_t group,
                                    uint64_t *sizes, uint32_t *desc, uint32_t *mrange);

/* Host preparation only (no GPU): same as the prepare entry point but stops
 * before device upload; the generation can be queried with dims/arrays of
 * the prepared problem (which 20..): 20 unit_rank 21 unit_count
 * 22 stratum_rank 23 stratum_count 24 cell_weight 25 unit_cell
 * 26 unit copy offset (observation deletion; zero under match deletion)
 * 27 the units of each cell, cells ascending (the device's gather list)
 * 28 the vertex count of every CMG level. */
int32_t vckss_mi355x_prepare_host_v1(const VckssEnginePrepareRequestInterruptV3 *request,
                                     const VckssEngineColumnsV3 *columns, uint64_t *output_handle);

#ifdef __cplusplus
}
#endif
#endif
