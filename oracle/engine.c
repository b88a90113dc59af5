/* Engine restatement (test infrastructure): two-way operator (operator.rs),
 * degree-4 hybrid Laplacian (cmg_impl.rs.in:78-363, 459-500, 550-570),
 * CMG_FULL_V2 batch solve with its complete-residual gate and warm-start
 * refinement ladder (full_cmg.rs:809-1111, 1158-1476) and the compressed
 * no-control match-deletion improved-JLA estimator (engine.rs:1060-1563,
 * 2218-3272, jla.rs:286-364). */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdio.h>
#include <time.h>

#include "orc_internal.h"
#include "sample.h"

#define ROUNDOFF_GATE (4096.0 * 2.220446049250313e-16)

typedef struct {
    double s, c;
} ssum; /* StableSum, engine.rs:327-352 */
static inline void ss_add(ssum *x, double v) { orc_neumaier(&x->s, &x->c, v); }
static inline double ss_fin(ssum x) { return x.s + x.c; }

/* ---------------- operator (operator.rs) ---------------- */

/* operator.rs:528-547 (center): Kahan compensated mean */
static void op_center(double *v, uint64_t n) {
    orc_kahan k = {0, 0};
    for (uint64_t i = 0; i < n; ++i) orc_kahan_add(&k, v[i]);
    double mean = k.sum / (double)n;
    for (uint64_t i = 0; i < n; ++i) v[i] -= mean;
}

/* operator.rs:550-566 (stable_dot, stable_norm) */
static double stable_dot(const double *a, const double *b, uint64_t n) {
    orc_kahan k = {0, 0};
    for (uint64_t i = 0; i < n; ++i) orc_kahan_add(&k, a[i] * b[i]);
    return k.sum;
}
static double stable_norm(const double *a, uint64_t n) {
    return sqrt(fmax(stable_dot(a, a, n), 0.0));
}

typedef struct {
    const orc_problem *p;
    double *wdiag, *fdiag;
    uint64_t n; /* hybrid vertices */
    orc_graph *graph;
    orc_cmg *cmg;
    /* auxiliary incidence for the harmonic lift (cmg_impl.rs.in:459-500) */
    uint64_t naux;
    uint64_t *aux_ptr;
    uint32_t *aux_nbr;
    double *aux_w;
} solver_t;

/* operator.rs:269-304 (schur_rhs) */
static void schur_rhs(const solver_t *s, const double *wr, const double *fr, double *out,
                      double *scratch_w) {
    const orc_problem *p = s->p;
    for (uint64_t w = 0; w < p->workers; ++w) scratch_w[w] = wr[w] / s->wdiag[w];
    memcpy(out, fr, p->firms * 8);
    for (uint64_t c = 0; c < p->cells; ++c)
        out[p->cell_firm[c]] -= p->cell_weight[c] * scratch_w[p->cell_worker[c]];
    op_center(out, p->firms);
}

/* operator.rs:210-262 (apply_full_schur) */
static void full_schur(const solver_t *s, const double *firm, double *out, double *wsum) {
    const orc_problem *p = s->p;
    memset(wsum, 0, p->workers * 8);
    for (uint64_t c = 0; c < p->cells; ++c)
        wsum[p->cell_worker[c]] += p->cell_weight[c] * firm[p->cell_firm[c]];
    for (uint64_t w = 0; w < p->workers; ++w) wsum[w] /= s->wdiag[w];
    for (uint64_t f = 0; f < p->firms; ++f) out[f] = s->fdiag[f] * firm[f];
    for (uint64_t c = 0; c < p->cells; ++c)
        out[p->cell_firm[c]] -= p->cell_weight[c] * wsum[p->cell_worker[c]];
}

/* full_cmg.rs:1425-1443 (reduced_relative_residual) via operator.rs:500-515 */
static double reduced_residual(const solver_t *s, const double *reduced, const double *rhs) {
    const orc_problem *p = s->p;
    uint64_t F = p->firms;
    double *firm = orc_alloc(F, 8), *act = orc_alloc(F, 8), *wsum = orc_alloc(p->workers, 8);
    memcpy(firm, reduced, F * 8);
    op_center(firm, F);
    full_schur(s, firm, act, wsum);
    op_center(act, F);
    for (uint64_t f = 0; f < F; ++f) act[f] = rhs[f] - act[f];
    double r = stable_norm(act, F), scale = stable_norm(rhs, F);
    free(firm); free(act); free(wsum);
    return scale == 0.0 ? r : r / scale;
}

/* operator.rs:311-343 (reconstruct_worker) */
static void reconstruct_worker(const solver_t *s, const double *wr, const double *firm,
                               double *worker) {
    const orc_problem *p = s->p;
    memcpy(worker, wr, p->workers * 8);
    for (uint64_t c = 0; c < p->cells; ++c)
        worker[p->cell_worker[c]] -= p->cell_weight[c] * firm[p->cell_firm[c]];
    for (uint64_t w = 0; w < p->workers; ++w) worker[w] /= s->wdiag[w];
}

/* operator.rs:386-458 (full_residual) */
static double full_residual(const solver_t *s, const double *worker, const double *firm,
                            const double *wr, const double *fr, double *rhs_norm_out) {
    const orc_problem *p = s->p;
    double *rw = orc_alloc(p->workers, 8), *rf = orc_alloc(p->firms, 8);
    for (uint64_t w = 0; w < p->workers; ++w) rw[w] = wr[w] - s->wdiag[w] * worker[w];
    for (uint64_t f = 0; f < p->firms; ++f) rf[f] = fr[f] - s->fdiag[f] * firm[f];
    for (uint64_t c = 0; c < p->cells; ++c) {
        uint32_t w = p->cell_worker[c], f = p->cell_firm[c];
        double weight = p->cell_weight[c];
        rw[w] -= weight * firm[f];
        rf[f] -= weight * worker[w];
    }
    double absn = sqrt(fmax(stable_dot(rw, rw, p->workers) + stable_dot(rf, rf, p->firms), 0.0));
    double rhsn = sqrt(fmax(stable_dot(wr, wr, p->workers) + stable_dot(fr, fr, p->firms), 0.0));
    free(rw);
    free(rf);
    if (rhs_norm_out) *rhs_norm_out = rhsn;
    return rhsn == 0.0 ? absn : absn / rhsn;
}

/* ---------------- hybrid graph (cmg_impl.rs.in) ---------------- */

typedef struct {
    uint32_t u, v;
    double w;
} wedge;

typedef struct {
    const orc_problem *p;
    uint64_t *primary;
} firm_sig_ctx;
/* cmg_impl.rs.in:573-632: firms ranked by their sorted (worker, weight bits) cells */
static int cmp_firm_signature(const void *a, const void *b, void *arg) {
    const orc_problem *p = ((const firm_sig_ctx *)arg)->p;
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    uint64_t i = p->firm_ptr[x], ie = p->firm_ptr[x + 1], j = p->firm_ptr[y], je = p->firm_ptr[y + 1];
    for (; i < ie && j < je; ++i, ++j) {
        uint64_t ci = p->firm_cells[i], cj = p->firm_cells[j];
        if (p->cell_worker[ci] != p->cell_worker[cj]) return p->cell_worker[ci] < p->cell_worker[cj] ? -1 : 1;
        uint64_t bi = orc_bits(p->cell_weight[ci]), bj = orc_bits(p->cell_weight[cj]);
        if (bi != bj) return bi < bj ? -1 : 1;
    }
    if (i < ie) return 1;
    if (j < je) return -1;
    return 0;
}

typedef struct {
    uint32_t nbr, edge;
    double w;
} aux_arc;
static uint64_t *g_firm_primary; /* sorting context for aux incidence */
static int cmp_aux_arc(const void *a, const void *b) {
    const aux_arc *x = a, *y = b;
    int c = orc_total_cmp(y->w, x->w); /* weight descending */
    if (c) return c;
    if (g_firm_primary[x->nbr] != g_firm_primary[y->nbr])
        return g_firm_primary[x->nbr] < g_firm_primary[y->nbr] ? -1 : 1;
    return x->edge < y->edge ? -1 : (x->edge > y->edge);
}

typedef struct {
    uint64_t key, idx;
} ki;
static int cmp_ki(const void *a, const void *b) {
    const ki *x = a, *y = b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* cmg_impl.rs.in:78-363 with direct_degree = 4 */
static int build_hybrid(solver_t *s) {
    const orc_problem *p = s->p;
    uint64_t F = p->firms, W = p->workers;
    uint64_t naux = 0;
    for (uint64_t w = 0; w < W; ++w) naux += (p->worker_ptr[w + 1] - p->worker_ptr[w]) > 4;
    uint64_t n = F + naux;
    wedge *raw = NULL;
    uint64_t m = 0, cap = 0;
    uint64_t aux = 0;
    for (uint64_t w = 0; w < W; ++w) {
        uint64_t b = p->worker_ptr[w], e = p->worker_ptr[w + 1], deg = e - b;
        /* cells of a worker are already in firm order */
        double mass = 0.0;
        for (uint64_t c = b; c < e; ++c) mass += p->cell_weight[c];
        if (deg <= 1) continue;
        uint64_t need = deg <= 4 ? deg * (deg - 1) / 2 : deg;
        if (m + need > cap) {
            cap = (m + need) * 2 + 1024;
            raw = realloc(raw, cap * sizeof *raw);
        }
        if (deg <= 4) {
            for (uint64_t l = b; l < e; ++l)
                for (uint64_t r = l + 1; r < e; ++r) {
                    raw[m].u = p->cell_firm[l];
                    raw[m].v = p->cell_firm[r];
                    raw[m].w = p->cell_weight[l] * p->cell_weight[r] / mass;
                    ++m;
                }
        } else {
            uint32_t av = (uint32_t)(F + aux++);
            for (uint64_t c = b; c < e; ++c) {
                raw[m].u = p->cell_firm[c];
                raw[m].v = av;
                raw[m].w = p->cell_weight[c];
                ++m;
            }
        }
    }
    /* BTreeMap accumulation: stable by key, contributions summed in insertion
     * order (cmg_impl.rs.in:634-665).  A stable merge sort keeps that order. */
    for (uint64_t i = 0; i < m; ++i)
        if (raw[i].u > raw[i].v) {
            uint32_t t = raw[i].u;
            raw[i].u = raw[i].v;
            raw[i].v = t;
        }
    /* stable: sort (key, original index) */
    ki *order = orc_alloc(m, sizeof *order);
    for (uint64_t i = 0; i < m; ++i) {
        order[i].key = ((uint64_t)raw[i].u << 32) | raw[i].v;
        order[i].idx = i;
    }
    qsort(order, m, sizeof *order, cmp_ki);
    uint32_t *eu = orc_alloc(m, 4), *ev = orc_alloc(m, 4);
    double *ew = orc_alloc(m, 8);
    uint64_t me = 0;
    for (uint64_t i = 0; i < m;) {
        uint64_t key = order[i].key;
        double sum = raw[order[i].idx].w;
        ++i;
        while (i < m && order[i].key == key) sum = sum + raw[order[i++].idx].w;
        eu[me] = (uint32_t)(key >> 32);
        ev[me] = (uint32_t)key;
        ew[me] = sum;
        ++me;
    }
    free(order);
    free(raw);
    s->n = n;
    s->graph = orc_graph_from_edges(n, me, eu, ev, ew);
    if (!s->graph) {
        free(eu); free(ev); free(ew);
        return ORC_CMG_SETUP;
    }
    /* aux incidence sorted by (weight desc, neighbour key, edge) */
    s->naux = naux;
    s->aux_ptr = orc_alloc(naux + 1, 8);
    for (uint64_t e = 0; e < me; ++e)
        if (ev[e] >= F) s->aux_ptr[ev[e] - F + 1]++;
    for (uint64_t a = 0; a < naux; ++a) s->aux_ptr[a + 1] += s->aux_ptr[a];
    aux_arc *arcs = orc_alloc(s->aux_ptr[naux], sizeof *arcs);
    uint64_t *cur = orc_alloc(naux + 1, 8);
    memcpy(cur, s->aux_ptr, naux * 8);
    for (uint64_t e = 0; e < me; ++e)
        if (ev[e] >= F) {
            aux_arc *a = &arcs[cur[ev[e] - F]++];
            a->nbr = eu[e];
            a->edge = (uint32_t)e;
            a->w = ew[e];
        }
    free(cur);
    uint64_t *primary = orc_alloc(F, 8);
    {
        uint64_t *fo = orc_alloc(F, 8);
        for (uint64_t f = 0; f < F; ++f) fo[f] = f;
        firm_sig_ctx ctx = {p, NULL};
        qsort_r(fo, F, 8, cmp_firm_signature, &ctx);
        uint64_t rank = 0;
        for (uint64_t i = 0; i < F; ++i) {
            if (i == 0 || cmp_firm_signature(&fo[i - 1], &fo[i], &ctx) != 0) ++rank;
            primary[fo[i]] = rank;
        }
        free(fo);
    }
    g_firm_primary = primary;
    for (uint64_t a = 0; a < naux; ++a)
        qsort(arcs + s->aux_ptr[a], s->aux_ptr[a + 1] - s->aux_ptr[a], sizeof *arcs, cmp_aux_arc);
    g_firm_primary = NULL;
    free(primary);
    s->aux_nbr = orc_alloc(s->aux_ptr[naux], 4);
    s->aux_w = orc_alloc(s->aux_ptr[naux], 8);
    for (uint64_t i = 0; i < s->aux_ptr[naux]; ++i) {
        s->aux_nbr[i] = arcs[i].nbr;
        s->aux_w[i] = arcs[i].w;
    }
    free(arcs);
    free(eu); free(ev); free(ew);
    return ORC_OK;
}

int orc_hybrid_graph(const orc_problem *p, orc_graph **out, uint64_t *auxiliary_count) {
    solver_t s = {0};
    s.p = p;
    int rc = build_hybrid(&s);
    free(s.aux_ptr); free(s.aux_nbr); free(s.aux_w);
    if (rc) return rc;
    *out = s.graph;
    if (auxiliary_count) *auxiliary_count = s.naux;
    return ORC_OK;
}

/* cmg_impl.rs.in:459-500 (lift_firm_into) */
static void lift_firm(const solver_t *s, const double *firm, double *lifted) {
    uint64_t F = s->p->firms;
    memset(lifted, 0, s->n * 8);
    memcpy(lifted, firm, F * 8);
    for (uint64_t a = 0; a < s->naux; ++a) {
        double num = 0.0, den = 0.0;
        for (uint64_t i = s->aux_ptr[a]; i < s->aux_ptr[a + 1]; ++i) {
            num += s->aux_w[i] * lifted[s->aux_nbr[i]];
            den += s->aux_w[i];
        }
        lifted[F + a] = num / den;
    }
}

static int solver_init(solver_t *s, const orc_problem *p) {
    memset(s, 0, sizeof *s);
    s->p = p;
    s->wdiag = orc_alloc(p->workers, 8);
    s->fdiag = orc_alloc(p->firms, 8);
    for (uint64_t c = 0; c < p->cells; ++c) {
        s->wdiag[p->cell_worker[c]] += p->cell_weight[c];
        s->fdiag[p->cell_firm[c]] += p->cell_weight[c];
    }
    int rc = build_hybrid(s);
    if (rc) return rc;
    /* full_cmg.rs:467-476: CMG options with max_levels 128 */
    orc_cmg_options opt = {700, 128, 5.0, 0.125};
    return orc_cmg_build(s->graph, opt, &s->cmg);
}

static void solver_free(solver_t *s) {
    free(s->wdiag); free(s->fdiag);
    orc_graph_free(s->graph);
    orc_cmg_free(s->cmg);
    free(s->aux_ptr); free(s->aux_nbr); free(s->aux_w);
}

/* ---------------- CMG_FULL_V2 batch solve (full_cmg.rs) ---------------- */

typedef struct {
    double *worker, *firm;
    uint64_t iterations, restarts;
    double reduced, complete, rhs_norm;
} column_t;

typedef struct {
    const solver_t *s;
    const double *wr, *fr;   /* k columns */
    const double *rhs;       /* hybrid rhs per column (n each) */
    const double *guess;     /* optional hybrid warm start per column */
    const uint64_t *which;   /* column indices to solve */
    uint64_t count;
    orc_pcg_options opt;
    column_t *cols;
    int *rc;
    uint64_t next;
    pthread_mutex_t lock;
} batch_job;

/* One column: PCG on the hybrid Laplacian, then full_cmg.rs:1241-1340
 * (finish_solved_column / extract_solved_column). */
static int solve_one(const batch_job *job, uint64_t k) {
    const solver_t *s = job->s;
    const orc_problem *p = s->p;
    uint64_t col = job->which[k], n = s->n, F = p->firms, W = p->workers;
    const double *rhs = job->rhs + k * n;
    double *x = orc_alloc(n, 8);
    orc_pcg_diag d;
    int rc = orc_pcg_solve(s->graph, s->cmg, rhs, job->guess ? job->guess + k * n : NULL, x,
                           job->opt, &d);
    if (rc) {
        free(x);
        return rc;
    }
    column_t *out = &job->cols[col];
    /* cmg_impl.rs.in:550-570 (normalize_firm_mean): plain sum */
    double sum = 0.0;
    for (uint64_t f = 0; f < F; ++f) sum += x[f];
    double mean = sum / (double)F;
    for (uint64_t i = 0; i < n; ++i) x[i] -= mean;
    memcpy(out->firm, x, F * 8);
    double *reduced = orc_alloc(F, 8);
    memcpy(reduced, x, F * 8);
    op_center(reduced, F);
    out->reduced = reduced_residual(s, reduced, rhs);
    reconstruct_worker(s, job->wr + col * W, out->firm, out->worker);
    out->complete = full_residual(s, out->worker, out->firm, job->wr + col * W, job->fr + col * F,
                                  &out->rhs_norm);
    out->iterations += d.iterations;
    out->restarts += d.restarts;
    free(reduced);
    free(x);
    return ORC_OK;
}

static void *batch_worker(void *arg) {
    batch_job *job = arg;
    for (;;) {
        pthread_mutex_lock(&job->lock);
        uint64_t k = job->next++;
        pthread_mutex_unlock(&job->lock);
        if (k >= job->count) break;
        job->rc[k] = solve_one(job, k);
    }
    return NULL;
}

static int run_batch(batch_job *job, uint32_t threads) {
    job->next = 0;
    pthread_mutex_init(&job->lock, NULL);
    uint32_t t = threads < 1 ? 1 : threads;
    if (t > job->count) t = (uint32_t)job->count;
    if (t <= 1) {
        batch_worker(job);
    } else {
        pthread_t *tid = orc_alloc(t, sizeof *tid);
        for (uint32_t i = 0; i < t; ++i) pthread_create(&tid[i], NULL, batch_worker, job);
        for (uint32_t i = 0; i < t; ++i) pthread_join(tid[i], NULL);
        free(tid);
    }
    pthread_mutex_destroy(&job->lock);
    for (uint64_t k = 0; k < job->count; ++k)
        if (job->rc[k]) return job->rc[k];
    return ORC_OK;
}

/* full_cmg.rs:1392-1402 (euclidean_norm, scaled) */
double orc_euclidean_norm(const double *v, uint64_t n) {
    double scale = 0.0;
    for (uint64_t i = 0; i < n; ++i) scale = fmax(scale, fabs(v[i]));
    if (scale == 0.0 || !isfinite(scale)) return scale;
    double s = 0.0;
    for (uint64_t i = 0; i < n; ++i) s += (v[i] / scale) * (v[i] / scale);
    return scale * sqrt(s);
}

/* full_cmg.rs:1375-1390 (refinement_pcg_options): the round's PCG stops at
 * the absolute residual factor * gate * (smallest failing rhs norm); when
 * that target is not a positive finite number, at the relative tolerance
 * pcg_tolerance * factor instead. */
void orc_refinement_options(double pcg_tolerance, double factor, double gate, double smallest_norm,
                            double *relative, double *absolute) {
    double target = factor * gate * smallest_norm;
    if (isfinite(target) && target > 0.0) {
        *relative = 0.0;
        *absolute = target;
    } else {
        *relative = pcg_tolerance * factor;
        *absolute = 0.0;
    }
}

/* full_cmg.rs:1445-1470 (failing_columns): ascending columns whose complete
 * or reduced residual is nonfinite or above the tolerance. */
uint64_t orc_failing_columns(const double *complete, const double *reduced, uint64_t k, double tolerance,
                             uint64_t *which) {
    uint64_t count = 0;
    for (uint64_t c = 0; c < k; ++c)
        if (!isfinite(complete[c]) || complete[c] > tolerance || !isfinite(reduced[c]) || reduced[c] > tolerance)
            which[count++] = c;
    return count;
}

static const double kRefinementFactors[3] = {0.1, 0.01, 0.001}; /* full_cmg.rs:37 */
double orc_refinement_factor(uint32_t round) { return round < 3 ? kRefinementFactors[round] : 0.0; }

/* full_cmg.rs:809-1111: k columns of (worker_rhs, firm_rhs) -> solutions.
 * rounds/refined (optional) count the refinement rounds and refined columns. */
static int solve_batch_counted(const solver_t *s, const double *wr, const double *fr, uint64_t k,
                               double effective_tol, double inner_ratio, const orc_engine_options *eo,
                               column_t *cols, uint64_t *rounds, uint64_t *refined) {
    const orc_problem *p = s->p;
    uint64_t n = s->n, F = p->firms, W = p->workers;
    double gate = fmax(10.0 * effective_tol, 1e-11);
    double *rhs = orc_alloc(k * n, 8), *scratch = orc_alloc(W, 8);
    for (uint64_t c = 0; c < k; ++c) {
        schur_rhs(s, wr + c * W, fr + c * F, rhs + c * n, scratch);
        cols[c].iterations = cols[c].restarts = 0;
    }
    free(scratch);
    uint64_t *which = orc_alloc(k, 8);
    int *rcs = orc_alloc(k, sizeof(int));
    for (uint64_t c = 0; c < k; ++c) which[c] = c;
    batch_job job = {0};
    job.s = s; job.wr = wr; job.fr = fr; job.rhs = rhs; job.which = which; job.count = k;
    job.opt.relative_tolerance = effective_tol * inner_ratio;
    job.opt.absolute_tolerance = 0.0;
    job.opt.max_iterations = eo->max_iterations;
    job.opt.recompute_interval = eo->recompute_interval;
    job.cols = cols; job.rc = rcs;
    int rc = run_batch(&job, eo->threads);
    double *complete = orc_alloc(k, 8), *reduced = orc_alloc(k, 8);
    for (int round = 0; rc == ORC_OK && round < 3; ++round) {
        for (uint64_t c = 0; c < k; ++c) {
            complete[c] = cols[c].complete;
            reduced[c] = cols[c].reduced;
        }
        uint64_t failing = orc_failing_columns(complete, reduced, k, gate, which);
        if (!failing) break;
        if (rounds) ++*rounds;
        if (refined) *refined += failing;
        /* full_cmg.rs:970-1088: warm start from the lifted firm solution */
        double *rrhs = orc_alloc(failing * n, 8), *guess = orc_alloc(failing * n, 8);
        double smallest = INFINITY;
        for (uint64_t i = 0; i < failing; ++i) {
            memcpy(rrhs + i * n, rhs + which[i] * n, n * 8);
            lift_firm(s, cols[which[i]].firm, guess + i * n);
            smallest = fmin(smallest, orc_euclidean_norm(rrhs + i * n, n));
        }
        batch_job rj = job;
        rj.rhs = rrhs; rj.guess = guess; rj.count = failing;
        orc_refinement_options(effective_tol * inner_ratio, kRefinementFactors[round], gate, smallest,
                               &rj.opt.relative_tolerance, &rj.opt.absolute_tolerance);
        rc = run_batch(&rj, eo->threads);
        free(rrhs);
        free(guess);
    }
    if (rc == ORC_OK)
        for (uint64_t c = 0; c < k; ++c)
            if (!isfinite(cols[c].complete) || cols[c].complete > gate ||
                !isfinite(cols[c].reduced) || cols[c].reduced > gate) {
                rc = orc_fail(ORC_FULL_RESIDUAL, "cmg_full_v2_refinement",
                              "column %llu: reduced %g complete %g exceeds %g",
                              (unsigned long long)c, cols[c].reduced, cols[c].complete, gate);
                break;
            }
    free(rhs);
    free(which);
    free(rcs);
    free(complete);
    free(reduced);
    return rc;
}

static int solve_batch(const solver_t *s, const double *wr, const double *fr, uint64_t k,
                       double effective_tol, double inner_ratio, const orc_engine_options *eo,
                       column_t *cols) {
    return solve_batch_counted(s, wr, fr, k, effective_tol, inner_ratio, eo, cols, NULL, NULL);
}

/* ---------------- estimator (engine.rs) ---------------- */

struct orc_engine_result {
    orc_engine_summary sum;
    uint64_t cells, units, probes;
    double *fitted_cell, *projection_share, *residual_share, *finite_bias, *finite_variance,
        *residual_mass, *deleted_mass, *cell_weight, *draws, *lev_iters;
    /* optional: every target probe's two FE solutions (worker side, firm
     * side), probe-major: target_worker[(2 t + side) W + w] */
    double *target_worker, *target_firm;
    uint64_t workers, firms;
};

typedef struct {
    ssum proj, res, proj4, res4, mixed;
} five_moments; /* engine.rs:354-374 */

static void five_add(five_moments *m, double projection, double residual) {
    double ps = projection * projection, rs = residual * residual;
    ss_add(&m->proj, ps);
    ss_add(&m->res, rs);
    ss_add(&m->proj4, ps * ps);
    ss_add(&m->res4, rs * rs);
    ss_add(&m->mixed, ps * rs);
}

/* jla.rs:232-269 (verify_accounting) */
static int verify_accounting(const double c[4], double tol) {
    double implied = c[0] + c[1] + 2.0 * c[2];
    double scale = fmax(fmax(fabs(c[0]), fabs(c[1])), fmax(fabs(2.0 * c[2]), fmax(fabs(c[3]), 1.0)));
    if (fabs(c[3] - implied) > tol * scale)
        return orc_fail(ORC_ACCOUNTING, "jla_components", "accounting identity failed");
    return ORC_OK;
}

/* jla.rs:286-364 (plugin_components) */
static int plugin_components(const orc_problem *p, const double *alpha, const double *psi,
                             double out[4]) {
    ssum wm = {0, 0}, fm = {0, 0};
    for (uint64_t r = 0; r < p->rows; ++r) {
        double weight = p->target[r];
        ss_add(&wm, weight * alpha[p->row_worker[r]]);
        ss_add(&fm, weight * psi[p->row_firm[r]]);
    }
    double wmean = ss_fin(wm) / p->target_total, fmean = ss_fin(fm) / p->target_total;
    ssum ws = {0, 0}, fs = {0, 0}, cv = {0, 0}, ts = {0, 0};
    for (uint64_t r = 0; r < p->rows; ++r) {
        double weight = p->target[r];
        double cw = alpha[p->row_worker[r]] - wmean, cf = psi[p->row_firm[r]] - fmean;
        ss_add(&ws, weight * cw * cw);
        ss_add(&fs, weight * cf * cf);
        ss_add(&cv, weight * cw * cf);
        double ct = cw + cf;
        ss_add(&ts, weight * ct * ct);
    }
    out[0] = ss_fin(ws) / p->target_total;
    out[1] = ss_fin(fs) / p->target_total;
    out[2] = ss_fin(cv) / p->target_total;
    out[3] = ss_fin(ts) / p->target_total;
    return verify_accounting(out, 1.0e-11);
}

/* engine.rs:2842-2881 (balance_score) */
static int balance_score(double *score, uint64_t n, double reference) {
    double original = score[n - 1];
    ssum pre = {0, 0};
    for (uint64_t i = 0; i + 1 < n; ++i) ss_add(&pre, score[i]);
    double last = -ss_fin(pre);
    if ((reference == 0.0 && last != original) ||
        (reference > 0.0 && fabs(last - original) > ROUNDOFF_GATE * reference))
        return orc_fail(ORC_TARGET_CENTERING, "jla_target", "compatibility repair exceeded roundoff");
    score[n - 1] = last;
    return ORC_OK;
}

static double now_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static column_t *alloc_cols(const orc_problem *p, uint64_t k) {
    column_t *c = orc_alloc(k, sizeof *c);
    for (uint64_t i = 0; i < k; ++i) {
        c[i].worker = orc_alloc(p->workers, 8);
        c[i].firm = orc_alloc(p->firms, 8);
    }
    return c;
}
static void free_cols(column_t *c, uint64_t k) {
    for (uint64_t i = 0; i < k; ++i) {
        free(c[i].worker);
        free(c[i].firm);
    }
    free(c);
}

/* engine.rs:2562-2619 (leverage_rhs): exact integer scatter */
static void leverage_rhs(const orc_problem *p, const int64_t *atoms, double *wr, double *fr) {
    memset(wr, 0, p->workers * 8);
    memset(fr, 0, p->firms * 8);
    for (uint64_t g = 0; g < p->units; ++g) {
        uint32_t c = p->unit_cell[g];
        double a = (double)atoms[g];
        wr[p->cell_worker[c]] += a;
        fr[p->cell_firm[c]] += a;
    }
}

/* engine.rs:2621-2703 + 2799-2971: centred target direction, transpose and
 * balance; column 0 = (score_w, 0), column 1 = (0, score_f) */
static int target_columns(const orc_problem *p, const int64_t *atoms, uint64_t probe, double *wr,
                          double *fr) {
    (void)probe;
    uint64_t C = p->cells, W = p->workers, F = p->firms;
    ssum *first = orc_alloc(C, sizeof *first);
    for (uint64_t s = 0; s < p->strata; ++s) {
        double scale = sqrt(p->stratum_per_copy[s] / p->target_total);
        ss_add(&first[p->stratum_cell[s]], scale * (double)atoms[s]);
    }
    double *d = orc_alloc(C, 8);
    for (uint64_t c = 0; c < C; ++c) d[c] = ss_fin(first[c]);
    free(first);
    ssum tot = {0, 0}, ab = {0, 0};
    for (uint64_t c = 0; c < C; ++c) {
        ss_add(&tot, d[c]);
        ss_add(&ab, fabs(d[c]));
    }
    double total = ss_fin(tot), reference = ss_fin(ab) + fabs(total);
    for (uint64_t c = 0; c < C; ++c) d[c] = d[c] - p->cell_target_sum[c] / p->target_total * total;
    double *sw = orc_alloc(W, 8), *sf = orc_alloc(F, 8);
    for (uint64_t c = 0; c < C; ++c) {
        sw[p->cell_worker[c]] += d[c];
        sf[p->cell_firm[c]] += d[c];
    }
    free(d);
    int rc = balance_score(sw, W, reference);
    if (!rc) rc = balance_score(sf, F, reference);
    memset(wr, 0, 2 * W * 8);
    memset(fr, 0, 2 * F * 8);
    memcpy(wr, sw, W * 8);
    memcpy(fr + F, sf, F * 8);
    free(sw);
    free(sf);
    return rc;
}

/* numerical attachment state (engine/numerical.rs:49-122) */
typedef struct {
    double *grad;  /* 5 per unit (observation units: g2, g4, offset) */
    double *offs;  /* per unit */
    ssum *folds;   /* 6 per unit: [fold][sw^2, sf^2, sw sf] */
    uint32_t status;
    /* observation units (generic_jla/numerical.rs:178-270): copy pullbacks
     * (q1, q3) per physical copy, the units in (class rank, row) order, the
     * copy ranges and the Counter-V1 words per replay column */
    double *copy;
    uint64_t *order;
    const uint64_t *phys;
    uint64_t class_words;
    double *ograd;  /* 3 per observation unit: g2, g4, offset */
    ssum *ofolds;   /* 6 per observation unit */
} num_state;

static int leverage_atoms(const orc_problem *p, uint64_t seed, uint64_t first, uint64_t k, int64_t *atoms);
static int copy_sign(uint64_t seed, uint64_t probe, uint64_t entity, uint64_t copy);
static uint64_t observation_row(const orc_problem *p, uint64_t g);

static int engine_run(const orc_problem *p, orc_engine_options opt, orc_engine_result **out,
                      orc_numerical_result *num, int keep_targets);

int orc_engine_run(const orc_problem *p, orc_engine_options opt, orc_engine_result **out) {
    return engine_run(p, opt, out, NULL, 0);
}

int orc_engine_run_targets(const orc_problem *p, orc_engine_options opt, orc_engine_result **out) {
    return engine_run(p, opt, out, NULL, 1);
}

int orc_engine_run_numerical(const orc_problem *p, orc_engine_options opt, orc_engine_result **out,
                             orc_numerical_result *num) {
    memset(num, 0, sizeof *num);
    return engine_run(p, opt, out, num, 0);
}

/* engine/numerical.rs:160-330 (replay): the leverage probes are solved again
 * in batches of at most 4; every probe's fold scores are
 * -sum_g folds[g][fold][j] * (gradient_g . u_g - offset_g). */
static int numerical_replay(const orc_problem *p, const orc_engine_options *opt, const solver_t *s,
                            num_state *ns, const double *draws, orc_numerical_result *num) {
    uint64_t G = p->units, W = p->workers, F = p->firms, R = opt->probes;
    double *prim = orc_alloc(3 * R, 8);
    for (uint64_t t = 0; t < R; ++t)
        for (int j = 0; j < 3; ++j) prim[3 * t + j] = draws[4 * t + j];
    double cond[9];
    int rc = orc_cross_covariance(prim, prim, R, cond);
    free(prim);
    if (rc) return rc;
    uint32_t status = ns->status;
    double *scores = orc_alloc(6 * R, 8); /* [fold][probe][3] */
    uint64_t batch = opt->leverage_width < 1 ? 1 : (opt->leverage_width > 4 ? 4 : opt->leverage_width);
    int64_t *atoms = orc_alloc(G * batch, 8);
    double *wr = orc_alloc(W * batch, 8), *fr = orc_alloc(F * batch, 8);
    column_t *cols = alloc_cols(p, batch);
    num->failed_replay_probe = UINT32_MAX;
    const uint64_t Gm = p->match_units, Go = G - Gm;
    if (status == 2) {
        for (uint64_t first = 0; first < R; first += batch) {
            uint64_t width = R - first < batch ? R - first : batch;
            if ((rc = leverage_atoms(p, opt->seed, first, width, atoms))) goto out;
            for (uint64_t g = 0; g < Gm; ++g)
                num->replay_generator_words += (p->unit_count[g] + 63) / 64 * width;
            /* generic_jla/numerical.rs:490-519: one evaluation per column for
             * every word a class's copies touch */
            num->replay_generator_words += ns->class_words * width;
            for (uint64_t c = 0; c < width; ++c) leverage_rhs(p, atoms + c * G, wr + c * W, fr + c * F);
            num->attempted_replay_rhs += width;
            rc = solve_batch(s, wr, fr, width, opt->probe_tolerance, 1.0, opt, cols);
            if (rc == ORC_PCG_MAX_ITERATIONS || rc == 62 || rc == ORC_PCG_CURVATURE ||
                rc == ORC_PCG_PRECONDITIONER || rc == ORC_FULL_RESIDUAL) {
                status = 7;
                num->failed_replay_probe = (uint32_t)first;
                num->replay_error_code = (uint32_t)rc;
                rc = ORC_OK;
                break;
            }
            if (rc) goto out;
            for (uint64_t c = 0; c < width; ++c) {
                uint64_t probe = first + c;
                num->certified_replay_rhs++;
                num->max_replay_complete_residual = fmax(num->max_replay_complete_residual, cols[c].complete);
                ssum sc[2][3] = {{{0, 0}, {0, 0}, {0, 0}}, {{0, 0}, {0, 0}, {0, 0}}};
                /* observation units (generic_jla/numerical.rs:617-641) */
                for (uint64_t k = 0; k < Go; ++k) {
                    uint64_t i = ns->order[k], g = Gm + i;
                    uint32_t cell = p->unit_cell[g];
                    double pp = cols[c].worker[p->cell_worker[cell]] + cols[c].firm[p->cell_firm[cell]];
                    ssum w1 = {0, 0}, w3 = {0, 0};
                    for (uint64_t j = 0; j < p->unit_count[g]; ++j) {
                        double q = (double)copy_sign(opt->seed, probe, p->unit_rank[g], p->unit_offset[g] + j);
                        ss_add(&w1, ns->copy[2 * (ns->phys[i] + j)] * q);
                        ss_add(&w3, ns->copy[2 * (ns->phys[i] + j) + 1] * q);
                    }
                    ssum alpha = {0, 0};
                    ss_add(&alpha, ns->ograd[3 * i] * pp * pp);
                    ss_add(&alpha, ns->ograd[3 * i + 1] * pp * pp * pp * pp);
                    ss_add(&alpha, -ns->ograd[3 * i + 2]);
                    ss_add(&alpha, ss_fin(w1) * pp);
                    ss_add(&alpha, ss_fin(w3) * pp * pp * pp);
                    double al = ss_fin(alpha);
                    for (int f = 0; f < 2; ++f)
                        for (int j = 0; j < 3; ++j) ss_add(&sc[f][j], -ss_fin(ns->ofolds[6 * i + 3 * f + j]) * al);
                }
                for (uint64_t g = 0; g < Gm; ++g) {
                    uint32_t cell = p->unit_cell[g];
                    double root = sqrt((double)p->unit_count[g]);
                    double pp = root * (cols[c].worker[p->cell_worker[cell]] + cols[c].firm[p->cell_firm[cell]]);
                    double m = (double)atoms[c * G + g] / root - pp;
                    double p2 = pp * pp, m2 = m * m;
                    double u[5] = {p2, m2, p2 * p2, m2 * m2, p2 * m2};
                    ssum alpha = {0, 0};
                    ss_add(&alpha, -ns->offs[g]);
                    for (int j = 0; j < 5; ++j) ss_add(&alpha, ns->grad[5 * g + j] * u[j]);
                    double al = ss_fin(alpha);
                    for (int f = 0; f < 2; ++f)
                        for (int j = 0; j < 3; ++j) ss_add(&sc[f][j], -ss_fin(ns->folds[6 * g + 3 * f + j]) * al);
                }
                for (int f = 0; f < 2; ++f)
                    for (int j = 0; j < 3; ++j) scores[(f * R + probe) * 3 + j] = ss_fin(sc[f][j]);
            }
        }
    }
    for (int f = 0; f < 2; ++f) {
        ssum m[3] = {{0, 0}, {0, 0}, {0, 0}};
        for (uint64_t t = 0; t < R; ++t)
            for (int j = 0; j < 3; ++j) ss_add(&m[j], scores[(f * R + t) * 3 + j] / (double)R);
        for (int j = 0; j < 3; ++j) num->score_means[3 * f + j] = ss_fin(m[j]);
    }
    double lev[9];
    if (status == 2) {
        if (orc_cross_covariance(scores, scores + 3 * R, R, lev) != ORC_OK) {
            status = 6;
            for (int k = 0; k < 9; ++k) lev[k] = NAN;
        }
    } else {
        for (int k = 0; k < 9; ++k) lev[k] = NAN;
    }
    uint32_t fin_status;
    orc_numerical_finalize(cond, lev, num->raw, num->usable, num->mcse, &fin_status, &num->psd_adjustment);
    memcpy(num->conditional, cond, sizeof cond);
    memcpy(num->leverage, lev, sizeof lev);
    num->status = status != 2 ? status : fin_status;
    num->leverage_probes = num->target_probes = (uint32_t)R;
    num->fold_a = (uint32_t)((R + 1) / 2);
    num->fold_b = (uint32_t)(R / 2);
out:
    free(scores);
    free(atoms);
    free(wr);
    free(fr);
    free_cols(cols, batch);
    return rc;
}

/* ---- observation deletion (generic engine, no controls) ----
 * One unit per stored row.  Its atom is the sum of the Rademacher signs of
 * the row's physical copies, copy j of a class being bit j % 64 of Counter-V1
 * word j / 64 of the class entity (generic_jla.rs:10321-10329, 5660-5706). */
static int copy_sign(uint64_t seed, uint64_t probe, uint64_t entity, uint64_t copy) {
    uint64_t word = orc_rng_word(seed, ORC_TAG_LEVERAGE, probe, entity, copy / 64);
    return ((word >> (copy % 64)) & 1) ? 1 : -1;
}

/* Units [match_units, units) are observation units: the last stored rows. */
static uint64_t observation_row(const orc_problem *p, uint64_t g) {
    return g - p->match_units + (p->rows - (p->units - p->match_units));
}

static const orc_problem *g_order_problem;
static int cmp_observation_units(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    const orc_problem *p = g_order_problem;
    uint64_t rx = p->unit_rank[p->match_units + x], ry = p->unit_rank[p->match_units + y];
    if (rx != ry) return rx < ry ? -1 : 1;
    return x < y ? -1 : (x > y);
}
/* Observation units (indices from match_units) in class order: ascending
 * class rank, rows ascending within a class (generic_jla.rs:4158-4215). */
static uint64_t *observation_order(const orc_problem *p) {
    uint64_t Go = p->units - p->match_units;
    uint64_t *order = orc_alloc(Go ? Go : 1, 8);
    for (uint64_t i = 0; i < Go; ++i) order[i] = i;
    g_order_problem = p;
    qsort(order, Go, 8, cmp_observation_units);
    return order;
}

static void observation_atoms(const orc_problem *p, uint64_t seed, uint64_t first, uint64_t k,
                              int64_t *atoms) {
    for (uint64_t c = 0; c < k; ++c)
        for (uint64_t g = p->match_units; g < p->units; ++g) {
            int64_t a = 0;
            for (uint64_t j = 0; j < p->unit_count[g]; ++j)
                a += copy_sign(seed, first + c, p->unit_rank[g], p->unit_offset[g] + j);
            atoms[c * p->units + g] = a;
        }
}

/* Leverage atoms of every unit, out[col * units + unit]: match units by the
 * compressed Counter-V1 sums (rng.rs:202-305), observation units by their
 * copies' signs. */
static int leverage_atoms(const orc_problem *p, uint64_t seed, uint64_t first, uint64_t k, int64_t *atoms) {
    uint64_t G = p->units, Gm = p->match_units;
    if (Gm) {
        int64_t *tmp = orc_alloc(Gm * k, 8);
        int rc = orc_rademacher_fill(seed, ORC_TAG_LEVERAGE, first, k, p->unit_rank, p->unit_count, Gm, tmp);
        for (uint64_t c = 0; !rc && c < k; ++c) memcpy(atoms + c * G, tmp + c * Gm, Gm * 8);
        free(tmp);
        if (rc) return rc;
    }
    if (G > Gm) observation_atoms(p, seed, first, k, atoms);
    return ORC_OK;
}

int orc_observation_atoms(const orc_problem *p, uint64_t seed, uint64_t first, uint64_t columns, int64_t *out) {
    return leverage_atoms(p, seed, first, columns, out);
}

/* ObservationMoments and ObservationCorrelations (generic_jla.rs:5540-5892):
 * per stored row the probe sums of p^2 and p^4, per physical copy those of
 * s p and s p^3, p the row's prediction and s the copy's sign. */
typedef struct {
    uint64_t *phys;          /* copies of observation unit i: [phys[i], phys[i+1]) */
    ssum *square, *fourth;   /* per observation unit */
    ssum *first, *third;     /* per physical copy */
} obs_state;

static void observation_state_init(const orc_problem *p, obs_state *o) {
    uint64_t Go = p->units - p->match_units;
    o->phys = orc_alloc(Go + 1, 8);
    for (uint64_t i = 0; i < Go; ++i) o->phys[i + 1] = o->phys[i] + p->unit_count[p->match_units + i];
    o->square = orc_alloc(Go, sizeof(ssum));
    o->fourth = orc_alloc(Go, sizeof(ssum));
    o->first = orc_alloc(o->phys[Go], sizeof(ssum));
    o->third = orc_alloc(o->phys[Go], sizeof(ssum));
}

static void observation_state_free(obs_state *o) {
    free(o->phys); free(o->square); free(o->fourth); free(o->first); free(o->third);
    o->phys = NULL; o->square = o->fourth = o->first = o->third = NULL;
}

/* statistical_batches.rs:79-206, probes in order */
static void observation_moments(const orc_problem *p, obs_state *o, uint64_t seed, uint64_t first,
                                uint64_t k, const column_t *cols) {
    for (uint64_t i = 0; i < p->units - p->match_units; ++i) {
        uint64_t g = p->match_units + i;
        uint32_t cell = p->unit_cell[g];
        for (uint64_t c = 0; c < k; ++c) {
            double pred = cols[c].worker[p->cell_worker[cell]] + cols[c].firm[p->cell_firm[cell]];
            ss_add(&o->square[i], pred * pred);
            ss_add(&o->fourth[i], pred * pred * pred * pred);
            for (uint64_t j = 0; j < p->unit_count[g]; ++j) {
                double s = (double)copy_sign(seed, first + c, p->unit_rank[g], p->unit_offset[g] + j);
                ss_add(&o->first[o->phys[i] + j], s * pred);
                ss_add(&o->third[o->phys[i] + j], s * pred * pred * pred);
            }
        }
    }
}

/* observation_deleted_adjustment (generic_jla.rs:5893-6017) with
 * FiveMoments::from_totals and finite (:441-513), no control leverage.  The
 * target contraction (statistical_batches.rs:318-331) sums
 * f y d p^2 over rows, d = e (sum over copies of the finite inverse) / f, and
 * p is one per cell: it is the cell-weight contraction of the match route
 * with weight[cell] = sum over the cell's rows of f y d.  Per row: projection
 * share, residual share, bias and variance are means over the copies;
 * residual mass = e; deleted mass = d. */
static int observation_adjustment(const orc_problem *p, const obs_state *o, const orc_engine_options *opt,
                                  orc_engine_result *res, ssum *cw, int *clipped) {
    double probes = (double)res->probes;
    int rc = ORC_OK;
    for (uint64_t i = 0; i < p->units - p->match_units && !rc; ++i) {
        uint64_t g = p->match_units + i, row = observation_row(p, g);
        double pf = ss_fin(o->square[i]), ps = ss_fin(o->fourth[i]);
        ssum inverse_sum = {0, 0}, lev = {0, 0}, rs = {0, 0}, bs = {0, 0}, vs = {0, 0};
        for (uint64_t q = o->phys[i]; q < o->phys[i + 1]; ++q) {
            double first = ss_fin(o->first[q]), third = ss_fin(o->third[q]);
            double tot[5] = {pf, probes + pf - 2.0 * first, ps,
                             probes + 6.0 * pf + ps - 4.0 * first - 4.0 * third, pf + ps - 2.0 * third};
            double total = (tot[0] + tot[1]) / probes;
            if (!isfinite(total) || total <= opt->block_tolerance) {
                rc = orc_fail(ORC_JLA_CONSTRAINT, "generic_jla_leverage", "projection and residual masses are not positive");
                break;
            }
            double projection = (tot[0] / probes) / total, residual = (tot[1] / probes) / total;
            double p4 = tot[2] / probes, m4 = tot[3] / probes, mixed = tot[4] / probes;
            double bias = (residual * p4 - projection * m4 + (residual - projection) * mixed) / probes;
            double variance = (residual * residual * p4 + projection * projection * m4 -
                               2.0 * projection * residual * mixed) / probes;
            if (!isfinite(projection) || !isfinite(residual) || !isfinite(bias) || !isfinite(variance) ||
                variance < -100.0 * opt->rank_tolerance) {
                rc = orc_fail(ORC_JLA_MOMENT, "generic_jla_leverage", "finite-projection moment failed");
                break;
            }
            if (variance < 0.0) *clipped = 1;
            variance = fmax(variance, 0.0);
            if (!isfinite(residual) || residual <= opt->block_tolerance) {
                rc = orc_fail(ORC_NONESTIMABLE, "generic_jla_observation_adjustment",
                              "estimated residual leverage is nonpositive");
                break;
            }
            double r = 1.0 / residual;
            double inverse = r + bias * r * r - variance * r * r * r;
            if (!isfinite(inverse) || inverse <= 0.0) {
                rc = orc_fail(ORC_BLOCK_INVERSE, "generic_jla_observation_adjustment",
                              "finite-projection observation inverse is nonpositive");
                break;
            }
            ss_add(&inverse_sum, inverse);
            ss_add(&lev, projection);
            ss_add(&rs, residual);
            ss_add(&bs, bias);
            ss_add(&vs, variance);
        }
        if (rc) break;
        double f = (double)p->unit_count[g];
        uint32_t cell = p->unit_cell[g];
        double e = p->outcome[row] - res->fitted_cell[cell];
        double deleted = e * (ss_fin(inverse_sum) / f);
        if (!isfinite(deleted)) {
            rc = orc_fail(ORC_CORRECTION_NONFINITE, "generic_jla_observation_adjustment", "observation deleted residual is nonfinite");
            break;
        }
        res->projection_share[g] = ss_fin(lev) / f;
        res->residual_share[g] = ss_fin(rs) / f;
        res->finite_bias[g] = ss_fin(bs) / f;
        res->finite_variance[g] = ss_fin(vs) / f;
        res->residual_mass[g] = e;
        res->deleted_mass[g] = deleted;
        ss_add(&cw[cell], f * p->outcome[row] * deleted);
    }
    return rc;
}

static int engine_run(const orc_problem *p, orc_engine_options opt, orc_engine_result **out,
                      orc_numerical_result *num, int keep_targets) {
    if (opt.probes < 2)
        return orc_fail(ORC_INVALID_INPUT, "jla_validate", "JLA probe count must be at least two");
    num_state ns = {NULL, NULL, NULL, 0};
    solver_t s;
    int rc = solver_init(&s, p);
    if (rc) {
        solver_free(&s);
        return rc;
    }
    orc_engine_result *res = orc_alloc(1, sizeof *res);
    uint64_t C = p->cells, G = p->units, W = p->workers, F = p->firms, S = p->strata;
    uint64_t R = opt.probes;
    res->cells = C; res->units = G; res->probes = R;
    res->workers = W; res->firms = F;
    if (keep_targets) {
        res->target_worker = orc_alloc(2 * R * W, 8);
        res->target_firm = orc_alloc(2 * R * F, 8);
    }
    double t0 = now_seconds();
    /* full fit: engine.rs:1108-1134 */
    uint64_t maxw = opt.leverage_width > opt.target_width ? opt.leverage_width : opt.target_width;
    if (maxw == 0) maxw = 1;
    double *wr = orc_alloc(2 * W * maxw, 8), *fr = orc_alloc(2 * F * maxw, 8);
    int64_t *atoms = orc_alloc(maxw * (G > S ? G : S), 8);
    column_t *cols = alloc_cols(p, 2 * maxw);
    for (uint64_t c = 0; c < C; ++c) {
        wr[p->cell_worker[c]] += p->cell_outcome_sum[c];
        fr[p->cell_firm[c]] += p->cell_outcome_sum[c];
    }
    column_t *fit = alloc_cols(p, 1);
    five_moments *mom = NULL;
    obs_state obs = {NULL, NULL, NULL, NULL, NULL};
    rc = solve_batch(&s, wr, fr, 1, opt.fit_tolerance, 0.01, &opt, fit);
    if (rc) {
        free_cols(fit, 1);
        goto fail;
    }
    res->sum.fit_iterations = fit[0].iterations;
    res->fitted_cell = orc_alloc(C, 8);
    for (uint64_t c = 0; c < C; ++c)
        res->fitted_cell[c] = fit[0].worker[p->cell_worker[c]] + fit[0].firm[p->cell_firm[c]];
    {
        ssum rss = {0, 0};
        for (uint64_t r = 0; r < p->rows; ++r) {
            double e = p->outcome[r] - res->fitted_cell[p->row_cell[r]];
            ss_add(&rss, (double)p->frequency[r] * e * e);
        }
        res->sum.weighted_rss = ss_fin(rss);
    }
    rc = plugin_components(p, fit[0].worker, fit[0].firm, res->sum.plugin);
    double max_complete = fit[0].complete;
    free_cols(fit, 1);
    if (rc) goto fail;

    /* leverage loop: engine.rs:1150-1246; observation deletion:
     * generic_jla.rs:5540-5892 */
    const uint64_t Gm = p->match_units, Go = G - Gm;
    if (Go) observation_state_init(p, &obs);
    mom = orc_alloc(G, sizeof *mom);
    res->lev_iters = orc_alloc(R, 8);
    uint64_t width = opt.leverage_width ? opt.leverage_width : 1;
    for (uint64_t first = 0; first < R; first += width) {
        uint64_t k = R - first < width ? R - first : width;
        if ((rc = leverage_atoms(p, opt.seed, first, k, atoms))) goto fail_mom;
        for (uint64_t c = 0; c < k; ++c) leverage_rhs(p, atoms + c * G, wr + c * W, fr + c * F);
        if ((rc = solve_batch(&s, wr, fr, k, opt.probe_tolerance, 1.0, &opt, cols))) goto fail_mom;
        for (uint64_t c = 0; c < k; ++c) {
            res->lev_iters[first + c] = (double)cols[c].iterations;
            res->sum.leverage_iterations += cols[c].iterations;
            max_complete = fmax(max_complete, cols[c].complete);
        }
        if (Go) observation_moments(p, &obs, opt.seed, first, k, cols);
        for (uint64_t g = 0; g < Gm; ++g) {
            uint32_t cell = p->unit_cell[g];
            double freq = (double)p->unit_count[g];
            for (uint64_t c = 0; c < k; ++c) {
                double pred = cols[c].worker[p->cell_worker[cell]] + cols[c].firm[p->cell_firm[cell]];
                double projection = sqrt(freq) * pred;
                double residual = (double)atoms[c * G + g] / sqrt(freq) - projection;
                if (!isfinite(projection) || !isfinite(residual)) {
                    rc = orc_fail(ORC_JLA_MOMENT, "jla_leverage", "nonfinite leverage moment");
                    goto fail_mom;
                }
                five_add(&mom[g], projection, residual);
            }
        }
    }
    /* leverage adjustment: engine.rs:2269-2410 (outcome mean 0, centering None) */
    int clipped = 0, obs_clipped = 0;
    {
        double probes = (double)R;
        res->projection_share = orc_alloc(G, 8); res->residual_share = orc_alloc(G, 8);
        res->finite_bias = orc_alloc(G, 8); res->finite_variance = orc_alloc(G, 8);
        res->residual_mass = orc_alloc(G, 8); res->deleted_mass = orc_alloc(G, 8);
        ssum *cw = orc_alloc(C, sizeof *cw);
        for (uint64_t g = 0; g < Gm; ++g) {
            double pf = ss_fin(mom[g].proj), mf = ss_fin(mom[g].res);
            double total_mean = (pf + mf) / probes;
            if (!isfinite(total_mean) || total_mean <= opt.block_tolerance) {
                free(cw);
                rc = orc_fail(ORC_JLA_CONSTRAINT, "jla_leverage", "projection/residual mass failed");
                goto fail_mom;
            }
            double projection = (pf / probes) / total_mean;
            double residual = (mf / probes) / total_mean;
            double p2 = ss_fin(mom[g].proj4) / probes, m2 = ss_fin(mom[g].res4) / probes;
            double mixed = ss_fin(mom[g].mixed) / probes;
            double bias = (residual * p2 - projection * m2 + (residual - projection) * mixed) / probes;
            double variance = (residual * residual * p2 + projection * projection * m2 -
                               2.0 * projection * residual * mixed) / probes;
            if (variance < -100.0 * opt.rank_tolerance) {
                free(cw);
                rc = orc_fail(ORC_JLA_MOMENT, "jla_leverage", "finite-projection variance negative");
                goto fail_mom;
            }
            if (variance < 0.0) clipped = 1;
            variance = fmax(variance, 0.0);
            if (1.0 - projection <= opt.block_tolerance) {
                free(cw);
                rc = orc_fail(ORC_NONESTIMABLE, "jla_leverage", "match residual block is singular");
                goto fail_mom;
            }
            double recip = 1.0 / residual;
            double rr = fabs(residual * recip - 1.0);
            if (!isfinite(recip) || rr > fmax(100.0 * opt.rank_tolerance, 1e-10)) {
                free(cw);
                rc = orc_fail(ORC_BLOCK_INVERSE, "jla_leverage", "residual inverse gate failed");
                goto fail_mom;
            }
            uint32_t cell = p->unit_cell[g];
            double resid_outcome = p->unit_outcome[g] - (double)p->unit_count[g] * res->fitted_cell[cell];
            double multiplier = recip + bias * (recip * recip) - variance * (recip * recip * recip);
            double deleted = resid_outcome * multiplier;
            ss_add(&cw[cell], (p->unit_outcome[g] - 0.0 * (double)p->unit_count[g]) * deleted);
            res->projection_share[g] = projection;
            res->residual_share[g] = residual;
            res->finite_bias[g] = bias;
            res->finite_variance[g] = variance;
            res->residual_mass[g] = resid_outcome;
            res->deleted_mass[g] = deleted;
        }
        /* observation units: generic_jla.rs:5893-6017 */
        if (Go && (rc = observation_adjustment(p, &obs, &opt, res, cw, &obs_clipped))) {
            free(cw);
            goto fail_mom;
        }
        res->cell_weight = orc_alloc(C, 8);
        for (uint64_t c = 0; c < C; ++c) res->cell_weight[c] = ss_fin(cw[c]);
        free(cw);
    }
    /* numerical state: engine/numerical.rs:62-122 (after the adjustment) */
    if (num && Gm) {
        ns.grad = orc_alloc(5 * G, 8);
        ns.offs = orc_alloc(G, 8);
        ns.folds = orc_alloc(6 * G, sizeof(ssum));
        ns.status = clipped ? 5 : 2;
        num->minimum_constrained = INFINITY;
        num->minimum_residual_margin = INFINITY;
        num->maximum_sensitivity_ratio = 0.0;
        double r = (double)R;
        for (uint64_t g = 0; g < Gm; ++g) {
            double u[5] = {ss_fin(mom[g].proj) / r, ss_fin(mom[g].res) / r, ss_fin(mom[g].proj4) / r,
                           ss_fin(mom[g].res4) / r, ss_fin(mom[g].mixed) / r};
            double fd[20], gr[5];
            if (!orc_finite_derivative(u, (uint32_t)R, fd)) {
                ns.status = 6;
                continue;
            }
            num->minimum_constrained = fmin(num->minimum_constrained, u[0] + u[1]);
            num->minimum_residual_margin = fmin(num->minimum_residual_margin, fd[1]);
            num->maximum_sensitivity_ratio = fmax(num->maximum_sensitivity_ratio, sqrt(fmax(fd[3], 0.0)) / fd[1]);
            if (fd[19] != 0.0) {
                ns.status = 5;
                continue;
            }
            if (!orc_fd_observation(fd, 0.0, gr)) {
                ns.status = 6;
                continue;
            }
            ssum off = {0, 0};
            for (int j = 0; j < 5; ++j) {
                ns.grad[5 * g + j] = gr[j];
                ss_add(&off, gr[j] * u[j]);
            }
            ns.offs[g] = ss_fin(off);
        }
    }
    /* numerical state of observation units: generic_jla/numerical.rs:172-270 */
    if (num && Go) {
        /* after the match units' state (generic_jla.rs:6053-6234 runs first):
         * the observation adjustment's clipping, then its events */
        if (!Gm) {
            ns.status = 2;
            num->minimum_constrained = INFINITY;
            num->minimum_residual_margin = INFINITY;
            num->maximum_sensitivity_ratio = 0.0;
        }
        if (obs_clipped) ns.status = 5;
        ns.ograd = orc_alloc(3 * Go, 8);
        ns.copy = orc_alloc(2 * obs.phys[Go], 8);
        ns.ofolds = orc_alloc(6 * Go, sizeof(ssum));
        ns.phys = obs.phys;
        ns.order = observation_order(p);
        for (uint64_t k = 0; k < Go; ++k)
            if (k == 0 || p->unit_rank[Gm + ns.order[k]] != p->unit_rank[Gm + ns.order[k - 1]]) {
                uint64_t mass = 0;
                for (uint64_t q = k; q < Go && p->unit_rank[Gm + ns.order[q]] == p->unit_rank[Gm + ns.order[k]]; ++q)
                    mass += p->unit_count[Gm + ns.order[q]];
                ns.class_words += (mass + 63) / 64;
            }
        double r = (double)R;
        for (uint64_t k = 0; k < Go; ++k) {
            uint64_t i = ns.order[k];
            double a2 = ss_fin(obs.square[i]) / r, a4 = ss_fin(obs.fourth[i]) / r;
            ssum g2 = {0, 0}, g4 = {0, 0}, off = {0, 0};
            for (uint64_t q = obs.phys[i]; q < obs.phys[i + 1]; ++q) {
                double c1 = ss_fin(obs.first[q]) / r, c3 = ss_fin(obs.third[q]) / r;
                const double x[4] = {a2, a4, c1, c3};
                double u[5], fd[20], pb[4], value;
                const int st = orc_copy_pullback(x, 0.0, (uint32_t)R, u, fd, &value, pb);
                num->minimum_constrained = fmin(num->minimum_constrained, u[0] + u[1]);
                if (st == 6 && !orc_finite_derivative(u, (uint32_t)R, fd)) {
                    ns.status = 6;
                    continue;
                }
                double ell = fd[1];
                num->minimum_residual_margin = fmin(num->minimum_residual_margin, ell);
                num->maximum_sensitivity_ratio = fmax(num->maximum_sensitivity_ratio, sqrt(fmax(fd[3], 0.0)) / ell);
                if (st != 0) {
                    ns.status = st;
                    continue;
                }
                ss_add(&g2, pb[0]);
                ss_add(&g4, pb[1]);
                ss_add(&off, pb[2] * c1);
                ss_add(&off, pb[3] * c3);
                ns.copy[2 * q] = pb[2];
                ns.copy[2 * q + 1] = pb[3];
            }
            ss_add(&off, ss_fin(g2) * a2);
            ss_add(&off, ss_fin(g4) * a4);
            ns.ograd[3 * i] = ss_fin(g2);
            ns.ograd[3 * i + 1] = ss_fin(g4);
            ns.ograd[3 * i + 2] = ss_fin(off);
        }
    }
    free(mom);
    mom = NULL;

    /* target loop: engine.rs:1373-1484 */
    res->draws = orc_alloc(4 * R, 8);
    width = opt.target_width ? opt.target_width : 1;
    for (uint64_t first = 0; first < R; first += width) {
        uint64_t k = R - first < width ? R - first : width;
        if ((rc = orc_rademacher_fill(opt.seed, ORC_TAG_TARGET, first, k, p->stratum_rank,
                                      p->stratum_count, S, atoms)))
            goto fail_mom;
        for (uint64_t c = 0; c < k; ++c)
            if ((rc = target_columns(p, atoms + c * S, first + c, wr + 2 * c * W, fr + 2 * c * F)))
                goto fail_mom;
        if ((rc = solve_batch(&s, wr, fr, 2 * k, opt.probe_tolerance, 1.0, &opt, cols))) goto fail_mom;
        for (uint64_t c = 0; c < k; ++c) {
            const column_t *a = &cols[2 * c], *b = &cols[2 * c + 1];
            res->sum.target_iterations += a->iterations + b->iterations;
            if (keep_targets) {
                uint64_t t = first + c;
                memcpy(res->target_worker + 2 * t * W, a->worker, W * 8);
                memcpy(res->target_worker + (2 * t + 1) * W, b->worker, W * 8);
                memcpy(res->target_firm + 2 * t * F, a->firm, F * 8);
                memcpy(res->target_firm + (2 * t + 1) * F, b->firm, F * 8);
            }
            max_complete = fmax(max_complete, fmax(a->complete, b->complete));
            /* engine.rs:3003-3046 (contract_target_solutions) */
            ssum ws = {0, 0}, fs = {0, 0}, cv = {0, 0};
            for (uint64_t cell = 0; cell < C; ++cell) {
                uint32_t w = p->cell_worker[cell], f = p->cell_firm[cell];
                double pw = a->worker[w] + a->firm[f], pf = b->worker[w] + b->firm[f];
                double wt = res->cell_weight[cell];
                ss_add(&ws, wt * pw * pw);
                ss_add(&fs, wt * pf * pf);
                ss_add(&cv, wt * pw * pf);
            }
            double *d = res->draws + 4 * (first + c);
            d[0] = ss_fin(ws);
            d[1] = ss_fin(fs);
            d[2] = ss_fin(cv);
            d[3] = d[0] + d[1] + 2.0 * d[2];
            /* engine/numerical.rs:124-157 (target folds) */
            if (num) {
                uint64_t probe = first + c;
                int fold = (int)(probe % 2);
                double count = fold == 0 ? (double)((R + 1) / 2) : (double)(R / 2);
                /* observation units: generic_jla/numerical.rs:377-389 */
                for (uint64_t i = 0; i < Go; ++i) {
                    uint64_t g = Gm + i, row = observation_row(p, g);
                    uint32_t cell = p->unit_cell[g];
                    uint32_t w = p->cell_worker[cell], f = p->cell_firm[cell];
                    double sw = a->worker[w] + a->firm[f], sf = b->worker[w] + b->firm[f];
                    double kappa[3] = {sw * sw, sf * sf, sw * sf};
                    for (int j = 0; j < 3; ++j)
                        ss_add(&ns.ofolds[6 * i + 3 * fold + j], p->outcome[row] * res->residual_mass[g] * kappa[j] / count);
                }
                for (uint64_t g = 0; g < Gm; ++g) {
                    uint32_t cell = p->unit_cell[g];
                    uint32_t w = p->cell_worker[cell], f = p->cell_firm[cell];
                    double sw = a->worker[w] + a->firm[f], sf = b->worker[w] + b->firm[f];
                    double mass = (p->unit_outcome[g] - 0.0 * (double)p->unit_count[g]) * res->residual_mass[g] / count;
                    double kappa[3] = {sw * sw, sf * sf, sw * sf};
                    for (int j = 0; j < 3; ++j) ss_add(&ns.folds[6 * g + 3 * fold + j], mass * kappa[j]);
                }
            }
        }
    }
    /* finalize: engine.rs:3108-3272 */
    {
        ssum m[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
        for (uint64_t t = 0; t < R; ++t)
            for (int j = 0; j < 4; ++j) ss_add(&m[j], res->draws[4 * t + j]);
        for (int j = 0; j < 4; ++j) res->sum.correction[j] = ss_fin(m[j]) / (double)R;
        for (int j = 0; j < 4; ++j)
            res->sum.corrected[j] = res->sum.plugin[j] - res->sum.correction[j];
        ssum v[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
        for (uint64_t t = 0; t < R; ++t)
            for (int j = 0; j < 4; ++j) {
                double e = res->draws[4 * t + j] - res->sum.correction[j];
                ss_add(&v[j], e * e);
            }
        double denom = (double)(R * (R - 1));
        for (int j = 0; j < 4; ++j) res->sum.mcse[j] = sqrt(ss_fin(v[j]) / denom);
    }
    res->sum.max_complete_residual = max_complete;
    if (num) {
        num->point_rhs = 1 + 3 * R;
        num->max_point_complete_residual = max_complete;
        rc = numerical_replay(p, &opt, &s, &ns, res->draws, num);
        free(ns.grad); free(ns.offs); free(ns.folds); free(ns.copy); free(ns.order); free(ns.ograd); free(ns.ofolds);
        ns.grad = NULL; ns.offs = NULL; ns.folds = NULL; ns.copy = NULL; ns.order = NULL; ns.ograd = NULL;
        ns.ofolds = NULL;
        if (rc) goto fail;
    }
    res->sum.solve_seconds = now_seconds() - t0;
    observation_state_free(&obs);
    free(atoms);
    free_cols(cols, 2 * maxw);
    free(wr);
    free(fr);
    solver_free(&s);
    *out = res;
    return ORC_OK;
fail_mom:
    free(mom);
fail:
    observation_state_free(&obs);
    free(ns.copy);
    free(ns.order);
    free(ns.ograd);
    free(ns.ofolds);
    free(ns.grad);
    free(ns.offs);
    free(ns.folds);
    free(atoms);
    free_cols(cols, 2 * maxw);
    free(wr);
    free(fr);
    solver_free(&s);
    orc_engine_result_free(res);
    return rc;
}

void orc_engine_result_free(orc_engine_result *r) {
    if (!r) return;
    free(r->fitted_cell); free(r->projection_share); free(r->residual_share);
    free(r->finite_bias); free(r->finite_variance); free(r->residual_mass);
    free(r->deleted_mass); free(r->cell_weight); free(r->draws); free(r->lev_iters);
    free(r->target_worker); free(r->target_firm);
    free(r);
}

void orc_engine_result_summary(const orc_engine_result *r, orc_engine_summary *s) { *s = r->sum; }

int orc_engine_result_array(const orc_engine_result *r, int which, double *dst, uint64_t cap) {
    const double *src;
    uint64_t n;
    switch (which) {
    case 0: src = r->fitted_cell; n = r->cells; break;
    case 1: src = r->projection_share; n = r->units; break;
    case 2: src = r->residual_share; n = r->units; break;
    case 3: src = r->finite_bias; n = r->units; break;
    case 4: src = r->finite_variance; n = r->units; break;
    case 5: src = r->residual_mass; n = r->units; break;
    case 6: src = r->deleted_mass; n = r->units; break;
    case 7: src = r->cell_weight; n = r->cells; break;
    case 8: src = r->draws; n = 4 * r->probes; break;
    case 9: src = r->lev_iters; n = r->probes; break;
    case 10: src = r->target_worker; n = 2 * r->probes * r->workers; break;
    case 11: src = r->target_firm; n = 2 * r->probes * r->firms; break;
    default: return orc_fail(ORC_INVALID_INPUT, "engine", "unknown array");
    }
    if (!src) return orc_fail(ORC_INVALID_INPUT, "engine", "array not recorded by this run");
    if (cap < n) return orc_fail(ORC_INVALID_INPUT, "engine", "capacity too small");
    memcpy(dst, src, n * 8);
    return ORC_OK;
}

/* CPU baseline: `count` leverage probe solves on the prepared problem. */
int orc_time_probe_solves(const orc_problem *p, orc_engine_options opt, uint32_t first,
                          uint32_t count, double *seconds, uint64_t *iterations) {
    solver_t s;
    int rc = solver_init(&s, p);
    if (rc) {
        solver_free(&s);
        return rc;
    }
    uint64_t G = p->units, W = p->workers, F = p->firms;
    int64_t *atoms = orc_alloc(G * count, 8);
    double *wr = orc_alloc(W * count, 8), *fr = orc_alloc(F * count, 8);
    column_t *cols = alloc_cols(p, count);
    double t0 = now_seconds();
    rc = orc_rademacher_fill(opt.seed, ORC_TAG_LEVERAGE, first, count, p->unit_rank, p->unit_count,
                             G, atoms);
    for (uint64_t c = 0; rc == ORC_OK && c < count; ++c)
        leverage_rhs(p, atoms + c * G, wr + c * W, fr + c * F);
    if (!rc) rc = solve_batch(&s, wr, fr, count, opt.probe_tolerance, 1.0, &opt, cols);
    *seconds = now_seconds() - t0;
    uint64_t it = 0;
    for (uint64_t c = 0; c < count; ++c) it += cols[c].iterations;
    *iterations = it;
    free_cols(cols, count);
    free(atoms); free(wr); free(fr);
    solver_free(&s);
    return rc;
}

/* Harmonic lift of a firm vector onto the hybrid vertices
 * (cmg_impl.rs.in:459-500); lifted has the hybrid dimension. */
int orc_hybrid_lift(const orc_problem *p, const double *firm, double *lifted, uint64_t capacity) {
    solver_t s = {0};
    s.p = p;
    int rc = build_hybrid(&s);
    if (!rc && capacity < s.n) rc = orc_fail(ORC_INVALID_INPUT, "cmg_graph", "lift capacity too small");
    if (!rc) lift_firm(&s, firm, lifted);
    orc_graph_free(s.graph);
    free(s.aux_ptr); free(s.aux_nbr); free(s.aux_w);
    return rc;
}

/* One CMG_FULL_V2 batch solve (full_cmg.rs:809-1111) of k columns of
 * (worker_rhs, firm_rhs), column-major: solutions, complete and reduced
 * residuals, PCG iterations per column, refinement rounds and refined
 * columns. */
int orc_solve_columns(const orc_problem *p, const double *wr, const double *fr, uint64_t k, double tolerance,
                      double inner_ratio, uint32_t max_iterations, uint32_t recompute_interval, uint32_t threads,
                      double *worker_out, double *firm_out, double *complete, double *reduced,
                      uint64_t *iterations, uint64_t *refine_rounds, uint64_t *refined_columns) {
    solver_t s;
    int rc = solver_init(&s, p);
    if (rc) {
        solver_free(&s);
        return rc;
    }
    orc_engine_options opt = {0};
    opt.threads = threads;
    opt.max_iterations = max_iterations;
    opt.recompute_interval = recompute_interval;
    column_t *cols = alloc_cols(p, k);
    *refine_rounds = *refined_columns = 0;
    rc = solve_batch_counted(&s, wr, fr, k, tolerance, inner_ratio, &opt, cols, refine_rounds, refined_columns);
    if (!rc)
        for (uint64_t c = 0; c < k; ++c) {
            memcpy(worker_out + c * p->workers, cols[c].worker, p->workers * 8);
            memcpy(firm_out + c * p->firms, cols[c].firm, p->firms * 8);
            complete[c] = cols[c].complete;
            reduced[c] = cols[c].reduced;
            iterations[c] = cols[c].iterations;
        }
    free_cols(cols, k);
    solver_free(&s);
    return rc;
}
