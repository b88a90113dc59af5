/* CMG restatement (test infrastructure): the serial, default-feature path of
 * the vendored CMG crate rust/vendor/cmg (revision e5613e81).  Arithmetic
 * order follows the reference exactly so that the FNV-1a digests in
 * rust/vendor/cmg/tests/fixtures/frozen_bits.txt [default] reproduce. */
#include <stdio.h>

#include "orc_internal.h"

static int g_empty_sum_negative = 1;
void orc_set_empty_sum_negative(int negative) { g_empty_sum_negative = negative; }
/* ldl.rs:78-80: the fold seed of Iterator::sum for f64. */
static double empty_sum(void) { return g_empty_sum_negative ? -0.0 : 0.0; }

/* ======================= Laplacian (graph.rs) ============================ */

typedef struct {
    uint64_t key;
    double w;
} raw_edge;

/* graph.rs:814-818 total order: endpoint key then weight total order */
static int raw_edge_cmp(const void *a, const void *b) {
    const raw_edge *x = a, *y = b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return orc_total_cmp(x->w, y->w);
}

/* graph.rs:279-414: sort, merge duplicates with a compensated sum (weights
 * ascending within a group), diagonal in canonical order. */
static orc_graph *canonical_from_raw(uint64_t n, raw_edge *raw, uint64_t m) {
    qsort(raw, m, sizeof *raw, raw_edge_cmp);
    orc_graph *g = orc_alloc(1, sizeof *g);
    g->n = n;
    g->diag = orc_alloc(n, sizeof(double));
    g->u = orc_alloc(m, sizeof(uint32_t));
    g->v = orc_alloc(m, sizeof(uint32_t));
    g->w = orc_alloc(m, sizeof(double));
    uint64_t read = 0, write = 0;
    while (read < m) {
        uint64_t key = raw[read].key;
        double sum = 0.0, corr = 0.0;
        while (read < m && raw[read].key == key) {
            orc_neumaier(&sum, &corr, raw[read].w);
            ++read;
        }
        double weight = sum + corr;
        uint32_t u = (uint32_t)(key >> 32), v = (uint32_t)key;
        g->u[write] = u;
        g->v[write] = v;
        g->w[write] = weight;
        g->diag[u] += weight;
        g->diag[v] += weight;
        ++write;
    }
    g->m = write;
    uint64_t diag_nnz = 0;
    double maxdeg = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        diag_nnz += g->diag[i] != 0.0;
        maxdeg = fmax(maxdeg, g->diag[i]);
    }
    g->nnz = diag_nnz + 2 * g->m;
    g->norm_bound = 2.0 * maxdeg;
    return g;
}

orc_graph *orc_graph_canonical(uint64_t n, uint64_t m, uint32_t *u, uint32_t *v,
                               double *w) {
    raw_edge *raw = orc_alloc(m, sizeof *raw);
    for (uint64_t e = 0; e < m; ++e) {
        uint32_t a = u[e] < v[e] ? u[e] : v[e], b = u[e] < v[e] ? v[e] : u[e];
        raw[e].key = ((uint64_t)a << 32) | b;
        raw[e].w = w[e];
    }
    orc_graph *g = canonical_from_raw(n, raw, m);
    free(raw);
    return g;
}

/* graph.rs:152-167 and map_edge graph.rs:746-790 */
orc_graph *orc_graph_from_edges(uint64_t n, uint64_t m, const uint32_t *u,
                                const uint32_t *v, const double *w) {
    raw_edge *raw = orc_alloc(m, sizeof *raw);
    for (uint64_t e = 0; e < m; ++e) {
        if (u[e] >= n || v[e] >= n || u[e] == v[e] || !isfinite(w[e]) || w[e] <= 0.0) {
            free(raw);
            orc_fail(ORC_CMG_SETUP, "cmg_graph", "invalid edge %llu", (unsigned long long)e);
            return NULL;
        }
        uint32_t a = u[e] < v[e] ? u[e] : v[e], b = u[e] < v[e] ? v[e] : u[e];
        raw[e].key = ((uint64_t)a << 32) | b;
        raw[e].w = w[e];
    }
    orc_graph *g = canonical_from_raw(n, raw, m);
    free(raw);
    return g;
}

void orc_graph_free(orc_graph *g) {
    if (!g) return;
    free(g->u); free(g->v); free(g->w); free(g->diag); free(g);
}
uint64_t orc_graph_vertices(const orc_graph *g) { return g->n; }
uint64_t orc_graph_edges(const orc_graph *g) { return g->m; }
void orc_graph_export(const orc_graph *g, uint32_t *u, uint32_t *v, double *w,
                      double *diag) {
    if (u) memcpy(u, g->u, g->m * sizeof *u);
    if (v) memcpy(v, g->v, g->m * sizeof *v);
    if (w) memcpy(w, g->w, g->m * sizeof *w);
    if (diag) memcpy(diag, g->diag, g->n * sizeof *diag);
}

/* graph.rs:664-705: edge-list scatter from zero in canonical order */
void orc_graph_matvec(const orc_graph *g, const double *x, double *y) {
    memset(y, 0, g->n * sizeof *y);
    for (uint64_t e = 0; e < g->m; ++e) {
        uint32_t u = g->u[e], v = g->v[e];
        double d = g->w[e] * (x[u] - x[v]);
        y[u] += d;
        y[v] -= d;
    }
}

/* ======================= components (components.rs) ===================== */

static uint64_t find_root64(uint64_t *parent, uint64_t vertex) {
    uint64_t root = vertex;
    while (parent[root] != root) root = parent[root];
    uint64_t cur = vertex;
    while (parent[cur] != cur) {
        uint64_t next = parent[cur];
        parent[cur] = root;
        cur = next;
    }
    return root;
}

/* components.rs:492-567: union-find, roots are minimum vertices, labels in
 * first-vertex order */
static void components_from_pairs(orc_components *c, uint64_t n, uint64_t m,
                                  const uint32_t *u, const uint32_t *v) {
    uint64_t *parent = orc_alloc(n, sizeof *parent);
    for (uint64_t i = 0; i < n; ++i) parent[i] = i;
    for (uint64_t e = 0; e < m; ++e) {
        uint64_t a = find_root64(parent, u[e]), b = find_root64(parent, v[e]);
        if (a != b) parent[a > b ? a : b] = a < b ? a : b;
    }
    for (uint64_t i = 0; i < n; ++i) parent[i] = find_root64(parent, i);
    uint64_t *root_label = orc_alloc(n, sizeof *root_label);
    for (uint64_t i = 0; i < n; ++i) root_label[i] = UINT64_MAX;
    c->n = n;
    c->labels = orc_alloc(n, sizeof *c->labels);
    c->count = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t r = parent[i];
        if (root_label[r] == UINT64_MAX) root_label[r] = c->count++;
        c->labels[i] = (uint32_t)root_label[r];
    }
    c->sizes = orc_alloc(c->count, sizeof *c->sizes);
    for (uint64_t i = 0; i < n; ++i) c->sizes[c->labels[i]]++;
    free(parent);
    free(root_label);
}

void orc_components_build(orc_components *c, const orc_graph *g) {
    components_from_pairs(c, g->n, g->m, g->u, g->v);
}

void orc_components_free(orc_components *c) {
    free(c->labels);
    free(c->sizes);
    memset(c, 0, sizeof *c);
}

/* components.rs:1069-1140 (compensated_sums_into) */
static int compensated_sums(const orc_components *c, const double *values,
                            double *sums, double *corr) {
    memset(sums, 0, c->count * sizeof *sums);
    memset(corr, 0, c->count * sizeof *corr);
    if (c->count == 1) {
        double s = 0.0, k = 0.0;
        for (uint64_t i = 0; i < c->n; ++i) {
            if (!isfinite(values[i]))
                return orc_fail(ORC_CMG_APPLY, "components", "nonfinite value at %llu",
                                (unsigned long long)i);
            orc_neumaier(&s, &k, values[i]);
        }
        sums[0] = s + k;
        corr[0] = k;
        return ORC_OK;
    }
    for (uint64_t i = 0; i < c->n; ++i) {
        if (!isfinite(values[i]))
            return orc_fail(ORC_CMG_APPLY, "components", "nonfinite value at %llu",
                            (unsigned long long)i);
        uint32_t l = c->labels[i];
        orc_neumaier(&sums[l], &corr[l], values[i]);
    }
    for (uint64_t k = 0; k < c->count; ++k) sums[k] += corr[k];
    return ORC_OK;
}

/* components.rs:751-815 (center_in_place_with_workspace, serial) */
int orc_components_center(const orc_components *c, double *values) {
    double *sums = orc_alloc(c->count, sizeof *sums);
    double *corr = orc_alloc(c->count, sizeof *corr);
    int rc = compensated_sums(c, values, sums, corr);
    if (rc == ORC_OK) {
        for (uint64_t k = 0; k < c->count; ++k) sums[k] = sums[k] / (double)c->sizes[k];
        if (c->count == 1) {
            double mean = sums[0];
            for (uint64_t i = 0; i < c->n; ++i) values[i] -= mean;
        } else {
            for (uint64_t i = 0; i < c->n; ++i) values[i] -= sums[c->labels[i]];
        }
    }
    free(sums);
    free(corr);
    return rc;
}

/* components.rs:669-744 (project_rhs_in_place_with_workspace) */
int orc_components_project(const orc_components *c, double *rhs, double *norm) {
    uint64_t k = c->count;
    double *sums = orc_alloc(k, 8), *corr = orc_alloc(k, 8), *scales = orc_alloc(k, 8),
           *scorr = orc_alloc(k, 8), *means = orc_alloc(k, 8), *pcorr = orc_alloc(k, 8);
    uint64_t *rep = orc_alloc(k, 8);
    int rc = ORC_OK;
    for (uint64_t i = 0; i < c->n; ++i) {
        double x = rhs[i];
        if (!isfinite(x)) {
            rc = orc_fail(ORC_CMG_APPLY, "components", "nonfinite rhs");
            goto done;
        }
        uint32_t l = c->labels[i];
        orc_neumaier(&sums[l], &corr[l], x);
        orc_neumaier(&scales[l], &scorr[l], fabs(x));
    }
    for (uint64_t j = 0; j < k; ++j) {
        sums[j] += corr[j];
        scales[j] += scorr[j];
    }
    for (uint64_t j = 0; j < k; ++j) {
        double tol = 1.0e-7 * fmax(scales[j], 1.0);
        if (fabs(sums[j]) > tol) {
            rc = orc_fail(ORC_CMG_APPLY, "components",
                          "incompatible Laplacian rhs in component %llu",
                          (unsigned long long)j);
            goto done;
        }
    }
    for (uint64_t j = 0; j < k; ++j) means[j] = sums[j] / (double)c->sizes[j];
    for (uint64_t i = 0; i < c->n; ++i) rhs[i] -= means[c->labels[i]];
    for (uint64_t j = 0; j < k; ++j) rep[j] = UINT64_MAX;
    for (uint64_t i = 0; i < c->n; ++i) {
        uint32_t l = c->labels[i];
        uint64_t cur = rep[l];
        if (cur == UINT64_MAX || fabs(rhs[i]) < fabs(rhs[cur]) ||
            (fabs(rhs[i]) == fabs(rhs[cur]) && i < cur))
            rep[l] = i;
    }
    for (int pass = 0; pass < 2; ++pass) {
        compensated_sums(c, rhs, sums, corr);
        for (uint64_t j = 0; j < k; ++j) {
            double residual = sums[j];
            rhs[rep[j]] -= residual;
            pcorr[j] += residual;
        }
    }
    {
        double scale = 0.0;
        for (uint64_t j = 0; j < k; ++j) {
            scale = fmax(scale, fabs(means[j]));
            scale = fmax(scale, fabs(means[j] + pcorr[j]));
        }
        if (scale == 0.0) {
            *norm = 0.0;
        } else {
            double sq = empty_sum();
            for (uint64_t j = 0; j < k; ++j) {
                double regular = means[j] / scale;
                double representative = (means[j] + pcorr[j]) / scale;
                sq += (double)(c->sizes[j] - 1) * regular * regular +
                      representative * representative;
            }
            *norm = scale * sqrt(sq);
        }
    }
done:
    free(sums); free(corr); free(scales); free(scorr); free(means); free(pcorr); free(rep);
    return rc;
}

/* components.rs:385-458 (center_lanes, one lane): zeroed sums, compensated
 * per-component sums in vertex order, (sum + correction) / size. */
static int center_plan(const orc_components *c, double *values, double *sums,
                       double *corr) {
    memset(sums, 0, c->count * sizeof *sums);
    memset(corr, 0, c->count * sizeof *corr);
    for (uint64_t i = 0; i < c->n; ++i) {
        if (!isfinite(values[i]))
            return orc_fail(ORC_CMG_APPLY, "cmg_center", "nonfinite value");
        uint32_t l = c->count <= 1 ? 0 : c->labels[i];
        orc_neumaier(&sums[l], &corr[l], values[i]);
    }
    for (uint64_t j = 0; j < c->count; ++j)
        sums[j] = (sums[j] + corr[j]) / (double)c->sizes[j];
    for (uint64_t i = 0; i < c->n; ++i)
        values[i] -= sums[c->count <= 1 ? 0 : c->labels[i]];
    return ORC_OK;
}

/* ======================= forest (forest.rs) ============================== */

/* forest.rs:137-152 and 325-337 */
static void heavy_forest(const orc_graph *g, uint64_t *parent, double *weight) {
    for (uint64_t i = 0; i < g->n; ++i) {
        parent[i] = i;
        weight[i] = 0.0;
    }
    for (uint64_t e = 0; e < g->m; ++e) {
        uint64_t a = g->u[e], b = g->v[e];
        double w = g->w[e];
        if (w > weight[a] || (w == weight[a] && b < parent[a])) {
            weight[a] = w;
            parent[a] = b;
        }
        if (w > weight[b] || (w == weight[b] && a < parent[b])) {
            weight[b] = w;
            parent[b] = a;
        }
    }
}

/* forest.rs:431-547 (split_forest_with): diameter then conductance cuts */
static void split_forest(uint64_t n, uint64_t *forest) {
    int64_t *anc = orc_alloc(n, sizeof *anc);
    uint64_t *indeg = orc_alloc(n, sizeof *indeg);
    uint8_t *visited = orc_alloc(n, 1);
    uint64_t *walk = orc_alloc(n + 2, sizeof *walk);
    int64_t *newanc = orc_alloc(n + 2, sizeof *newanc);
    for (uint64_t i = 0; i < n; ++i) indeg[forest[i]]++;
    for (uint64_t start = 0; start < n; ++start) {
        uint64_t current = start;
        int cont = 1;
        for (;;) {
            if (!(cont && indeg[current] == 0 && !visited[current])) break;
            cont = 0;
            int64_t in_path = 0;
            uint64_t len = 0;
            walk[len] = current;
            newanc[len] = 0;
            ++len;
            uint64_t k = 0;
            for (;;) {
                if (!(k <= 5 || visited[current])) break;
                current = forest[current];
                int terminated = current == walk[k] || (k > 0 && current == walk[k - 1]);
                if (terminated) break;
                ++k;
                walk[len] = current;
                in_path += visited[current] ? 0 : 1;
                newanc[len] = in_path;
                ++len;
            }
            if (k > 5) {
                uint64_t middle = k / 2;
                forest[walk[middle]] = walk[middle];
                uint64_t next = walk[middle + 1];
                indeg[next]--;
                int64_t removed = anc[walk[middle]];
                for (uint64_t i = middle + 1; i <= k; ++i) anc[walk[i]] -= removed;
                for (uint64_t i = 0; i <= middle; ++i) {
                    visited[walk[i]] = 1;
                    anc[walk[i]] += newanc[i];
                }
                current = next;
                cont = 1;
            }
            if (!cont) {
                for (uint64_t i = 0; i <= k; ++i) {
                    anc[walk[i]] += newanc[i];
                    visited[walk[i]] = 1;
                }
            }
        }
    }
    for (uint64_t start = 0; start < n; ++start) {
        uint64_t current = start;
        for (;;) {
            if (indeg[current] != 0) break;
            uint64_t previous = current;
            int restart = 0;
            for (;;) {
                uint64_t next = forest[current];
                if (next == current || next == previous) break;
                if (anc[current] > 2 && anc[next] - anc[current] > 2) {
                    forest[current] = current;
                    indeg[next]--;
                    int64_t removed = anc[current];
                    uint64_t ap = current, ac = next;
                    for (;;) {
                        anc[ac] -= removed;
                        uint64_t an = forest[ac];
                        if (an == ac || an == ap) break;
                        ap = ac;
                        ac = an;
                    }
                    current = next;
                    restart = 1;
                    break;
                }
                previous = current;
                current = next;
            }
            if (!restart) break;
        }
    }
    free(anc); free(indeg); free(visited); free(walk); free(newanc);
}

/* forest.rs:339-375 */
static void low_degree_correction(const orc_graph *g, double threshold,
                                  const double *selected, uint64_t *parent) {
    int any = 0;
    for (uint64_t i = 0; i < g->n; ++i) {
        if (g->diag[i] > 0.0 && selected[i] / g->diag[i] < threshold) {
            any = 1;
            break;
        }
    }
    if (!any) return;
    double *incident = orc_alloc(g->n, sizeof *incident);
    for (uint64_t i = 0; i < g->n; ++i) {
        uint64_t t = parent[i];
        if (t != i) {
            incident[i] += selected[i];
            incident[t] += selected[i];
        }
    }
    for (uint64_t i = 0; i < g->n; ++i)
        if (g->diag[i] > 0.0 && incident[i] / g->diag[i] < threshold) parent[i] = i;
    free(incident);
}

/* forest.rs:550-607 (component_labels): union to the smaller root, labels in
 * first-occurrence order of roots */
static uint64_t forest_labels(uint64_t n, const uint64_t *parent, uint32_t *labels) {
    uint64_t *set = orc_alloc(n, sizeof *set);
    for (uint64_t i = 0; i < n; ++i) set[i] = i;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t a = find_root64(set, i), b = find_root64(set, parent[i]);
        if (a != b) {
            if (a < b) set[b] = a;
            else set[a] = b;
        }
    }
    for (uint64_t i = 0; i < n; ++i) set[i] = find_root64(set, i);
    uint64_t *root_label = orc_alloc(n, sizeof *root_label);
    for (uint64_t i = 0; i < n; ++i) root_label[i] = UINT64_MAX;
    uint64_t count = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t r = set[i];
        if (root_label[r] == UINT64_MAX) root_label[r] = count++;
        labels[i] = (uint32_t)root_label[r];
    }
    free(set);
    free(root_label);
    return count;
}

/* forest.rs:94-112 */
static uint64_t aggregation_labels(const orc_graph *g, double threshold, uint32_t *labels) {
    uint64_t *parent = orc_alloc(g->n, sizeof *parent);
    double *weight = orc_alloc(g->n, sizeof *weight);
    heavy_forest(g, parent, weight);
    split_forest(g->n, parent);
    low_degree_correction(g, threshold, weight, parent);
    uint64_t count = forest_labels(g->n, parent, labels);
    free(parent);
    free(weight);
    return count;
}

/* coarsen.rs:359-401 (contract_with, serial) */
static orc_graph *contract(const orc_graph *g, const uint32_t *labels, uint64_t coarse) {
    raw_edge *raw = orc_alloc(g->m, sizeof *raw);
    uint64_t m = 0;
    for (uint64_t e = 0; e < g->m; ++e) {
        uint32_t a = labels[g->u[e]], b = labels[g->v[e]];
        if (a == b) continue;
        uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        raw[m].key = ((uint64_t)lo << 32) | hi;
        raw[m].w = g->w[e];
        ++m;
    }
    orc_graph *c = canonical_from_raw(coarse, raw, m);
    free(raw);
    return c;
}

/* ======================= LDL terminal (ldl.rs) =========================== */

static uint64_t perm_key_nnz_global_dummy;
typedef struct {
    uint64_t pattern, vertex;
} perm_key;
static int perm_key_cmp(const void *a, const void *b) {
    const perm_key *x = a, *y = b;
    if (x->pattern != y->pattern) return x->pattern < y->pattern ? -1 : 1;
    return x->vertex < y->vertex ? -1 : (x->vertex > y->vertex);
}

/* ldl.rs:677-803 (factor_metered, whole graph) and ldl.rs:999-1041
 * (factor_dense), ldl.rs:185-281 (LowerFactor::from_dense) */
static int ldl_factor(const orc_graph *g, orc_ldl **out) {
    (void)perm_key_nnz_global_dummy;
    orc_ldl *f = orc_alloc(1, sizeof *f);
    uint64_t n = g->n;
    f->n = n;
    orc_components_build(&f->comps, g);
    f->anchors = orc_alloc(f->comps.count, sizeof *f->anchors);
    for (uint64_t i = 0; i < n; ++i) f->anchors[f->comps.labels[i]] = i;
    uint8_t *is_anchor = orc_alloc(n, 1);
    for (uint64_t c = 0; c < f->comps.count; ++c) is_anchor[f->anchors[c]] = 1;
    uint64_t *pattern = orc_alloc(n, sizeof *pattern);
    perm_key *keys = orc_alloc(n, sizeof *keys);
    uint64_t dim = 0;
    for (uint64_t i = 0; i < n; ++i)
        if (!is_anchor[i]) pattern[i] = 1;
    for (uint64_t e = 0; e < g->m; ++e)
        if (!is_anchor[g->u[e]] && !is_anchor[g->v[e]]) {
            pattern[g->u[e]]++;
            pattern[g->v[e]]++;
        }
    for (uint64_t i = 0; i < n; ++i)
        if (!is_anchor[i]) {
            keys[dim].pattern = pattern[i];
            keys[dim].vertex = i;
            ++dim;
        }
    qsort(keys, dim, sizeof *keys, perm_key_cmp);
    f->dim = dim;
    f->perm = orc_alloc(dim, sizeof *f->perm);
    for (uint64_t i = 0; i < dim; ++i) f->perm[i] = keys[i].vertex;
    free(keys);
    free(pattern);

    double *a = orc_alloc(dim * dim, sizeof *a);   /* matrix, row-major */
    double *l = orc_alloc(dim * dim, sizeof *l);   /* dense lower */
    uint64_t *fidx = orc_alloc(n, sizeof *fidx);
    for (uint64_t i = 0; i < n; ++i) fidx[i] = UINT64_MAX;
    for (uint64_t k = 0; k < dim; ++k) {
        fidx[f->perm[k]] = k;
        a[k * dim + k] = g->diag[f->perm[k]];
    }
    for (uint64_t e = 0; e < g->m; ++e) {
        uint64_t x = fidx[g->u[e]], y = fidx[g->v[e]];
        if (x == UINT64_MAX || y == UINT64_MAX) continue;
        a[x * dim + y] -= g->w[e];
        a[y * dim + x] -= g->w[e];
    }
    f->diag = orc_alloc(dim, sizeof *f->diag);
    for (uint64_t r = 0; r < dim; ++r) l[r * dim + r] = 1.0;
    int rc = ORC_OK;
    for (uint64_t col = 0; col < dim; ++col) {
        double pivot = a[col * dim + col];
        for (uint64_t p = 0; p < col; ++p) {
            double value = l[col * dim + p];
            pivot -= value * value * f->diag[p];
        }
        if (!isfinite(pivot) || pivot <= 0.0) {
            rc = orc_fail(ORC_CMG_SETUP, "ldl", "nonpositive pivot at vertex %llu",
                          (unsigned long long)f->perm[col]);
            break;
        }
        f->diag[col] = pivot;
        for (uint64_t row = col + 1; row < dim; ++row) {
            double value = a[row * dim + col];
            for (uint64_t p = 0; p < col; ++p)
                value -= l[row * dim + p] * l[col * dim + p] * f->diag[p];
            l[row * dim + col] = value / pivot;
        }
    }
    free(a);
    free(fidx);
    free(is_anchor);
    if (rc != ORC_OK) {
        free(l);
        *out = f;
        return rc;
    }
    uint64_t strict = 0;
    for (uint64_t r = 0; r < dim; ++r)
        for (uint64_t c = 0; c < r; ++c) strict += l[r * dim + c] != 0.0;
    f->factor_nnz = dim + strict;
    uint64_t packed_slots = dim * (dim ? dim - 1 : 0) / 2;
    uint64_t packed_bytes = packed_slots * 8;
    uint64_t sparse_bytes = strict * (2 * 4 + 2 * 8) + 2 * (dim + 1) * 8;
    if (sparse_bytes < packed_bytes) {
        f->sparse = 1;
        f->row_off = orc_alloc(dim + 1, 8);
        f->row_col = orc_alloc(strict, 4);
        f->row_val = orc_alloc(strict, 8);
        uint64_t k = 0;
        for (uint64_t r = 0; r < dim; ++r) {
            for (uint64_t c = 0; c < r; ++c)
                if (l[r * dim + c] != 0.0) {
                    f->row_col[k] = (uint32_t)c;
                    f->row_val[k] = l[r * dim + c];
                    ++k;
                }
            f->row_off[r + 1] = k;
        }
        /* ldl.rs:233-281: transpose, rows ascending within each column */
        f->col_off = orc_alloc(dim + 1, 8);
        f->col_row = orc_alloc(strict, 4);
        f->col_val = orc_alloc(strict, 8);
        for (uint64_t i = 0; i < strict; ++i) f->col_off[f->row_col[i] + 1]++;
        for (uint64_t c = 0; c < dim; ++c) f->col_off[c + 1] += f->col_off[c];
        uint64_t *next = orc_alloc(dim + 1, 8);
        memcpy(next, f->col_off, dim * 8);
        for (uint64_t r = 0; r < dim; ++r)
            for (uint64_t i = f->row_off[r]; i < f->row_off[r + 1]; ++i) {
                uint64_t d = next[f->row_col[i]]++;
                f->col_row[d] = (uint32_t)r;
                f->col_val[d] = f->row_val[i];
            }
        free(next);
    } else {
        f->packed = orc_alloc(packed_slots, 8);
        uint64_t k = 0;
        for (uint64_t r = 0; r < dim; ++r)
            for (uint64_t c = 0; c < r; ++c) f->packed[k++] = l[r * dim + c];
    }
    free(l);
    *out = f;
    return ORC_OK;
}

static void ldl_free(orc_ldl *f) {
    if (!f) return;
    orc_components_free(&f->comps);
    free(f->anchors); free(f->perm); free(f->packed);
    free(f->row_off); free(f->row_col); free(f->row_val);
    free(f->col_off); free(f->col_row); free(f->col_val);
    free(f->diag); free(f);
}

/* ldl.rs:892-997 (solve_into_compatible, default features) */
static void ldl_solve(const orc_ldl *f, const double *rhs, double *solution,
                      double *forward, double *fsol) {
    uint64_t dim = f->dim;
    for (uint64_t row = 0; row < dim; ++row) {
        double corr = empty_sum();
        if (f->sparse) {
            for (uint64_t i = f->row_off[row]; i < f->row_off[row + 1]; ++i)
                corr = corr + f->row_val[i] * forward[f->row_col[i]];
        } else {
            uint64_t start = row * (row ? row - 1 : 0) / 2;
            for (uint64_t c = 0; c < row; ++c) corr = corr + f->packed[start + c] * forward[c];
        }
        forward[row] = rhs[f->perm[row]] - corr;
    }
    for (uint64_t row = 0; row < dim; ++row) forward[row] /= f->diag[row];
    for (uint64_t r = dim; r-- > 0;) {
        double corr = empty_sum();
        if (f->sparse) {
            for (uint64_t i = f->col_off[r]; i < f->col_off[r + 1]; ++i)
                corr = corr + f->col_val[i] * fsol[f->col_row[i]];
        } else {
            uint64_t index = r * (r + 1) / 2 + r;
            for (uint64_t later = r + 1; later < dim; ++later) {
                corr += f->packed[index] * fsol[later];
                index += later;
            }
        }
        fsol[r] = forward[r] - corr;
    }
    memset(solution, 0, f->n * sizeof *solution);
    for (uint64_t k = 0; k < dim; ++k) solution[f->perm[k]] = fsol[k];
}

/* ======================= hierarchy (hierarchy.rs) ======================== */

static uint64_t repeat_from_nonzeros(uint64_t fine, uint64_t coarse) {
    if (coarse == 0) return 1;
    uint64_t q = fine / coarse;
    q = q > 0 ? q - 1 : 0;
    return q > 1 ? q : 1;
}

enum { TR_DIRECT = 0, TR_FULL = 1, TR_STAGNATED = 2, TR_FILL = 3, TR_MAXLEVELS = 4 };

static orc_graph *graph_clone(const orc_graph *g) {
    orc_graph *c = orc_alloc(1, sizeof *c);
    *c = *g;
    c->u = orc_alloc(g->m, 4); memcpy(c->u, g->u, g->m * 4);
    c->v = orc_alloc(g->m, 4); memcpy(c->v, g->v, g->m * 4);
    c->w = orc_alloc(g->m, 8); memcpy(c->w, g->w, g->m * 8);
    c->diag = orc_alloc(g->n, 8); memcpy(c->diag, g->diag, g->n * 8);
    return c;
}

/* hierarchy.rs:662-683 */
static void make_level(orc_level *level, orc_graph *g, uint32_t *labels,
                       uint64_t coarse, uint64_t repeat, int terminal) {
    level->graph = g;
    level->inv_diag = orc_alloc(g->n, sizeof(double));
    for (uint64_t i = 0; i < g->n; ++i)
        level->inv_diag[i] = g->diag[i] > 0.0 ? 0.5 / g->diag[i] : 0.0;
    level->labels = labels;
    level->coarse = coarse;
    level->repeat = repeat;
    level->terminal = terminal;
}

/* hierarchy.rs:437-618 (build_metered) + preconditioner.rs:874-957 (finalize) */
int orc_cmg_build(const orc_graph *graph, orc_cmg_options opt, orc_cmg **out) {
    if (opt.direct_threshold == 0 || opt.max_levels == 0)
        return orc_fail(ORC_CMG_SETUP, "cmg_options", "invalid options");
    orc_cmg *c = orc_alloc(1, sizeof *c);
    c->levels = orc_alloc(opt.max_levels + 1, sizeof *c->levels);
    uint64_t initial_nnz = graph->nnz, cumulative = 0;
    orc_graph *current = graph_clone(graph);
    int reason;
    for (;;) {
        uint64_t n = current->n;
        if (n == 0 || n <= 1 || n < opt.direct_threshold) {
            reason = TR_DIRECT;
            make_level(&c->levels[c->nlevels++], current, NULL, 0, 0, reason + 1);
            break;
        }
        uint32_t *labels = orc_alloc(n, sizeof *labels);
        uint64_t coarse = aggregation_labels(current, opt.low_degree_threshold, labels);
        int terminal = -1;
        if (coarse == 1) {
            terminal = TR_FULL;
        } else {
            cumulative += current->nnz;
            if (coarse >= (n > 0 ? n - 1 : 0)) terminal = TR_STAGNATED;
            else if ((double)cumulative > opt.max_hierarchy_nnz_factor * (double)initial_nnz)
                terminal = TR_FILL;
            else if (c->nlevels + 1 >= opt.max_levels) terminal = TR_MAXLEVELS;
        }
        if (terminal >= 0) {
            reason = terminal;
            make_level(&c->levels[c->nlevels++], current, labels, coarse, 0, reason + 1);
            break;
        }
        orc_graph *coarse_graph = contract(current, labels, coarse);
        uint64_t repeat = repeat_from_nonzeros(current->nnz, coarse_graph->nnz);
        make_level(&c->levels[c->nlevels++], current, labels, coarse, repeat, 0);
        current = coarse_graph;
    }
    c->terminal_reason = reason;
    c->cumulative_nnz = cumulative;
    c->vertex_counts = orc_alloc(c->nlevels, 8);
    c->nnz_counts = orc_alloc(c->nlevels, 8);
    for (uint64_t i = 0; i < c->nlevels; ++i) {
        c->vertex_counts[i] = c->levels[i].graph->n;
        c->nnz_counts[i] = c->levels[i].graph->nnz;
    }
    orc_components_build(&c->finest, c->levels[0].graph);
    c->coarse_plans = orc_alloc(c->nlevels, sizeof *c->coarse_plans);
    for (uint64_t i = 1; i < c->nlevels; ++i)
        orc_components_build(&c->coarse_plans[i - 1], c->levels[i].graph);
    if (reason == TR_DIRECT) {
        int rc = ldl_factor(c->levels[c->nlevels - 1].graph, &c->ldl);
        if (rc != ORC_OK) {
            orc_cmg_free(c);
            return rc;
        }
    }
    c->repeat_counts = orc_alloc(c->nlevels, 8);
    for (uint64_t i = 0; i < c->nlevels; ++i) c->repeat_counts[i] = c->levels[i].repeat;
    if (c->nlevels >= 2 && c->ldl) {
        uint64_t pen = c->nlevels - 2;
        uint64_t repeat = repeat_from_nonzeros(c->levels[pen].graph->nnz, c->ldl->factor_nnz);
        c->repeat_counts[pen] = repeat;
        c->levels[pen].repeat = repeat;
    }
    *out = c;
    return ORC_OK;
}

void orc_cmg_free(orc_cmg *c) {
    if (!c) return;
    for (uint64_t i = 0; i < c->nlevels; ++i) {
        orc_graph_free(c->levels[i].graph);
        free(c->levels[i].inv_diag);
        free(c->levels[i].labels);
    }
    for (uint64_t i = 1; i < c->nlevels; ++i) orc_components_free(&c->coarse_plans[i - 1]);
    orc_components_free(&c->finest);
    free(c->coarse_plans);
    free(c->levels);
    free(c->vertex_counts);
    free(c->nnz_counts);
    free(c->repeat_counts);
    ldl_free(c->ldl);
    free(c);
}

uint64_t orc_cmg_levels(const orc_cmg *c) { return c->nlevels; }
void orc_cmg_level_info(const orc_cmg *c, uint64_t level, uint64_t info[6]) {
    const orc_level *l = &c->levels[level];
    info[0] = l->graph->n;
    info[1] = l->graph->m;
    info[2] = l->graph->nnz;
    info[3] = c->repeat_counts[level];
    info[4] = (uint64_t)l->terminal;
    info[5] = l->coarse;
}
int orc_cmg_level_labels(const orc_cmg *c, uint64_t level, uint32_t *labels) {
    const orc_level *l = &c->levels[level];
    if (!l->labels) return ORC_INVALID_INPUT;
    memcpy(labels, l->labels, l->graph->n * sizeof *labels);
    return ORC_OK;
}
const orc_graph *orc_cmg_level_graph(const orc_cmg *c, uint64_t level) {
    return c->levels[level].graph;
}

/* ======================= cycle (preconditioner/cycle.rs) ================= */

typedef struct {
    double *residual, *coarse_rhs, *coarse_corr, *sums, *corrs;
    double *forward, *fsol;
} level_ws;

typedef struct {
    const orc_cmg *c;
    level_ws *ws;
} cycle_ctx;

static void cycle_ws_alloc(const orc_cmg *c, level_ws **out) {
    level_ws *ws = orc_alloc(c->nlevels, sizeof *ws);
    for (uint64_t i = 0; i < c->nlevels; ++i) {
        uint64_t n = c->levels[i].graph->n;
        ws[i].residual = orc_alloc(n, 8);
        if (i + 1 < c->nlevels) {
            uint64_t nc = c->levels[i + 1].graph->n;
            ws[i].coarse_rhs = orc_alloc(nc, 8);
            ws[i].coarse_corr = orc_alloc(nc, 8);
            ws[i].sums = orc_alloc(c->coarse_plans[i].count, 8);
            ws[i].corrs = orc_alloc(c->coarse_plans[i].count, 8);
        }
    }
    if (c->ldl) {
        ws[c->nlevels - 1].forward = orc_alloc(c->ldl->dim, 8);
        ws[c->nlevels - 1].fsol = orc_alloc(c->ldl->dim, 8);
    }
    *out = ws;
}

static void cycle_ws_free(const orc_cmg *c, level_ws *ws) {
    for (uint64_t i = 0; i < c->nlevels; ++i) {
        free(ws[i].residual); free(ws[i].coarse_rhs); free(ws[i].coarse_corr);
        free(ws[i].sums); free(ws[i].corrs); free(ws[i].forward); free(ws[i].fsol);
    }
    free(ws);
}

/* cycle.rs:136-311 */
static int cycle_level(cycle_ctx *ctx, uint64_t index, const double *rhs, double *out,
                       uint64_t iterations) {
    const orc_cmg *c = ctx->c;
    const orc_level *level = &c->levels[index];
    uint64_t n = level->graph->n;
    level_ws *ws = &ctx->ws[index];
    if (level->terminal) {
        if (level->terminal == TR_DIRECT + 1) {
            ldl_solve(c->ldl, rhs, out, ws->forward, ws->fsol);
        } else {
            for (uint64_t i = 0; i < n; ++i) out[i] = level->inv_diag[i] * rhs[i];
        }
        return ORC_OK;
    }
    uint64_t child = c->repeat_counts[index];
    uint64_t nc = c->levels[index + 1].graph->n;
    for (uint64_t it = 0; it < iterations; ++it) {
        if (it == 0) {
            for (uint64_t i = 0; i < n; ++i) out[i] = level->inv_diag[i] * rhs[i];
        } else {
            orc_graph_matvec(level->graph, out, ws->residual);
            for (uint64_t i = 0; i < n; ++i)
                out[i] += level->inv_diag[i] * (rhs[i] - ws->residual[i]);
        }
        orc_graph_matvec(level->graph, out, ws->residual);
        memset(ws->coarse_rhs, 0, nc * 8);
        for (uint64_t i = 0; i < n; ++i)
            ws->coarse_rhs[level->labels[i]] += rhs[i] - ws->residual[i];
        int rc = center_plan(&c->coarse_plans[index], ws->coarse_rhs, ws->sums, ws->corrs);
        if (rc != ORC_OK) return rc;
        rc = cycle_level(ctx, index + 1, ws->coarse_rhs, ws->coarse_corr, child);
        if (rc != ORC_OK) return rc;
        for (uint64_t i = 0; i < n; ++i) out[i] += ws->coarse_corr[level->labels[i]];
        orc_graph_matvec(level->graph, out, ws->residual);
        for (uint64_t i = 0; i < n; ++i)
            out[i] += level->inv_diag[i] * (rhs[i] - ws->residual[i]);
    }
    return ORC_OK;
}

static int cycle_apply(const orc_cmg *c, level_ws *ws, const double *rhs, double *out) {
    cycle_ctx ctx = {c, ws};
    return cycle_level(&ctx, 0, rhs, out, 1);
}

/* preconditioner.rs:1041-1110: apply_into projects, apply_compatible_into does not */
int orc_cmg_apply(const orc_cmg *c, const double *rhs, double *out, int project) {
    level_ws *ws;
    cycle_ws_alloc(c, &ws);
    uint64_t n = c->levels[0].graph->n;
    double *b = orc_alloc(n, 8);
    memcpy(b, rhs, n * 8);
    int rc = ORC_OK;
    if (project) {
        double norm;
        rc = orc_components_project(&c->finest, b, &norm);
    }
    if (rc == ORC_OK) rc = cycle_apply(c, ws, b, out);
    free(b);
    cycle_ws_free(c, ws);
    return rc;
}

/* ======================= PCG (pcg.rs) ==================================== */

/* pcg.rs:2124-2158 (serial_norm) */
double orc_serial_norm(const double *v, uint64_t n) {
    double scale = 0.0;
    for (uint64_t i = 0; i < n; ++i) scale = fmax(scale, fabs(v[i]));
    if (scale == 0.0) return 0.0;
    double s = 0.0, k = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        double x = v[i] / scale;
        orc_neumaier(&s, &k, x * x);
    }
    return scale * sqrt(s + k);
}

/* pcg.rs:2014-2032 (serial compensated dot) */
double orc_dot(const double *a, const double *b, uint64_t n) {
    double s = 0.0, k = 0.0;
    for (uint64_t i = 0; i < n; ++i) orc_neumaier(&s, &k, a[i] * b[i]);
    return s + k;
}

/* pcg.rs:2209-2236 */
static double original_residual_norm(const double *rhs, const double *proj,
                                     const double *res, uint64_t n) {
    double scale = 0.0;
    for (uint64_t i = 0; i < n; ++i) scale = fmax(scale, fabs(res[i] + (rhs[i] - proj[i])));
    if (scale == 0.0) return 0.0;
    double s = 0.0, k = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        double x = (res[i] + (rhs[i] - proj[i])) / scale;
        orc_neumaier(&s, &k, x * x);
    }
    return scale * sqrt(s + k);
}

static double allowed(orc_pcg_options o, double rhs_norm, double bound, double x_norm) {
    return o.absolute_tolerance + o.relative_tolerance * (rhs_norm + bound * x_norm);
}

static void fill_diag(orc_pcg_diag *d, const double *x, uint64_t n, uint64_t it,
                      double init, double res, double rhs_norm, double tol, double bound,
                      uint64_t restarts, double proj_norm) {
    double xn = orc_serial_norm(x, n);
    double denom = rhs_norm + bound * xn;
    d->iterations = it;
    d->initial_residual_norm = init;
    d->residual_norm = res;
    d->relative_residual = rhs_norm > 0.0 ? res / rhs_norm : res;
    d->backward_error = denom > 0.0 ? res / denom : 0.0;
    d->tolerance = tol;
    d->restarts = restarts;
    d->rhs_projection_norm = proj_norm;
}

/* pcg.rs:1537-1909 (kernel, serial, connected-or-not default features) */
int orc_pcg_solve(const orc_graph *g, const orc_cmg *c, const double *rhs,
                  const double *guess, double *solution, orc_pcg_options opt,
                  orc_pcg_diag *diag) {
    uint64_t n = g->n;
    double *proj = orc_alloc(n, 8), *x = orc_alloc(n, 8), *r = orc_alloc(n, 8),
           *z = orc_alloc(n, 8), *p = orc_alloc(n, 8), *ap = orc_alloc(n, 8);
    level_ws *ws;
    cycle_ws_alloc(c, &ws);
    const orc_components *comps = &c->finest;
    int rc = ORC_OK;
    memcpy(proj, rhs, n * 8);
    double proj_norm = 0.0;
    rc = orc_components_project(comps, proj, &proj_norm);
    if (rc) goto done;
    if (guess) {
        memcpy(x, guess, n * 8);
        if ((rc = orc_components_center(comps, x))) goto done;
        orc_graph_matvec(g, x, r);
        for (uint64_t i = 0; i < n; ++i) r[i] = proj[i] - r[i];
    } else {
        memcpy(r, proj, n * 8);
    }
    double rhs_norm = orc_serial_norm(rhs, n);
    double init_norm = guess ? original_residual_norm(rhs, proj, r, n) : rhs_norm;
    double proj_init = guess ? orc_serial_norm(r, n) : orc_serial_norm(proj, n);
    double bound = g->norm_bound;
    double x0_norm = guess ? orc_serial_norm(x, n) : 0.0;
    double tol = allowed(opt, rhs_norm, bound, x0_norm);
    if (init_norm <= tol) {
        fill_diag(diag, x, n, 0, init_norm, init_norm, rhs_norm, tol, bound, 0, proj_norm);
        goto done;
    }
    if (proj_init == 0.0) {
        rc = orc_fail(ORC_FULL_RESIDUAL, "pcg", "residual verification failed at 0");
        goto done;
    }
    if ((rc = cycle_apply(c, ws, r, z))) goto done;
    if ((rc = orc_components_center(comps, z))) goto done;
    double rho = orc_dot(r, z, n);
    if (!isfinite(rho) || rho <= 0.0) {
        rc = orc_fail(ORC_PCG_PRECONDITIONER, "pcg", "nonpositive r^T M r at 0");
        goto done;
    }
    memcpy(p, z, n * 8);
    uint64_t restarts = 0;
    double last_tol = tol;
    for (uint64_t it = 1; it <= opt.max_iterations; ++it) {
        orc_graph_matvec(g, p, ap);
        double curv = orc_dot(p, ap, n);
        if (!isfinite(curv) || curv <= 0.0) {
            rc = orc_fail(ORC_PCG_CURVATURE, "pcg", "nonpositive p^T A p at %llu",
                          (unsigned long long)it);
            goto done;
        }
        double alpha = rho / curv;
        if (!isfinite(alpha)) {
            rc = orc_fail(ORC_PCG_CURVATURE, "pcg", "nonfinite alpha");
            goto done;
        }
        for (uint64_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * ap[i];
        }
        if ((rc = orc_components_center(comps, x))) goto done;
        double xn = orc_serial_norm(x, n), rn = orc_serial_norm(r, n);
        last_tol = allowed(opt, rhs_norm, bound, xn);
        int candidate = rn <= last_tol;
        int scheduled = it % opt.recompute_interval == 0;
        int restarted = 0;
        if (candidate || scheduled) {
            orc_graph_matvec(g, x, ap);
            for (uint64_t i = 0; i < n; ++i) ap[i] = proj[i] - ap[i];
            double fresh = orc_serial_norm(ap, n);
            memcpy(r, ap, n * 8);
            restarted = 1;
            ++restarts;
            if (fresh <= last_tol) {
                double orig = original_residual_norm(rhs, proj, ap, n);
                if (orig <= last_tol) {
                    fill_diag(diag, x, n, it, init_norm, orig, rhs_norm, last_tol, bound,
                              restarts, proj_norm);
                    goto done;
                }
                if (it == opt.max_iterations) {
                    rc = orc_fail(ORC_FULL_RESIDUAL, "pcg", "residual verification failed");
                    goto done;
                }
            }
        }
        if (it == opt.max_iterations) break;
        if ((rc = orc_components_center(comps, r))) goto done;
        if ((rc = cycle_apply(c, ws, r, z))) goto done;
        if ((rc = orc_components_center(comps, z))) goto done;
        double new_rho = orc_dot(r, z, n);
        if (!isfinite(new_rho) || new_rho <= 0.0) {
            rc = orc_fail(ORC_PCG_PRECONDITIONER, "pcg", "nonpositive new r^T M r at %llu",
                          (unsigned long long)it);
            goto done;
        }
        if (restarted) {
            memcpy(p, z, n * 8);
        } else {
            double beta = new_rho / rho;
            for (uint64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        }
        rho = new_rho;
    }
    rc = orc_fail(ORC_PCG_MAX_ITERATIONS, "pcg", "maximum iterations reached");
done:
    if (solution) memcpy(solution, x, n * 8);
    cycle_ws_free(c, ws);
    free(proj); free(x); free(r); free(z); free(p); free(ap);
    return rc;
}

/* ======================= frozen digests (tests/frozen_bits.rs) =========== */

typedef struct {
    uint64_t s;
} lcg;
static uint64_t lcg_next(lcg *r) {
    r->s = r->s * 6364136223846793005ULL + 1442695040888963407ULL;
    return r->s >> 11;
}
static uint64_t lcg_below(lcg *r, uint64_t bound) { return lcg_next(r) % bound; }
static double lcg_weight(lcg *r) {
    static const double scale[4] = {1.0, 0.5, 4.0, 1.0 / 64.0};
    double base = (double)(1 + lcg_below(r, 64)) / 8.0;
    return base * scale[lcg_below(r, 4)];
}

typedef struct {
    uint64_t n, m, cap;
    uint32_t *u, *v;
    double *w;
} edge_list;
static void el_push(edge_list *e, uint64_t u, uint64_t v, double w) {
    if (e->m == e->cap) {
        e->cap = e->cap ? 2 * e->cap : 1024;
        e->u = realloc(e->u, e->cap * 4);
        e->v = realloc(e->v, e->cap * 4);
        e->w = realloc(e->w, e->cap * 8);
    }
    e->u[e->m] = (uint32_t)u;
    e->v[e->m] = (uint32_t)v;
    e->w[e->m] = w;
    e->m++;
}

/* frozen_bits.rs:104-122 */
static void chain_edges(edge_list *e) {
    for (uint64_t v = 1; v < 300; ++v) {
        double w = v == 150 ? 1.0 / 4096.0 : 1.0 + (double)(v % 7) / 4.0;
        el_push(e, v - 1, v, w);
        if (v >= 3 && v % 5 == 0) el_push(e, v - 3, v, 0.125);
    }
    for (uint64_t v = 301; v < 360; ++v) el_push(e, v - 1, v, 2.0 + (double)(v % 3));
    e->n = 364;
}

/* frozen_bits.rs:125-145 */
static void random_edges(edge_list *e, uint64_t vertices, uint64_t count, uint64_t seed) {
    lcg r = {seed};
    while (e->m < count) {
        uint64_t u = lcg_below(&r, vertices - 7);
        uint64_t v;
        if (lcg_below(&r, 4) == 0) v = (u + 1 + lcg_below(&r, 3)) % (vertices - 7);
        else v = lcg_below(&r, vertices - 7);
        if (u != v) {
            double w = lcg_weight(&r);
            el_push(e, u, v, w);
            if (lcg_below(&r, 9) == 0) el_push(e, v, u, w * 0.75);
        }
    }
    e->m = count;
    e->n = vertices;
}

/* frozen_bits.rs:148-166 */
static void grid_edges(edge_list *e, uint64_t rows, uint64_t cols) {
    for (uint64_t r = 0; r < rows; ++r)
        for (uint64_t c = 0; c < cols; ++c) {
            if (c + 1 < cols) el_push(e, r * cols + c, r * cols + c + 1, 1.0 + (double)((r * 7 + c) % 5) / 8.0);
            if (r + 1 < rows) el_push(e, r * cols + c, (r + 1) * cols + c, 0.5 + (double)((r + 3 * c) % 11) / 16.0);
        }
    e->n = rows * cols;
}

static void dg_usize(uint64_t *h, uint64_t v) { orc_fnv_word(h, v); }
static void dg_f64(uint64_t *h, double v) { orc_fnv_word(h, orc_bits(v)); }
static void dg_values(uint64_t *h, const double *v, uint64_t n) {
    dg_usize(h, n);
    for (uint64_t i = 0; i < n; ++i) dg_f64(h, v[i]);
}

/* frozen_bits.rs:184-195 */
static void graph_digest(uint64_t *h, const orc_graph *g) {
    dg_usize(h, g->n);
    dg_usize(h, g->m);
    for (uint64_t e = 0; e < g->m; ++e) {
        dg_usize(h, g->u[e]);
        dg_usize(h, g->v[e]);
        dg_f64(h, g->w[e]);
    }
    dg_values(h, g->diag, g->n);
    dg_usize(h, g->nnz);
    dg_f64(h, g->norm_bound);
}

/* frozen_bits.rs:197-203 */
static double *compatible_rhs(const orc_graph *g, uint64_t seed) {
    lcg r = {seed};
    double *known = orc_alloc(g->n, 8), *out = orc_alloc(g->n, 8);
    for (uint64_t i = 0; i < g->n; ++i) known[i] = (double)lcg_below(&r, 2001) / 1000.0 - 1.0;
    orc_graph_matvec(g, known, out);
    free(known);
    return out;
}

static int preconditioner_digest(const orc_graph *g, const orc_cmg *c, uint64_t *out) {
    uint64_t h = 0xcbf29ce484222325ULL;
    dg_usize(&h, (uint64_t)c->terminal_reason);
    dg_usize(&h, c->nlevels);
    for (uint64_t i = 0; i < c->nlevels; ++i) dg_usize(&h, c->vertex_counts[i]);
    dg_usize(&h, c->nlevels);
    for (uint64_t i = 0; i < c->nlevels; ++i) dg_usize(&h, c->nnz_counts[i]);
    dg_usize(&h, c->cumulative_nnz);
    for (uint64_t i = 0; i < c->nlevels; ++i) {
        const orc_level *l = &c->levels[i];
        graph_digest(&h, l->graph);
        dg_values(&h, l->inv_diag, l->graph->n);
        dg_usize(&h, l->repeat);
        if (l->labels) {
            dg_usize(&h, l->graph->n);
            for (uint64_t k = 0; k < l->graph->n; ++k) dg_usize(&h, l->labels[k]);
            dg_usize(&h, l->coarse);
        }
    }
    dg_usize(&h, c->nlevels);
    for (uint64_t i = 0; i < c->nlevels; ++i) dg_usize(&h, c->repeat_counts[i]);
    int rc = ORC_OK;
    if (c->ldl) {
        const orc_ldl *f = c->ldl;
        dg_usize(&h, f->comps.count);
        for (uint64_t k = 0; k < f->comps.count; ++k) dg_usize(&h, f->anchors[k]);
        dg_usize(&h, f->dim);
        for (uint64_t k = 0; k < f->dim; ++k) dg_usize(&h, f->perm[k]);
        dg_usize(&h, f->dim);
        dg_usize(&h, f->factor_nnz);
        const orc_graph *t = c->levels[c->nlevels - 1].graph;
        double *b = compatible_rhs(t, 91);
        /* ldl.rs:866-890: validate_rhs then solve */
        double *bb = orc_alloc(t->n, 8);
        memcpy(bb, b, t->n * 8);
        double *x = orc_alloc(t->n, 8), *fw = orc_alloc(f->dim, 8), *fs = orc_alloc(f->dim, 8);
        ldl_solve(f, bb, x, fw, fs);
        dg_values(&h, x, t->n);
        free(b); free(bb); free(x); free(fw); free(fs);
    }
    double *o = orc_alloc(g->n, 8);
    for (uint64_t seed = 1; seed <= 3; ++seed) {
        double *b = compatible_rhs(g, seed);
        rc = orc_cmg_apply(c, b, o, 0);
        free(b);
        if (rc) break;
        dg_values(&h, o, g->n);
    }
    if (!rc) {
        double *b = compatible_rhs(g, 4);
        for (uint64_t i = 0; i < g->n; ++i) b[i] += 1e-14 * (double)(c->finest.labels[i] % 3);
        rc = orc_cmg_apply(c, b, o, 1);
        free(b);
        if (!rc) dg_values(&h, o, g->n);
    }
    free(o);
    *out = h;
    return rc;
}

static void diag_digest(uint64_t *h, const orc_pcg_diag *d) {
    dg_usize(h, d->iterations);
    dg_f64(h, d->initial_residual_norm);
    dg_f64(h, d->residual_norm);
    dg_f64(h, d->relative_residual);
    dg_f64(h, d->backward_error);
    dg_f64(h, d->tolerance);
    dg_f64(h, d->rhs_projection_norm);
    dg_usize(h, d->restarts);
}

int orc_frozen_digest(const char *family, const char *kind, uint64_t *out) {
    edge_list e = {0};
    uint64_t threshold;
    if (!strcmp(family, "chain")) { chain_edges(&e); threshold = 2; }
    else if (!strcmp(family, "chain_direct")) { chain_edges(&e); threshold = 128; }
    else if (!strcmp(family, "random")) { random_edges(&e, 3000, 12000, 7); threshold = 64; }
    else if (!strcmp(family, "grid")) { grid_edges(&e, 260, 262); threshold = 700; }
    else return orc_fail(ORC_INVALID_INPUT, "digest", "unknown family %s", family);
    orc_graph *g = orc_graph_from_edges(e.n, e.m, e.u, e.v, e.w);
    free(e.u); free(e.v); free(e.w);
    if (!g) return ORC_CMG_SETUP;
    int rc = ORC_OK;
    uint64_t h = 0xcbf29ce484222325ULL;
    if (!strcmp(kind, "graph_from_edges")) {
        graph_digest(&h, g);
        *out = h;
        orc_graph_free(g);
        return ORC_OK;
    }
    orc_cmg_options opt = {threshold, 32, 5.0, 0.125};
    orc_cmg *c = NULL;
    if ((rc = orc_cmg_build(g, opt, &c))) {
        orc_graph_free(g);
        return rc;
    }
    if (!strcmp(kind, "build")) {
        rc = preconditioner_digest(g, c, out);
    } else if (!strcmp(kind, "pcg") || !strcmp(kind, "pcg_guess")) {
        orc_pcg_options po = {1e-10, 0.0, 500, 25};
        double *b = compatible_rhs(g, 21);
        double *x = orc_alloc(g->n, 8), *guess = NULL;
        if (!strcmp(kind, "pcg_guess")) {
            guess = orc_alloc(g->n, 8);
            for (uint64_t i = 0; i < g->n; ++i) guess[i] = (double)(i % 3) * 0.1;
        }
        orc_pcg_diag d;
        rc = orc_pcg_solve(g, c, b, guess, x, po, &d);
        if (!rc) {
            dg_values(&h, x, g->n);
            diag_digest(&h, &d);
            *out = h;
        }
        free(b); free(x); free(guess);
    } else {
        rc = orc_fail(ORC_INVALID_INPUT, "digest", "unknown kind %s", kind);
    }
    orc_cmg_free(c);
    orc_graph_free(g);
    return rc;
}
