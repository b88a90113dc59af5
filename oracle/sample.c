/* Sample pipeline restatement (test infrastructure): validation
 * (types.rs:164-263), canonicalization (problem.rs:90-175, 696-746,
 * 1213-1248), match-deletion leave-out graph selection (graph.rs:262-512,
 * 625-1100), compression (problem.rs:194-606, 748-957) and the no-control JLA
 * plan (jla.rs:48-105, 413-790). */
#define _GNU_SOURCE
#include <stdio.h>

#include "orc_internal.h"
#include "sample.h"

/* ---------------- small utilities ---------------- */

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y);
}

/* problem.rs:696-746 (redense): sorted distinct levels, dense u32 ids */
static uint32_t *redense(const uint64_t *values, uint64_t n, uint64_t *levels_out,
                         uint64_t **levels_ptr) {
    uint64_t *levels = orc_alloc(n, 8);
    memcpy(levels, values, n * 8);
    qsort(levels, n, 8, cmp_u64);
    uint64_t k = 0;
    for (uint64_t i = 0; i < n; ++i)
        if (i == 0 || levels[i] != levels[k - 1]) levels[k++] = levels[i];
    uint32_t *dense = orc_alloc(n, 4);
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t lo = 0, hi = k;
        while (lo < hi) {
            uint64_t mid = (lo + hi) / 2;
            if (levels[mid] < values[i]) lo = mid + 1;
            else hi = mid;
        }
        dense[i] = (uint32_t)lo;
    }
    *levels_out = k;
    if (levels_ptr) *levels_ptr = levels;
    else free(levels);
    return dense;
}

static uint64_t uf_find(uint64_t *parent, uint64_t x) {
    uint64_t r = x;
    while (parent[r] != r) r = parent[r];
    while (parent[x] != r) {
        uint64_t nx = parent[x];
        parent[x] = r;
        x = nx;
    }
    return r;
}

/* ---------------- graph selection (graph.rs) ---------------- */

typedef struct {
    uint64_t rows, W, F, D;
    const uint32_t *w, *f, *d;
    const uint64_t *freq;
    uint32_t *del_worker, *del_firm; /* coordinate of each deletion id */
} sel_ctx;

/* graph.rs:625-731 (largest_component): components of the coordinate graph
 * ranked by (firm count, physical mass); a tie is GraphUnidentified. */
static int largest_component(const sel_ctx *s, uint8_t *active, uint64_t *components) {
    uint64_t nodes = s->W + s->F;
    uint64_t *parent = orc_alloc(nodes, 8);
    uint8_t *has_edge = orc_alloc(nodes, 1);
    for (uint64_t i = 0; i < nodes; ++i) parent[i] = i;
    for (uint64_t r = 0; r < s->rows; ++r) {
        if (!active[r]) continue;
        uint64_t a = uf_find(parent, s->w[r]), b = uf_find(parent, s->W + s->f[r]);
        has_edge[s->w[r]] = has_edge[s->W + s->f[r]] = 1;
        if (a != b) parent[a > b ? a : b] = a < b ? a : b;
    }
    uint64_t *label = orc_alloc(nodes, 8);
    uint64_t count = 0;
    for (uint64_t i = 0; i < nodes; ++i) label[i] = UINT64_MAX;
    for (uint64_t i = 0; i < nodes; ++i) {
        if (!has_edge[i]) continue;
        uint64_t r = uf_find(parent, i);
        if (label[r] == UINT64_MAX) label[r] = count++;
        label[i] = label[r];
    }
    if (count == 0) {
        free(parent); free(has_edge); free(label);
        return orc_fail(ORC_GRAPH_EMPTY, "graph", "no connected worker-firm component remains");
    }
    uint64_t *firms = orc_alloc(count, 8), *mass = orc_alloc(count, 8);
    for (uint64_t f = 0; f < s->F; ++f)
        if (label[s->W + f] != UINT64_MAX) firms[label[s->W + f]]++;
    for (uint64_t r = 0; r < s->rows; ++r)
        if (active[r]) mass[label[s->w[r]]] += s->freq[r];
    uint64_t best = 0, winners = 0;
    for (uint64_t c = 1; c < count; ++c)
        if (firms[c] > firms[best] || (firms[c] == firms[best] && mass[c] > mass[best])) best = c;
    for (uint64_t c = 0; c < count; ++c) winners += firms[c] == firms[best] && mass[c] == mass[best];
    int rc = ORC_OK;
    if (winners != 1) {
        rc = orc_fail(ORC_GRAPH_UNIDENTIFIED, "graph",
                      "multiple connected components tie for largest-component selection");
    } else {
        for (uint64_t r = 0; r < s->rows; ++r)
            if (active[r] && label[s->w[r]] != best) active[r] = 0;
    }
    if (components) *components = count;
    free(parent); free(has_edge); free(label); free(firms); free(mass);
    return rc;
}

/* graph.rs:793-808 (distinct_deletion_counts) */
static uint64_t *worker_units(const sel_ctx *s, const uint8_t *active) {
    uint8_t *seen = orc_alloc(s->D, 1);
    uint64_t *count = orc_alloc(s->W, 8);
    for (uint64_t r = 0; r < s->rows; ++r)
        if (active[r] && !seen[s->d[r]]) {
            seen[s->d[r]] = 1;
            count[s->w[r]]++;
        }
    free(seen);
    return count;
}

typedef struct {
    uint64_t *off;
    uint64_t *to;
    uint64_t *edge;
} adjacency;

static void adjacency_free(adjacency *a) { free(a->off); free(a->to); free(a->edge); }

/* adjacency with neighbours sorted by (to, edge): graph.rs:733-791 */
static void build_adjacency(uint64_t nodes, uint64_t m, const uint64_t *eu,
                            const uint64_t *ev, adjacency *a) {
    a->off = orc_alloc(nodes + 1, 8);
    a->to = orc_alloc(2 * m, 8);
    a->edge = orc_alloc(2 * m, 8);
    for (uint64_t e = 0; e < m; ++e) {
        a->off[eu[e] + 1]++;
        a->off[ev[e] + 1]++;
    }
    for (uint64_t i = 0; i < nodes; ++i) a->off[i + 1] += a->off[i];
    uint64_t *cur = orc_alloc(nodes, 8);
    memcpy(cur, a->off, nodes * 8);
    for (uint64_t e = 0; e < m; ++e) {
        uint64_t p = cur[eu[e]]++;
        a->to[p] = ev[e];
        a->edge[p] = e;
        p = cur[ev[e]]++;
        a->to[p] = eu[e];
        a->edge[p] = e;
    }
    free(cur);
}

/* Iterative DFS low-link.  mode 0: articulation vertices (graph.rs:872-936),
 * mode 1: bridge edges (graph.rs:1023-1084). */
static void lowlink(uint64_t nodes, const adjacency *a, uint64_t m, int mode,
                    uint8_t *flag_out) {
    uint64_t *disc = orc_alloc(nodes, 8), *low = orc_alloc(nodes, 8);
    uint64_t *parent = orc_alloc(nodes, 8), *children = orc_alloc(nodes, 8);
    uint64_t *stack_node = orc_alloc(nodes + 1, 8), *stack_pedge = orc_alloc(nodes + 1, 8),
             *stack_next = orc_alloc(nodes + 1, 8);
    uint64_t clock = 0;
    for (uint64_t i = 0; i < nodes; ++i) parent[i] = UINT64_MAX;
    (void)m;
    for (uint64_t root = 0; root < nodes; ++root) {
        if (a->off[root] == a->off[root + 1] || disc[root]) continue;
        disc[root] = low[root] = ++clock;
        uint64_t top = 0;
        stack_node[0] = root;
        stack_pedge[0] = UINT64_MAX;
        stack_next[0] = a->off[root];
        for (;;) {
            uint64_t node = stack_node[top];
            if (stack_next[top] < a->off[node + 1]) {
                uint64_t p = stack_next[top]++;
                uint64_t to = a->to[p], edge = a->edge[p];
                if (edge == stack_pedge[top]) continue;
                if (!disc[to]) {
                    parent[to] = node;
                    children[node]++;
                    disc[to] = low[to] = ++clock;
                    ++top;
                    stack_node[top] = to;
                    stack_pedge[top] = edge;
                    stack_next[top] = a->off[to];
                } else if (disc[to] < low[node]) {
                    low[node] = disc[to];
                }
            } else {
                uint64_t pedge = stack_pedge[top];
                uint64_t par = parent[node];
                if (par != UINT64_MAX) {
                    if (mode == 1 && low[node] > disc[par]) flag_out[pedge] = 1;
                    if (low[node] < low[par]) low[par] = low[node];
                    if (mode == 0 && parent[par] != UINT64_MAX && low[node] >= disc[par])
                        flag_out[par] = 1;
                }
                if (top == 0) break;
                --top;
            }
        }
        if (mode == 0) flag_out[root] = children[root] > 1;
    }
    free(disc); free(low); free(parent); free(children);
    free(stack_node); free(stack_pedge); free(stack_next);
}

/* distinct active coordinates (w, f) as edges of the coordinate graph */
static uint64_t active_cells(const sel_ctx *s, const uint8_t *active, uint64_t **keys_out) {
    uint64_t *keys = orc_alloc(s->rows, 8), k = 0;
    for (uint64_t r = 0; r < s->rows; ++r)
        if (active[r]) keys[k++] = (uint64_t)s->w[r] * s->F + s->f[r];
    qsort(keys, k, 8, cmp_u64);
    uint64_t u = 0;
    for (uint64_t i = 0; i < k; ++i)
        if (i == 0 || keys[i] != keys[u - 1]) keys[u++] = keys[i];
    *keys_out = keys;
    return u;
}

/* graph.rs:872-886 (worker_articulations) */
static uint8_t *worker_articulations(const sel_ctx *s, const uint8_t *active) {
    uint64_t *keys;
    uint64_t m = active_cells(s, active, &keys);
    uint64_t *eu = orc_alloc(m, 8), *ev = orc_alloc(m, 8);
    for (uint64_t e = 0; e < m; ++e) {
        eu[e] = keys[e] / s->F;
        ev[e] = s->W + keys[e] % s->F;
    }
    uint64_t nodes = s->W + s->F;
    adjacency a;
    build_adjacency(nodes, m, eu, ev, &a);
    uint8_t *flag = orc_alloc(nodes, 1);
    lowlink(nodes, &a, m, 0, flag);
    adjacency_free(&a);
    free(keys); free(eu); free(ev);
    return flag; /* first W entries are worker flags */
}

/* graph.rs:938-1021 (deletion_bridges): edges are active deletion units */
static uint64_t deletion_bridges(const sel_ctx *s, const uint8_t *active, uint8_t *is_bridge) {
    uint8_t *seen = orc_alloc(s->D, 1);
    for (uint64_t r = 0; r < s->rows; ++r)
        if (active[r]) seen[s->d[r]] = 1;
    uint64_t m = 0;
    for (uint64_t d = 0; d < s->D; ++d) m += seen[d];
    uint64_t *eu = orc_alloc(m, 8), *ev = orc_alloc(m, 8), *ed = orc_alloc(m, 8);
    uint64_t e = 0;
    for (uint64_t d = 0; d < s->D; ++d)
        if (seen[d]) {
            eu[e] = s->del_worker[d];
            ev[e] = s->W + s->del_firm[d];
            ed[e] = d;
            ++e;
        }
    uint64_t nodes = s->W + s->F, bridges = 0;
    memset(is_bridge, 0, s->D);
    if (m) {
        adjacency a;
        build_adjacency(nodes, m, eu, ev, &a);
        uint8_t *flag = orc_alloc(m, 1);
        lowlink(nodes, &a, m, 1, flag);
        for (uint64_t k = 0; k < m; ++k)
            if (flag[k]) {
                is_bridge[ed[k]] = 1;
                ++bridges;
            }
        adjacency_free(&a);
        free(flag);
    }
    free(seen); free(eu); free(ev); free(ed);
    return bridges;
}

/* graph.rs:262-512 (select_match_deletion_graph_standard_with_interrupt) */
static int select_match(const sel_ctx *s, uint8_t *active, orc_selection_receipt *rec) {
    memset(rec, 0, sizeof *rec);
    rec->input_rows = s->rows;
    for (uint64_t r = 0; r < s->rows; ++r) active[r] = 1;
    int rc;
    uint64_t comps;
    if ((rc = largest_component(s, active, &comps))) return rc;
    uint64_t *units = worker_units(s, active);
    for (uint64_t r = 0; r < s->rows; ++r)
        if (active[r]) active[r] = units[s->w[r]] > 1;
    free(units);
    uint64_t movers = 0;
    for (uint64_t r = 0; r < s->rows; ++r) movers += active[r];
    if (movers == 0)
        return orc_fail(ORC_GRAPH_EMPTY, "graph",
                        "no observations remain in the mover target population");
    if ((rc = largest_component(s, active, &comps))) return rc;
    uint8_t *is_bridge = orc_alloc(s->D, 1);
    uint64_t initial_edges = 0;
    {
        uint8_t *seen = orc_alloc(s->D, 1);
        for (uint64_t r = 0; r < s->rows; ++r)
            if (active[r] && !seen[s->d[r]]) { seen[s->d[r]] = 1; ++initial_edges; }
        free(seen);
    }
    uint64_t bound = s->W + initial_edges + 1;
    for (;;) {
        if (rec->fixed_point_iterations > bound) {
            free(is_bridge);
            return orc_fail(ORC_GRAPH_CERTIFICATE, "graph", "fixed point exceeded its bound");
        }
        if ((rc = largest_component(s, active, &comps))) { free(is_bridge); return rc; }
        units = worker_units(s, active);
        uint64_t removed = 0;
        for (uint64_t w = 0; w < s->W; ++w) removed += units[w] == 1;
        if (removed) {
            for (uint64_t r = 0; r < s->rows; ++r)
                if (active[r] && units[s->w[r]] == 1) active[r] = 0;
            free(units);
            rec->insufficient_workers_removed += removed;
            rec->degree_iterations++;
            rec->fixed_point_iterations++;
            continue;
        }
        free(units);
        uint8_t *art = worker_articulations(s, active);
        uint64_t arts = 0;
        for (uint64_t w = 0; w < s->W; ++w) arts += art[w];
        if (arts) {
            for (uint64_t r = 0; r < s->rows; ++r)
                if (active[r] && art[s->w[r]]) active[r] = 0;
            free(art);
            rec->articulation_workers_removed += arts;
            rec->articulation_iterations++;
            rec->fixed_point_iterations++;
            continue;
        }
        free(art);
        uint64_t bridges = deletion_bridges(s, active, is_bridge);
        if (!bridges) break;
        uint64_t removed_rows = 0;
        for (uint64_t r = 0; r < s->rows; ++r)
            if (active[r] && is_bridge[s->d[r]]) {
                active[r] = 0;
                ++removed_rows;
            }
        rec->bridge_units_removed += bridges;
        rec->bridge_rows_removed += removed_rows;
        rec->bridge_iterations++;
        rec->fixed_point_iterations++;
    }
    if ((rc = largest_component(s, active, &comps))) { free(is_bridge); return rc; }
    if (deletion_bridges(s, active, is_bridge)) {
        free(is_bridge);
        return orc_fail(ORC_GRAPH_CERTIFICATE, "graph", "final graph still contains a bridge");
    }
    free(is_bridge);
    uint64_t retained = 0, mass = 0;
    for (uint64_t r = 0; r < s->rows; ++r)
        if (active[r]) {
            ++retained;
            mass += s->freq[r];
        }
    if (!retained) return orc_fail(ORC_GRAPH_EMPTY, "graph", "no component remains");
    rec->retained_rows = retained;
    rec->retained_physical_mass = mass;
    return ORC_OK;
}

/* graph.rs:82-242 (select_observation_deletion_graph_with_interrupt): the
 * largest component, then to a fixed point drop workers with one physical
 * observation, else the articulation workers of the coordinate graph. */
static int select_observation(const sel_ctx *s, uint8_t *active, orc_selection_receipt *rec) {
    memset(rec, 0, sizeof *rec);
    rec->input_rows = s->rows;
    for (uint64_t r = 0; r < s->rows; ++r) active[r] = 1;
    int rc;
    uint64_t comps;
    if ((rc = largest_component(s, active, &comps))) return rc;
    const uint64_t bound = s->W + 1;
    uint64_t *phys = orc_alloc(s->W, 8);
    for (;;) {
        if (rec->fixed_point_iterations > bound) {
            free(phys);
            return orc_fail(ORC_GRAPH_CERTIFICATE, "graph", "observation pruning exceeded its bound");
        }
        if ((rc = largest_component(s, active, &comps))) { free(phys); return rc; }
        memset(phys, 0, s->W * 8);
        for (uint64_t r = 0; r < s->rows; ++r)
            if (active[r]) phys[s->w[r]] += s->freq[r];
        uint64_t removed = 0;
        for (uint64_t w = 0; w < s->W; ++w) removed += phys[w] == 1;
        if (removed) {
            for (uint64_t r = 0; r < s->rows; ++r)
                if (active[r] && phys[s->w[r]] == 1) active[r] = 0;
            rec->insufficient_workers_removed += removed;
            rec->degree_iterations++;
            rec->fixed_point_iterations++;
            continue;
        }
        uint8_t *art = worker_articulations(s, active);
        uint64_t arts = 0;
        for (uint64_t w = 0; w < s->W; ++w) arts += art[w];
        if (!arts) { free(art); break; }
        for (uint64_t r = 0; r < s->rows; ++r)
            if (active[r] && art[s->w[r]]) active[r] = 0;
        free(art);
        rec->articulation_workers_removed += arts;
        rec->articulation_iterations++;
        rec->fixed_point_iterations++;
    }
    free(phys);
    if ((rc = largest_component(s, active, &comps))) return rc;
    uint64_t retained = 0, mass = 0;
    for (uint64_t r = 0; r < s->rows; ++r)
        if (active[r]) {
            ++retained;
            mass += s->freq[r];
        }
    if (!retained) return orc_fail(ORC_GRAPH_EMPTY, "graph", "no component remains");
    rec->retained_rows = retained;
    rec->retained_physical_mass = mass;
    return ORC_OK;
}

/* ---------------- compression (problem.rs) ---------------- */

typedef struct {
    const orc_problem *p;
    const uint64_t *src;      /* retained source rows */
    const double *outcome, *target;
    const uint64_t *freq;
    const double *per_copy;
    const uint64_t *rank;
} sort_ctx;

static int cmp_cell_rows(const void *a, const void *b, void *arg) {
    const orc_problem *p = arg;
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    if (p->row_worker[x] != p->row_worker[y]) return p->row_worker[x] < p->row_worker[y] ? -1 : 1;
    if (p->row_firm[x] != p->row_firm[y]) return p->row_firm[x] < p->row_firm[y] ? -1 : 1;
    return x < y ? -1 : (x > y);
}

static int cmp_firm_cells(const void *a, const void *b, void *arg) {
    const orc_problem *p = arg;
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    if (p->cell_firm[x] != p->cell_firm[y]) return p->cell_firm[x] < p->cell_firm[y] ? -1 : 1;
    if (p->cell_worker[x] != p->cell_worker[y]) return p->cell_worker[x] < p->cell_worker[y] ? -1 : 1;
    return x < y ? -1 : (x > y);
}

/* problem.rs:1189-1211 (compare_target_rows, no controls) */
static int cmp_target_rows_key(const orc_problem *p, uint64_t x, uint64_t y) {
    if (p->row_worker[x] != p->row_worker[y]) return p->row_worker[x] < p->row_worker[y] ? -1 : 1;
    if (p->row_firm[x] != p->row_firm[y]) return p->row_firm[x] < p->row_firm[y] ? -1 : 1;
    if (p->row_deletion[x] != p->row_deletion[y]) return p->row_deletion[x] < p->row_deletion[y] ? -1 : 1;
    int c = orc_total_cmp(p->outcome[x], p->outcome[y]);
    if (c) return c;
    if (p->frequency[x] != p->frequency[y]) return p->frequency[x] < p->frequency[y] ? -1 : 1;
    return orc_total_cmp(p->target[x], p->target[y]);
}
static int cmp_target_rows(const void *a, const void *b, void *arg) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    int c = cmp_target_rows_key(arg, x, y);
    if (c) return c;
    return x < y ? -1 : (x > y);
}

/* jla.rs:518-533 (compare_semantic_rows) */
static int cmp_semantic_key(const orc_problem *p, uint64_t x, uint64_t y) {
    if (p->row_worker[x] != p->row_worker[y]) return p->row_worker[x] < p->row_worker[y] ? -1 : 1;
    if (p->row_firm[x] != p->row_firm[y]) return p->row_firm[x] < p->row_firm[y] ? -1 : 1;
    if (p->row_deletion[x] != p->row_deletion[y]) return p->row_deletion[x] < p->row_deletion[y] ? -1 : 1;
    int c = orc_total_cmp(orc_canonical_zero(p->per_copy[x]), orc_canonical_zero(p->per_copy[y]));
    if (c) return c;
    return orc_total_cmp(orc_canonical_zero(p->outcome[x]), orc_canonical_zero(p->outcome[y]));
}
static int cmp_semantic_rows(const void *a, const void *b, void *arg) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    int c = cmp_semantic_key(arg, x, y);
    if (c) return c;
    return x < y ? -1 : (x > y);
}

/* generic_jla.rs:4305-4334 (compare_observation_rows), no controls and no
 * probe order: worker, firm, canonical per-copy target mass, outcome */
static int cmp_observation_key(const orc_problem *p, uint64_t x, uint64_t y) {
    if (p->row_worker[x] != p->row_worker[y]) return p->row_worker[x] < p->row_worker[y] ? -1 : 1;
    if (p->row_firm[x] != p->row_firm[y]) return p->row_firm[x] < p->row_firm[y] ? -1 : 1;
    int c = orc_total_cmp(orc_canonical_zero(p->per_copy[x]), orc_canonical_zero(p->per_copy[y]));
    if (c) return c;
    return orc_total_cmp(p->outcome[x], p->outcome[y]);
}
static int cmp_observation_rows(const void *a, const void *b, void *arg) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    int c = cmp_observation_key(arg, x, y);
    if (c) return c;
    return x < y ? -1 : (x > y);
}

/* Dense 1-based observation ranks of every stored row; `order` returns the
 * rows sorted by (rank, row) (generic_jla.rs:4267-4303). */
static void observation_ranks(const orc_problem *p, uint64_t *order, uint64_t *rank) {
    for (uint64_t i = 0; i < p->rows; ++i) order[i] = i;
    qsort_r(order, p->rows, 8, cmp_observation_rows, (void *)p);
    uint64_t current = 0;
    for (uint64_t i = 0; i < p->rows; ++i) {
        if (i == 0 || cmp_observation_key(p, order[i - 1], order[i]) != 0) ++current;
        rank[order[i]] = current;
    }
}

/* jla.rs:636-646 (target_plan ordering) */
static int cmp_target_plan(const void *a, const void *b, void *arg) {
    const orc_problem *p = arg;
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    if (p->row_cell[x] != p->row_cell[y]) return p->row_cell[x] < p->row_cell[y] ? -1 : 1;
    int c = orc_total_cmp(orc_canonical_zero(p->per_copy[x]), orc_canonical_zero(p->per_copy[y]));
    if (c) return c;
    if (p->row_rank[x] != p->row_rank[y]) return p->row_rank[x] < p->row_rank[y] ? -1 : 1;
    return x < y ? -1 : (x > y);
}

static void stable_add(double *sum, double *corr, double v) { orc_neumaier(sum, corr, v); }

/* topology checksum: problem.rs:1213-1241 */
typedef struct {
    uint32_t w, f, d;
    uint64_t q;
} topo_rec;
static int cmp_topo(const void *a, const void *b) {
    const topo_rec *x = a, *y = b;
    if (x->w != y->w) return x->w < y->w ? -1 : 1;
    if (x->f != y->f) return x->f < y->f ? -1 : 1;
    if (x->d != y->d) return x->d < y->d ? -1 : 1;
    return x->q < y->q ? -1 : (x->q > y->q);
}
static uint64_t topology_checksum(uint64_t n, const uint32_t *w, const uint32_t *f,
                                  const uint32_t *d, const uint64_t *q) {
    topo_rec *r = orc_alloc(n, sizeof *r);
    for (uint64_t i = 0; i < n; ++i) {
        r[i].w = w[i]; r[i].f = f[i]; r[i].d = d[i]; r[i].q = q[i];
    }
    qsort(r, n, sizeof *r, cmp_topo);
    uint64_t h = 0xcbf29ce484222325ULL;
    for (uint64_t i = 0; i < n; ++i) {
        orc_fnv_word(&h, r[i].w);
        orc_fnv_word(&h, r[i].f);
        orc_fnv_word(&h, r[i].d);
        orc_fnv_word(&h, r[i].q);
    }
    free(r);
    return h;
}

static uint32_t *redense_selected(const uint32_t *values, const uint64_t *src, uint64_t n,
                                  uint64_t levels, uint64_t *count) {
    uint8_t *present = orc_alloc(levels, 1);
    for (uint64_t i = 0; i < n; ++i) present[values[src[i]]] = 1;
    uint32_t *map = orc_alloc(levels, 4);
    uint32_t next = 0;
    for (uint64_t l = 0; l < levels; ++l) map[l] = present[l] ? next++ : UINT32_MAX;
    *count = next;
    free(present);
    return map;
}

int orc_prepare(uint64_t rows, const uint64_t *worker, const uint64_t *firm,
                const uint64_t *deletion, const double *outcome,
                const uint64_t *frequency, const double *target, orc_problem **out) {
    return orc_prepare_masked(rows, worker, firm, deletion, outcome, frequency, target, NULL, out);
}

static int prepare_impl(uint64_t rows, const uint64_t *worker, const uint64_t *firm,
                        const uint64_t *deletion, const double *outcome, const uint64_t *frequency,
                        const double *target, const uint8_t *mask, int observation, orc_problem **out);

/* As orc_prepare; a non-NULL mask replaces graph selection (problem.rs:194
 * `compress(active)` called with a caller mask, as the reference unit tests do). */
int orc_prepare_masked(uint64_t rows, const uint64_t *worker, const uint64_t *firm,
                       const uint64_t *deletion, const double *outcome,
                       const uint64_t *frequency, const double *target, const uint8_t *mask,
                       orc_problem **out) {
    return prepare_impl(rows, worker, firm, deletion, outcome, frequency, target, mask, 0, out);
}

/* As orc_prepare with the observation-deletion graph selection. */
int orc_prepare_observation(uint64_t rows, const uint64_t *worker, const uint64_t *firm,
                            const uint64_t *deletion, const double *outcome,
                            const uint64_t *frequency, const double *target, orc_problem **out) {
    return prepare_impl(rows, worker, firm, deletion, outcome, frequency, target, NULL, 1, out);
}

/* stayer_hybrid.rs:71-348 (prepare_exact_stayer_hybrid): eligible stayer
 * rows, 1-based dense retained-mover firm and stayer-worker indices, appended
 * to the prepared mover problem m (identifiers of m's dense levels; one
 * deletion unit per stayer row) and recompressed with every row retained.
 * Stayer rows become observation units (generic_jla.rs:2170-2213): entity
 * the row's observation class ranked over all rows, copies numbered on over
 * the class's stayer rows (:4217-4265). */
int orc_prepare_stayers(const orc_problem *m, uint64_t rows, const uint64_t *firm, const uint64_t *worker,
                        const double *outcome, const uint64_t *frequency, const double *target,
                        orc_problem **out) {
    if (m->observation)
        return orc_fail(ORC_UNSUPPORTED, "session_stayer_augmentation",
                        "the stayer hybrid requires match-deletion mover preparation");
    uint64_t SW = 0;
    for (uint64_t r = 0; r < rows; ++r) SW = worker[r] > SW ? worker[r] : SW;
    if ((rows == 0) != (SW == 0))
        return orc_fail(ORC_INVALID_IDENTIFIER, "stayer_augmentation",
                        "eligible stayer workers must be a nonempty dense map when rows are supplied");
    uint64_t *wf = orc_alloc(SW + 1, 8), *wp = orc_alloc(SW + 1, 8);
    int rc = ORC_OK;
    for (uint64_t r = 0; r < rows && !rc; ++r) {
        if (firm[r] == 0 || firm[r] > m->firms)
            rc = orc_fail(ORC_INVALID_IDENTIFIER, "stayer_augmentation", "stayer row is not attached to a retained mover firm");
        else if (worker[r] == 0 || worker[r] > SW)
            rc = orc_fail(ORC_INVALID_IDENTIFIER, "stayer_augmentation", "stayer row has a non-dense worker identifier");
        else if (wf[worker[r]] && wf[worker[r]] != firm[r])
            rc = orc_fail(ORC_INVALID_IDENTIFIER, "stayer_augmentation", "eligible stayer worker spans more than one firm");
        else {
            wf[worker[r]] = firm[r];
            wp[worker[r]] += frequency[r];
        }
    }
    for (uint64_t w = 1; w <= SW && !rc; ++w)
        if (!wf[w]) rc = orc_fail(ORC_INVALID_IDENTIFIER, "stayer_augmentation", "eligible stayer worker identifiers are not dense");
    for (uint64_t w = 1; w <= SW && !rc; ++w)
        if (wp[w] < 2)
            rc = orc_fail(ORC_INVALID_INPUT, "stayer_augmentation",
                          "every eligible stayer worker must represent at least two physical observations");
    free(wf);
    free(wp);
    if (rc) return rc;
    uint64_t RM = m->rows, R = RM + rows;
    uint64_t *cw = orc_alloc(R, 8), *cf = orc_alloc(R, 8), *cd = orc_alloc(R, 8), *cq = orc_alloc(R, 8);
    double *cy = orc_alloc(R, 8), *ct = orc_alloc(R, 8);
    uint8_t *all = orc_alloc(R, 1);
    for (uint64_t r = 0; r < RM; ++r) {
        cw[r] = (uint64_t)m->row_worker[r] + 1; cf[r] = (uint64_t)m->row_firm[r] + 1;
        cd[r] = (uint64_t)m->row_deletion[r] + 1; cy[r] = m->outcome[r]; cq[r] = m->frequency[r];
        ct[r] = m->target[r];
    }
    for (uint64_t r = 0; r < rows; ++r) {
        cw[RM + r] = m->workers + worker[r]; cf[RM + r] = firm[r]; cd[RM + r] = m->units + r + 1;
        cy[RM + r] = outcome[r]; cq[RM + r] = frequency[r]; ct[RM + r] = target[r];
    }
    memset(all, 1, R);
    orc_problem *p = NULL;
    rc = prepare_impl(R, cw, cf, cd, cy, cq, ct, all, 0, &p);
    free(cw); free(cf); free(cd); free(cq); free(cy); free(ct); free(all);
    if (rc) return rc;
    if (p->rows != R || p->workers != m->workers + SW || p->firms != m->firms || p->units != m->units + rows) {
        orc_problem_free(p);
        return orc_fail(ORC_INVARIANT, "stayer_augmentation", "combined exact problem dimensions do not reconcile");
    }
    p->stayers = 1;
    p->match_units = m->units;
    p->unit_offset = orc_alloc(p->units, 8);
    if (rows) {
        uint64_t *order = orc_alloc(R, 8), *rank = orc_alloc(R, 8);
        observation_ranks(p, order, rank);
        uint64_t offset = 0, last = 0;
        for (uint64_t i = 0; i < R; ++i) {
            uint64_t row = order[i];
            if (row < RM) continue;
            if (rank[row] != last) offset = 0;
            last = rank[row];
            uint64_t u = m->units + (row - RM);
            p->unit_rank[u] = rank[row];
            p->unit_offset[u] = offset;
            offset += p->frequency[row];
        }
        free(order);
        free(rank);
    }
    *out = p;
    return ORC_OK;
}

static int prepare_impl(uint64_t rows, const uint64_t *worker, const uint64_t *firm,
                        const uint64_t *deletion, const double *outcome, const uint64_t *frequency,
                        const double *target, const uint8_t *mask, int observation, orc_problem **out) {
    /* types.rs:185-263 (validate) */
    if (rows == 0) return orc_fail(ORC_INVALID_INPUT, "ingest", "input sample is empty");
    uint64_t physical = 0;
    double target_total = 0.0;
    for (uint64_t r = 0; r < rows; ++r) {
        if (!isfinite(outcome[r]))
            return orc_fail(ORC_INVALID_INPUT, "ingest", "outcome is nonfinite at row %llu",
                            (unsigned long long)r);
        if (frequency[r] == 0)
            return orc_fail(ORC_INVALID_WEIGHT, "ingest", "frequency is zero at row %llu",
                            (unsigned long long)r);
        if (!isfinite(target[r]) || target[r] < 0.0)
            return orc_fail(ORC_INVALID_TARGET_WEIGHT, "ingest",
                            "target weight is invalid at row %llu", (unsigned long long)r);
        physical += frequency[r];
        if (physical > (1ULL << 53))
            return orc_fail(ORC_RESOURCE_LIMIT, "ingest", "physical total exceeds 2^53");
        target_total += target[r];
    }
    if (!isfinite(target_total) || target_total <= 0.0)
        return orc_fail(ORC_INVALID_TARGET_WEIGHT, "ingest",
                        "target weights must have positive finite total mass");
    orc_problem *p = orc_alloc(1, sizeof *p);
    p->input_rows = rows;
    uint64_t W0, F0, D0;
    uint32_t *cw = redense(worker, rows, &W0, NULL);
    uint32_t *cf = redense(firm, rows, &F0, NULL);
    uint32_t *cd = redense(deletion, rows, &D0, NULL);
    /* problem.rs:124-145: a deletion id stays within one coordinate */
    uint32_t *dw = orc_alloc(D0, 4), *df = orc_alloc(D0, 4);
    uint8_t *dseen = orc_alloc(D0, 1);
    int rc = ORC_OK;
    for (uint64_t r = 0; r < rows; ++r) {
        uint32_t d = cd[r];
        if (!dseen[d]) {
            dseen[d] = 1;
            dw[d] = cw[r];
            df[d] = cf[r];
        } else if (dw[d] != cw[r] || df[d] != cf[r]) {
            rc = orc_fail(ORC_INVALID_IDENTIFIER, "canonicalize",
                          "each deletion identifier must stay within one worker-firm coordinate");
            break;
        }
    }
    free(dseen);
    if (rc) goto fail_early;
    p->input_topology_checksum = topology_checksum(rows, cw, cf, cd, frequency);

    sel_ctx s = {rows, W0, F0, D0, cw, cf, cd, frequency, dw, df};
    p->mask = orc_alloc(rows, 1);
    if (mask) {
        memcpy(p->mask, mask, rows);
    } else if ((rc = observation ? select_observation(&s, p->mask, &p->selection)
                                 : select_match(&s, p->mask, &p->selection))) {
        goto fail_early;
    }

    /* problem.rs:206-606 (compress) */
    uint64_t R = 0;
    for (uint64_t r = 0; r < rows; ++r) R += p->mask[r];
    p->rows = R;
    p->retained = orc_alloc(R, 8);
    for (uint64_t r = 0, k = 0; r < rows; ++r)
        if (p->mask[r]) p->retained[k++] = r;
    uint64_t W, F, G;
    uint32_t *wmap = redense_selected(cw, p->retained, R, W0, &W);
    uint32_t *fmap = redense_selected(cf, p->retained, R, F0, &F);
    uint32_t *dmap = redense_selected(cd, p->retained, R, D0, &G);
    if (F < 2) {
        free(wmap); free(fmap); free(dmap);
        rc = orc_fail(ORC_GRAPH_UNIDENTIFIED, "compression",
                      "retained graph must contain at least two firms");
        goto fail_early;
    }
    p->workers = W; p->firms = F; p->units = G;
    p->row_worker = orc_alloc(R, 4); p->row_firm = orc_alloc(R, 4);
    p->row_deletion = orc_alloc(R, 4); p->row_cell = orc_alloc(R, 4);
    p->outcome = orc_alloc(R, 8); p->frequency = orc_alloc(R, 8); p->target = orc_alloc(R, 8);
    for (uint64_t i = 0; i < R; ++i) {
        uint64_t r = p->retained[i];
        p->row_worker[i] = wmap[cw[r]];
        p->row_firm[i] = fmap[cf[r]];
        p->row_deletion[i] = dmap[cd[r]];
        p->outcome[i] = outcome[r];
        p->frequency[i] = frequency[r];
        p->target[i] = target[r];
    }
    free(wmap); free(fmap); free(dmap);

    uint64_t *order = orc_alloc(R, 8);
    for (uint64_t i = 0; i < R; ++i) order[i] = i;
    qsort_r(order, R, 8, cmp_cell_rows, p);
    /* count cells */
    uint64_t C = 0;
    for (uint64_t i = 0; i < R; ++i)
        if (i == 0 || p->row_worker[order[i]] != p->row_worker[order[i - 1]] ||
            p->row_firm[order[i]] != p->row_firm[order[i - 1]])
            ++C;
    p->cells = C;
    p->cell_worker = orc_alloc(C, 4); p->cell_firm = orc_alloc(C, 4);
    p->cell_weight = orc_alloc(C, 8); p->cell_outcome_sum = orc_alloc(C, 8);
    p->cell_target_sum = orc_alloc(C, 8);
    for (uint64_t i = 0, cell = 0; i < R;) {
        uint64_t begin = i;
        uint32_t w = p->row_worker[order[i]], f = p->row_firm[order[i]];
        while (i < R && p->row_worker[order[i]] == w && p->row_firm[order[i]] == f) ++i;
        orc_kahan kw = {0, 0}, ko = {0, 0}, kt = {0, 0};
        for (uint64_t j = begin; j < i; ++j) {
            uint64_t row = order[j];
            double weight = (double)p->frequency[row];
            orc_kahan_add(&kw, weight);
            orc_kahan_add(&ko, weight * p->outcome[row]);
            orc_kahan_add(&kt, p->target[row]);
            p->row_cell[row] = (uint32_t)cell;
        }
        p->cell_worker[cell] = w;
        p->cell_firm[cell] = f;
        p->cell_weight[cell] = kw.sum;
        p->cell_outcome_sum[cell] = ko.sum;
        p->cell_target_sum[cell] = kt.sum;
        ++cell;
    }
    /* worker_index: cells 0..C are grouped by worker (problem.rs:402) */
    p->worker_ptr = orc_alloc(W + 1, 8);
    for (uint64_t c = 0; c < C; ++c) p->worker_ptr[p->cell_worker[c] + 1]++;
    for (uint64_t w = 0; w < W; ++w) p->worker_ptr[w + 1] += p->worker_ptr[w];
    /* firm_index: cells sorted by (firm, worker, cell) (problem.rs:403-424) */
    p->firm_cells = orc_alloc(C, 8);
    for (uint64_t c = 0; c < C; ++c) p->firm_cells[c] = c;
    qsort_r(p->firm_cells, C, 8, cmp_firm_cells, p);
    p->firm_ptr = orc_alloc(F + 1, 8);
    for (uint64_t c = 0; c < C; ++c) p->firm_ptr[p->cell_firm[c] + 1]++;
    for (uint64_t f = 0; f < F; ++f) p->firm_ptr[f + 1] += p->firm_ptr[f];
    /* exact target strata count (problem.rs:959-1035) */
    for (uint64_t i = 0; i < R; ++i) order[i] = i;
    qsort_r(order, R, 8, cmp_target_rows, p);
    p->exact_strata = 0;
    for (uint64_t i = 0; i < R; ++i)
        if (i == 0 || cmp_target_rows_key(p, order[i - 1], order[i]) != 0) p->exact_strata++;
    /* deletion_index (problem.rs:890-957): rows ascending within each unit */
    p->unit_ptr = orc_alloc(G + 1, 8);
    p->unit_rows = orc_alloc(R, 8);
    for (uint64_t i = 0; i < R; ++i) p->unit_ptr[p->row_deletion[i] + 1]++;
    for (uint64_t g = 0; g < G; ++g) p->unit_ptr[g + 1] += p->unit_ptr[g];
    {
        uint64_t *cur = orc_alloc(G + 1, 8);
        memcpy(cur, p->unit_ptr, G * 8);
        for (uint64_t i = 0; i < R; ++i) p->unit_rows[cur[p->row_deletion[i]]++] = i;
        free(cur);
    }
    p->physical_total = 0;
    orc_kahan kt = {0, 0};
    for (uint64_t i = 0; i < R; ++i) {
        p->physical_total += p->frequency[i];
        orc_kahan_add(&kt, p->target[i]);
    }
    p->target_total = kt.sum;
    if (!isfinite(p->target_total) || p->target_total <= 0.0) {
        free(order);
        rc = orc_fail(ORC_INVALID_TARGET_WEIGHT, "compression",
                      "retained target mass is not positive and finite");
        goto fail_early;
    }
    p->topology_checksum = topology_checksum(R, p->row_worker, p->row_firm, p->row_deletion,
                                             p->frequency);

    /* ---------------- JLA plan (jla.rs) ---------------- */
    p->per_copy = orc_alloc(R, 8);
    for (uint64_t i = 0; i < R; ++i) {
        double v = orc_canonical_zero(p->target[i] / (double)p->frequency[i]);
        if (!isfinite(v) || v < 0.0) {
            free(order);
            rc = orc_fail(ORC_INVALID_TARGET_WEIGHT, "jla_plan", "per-copy target mass invalid");
            goto fail_early;
        }
        p->per_copy[i] = v;
    }
    p->observation = observation;
    if (observation) {
        /* generic_jla.rs:4267-4303 (observation_row_ranks), 4158-4215
         * (observation_classes: rows of equal rank, ascending) and
         * statistical_batches.rs:15-53 (per-row copy offsets) */
        for (uint64_t i = 0; i < R; ++i) order[i] = i;
        p->row_rank = orc_alloc(R, 8);
        observation_ranks(p, order, p->row_rank);
        p->unit_offset = orc_alloc(R, 8);
        uint64_t offset = 0;
        for (uint64_t i = 0; i < R; ++i) {
            if (i == 0 || p->row_rank[order[i - 1]] != p->row_rank[order[i]]) offset = 0;
            p->unit_offset[order[i]] = offset;
            offset += p->frequency[order[i]];
        }
        G = R;
        p->units = R;
        free(p->unit_ptr);
        p->unit_ptr = orc_alloc(R + 1, 8);
        for (uint64_t i = 0; i <= R; ++i) p->unit_ptr[i] = i;
        for (uint64_t i = 0; i < R; ++i) p->unit_rows[i] = i;
        p->unit_cell = orc_alloc(R, 4); p->unit_count = orc_alloc(R, 8);
        p->unit_outcome = orc_alloc(R, 8); p->unit_target = orc_alloc(R, 8);
        p->unit_rank = orc_alloc(R, 8);
        for (uint64_t i = 0; i < R; ++i) {
            p->unit_cell[i] = p->row_cell[i];
            p->unit_count[i] = p->frequency[i];
            p->unit_outcome[i] = (double)p->frequency[i] * p->outcome[i];
            p->unit_target[i] = p->target[i];
            p->unit_rank[i] = p->row_rank[i];
        }
    } else {
    /* jla.rs:413-480 (semantic_row_ranks) */
    for (uint64_t i = 0; i < R; ++i) order[i] = i;
    qsort_r(order, R, 8, cmp_semantic_rows, p);
    p->row_rank = orc_alloc(R, 8);
    {
        uint64_t current = 0;
        for (uint64_t i = 0; i < R; ++i) {
            if (i == 0 || cmp_semantic_key(p, order[i - 1], order[i]) != 0) ++current;
            p->row_rank[order[i]] = current;
        }
    }
    /* jla.rs:574-648 (deletion_plan) */
    p->unit_cell = orc_alloc(G, 4); p->unit_count = orc_alloc(G, 8);
    p->unit_outcome = orc_alloc(G, 8); p->unit_target = orc_alloc(G, 8);
    p->unit_rank = orc_alloc(G, 8);
    for (uint64_t g = 0; g < G; ++g) {
        uint64_t physical_g = 0, min_rank = UINT64_MAX;
        double os = 0, oc = 0, ts = 0, tc = 0;
        uint32_t cell = UINT32_MAX;
        for (uint64_t k = p->unit_ptr[g]; k < p->unit_ptr[g + 1]; ++k) {
            uint64_t row = p->unit_rows[k];
            physical_g += p->frequency[row];
            if (cell == UINT32_MAX) cell = p->row_cell[row];
            else if (cell != p->row_cell[row]) {
                free(order);
                rc = orc_fail(ORC_INVALID_IDENTIFIER, "jla_plan",
                              "one deletion unit crosses coefficient cells");
                goto fail_early;
            }
            stable_add(&os, &oc, (double)p->frequency[row] * p->outcome[row]);
            stable_add(&ts, &tc, p->target[row]);
            if (p->row_rank[row] < min_rank) min_rank = p->row_rank[row];
        }
        p->unit_cell[g] = cell;
        p->unit_count[g] = physical_g;
        p->unit_outcome[g] = os + oc;
        p->unit_target[g] = ts + tc;
        p->unit_rank[g] = min_rank;
    }
    }
    /* jla.rs:651-790 (target_plan) */
    for (uint64_t i = 0; i < R; ++i) order[i] = i;
    qsort_r(order, R, 8, cmp_target_plan, p);
    {
        uint64_t S = 0;
        for (uint64_t i = 0; i < R; ++i)
            if (i == 0 || p->row_cell[order[i]] != p->row_cell[order[i - 1]] ||
                orc_bits(p->per_copy[order[i]]) != orc_bits(p->per_copy[order[i - 1]]))
                ++S;
        p->strata = S;
        p->stratum_cell = orc_alloc(S, 4); p->stratum_per_copy = orc_alloc(S, 8);
        p->stratum_count = orc_alloc(S, 8); p->stratum_target = orc_alloc(S, 8);
        p->stratum_rank = orc_alloc(S, 8);
        uint64_t s = 0;
        for (uint64_t i = 0; i < R;) {
            uint64_t begin = i;
            uint32_t cell = p->row_cell[order[i]];
            uint64_t bits = orc_bits(p->per_copy[order[i]]);
            while (i < R && p->row_cell[order[i]] == cell && orc_bits(p->per_copy[order[i]]) == bits)
                ++i;
            uint64_t physical_s = 0, min_rank = UINT64_MAX;
            double ts = 0, tc = 0;
            for (uint64_t j = begin; j < i; ++j) {
                uint64_t row = order[j];
                physical_s += p->frequency[row];
                stable_add(&ts, &tc, p->target[row]);
                if (p->row_rank[row] < min_rank) min_rank = p->row_rank[row];
            }
            p->stratum_cell[s] = cell;
            p->stratum_per_copy[s] = p->per_copy[order[begin]];
            p->stratum_count[s] = physical_s;
            p->stratum_target[s] = ts + tc;
            p->stratum_rank[s] = min_rank;
            ++s;
        }
    }
    free(order);
    free(cw); free(cf); free(cd); free(dw); free(df);
    p->match_units = observation ? 0 : p->units;
    *out = p;
    return ORC_OK;
fail_early:
    free(cw); free(cf); free(cd); free(dw); free(df);
    orc_problem_free(p);
    return rc;
}

void orc_problem_free(orc_problem *p) {
    if (!p) return;
    void *ptrs[] = {p->mask, p->retained, p->row_worker, p->row_firm, p->row_deletion,
                    p->row_cell, p->outcome, p->frequency, p->target, p->cell_worker,
                    p->cell_firm, p->cell_weight, p->cell_outcome_sum, p->cell_target_sum,
                    p->worker_ptr, p->firm_ptr, p->firm_cells, p->unit_ptr, p->unit_rows,
                    p->per_copy, p->row_rank, p->unit_cell, p->unit_count, p->unit_outcome,
                    p->unit_target, p->unit_rank, p->stratum_cell, p->stratum_per_copy,
                    p->stratum_count, p->stratum_target, p->stratum_rank, p->unit_offset};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
    free(p);
}

void orc_problem_dims(const orc_problem *p, uint64_t dims[9]) {
    dims[0] = p->input_rows;
    dims[1] = p->rows;
    dims[2] = p->physical_total;
    dims[3] = p->workers;
    dims[4] = p->firms;
    dims[5] = p->cells;
    dims[6] = p->units;
    dims[7] = p->exact_strata;
    dims[8] = p->strata;
}
double orc_problem_target_total(const orc_problem *p) { return p->target_total; }
uint64_t orc_problem_topology_checksum(const orc_problem *p) { return p->topology_checksum; }
void orc_problem_mask(const orc_problem *p, uint8_t *mask) { memcpy(mask, p->mask, p->input_rows); }

int orc_problem_array(const orc_problem *p, int which, void *dst) {
    switch (which) {
    case 0: memcpy(dst, p->cell_worker, p->cells * 4); break;
    case 1: memcpy(dst, p->cell_firm, p->cells * 4); break;
    case 2: memcpy(dst, p->cell_weight, p->cells * 8); break;
    case 3: memcpy(dst, p->unit_cell, p->units * 4); break;
    case 4: memcpy(dst, p->unit_count, p->units * 8); break;
    case 5: memcpy(dst, p->unit_rank, p->units * 8); break;
    case 6: memcpy(dst, p->stratum_cell, p->strata * 4); break;
    case 7: memcpy(dst, p->stratum_count, p->strata * 8); break;
    case 8: memcpy(dst, p->stratum_rank, p->strata * 8); break;
    case 9: memcpy(dst, p->stratum_per_copy, p->strata * 8); break;
    case 10: memcpy(dst, p->cell_outcome_sum, p->cells * 8); break;
    case 11: memcpy(dst, p->cell_target_sum, p->cells * 8); break;
    case 12: memcpy(dst, p->unit_outcome, p->units * 8); break;
    case 13:
        if (p->unit_offset) memcpy(dst, p->unit_offset, p->units * 8);
        else memset(dst, 0, p->units * 8);
        break;
    case 14: memcpy(dst, p->retained, p->rows * 8); break;
    case 15: memcpy(dst, p->row_cell, p->rows * 4); break;
    case 16: memcpy(dst, p->unit_ptr, (p->units + 1) * 8); break;
    case 17: memcpy(dst, p->unit_rows, p->rows * 8); break;
    default: return orc_fail(ORC_INVALID_INPUT, "problem", "unknown array %d", which);
    }
    return ORC_OK;
}
