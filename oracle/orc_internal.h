/* Internal helpers shared by the oracle translation units (test infrastructure). */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

int orc_fail(int code, const char *phase, const char *fmt, ...);

static inline void *orc_alloc(size_t count, size_t size) {
    void *p = calloc(count ? count : 1, size);
    return p;
}

/* Neumaier compensated add: vendor/cmg/src/graph.rs:863-871,
 * vckss-core engine.rs:337-348 (StableSum::add). */
static inline void orc_neumaier(double *sum, double *corr, double value) {
    double next = *sum + value;
    *corr += (fabs(*sum) >= fabs(value)) ? ((*sum - next) + value)
                                          : ((value - next) + *sum);
    *sum = next;
}

/* Kahan sum returning the running sum only: problem.rs:1250-1267 and
 * parallel.rs:122-132 (compensated_sum). */
typedef struct {
    double sum, corr;
} orc_kahan;
static inline void orc_kahan_add(orc_kahan *k, double value) {
    double adjusted = value - k->corr;
    double next = k->sum + adjusted;
    k->corr = (next - k->sum) - adjusted;
    k->sum = next;
}

/* f64::total_cmp as a signed 64-bit key. */
static inline int64_t orc_total_key(double x) {
    int64_t bits;
    memcpy(&bits, &x, 8);
    bits ^= (int64_t)(((uint64_t)(bits >> 63)) >> 1);
    return bits;
}
static inline int orc_total_cmp(double a, double b) {
    int64_t ka = orc_total_key(a), kb = orc_total_key(b);
    return ka < kb ? -1 : (ka > kb ? 1 : 0);
}
static inline double orc_canonical_zero(double x) { return x == 0.0 ? 0.0 : x; }
static inline uint64_t orc_bits(double x) {
    uint64_t b;
    memcpy(&b, &x, 8);
    return b;
}

/* FNV-1a over little-endian words: problem.rs:1243-1248, frozen_bits.rs:19-59 */
static inline void orc_fnv_word(uint64_t *hash, uint64_t value) {
    for (int i = 0; i < 8; ++i) {
        *hash ^= (value >> (8 * i)) & 0xffu;
        *hash *= 0x00000100000001b3ULL;
    }
}

/* ---- CMG internals used by the engine ---------------------------------- */
struct orc_graph {
    uint64_t n, m;
    uint32_t *u, *v;
    double *w;
    double *diag;
    uint64_t nnz;
    double norm_bound;
};

typedef struct {
    uint64_t n;
    uint32_t *labels; /* component label of each vertex, first-vertex order */
    uint64_t count;
    uint64_t *sizes;
} orc_components;

typedef struct {
    orc_graph *graph;
    double *inv_diag;
    uint32_t *labels; /* aggregation labels (NULL when none) */
    uint64_t coarse;
    uint64_t repeat;
    int terminal;     /* 0 none, else TerminalReason + 1 */
} orc_level;

typedef struct {
    uint64_t n;             /* graph vertices */
    orc_components comps;
    uint64_t *anchors;
    uint64_t *perm;
    uint64_t dim;
    int sparse;
    double *packed;         /* dim*(dim-1)/2 strict lower, row-major */
    uint64_t *row_off; uint32_t *row_col; double *row_val;
    uint64_t *col_off; uint32_t *col_row; double *col_val;
    double *diag;
    uint64_t factor_nnz;
} orc_ldl;

struct orc_cmg {
    uint64_t nlevels;
    orc_level *levels;
    int terminal_reason;
    uint64_t *vertex_counts, *nnz_counts;
    uint64_t cumulative_nnz;
    uint64_t *repeat_counts;
    orc_components finest;
    orc_components *coarse_plans; /* nlevels - 1 */
    orc_ldl *ldl;                 /* NULL if not direct */
};

orc_graph *orc_graph_canonical(uint64_t n, uint64_t m, uint32_t *u, uint32_t *v,
                               double *w);
void orc_components_build(orc_components *c, const orc_graph *g);
void orc_components_free(orc_components *c);
int orc_components_center(const orc_components *c, double *values);
int orc_components_project(const orc_components *c, double *rhs, double *norm);
double orc_serial_norm(const double *v, uint64_t n);
double orc_dot(const double *a, const double *b, uint64_t n);

#endif
