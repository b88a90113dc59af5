/*
 * CPU oracle: local point-probe linearisation of the numerical MCSE
 * (rust/crates/vckss-core/src/numerical_mc.rs).  TEST INFRASTRUCTURE ONLY.
 *
 * Pins (tests/test_oracle_numerical.py): the derivative fields, the
 * observation gradient and the copy pullback against central finite
 * differences of the functions they differentiate; the covariance policy cases of
 * rust/crates/vckss-core/tests/numerical_mc.rs:7-50 (scaled covariance,
 * signed stage, exact zero, clipped adjustment); the fold estimator's exact
 * unbiasedness by enumeration (tests/numerical_mc.rs:52-107); finalize
 * against an eigendecomposition on PSD and indefinite matrices.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "orc.h"
#include "orc_internal.h"

/* numerical_mc.rs:54-112 (FiniteDerivative::new).  out: h, m, bias,
 * variance, dh[5], db[5], dv[5], nonsmooth.  Returns 1 when defined. */
int orc_finite_derivative(const double u[5], uint32_t r, double out[20]) {
    double p = u[0], mm = u[1], a = u[2], b = u[3], c = u[4];
    double s = p + mm;
    if (r < 2 || !(s > 0.0)) return 0;
    for (int j = 0; j < 5; ++j)
        if (!isfinite(u[j])) return 0;
    double h = p / s, m = mm / s, rr = (double)r;
    double bias = (m * a - h * b + (m - h) * c) / rr;
    double variance = (m * m * a + h * h * b - 2.0 * h * m * c) / rr;
    double dh[5] = {m / s, -h / s, 0.0, 0.0, 0.0};
    double db[5], dv[5];
    for (int j = 0; j < 5; ++j) {
        double da = j == 2, dbb = j == 3, dc = j == 4;
        db[j] = (-(a + b + 2.0 * c) * dh[j] + m * da - h * dbb + (m - h) * dc) / rr;
        dv[j] = ((-2.0 * m * a + 2.0 * h * b - 2.0 * c * (m - h)) * dh[j] + m * m * da + h * h * dbb -
                 2.0 * h * m * dc) / rr;
    }
    out[0] = h; out[1] = m; out[2] = bias; out[3] = variance;
    for (int j = 0; j < 5; ++j) { out[4 + j] = dh[j]; out[9 + j] = db[j]; out[14 + j] = dv[j]; }
    out[19] = variance < 0.0 ? 1.0 : 0.0;
    for (int k = 0; k < 19; ++k)
        if (!isfinite(out[k])) return 0;
    return 1;
}

/* numerical_mc.rs:114-125 (observation): the gradient of
 * 1/l + bias/l^2 - variance/l^3 with l = m - control.  Returns 1 when
 * defined. */
int orc_fd_observation(const double fd[20], double control, double grad[5]) {
    double ell = fd[1] - control;
    if (ell <= 0.0 || fd[19] != 0.0) return 0;
    double inv = 1.0 / ell, bias = fd[2], var = fd[3];
    double inv2 = inv * inv, inv3 = inv2 * inv, inv4 = inv2 * inv2;
    double factor = -inv2 - 2.0 * bias * inv3 + 3.0 * var * inv4;
    for (int j = 0; j < 5; ++j) {
        grad[j] = -factor * fd[4 + j] + inv2 * fd[9 + j] - inv3 * fd[14 + j];
        if (!isfinite(grad[j])) return 0;
    }
    return 1;
}

/* One copy of an observation unit (generic_jla/numerical.rs:213-256): the
 * moment map u(a2, a4, c1, c3) of its row's mean projection square and
 * fourth power and the copy's first and third correlations, the linearised
 * correction value 1/l + bias/l^2 - variance/l^3 with l = m - control, and
 * its pullback to (a2, a4, c1, c3) (numerical_mc.rs:137-145, copy_pullback).
 * fd receives the derivative fields.  Returns 0, 5 (nonsmooth adjustment)
 * or 6 (nonfinite derivative). */
int orc_copy_pullback(const double x[4], double control, uint32_t r, double u[5], double fd[20], double *value,
                      double pb[4]) {
    const double a2 = x[0], a4 = x[1], c1 = x[2], c3 = x[3];
    u[0] = a2;
    u[1] = 1.0 + a2 - 2.0 * c1;
    u[2] = a4;
    u[3] = 1.0 + 6.0 * a2 + a4 - 4.0 * c1 - 4.0 * c3;
    u[4] = a2 + a4 - 2.0 * c3;
    if (!orc_finite_derivative(u, r, fd)) return 6;
    if (fd[19] != 0.0) return 5;
    double g[5];
    if (!orc_fd_observation(fd, control, g)) return 6;
    const double ell = fd[1] - control;
    *value = 1.0 / ell + fd[2] / (ell * ell) - fd[3] / (ell * ell * ell);
    pb[0] = g[0] + g[1] + 6.0 * g[3] + g[4];
    pb[1] = g[2] + g[3] + g[4];
    pb[2] = -2.0 * g[1] - 4.0 * g[3];
    pb[3] = -4.0 * g[3] - 2.0 * g[4];
    return 0;
}

static void means3(const double *x, uint64_t n, double out[3]) {
    double s[3] = {0, 0, 0}, c[3] = {0, 0, 0};
    for (uint64_t i = 0; i < n; ++i)
        for (int j = 0; j < 3; ++j) orc_neumaier(&s[j], &c[j], x[3 * i + j] / (double)n);
    for (int j = 0; j < 3; ++j) out[j] = s[j] + c[j];
}

/* numerical_mc.rs:180-226 (cross_covariance): a, b are n x 3 scores.
 * Returns ORC_OK, or ORC_INVALID_INPUT for bad counts or nonfinite scores. */
int orc_cross_covariance(const double *a, const double *b, uint64_t n, double out[9]) {
    if (n < 2) return orc_fail(ORC_INVALID_INPUT, "numerical_mc", "invalid score counts");
    double ma[3], mb[3];
    means3(a, n, ma);
    means3(b, n, mb);
    double scale = 0.0;
    for (uint64_t i = 0; i < 3 * n; ++i) {
        if (!isfinite(a[i]) || !isfinite(b[i]))
            return orc_fail(ORC_INVALID_INPUT, "numerical_mc", "nonfinite scores");
        scale = fmax(scale, fmax(fabs(a[i]), fabs(b[i])));
    }
    memset(out, 0, 9 * sizeof(double));
    if (scale == 0.0) return ORC_OK;
    double s[3][3] = {{0}}, c[3][3] = {{0}};
    double divisor = (double)n * (double)(n - 1);
    for (uint64_t k = 0; k < n; ++k) {
        const double *x = a + 3 * k, *y = b + 3 * k;
        for (int i = 0; i < 3; ++i)
            for (int j = i; j < 3; ++j) {
                double ai = x[i] / scale - ma[i] / scale, aj = x[j] / scale - ma[j] / scale;
                double bi = y[i] / scale - mb[i] / scale, bj = y[j] / scale - mb[j] / scale;
                orc_neumaier(&s[i][j], &c[i][j], (ai * bj + bi * aj) / (2.0 * divisor));
            }
    }
    for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j) {
            out[3 * i + j] = ((s[i][j] + c[i][j]) * scale) * scale;
            out[3 * j + i] = out[3 * i + j];
        }
    return ORC_OK;
}

static double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : (signbit(x) ? -1.0 : 1.0)); }

/* numerical_mc.rs:232-354 (finalize).  status: 2 ok_local, 3 psd adjusted,
 * 4 unstable non-PSD, 6 nonfinite (usable and mcse NaN unless status 2/3). */
void orc_numerical_finalize(const double cond[9], const double lev[9], double raw[9], double usable[9],
                            double mcse[4], uint32_t *status, double *psd_adjustment) {
    for (int k = 0; k < 9; ++k) raw[k] = cond[k] + lev[k];
    for (int k = 0; k < 9; ++k) usable[k] = NAN;
    for (int k = 0; k < 4; ++k) mcse[k] = NAN;
    *status = 6;
    *psd_adjustment = 0.0;
    for (int k = 0; k < 9; ++k)
        if (!isfinite(cond[k]) || !isfinite(lev[k]) || !isfinite(raw[k])) return;
    double scale = 0.0;
    for (int k = 0; k < 9; ++k) scale = fmax(scale, fmax(fabs(cond[k]), fabs(lev[k])));
    if (scale == 0.0) {
        *status = 2;
        memcpy(usable, raw, 9 * sizeof(double));
        for (int k = 0; k < 4; ++k) mcse[k] = 0.0;
        return;
    }
    double nc = 0.0, nl = 0.0;
    for (int k = 0; k < 9; ++k) {
        nc += (cond[k] / scale) * (cond[k] / scale);
        nl += (lev[k] / scale) * (lev[k] / scale);
    }
    double tau = 1e-12 * (sqrt(nc) + sqrt(nl));
    double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = (raw[3 * i + j] / scale + raw[3 * j + i] / scale) * 0.5;
    static const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int it = 0; it < 32; ++it) {
        int best = 0;
        for (int k = 1; k < 3; ++k) /* max_by keeps the last of equal maxima */
            if (fabs(a[pairs[k][0]][pairs[k][1]]) >= fabs(a[pairs[best][0]][pairs[best][1]])) best = k;
        int p = pairs[best][0], q = pairs[best][1];
        if (fabs(a[p][q]) <= 1e-16) break;
        double delta = (a[q][q] - a[p][p]) * 0.5;
        double t = delta == 0.0 ? sgn(a[p][q]) : a[p][q] / (delta + sgn(delta) * hypot(delta, a[p][q]));
        double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        double off = a[p][q];
        a[p][p] -= t * off;
        a[q][q] += t * off;
        a[p][q] = 0.0;
        a[q][p] = 0.0;
        for (int k = 0; k < 3; ++k) {
            if (k != p && k != q) {
                double ap = a[k][p], aq = a[k][q];
                a[k][p] = c * ap - s * aq;
                a[p][k] = a[k][p];
                a[k][q] = s * ap + c * aq;
                a[q][k] = a[k][q];
            }
            double vp = v[k][p], vq = v[k][q];
            v[k][p] = c * vp - s * vq;
            v[k][q] = s * vp + c * vq;
        }
    }
    for (int i = 0; i < 3; ++i)
        if (a[i][i] < -tau) {
            *status = 4;
            return;
        }
    int adjusted = a[0][0] < 0.0 || a[1][1] < 0.0 || a[2][2] < 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            if (adjusted) {
                double sum = 0.0;
                for (int k = 0; k < 3; ++k) sum += v[i][k] * fmax(a[k][k], 0.0) * v[j][k];
                usable[3 * i + j] = sum * scale;
            } else {
                usable[3 * i + j] = raw[3 * i + j] * 0.5 + raw[3 * j + i] * 0.5;
            }
        }
    double pa = 0.0;
    for (int k = 0; k < 9; ++k) pa += ((usable[k] - raw[k]) / scale) * ((usable[k] - raw[k]) / scale);
    *psd_adjustment = sqrt(pa) * scale;
    static const double contrast[3] = {1.0, 1.0, 2.0};
    double total = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) total += contrast[i] * (usable[3 * i + j] / scale) * contrast[j];
    mcse[0] = sqrt(fmax(usable[0], 0.0));
    mcse[1] = sqrt(fmax(usable[4], 0.0));
    mcse[2] = sqrt(fmax(usable[8], 0.0));
    mcse[3] = sqrt(fmax(total, 0.0)) * sqrt(scale);
    *status = adjusted ? 3 : 2;
}
