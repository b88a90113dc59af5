// Numerical MCSE covariance (numerical_mc.rs:180-354) on the host: R-sized
// score sets and 3 x 3 matrices, no device work.
#include "numerical.hpp"

#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace vk {

std::array<double, 3> score_means(const std::vector<double> &x) {
  const size_t n = x.size() / 3;
  Neumaier s[3];
  for (size_t i = 0; i < n; ++i)
    for (int j = 0; j < 3; ++j) s[j].add(x[3 * i + j] / (double)n);
  return {s[0].value(), s[1].value(), s[2].value()};
}

bool cross_covariance(const std::vector<double> &a, const std::vector<double> &b, Matrix3 &out) {
  const size_t n = a.size() / 3;
  if (n < 2 || a.size() != b.size()) return false;
  out.fill(0.0);
  double scale = 0.0;
  for (size_t i = 0; i < a.size(); ++i) {
    if (!std::isfinite(a[i]) || !std::isfinite(b[i])) return false;
    scale = std::max({scale, std::fabs(a[i]), std::fabs(b[i])});
  }
  if (scale == 0.0) return true;
  const auto ma = score_means(a), mb = score_means(b);
  Neumaier acc[3][3];
  const double divisor = (double)n * (double)(n - 1);
  for (size_t k = 0; k < n; ++k) {
    const double *x = &a[3 * k], *y = &b[3 * k];
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j) {
        const double ai = x[i] / scale - ma[i] / scale, aj = x[j] / scale - ma[j] / scale;
        const double bi = y[i] / scale - mb[i] / scale, bj = y[j] / scale - mb[j] / scale;
        acc[i][j].add((ai * bj + bi * aj) / (2.0 * divisor));
      }
  }
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j) {
      out[3 * i + j] = (acc[i][j].value() * scale) * scale;
      out[3 * j + i] = out[3 * i + j];
    }
  return true;
}

namespace {
double signum(double x) { return std::signbit(x) ? -1.0 : 1.0; }
}  // namespace

// Symmetric part of raw = conditional + leverage, scaled; eigenvalues by
// Jacobi rotations (largest off-diagonal first, at most 32 sweeps); a small
// negative spectrum (within 1e-12 of the inputs' norm) is clipped to zero,
// a larger one is unstable.
Covariance finalize_covariance(const Matrix3 &conditional, const Matrix3 &leverage) {
  Covariance r;
  r.conditional = conditional;
  r.leverage = leverage;
  for (int k = 0; k < 9; ++k) r.raw[k] = conditional[k] + leverage[k];
  r.usable.fill(NAN);
  r.mcse.fill(NAN);
  for (int k = 0; k < 9; ++k)
    if (!std::isfinite(conditional[k]) || !std::isfinite(leverage[k]) || !std::isfinite(r.raw[k])) return r;
  double scale = 0.0;
  for (int k = 0; k < 9; ++k) scale = std::max({scale, std::fabs(conditional[k]), std::fabs(leverage[k])});
  if (scale == 0.0) {
    r.status = 2;
    r.usable = r.raw;
    r.mcse.fill(0.0);
    return r;
  }
  double nc = 0.0, nl = 0.0;
  for (int k = 0; k < 9; ++k) {
    nc += (conditional[k] / scale) * (conditional[k] / scale);
    nl += (leverage[k] / scale) * (leverage[k] / scale);
  }
  const double tau = 1e-12 * (std::sqrt(nc) + std::sqrt(nl));
  double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = (r.raw[3 * i + j] / scale + r.raw[3 * j + i] / scale) * 0.5;
  static const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int sweep = 0; sweep < 32; ++sweep) {
    int best = 0;
    for (int k = 1; k < 3; ++k)  // the last of equal maxima, as max_by keeps it
      if (std::fabs(a[pairs[k][0]][pairs[k][1]]) >= std::fabs(a[pairs[best][0]][pairs[best][1]])) best = k;
    const int p = pairs[best][0], q = pairs[best][1];
    if (std::fabs(a[p][q]) <= 1e-16) break;
    const double delta = (a[q][q] - a[p][p]) * 0.5;
    const double t = delta == 0.0 ? signum(a[p][q])
                                  : a[p][q] / (delta + signum(delta) * std::hypot(delta, a[p][q]));
    const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c, off = a[p][q];
    a[p][p] -= t * off;
    a[q][q] += t * off;
    a[p][q] = a[q][p] = 0.0;
    for (int k = 0; k < 3; ++k) {
      if (k != p && k != q) {
        const double ap = a[k][p], aq = a[k][q];
        a[k][p] = a[p][k] = c * ap - s * aq;
        a[k][q] = a[q][k] = s * ap + c * aq;
      }
      const double vp = v[k][p], vq = v[k][q];
      v[k][p] = c * vp - s * vq;
      v[k][q] = s * vp + c * vq;
    }
  }
  for (int i = 0; i < 3; ++i)
    if (a[i][i] < -tau) {
      r.status = 4;
      return r;
    }
  const bool adjusted = a[0][0] < 0.0 || a[1][1] < 0.0 || a[2][2] < 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      if (adjusted) {
        double sum = 0.0;
        for (int k = 0; k < 3; ++k) sum += v[i][k] * std::max(a[k][k], 0.0) * v[j][k];
        r.usable[3 * i + j] = sum * scale;
      } else {
        r.usable[3 * i + j] = r.raw[3 * i + j] * 0.5 + r.raw[3 * j + i] * 0.5;
      }
    }
  double pa = 0.0;
  for (int k = 0; k < 9; ++k) pa += ((r.usable[k] - r.raw[k]) / scale) * ((r.usable[k] - r.raw[k]) / scale);
  r.psd_adjustment = std::sqrt(pa) * scale;
  static const double contrast[3] = {1.0, 1.0, 2.0};
  double total = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) total += contrast[i] * (r.usable[3 * i + j] / scale) * contrast[j];
  r.mcse = {std::sqrt(std::max(r.usable[0], 0.0)), std::sqrt(std::max(r.usable[4], 0.0)),
            std::sqrt(std::max(r.usable[8], 0.0)), std::sqrt(std::max(total, 0.0)) * std::sqrt(scale)};
  r.status = adjusted ? 3 : 2;
  return r;
}

}  // namespace vk
