// Dense exact estimator on the device (exact_estimator.rs:490-1244 for match
// deletion, joint nuisance, no controls and no centering).
//
// The information of the permutation-equivariant W + F embedding (worker
// dummy, firm coordinates e_f - 1/F: exact_estimator.rs:1817-1848) is filled
// densely, diagonally scaled (D = diag(A_ii^-1/2)) and shifted by the unit
// projector onto its scaled null direction (the firm constant), S = D A D +
// n n'.  S then has the quotient spectrum with the null eigenvalue replaced
// by 1, so its reciprocal condition number is the quotient's
// (dense.rs:89-160 certifies the same quantity on the unshifted matrix).
//
//   blocked Cholesky S = L L'       k_potrf_diag, k_trsm_panel, k_syrk_update
//   inverse S^-1 = L^-T L^-1        k_forward_identity, k_backward (column slabs)
//   residual gate ||S S^-1 - I||    k_residual_colsq (tiled product)
//   fit beta = D S^-1 D X'y         k_gemv
//   per cell c = (w, f)             k_exact_cell: u = D S^-1 D x_c, the exact
//                                   leverage h_c = x_c'u and the target
//                                   bilinear form of u (exact_estimator.rs:1698-1787)
//
// A match block's rows share one cell, so its maker I - h_c q q' (q = the
// rows' square-root frequencies, exact_estimator.rs:1448-1696) has smallest
// eigenvalue 1 - h_c n_g and its deleted residual sums to
// (sum f e)/(1 - h_c n_g): the host combines the blocks in closed form.  Under
// observation deletion every physical copy is a block of its own (maker
// 1 - h_c, exact_estimator.rs:914-987).
#include <chrono>
#include <limits>

#include "device_impl.hpp"

namespace vk {
namespace {

using gpu::DBuf;
using gpu::grid_for;

constexpr int kNb = 64;          // Cholesky block
constexpr int kPanel = 32;       // panel rows per triangular-solve workgroup
constexpr int kSlab = 16;        // right-hand-side columns per substitution workgroup
constexpr uint32_t kExactMaxDimension = 8000;  // the cell pass keeps one u (p doubles) in LDS

// S_ij = q_i q_j a_ij + n_i n_j without the cell terms (added by
// k_exact_fill_cells): a_ww = d_w, a_wg = -d_w/F,
// a_gh = delta_gh d_g - (d_g + d_h)/F + N/F^2.
static __global__ void k_exact_fill(uint32_t p, uint32_t W, uint32_t F, const double *__restrict__ d, double N,
                                    const double *__restrict__ q, const double *__restrict__ nvec,
                                    double *__restrict__ S) {
  const uint64_t total = (uint64_t)p * p;
  const double invF = 1.0 / F;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = t / p, j = t % p;
    double a;
    if (i < W && j < W)
      a = i == j ? d[i] : 0.0;
    else if (i < W)
      a = -d[i] * invF;
    else if (j < W)
      a = -d[j] * invF;
    else
      a = (i == j ? d[i] : 0.0) - (d[i] + d[j]) * invF + N * invF * invF;
    S[t] = q[i] * q[j] * a + nvec[i] * nvec[j];
  }
}

// Cell terms: the (worker, firm) entries of distinct cells, both triangles.
static __global__ void k_exact_fill_cells(uint32_t C, uint32_t p, uint32_t W, const uint32_t *__restrict__ cw,
                                          const uint32_t *__restrict__ cf, const double *__restrict__ wt,
                                          const double *__restrict__ q, double *__restrict__ S) {
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < C; c += gridDim.x * blockDim.x) {
    const uint32_t w = cw[c], f = W + cf[c];
    const double v = q[w] * q[f] * wt[c];
    S[(uint64_t)w * p + f] += v;
    S[(uint64_t)f * p + w] += v;
  }
}

// Factor the diagonal block [k, k + n) in LDS.  A nonpositive or nonfinite
// pivot sets status and continues with pivot 1 (the host then fails).
static __global__ void __launch_bounds__(256) k_potrf_diag(double *__restrict__ S, uint32_t p, uint32_t k,
                                                           int *__restrict__ status) {
  __shared__ double a[kNb][kNb + 1];
  const int n = min(kNb, (int)(p - k));
  for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
    const int i = t / n, j = t % n;
    a[i][j] = j <= i ? S[(uint64_t)(k + i) * p + k + j] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    if (threadIdx.x == 0) {
      double piv = a[j][j];
      if (!(piv > 0.0) || !isfinite(piv)) {
        status[0] = 1;
        piv = 1.0;
      }
      a[j][j] = sqrt(piv);
    }
    __syncthreads();
    const double djj = a[j][j];
    for (int i = j + 1 + threadIdx.x; i < n; i += blockDim.x) a[i][j] /= djj;
    __syncthreads();
    const int m = n - j - 1;
    for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
      const int i = j + 1 + t / m, l = j + 1 + t % m;
      if (l <= i) a[i][l] -= a[i][j] * a[l][j];
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
    const int i = t / n, j = t % n;
    if (j <= i) S[(uint64_t)(k + i) * p + k + j] = a[i][j];
  }
}

// Panel below the diagonal block: X L11' = A21 for kPanel rows per workgroup.
static __global__ void __launch_bounds__(256) k_trsm_panel(double *__restrict__ S, uint32_t p, uint32_t k) {
  __shared__ double l[kNb][kNb + 1];
  __shared__ double x[kPanel][kNb + 1];
  const int n = min(kNb, (int)(p - k));
  const uint32_t r0 = k + n + blockIdx.x * kPanel;
  const int rows = min(kPanel, (int)(p - r0));
  for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
    const int i = t / n, j = t % n;
    l[i][j] = j <= i ? S[(uint64_t)(k + i) * p + k + j] : 0.0;
  }
  for (int t = threadIdx.x; t < rows * n; t += blockDim.x) {
    const int r = t / n, j = t % n;
    x[r][j] = S[(uint64_t)(r0 + r) * p + k + j];
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    for (int r = threadIdx.x; r < rows; r += blockDim.x) x[r][j] /= l[j][j];
    __syncthreads();
    const int m = n - j - 1;
    for (int t = threadIdx.x; t < rows * m; t += blockDim.x) {
      const int r = t / m, c = j + 1 + t % m;
      x[r][c] -= x[r][j] * l[c][j];
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < rows * n; t += blockDim.x) {
    const int r = t / n, j = t % n;
    S[(uint64_t)(r0 + r) * p + k + j] = x[r][j];
  }
}

// Trailing update of the lower triangle: S22 -= L21 L21', one 64 x 64 tile
// per workgroup (tiles above the diagonal return), 4 x 4 outputs per thread.
static __global__ void __launch_bounds__(256) k_syrk_update(double *__restrict__ S, uint32_t p, uint32_t k) {
  if (blockIdx.y > blockIdx.x) return;
  __shared__ double As[32][64 + 1], Bs[32][64 + 1];
  const int n = min(kNb, (int)(p - k));
  const uint32_t base = k + n;
  const uint32_t i0 = base + blockIdx.x * 64, j0 = base + blockIdx.y * 64;
  const int ti = threadIdx.x / 16, tj = threadIdx.x % 16;
  double acc[4][4] = {};
  for (int m0 = 0; m0 < n; m0 += 32) {
    for (int t = threadIdx.x; t < 64 * 32; t += blockDim.x) {
      const int r = t / 32, m = t % 32;
      const bool in = m0 + m < n;
      As[m][r] = (in && i0 + r < p) ? S[(uint64_t)(i0 + r) * p + k + m0 + m] : 0.0;
      Bs[m][r] = (in && j0 + r < p) ? S[(uint64_t)(j0 + r) * p + k + m0 + m] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int m = 0; m < 32; ++m) {
      double a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = As[m][ti * 4 + r];
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = Bs[m][tj * 4 + c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] += a[r] * b[c];
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint32_t i = i0 + ti * 4 + r, j = j0 + tj * 4 + c;
      if (i < p && j < p && j <= i) S[(uint64_t)i * p + j] -= acc[r][c];
    }
}

// Lt = L' of the lower triangle (zeros above), 32 x 32 tiles through LDS.
static __global__ void __launch_bounds__(256) k_transpose_lower(const double *__restrict__ L, double *__restrict__ Lt,
                                                                uint32_t p) {
  __shared__ double tile[32][33];
  const uint32_t bx = blockIdx.x * 32, by = blockIdx.y * 32;  // L columns bx.., rows by..
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  for (int r = ty; r < 32; r += 8) {
    const uint32_t i = by + r, j = bx + tx;
    tile[r][tx] = (i < p && j < p && j <= i) ? L[(uint64_t)i * p + j] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const uint32_t i = bx + r, j = by + tx;  // Lt row i = L column i
    if (i < p && j < p) Lt[(uint64_t)i * p + j] = tile[tx][r];
  }
}

// Y = L^-1 by forward substitution on the identity, kSlab columns per
// workgroup (16 threads per column).  Thread q of a column owns the rows
// m = c0 + q (mod kSlab): it writes y_m and is the only thread to read it
// back, so the sequential recurrence needs no barrier.  Y is zero above.
static __global__ void __launch_bounds__(256) k_forward_identity(const double *__restrict__ L,
                                                                 double *__restrict__ Y, uint32_t p) {
  const uint32_t c0 = blockIdx.x * kSlab;
  const uint32_t c = c0 + threadIdx.x / kSlab, q = threadIdx.x % kSlab;
  if (c >= p) return;  // whole 16-lane groups leave together
  for (uint32_t i = c0; i < p; ++i) {
    const double *Li = L + (uint64_t)i * p;
    double s = 0.0;
    for (uint32_t m = c0 + q; m < i; m += kSlab) s += Li[m] * Y[(uint64_t)m * p + c];
    s += __shfl_xor(s, 8);
    s += __shfl_xor(s, 4);
    s += __shfl_xor(s, 2);
    s += __shfl_xor(s, 1);
    if (q == (i - c0) % kSlab) Y[(uint64_t)i * p + c] = ((i == c ? 1.0 : 0.0) - s) / Li[i];
  }
}

// X = L^-T Y in place (backward substitution), Lt = L'.  Thread q owns the
// rows m = q (mod kSlab).
static __global__ void __launch_bounds__(256) k_backward(const double *__restrict__ Lt, double *__restrict__ X,
                                                         uint32_t p) {
  const uint32_t c0 = blockIdx.x * kSlab;
  const uint32_t c = c0 + threadIdx.x / kSlab, q = threadIdx.x % kSlab;
  if (c >= p) return;
  for (uint32_t ii = p; ii-- > 0;) {
    const double *Ri = Lt + (uint64_t)ii * p;
    double s = 0.0;
    const uint32_t first = ii + 1 + (q + kSlab - (ii + 1) % kSlab) % kSlab;
    for (uint32_t m = first; m < p; m += kSlab) s += Ri[m] * X[(uint64_t)m * p + c];
    s += __shfl_xor(s, 8);
    s += __shfl_xor(s, 4);
    s += __shfl_xor(s, 2);
    s += __shfl_xor(s, 1);
    if (q == ii % kSlab) X[(uint64_t)ii * p + c] = (X[(uint64_t)ii * p + c] - s) / Ri[ii];
  }
}

// G = (X + X')/2, 32 x 32 tiles.
static __global__ void __launch_bounds__(256) k_symmetrize(const double *__restrict__ X, double *__restrict__ G,
                                                           uint32_t p) {
  __shared__ double tile[32][33];
  const uint32_t bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  for (int r = ty; r < 32; r += 8) {
    const uint32_t i = bx + r, j = by + tx;  // X' tile: rows of X from bx
    tile[r][tx] = (i < p && j < p) ? X[(uint64_t)i * p + j] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const uint32_t i = by + r, j = bx + tx;
    if (i < p && j < p) G[(uint64_t)i * p + j] = 0.5 * (X[(uint64_t)i * p + j] + tile[tx][r]);
  }
}

// R = A B - I for 64 x 64 output tiles; per tile and column the sum of
// squares of R over the tile's rows -> partial[tile_row * p + column].
static __global__ void __launch_bounds__(256) k_residual_colsq(const double *__restrict__ A,
                                                               const double *__restrict__ B, uint32_t p,
                                                               double *__restrict__ partial) {
  __shared__ double As[16][64 + 1], Bs[16][64 + 1];
  __shared__ double red[16][64];
  const uint32_t i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
  const int ti = threadIdx.x / 16, tj = threadIdx.x % 16;
  double acc[4][4] = {};
  for (uint32_t k0 = 0; k0 < p; k0 += 16) {
    for (int t = threadIdx.x; t < 16 * 64; t += blockDim.x) {
      const int r = t / 16, kk = t % 16;
      As[kk][r] = (i0 + r < p && k0 + kk < p) ? A[(uint64_t)(i0 + r) * p + k0 + kk] : 0.0;
      const int kb = t / 64, col = t % 64;
      Bs[kb][col] = (k0 + kb < p && j0 + col < p) ? B[(uint64_t)(k0 + kb) * p + j0 + col] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 16; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = As[kk][ti * 4 + r];
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = Bs[kk][tj * 4 + c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] += a[r] * b[c];
    }
    __syncthreads();
  }
  for (int c = 0; c < 4; ++c) {
    double sq = 0.0;
    for (int r = 0; r < 4; ++r) {
      const uint32_t i = i0 + ti * 4 + r, j = j0 + tj * 4 + c;
      if (i < p && j < p) {
        const double v = acc[r][c] - (i == j ? 1.0 : 0.0);
        sq += v * v;
      }
    }
    red[ti][tj * 4 + c] = sq;
  }
  __syncthreads();
  if (threadIdx.x < 64 && j0 + threadIdx.x < p) {
    double s = 0.0;
    for (int r = 0; r < 16; ++r) s += red[r][threadIdx.x];
    partial[(uint64_t)blockIdx.y * p + j0 + threadIdx.x] = s;
  }
}

// Column norms of the residual, the Frobenius norm of G (per column) and the
// absolute row sums of S (Gershgorin bound of its largest eigenvalue).
static __global__ void k_exact_norms(const double *__restrict__ partial, uint32_t tiles, const double *__restrict__ G,
                                     const double *__restrict__ S, uint32_t p, double *__restrict__ out) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < p; j += gridDim.x * blockDim.x) {
    double r = 0.0, g = 0.0, s = 0.0;
    for (uint32_t t = 0; t < tiles; ++t) r += partial[(uint64_t)t * p + j];
    for (uint32_t i = 0; i < p; ++i) {
      const double v = G[(uint64_t)i * p + j];  // column j (G symmetric: coalesced across j)
      g += v * v;
      s += fabs(S[(uint64_t)i * p + j]);
    }
    out[j] = sqrt(r);
    out[p + j] = g;
    out[2 * p + j] = s;
  }
}

// out[i] = sum_j M[i][j] v[j], one wavefront per row.
static __global__ void __launch_bounds__(256) k_gemv(const double *__restrict__ M, const double *__restrict__ v,
                                                     uint32_t p, double *__restrict__ out) {
  const uint32_t row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64, lane = threadIdx.x % 64;
  if (row >= p) return;
  double s = 0.0;
  for (uint32_t j = lane; j < p; j += 64) s += M[(uint64_t)row * p + j] * v[j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) out[row] = s;
}

// Per cell c = (w, f): u = D S^-1 D x_c (x_c = e_w + e_{W+f} - 1/F on the
// firm block) in LDS, the exact leverage h = x_c'u, and the target bilinear
// form of u: worker and firm share-weighted variances and the cell-mass
// covariance (exact_estimator.rs:1736-1786).  out[4c..4c+3] = h, worker,
// firm, covariance.
static __global__ void __launch_bounds__(256) k_exact_cell(const double *__restrict__ G, const double *__restrict__ q,
                                                           const double *__restrict__ mvec, uint32_t p, uint32_t W,
                                                           uint32_t F, uint32_t C, const uint32_t *__restrict__ cw,
                                                           const uint32_t *__restrict__ cf,
                                                           const double *__restrict__ ws,
                                                           const double *__restrict__ fs,
                                                           const double *__restrict__ cm, double *__restrict__ out) {
  extern __shared__ double u[];  // p values, then 6 x 4 wave partials
  double *red = u + p;
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  for (uint32_t c = blockIdx.x; c < C; c += gridDim.x) {
    const uint32_t w = cw[c], f = W + cf[c];
    const double qw = q[w], qf = q[f];
    for (uint32_t v = threadIdx.x; v < p; v += blockDim.x)
      u[v] = q[v] * (qw * G[(uint64_t)w * p + v] + qf * G[(uint64_t)f * p + v]) - mvec[v];
    __syncthreads();
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t v = threadIdx.x; v < W; v += blockDim.x) {
      const double x = u[v];
      s[0] += ws[v] * x;
      s[1] += ws[v] * x * x;
    }
    for (uint32_t g = threadIdx.x; g < F; g += blockDim.x) {
      const double x = u[W + g];
      s[2] += fs[g] * x;
      s[3] += fs[g] * x * x;
      s[4] += x;
    }
    for (uint32_t e = threadIdx.x; e < C; e += blockDim.x) s[5] += cm[e] * u[cw[e]] * u[W + cf[e]];
#pragma unroll
    for (int k = 0; k < 6; ++k)
      for (int off = 32; off > 0; off >>= 1) s[k] += __shfl_xor(s[k], off);
    if (lane == 0)
      for (int k = 0; k < 6; ++k) red[k * 4 + wave] = s[k];
    __syncthreads();
    if (threadIdx.x == 0) {
      double t[6];
      for (int k = 0; k < 6; ++k) t[k] = red[k * 4] + red[k * 4 + 1] + red[k * 4 + 2] + red[k * 4 + 3];
      out[4 * (uint64_t)c] = u[w] + u[f] - t[4] / F;
      out[4 * (uint64_t)c + 1] = t[1] - t[0] * t[0];
      out[4 * (uint64_t)c + 2] = t[3] - t[2] * t[2];
      out[4 * (uint64_t)c + 3] = t[5] - t[0] * t[2];
    }
    __syncthreads();
  }
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

SolveResult DeviceEngine::run_exact(const SolveOptions &opt) {
  Impl &E = *impl_;
  const Prepared &P = *E.P;
  const auto t0 = std::chrono::steady_clock::now();
  opt.checkpoint("exact_estimator");
  const uint32_t W = (uint32_t)P.workers, F = (uint32_t)P.firms, C = (uint32_t)P.cells;
  if (P.rows == 0 || W == 0 || F < 2)
    fail(kInvalidInput, "exact_estimator", "the retained exact problem has invalid dimensions");
  const uint64_t p64 = (uint64_t)W + F;
  if (p64 > kExactMaxDimension)
    fail(kResourceLimit, "exact_estimator",
         "the embedded exact dimension " + std::to_string(p64) + " exceeds the device limit " +
             std::to_string(kExactMaxDimension));
  const uint32_t p = (uint32_t)p64;
  // host: degrees, scaling, null direction, target shares, right-hand side
  std::vector<double> d(p, 0.0), q(p), nvec(p, 0.0), ws(W, 0.0), fs(F, 0.0), cm(C), rhs(p, 0.0);
  double N = 0.0, ytot = 0.0;
  for (uint32_t c = 0; c < C; ++c) {
    d[P.cell_worker[c]] += P.cell_weight[c];
    d[W + P.cell_firm[c]] += P.cell_weight[c];
    N += P.cell_weight[c];
    ws[P.cell_worker[c]] += P.cell_target[c] / P.target_total;
    fs[P.cell_firm[c]] += P.cell_target[c] / P.target_total;
    cm[c] = P.cell_target[c] / P.target_total;
    rhs[P.cell_worker[c]] += P.cell_outcome[c];
    rhs[W + P.cell_firm[c]] += P.cell_outcome[c];
    ytot += P.cell_outcome[c];
  }
  double null_sq = 0.0;
  for (uint32_t i = 0; i < p; ++i) {
    const double aii = i < W ? d[i] : d[i] - 2.0 * d[i] / F + N / ((double)F * F);
    if (!(aii > 0.0) || !std::isfinite(aii))
      fail(kSingularInformation, "dense_symmetric", "symmetric information has a nonpositive diagonal");
    q[i] = 1.0 / std::sqrt(aii);
    if (i >= W) {
      nvec[i] = std::sqrt(aii);
      null_sq += aii;
    }
  }
  for (uint32_t i = W; i < p; ++i) nvec[i] /= std::sqrt(null_sq);
  for (uint32_t g = 0; g < F; ++g) rhs[W + g] -= ytot / F;

  const uint64_t pp = (uint64_t)p * p;
  DBuf<double> S, Saug, T, X, dd, dq, dn, dv, dout, dm, dws, dfs, dcm, dcell;
  DBuf<int> status;
  S.alloc(pp); Saug.alloc(pp); T.alloc(pp); X.alloc(pp);
  dd.upload(d); dq.upload(q); dn.upload(nvec); dws.upload(ws); dfs.upload(fs); dcm.upload(cm);
  status.alloc(1);
  hipStream_t st = E.stream;
  VK_HIP(hipMemsetAsync(status.p, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_exact_fill, grid_for(pp), kBlock, 0, st, p, W, F, dd.p, N, dq.p, dn.p, S.p);
  hipLaunchKernelGGL(k_exact_fill_cells, grid_for(C), kBlock, 0, st, C, p, W, E.cell_worker.p, E.cell_firm.p,
                     E.cell_weight.p, dq.p, S.p);
  VK_HIP(hipMemcpyAsync(Saug.p, S.p, pp * sizeof(double), hipMemcpyDeviceToDevice, st));
  // blocked Cholesky, in place (lower)
  for (uint32_t k = 0; k < p; k += kNb) {
    const uint32_t n = std::min<uint32_t>(kNb, p - k);
    hipLaunchKernelGGL(k_potrf_diag, 1, kBlock, 0, st, S.p, p, k, status.p);
    const uint32_t rest = p - k - n;
    if (rest == 0) continue;
    hipLaunchKernelGGL(k_trsm_panel, (rest + kPanel - 1) / kPanel, kBlock, 0, st, S.p, p, k);
    const uint32_t tiles = (rest + 63) / 64;
    hipLaunchKernelGGL(k_syrk_update, dim3(tiles, tiles), kBlock, 0, st, S.p, p, k);
  }
  int bad = 0;
  VK_HIP(hipMemcpyAsync(&bad, status.p, sizeof(int), hipMemcpyDeviceToHost, st));
  VK_HIP(hipStreamSynchronize(st));
  if (bad)
    fail(kSingularInformation, "exact_estimator", "full-firm zero-sum quotient is singular or numerically unidentified");
  // inverse: Y = L^-1 (in X), X = L^-T Y, G = (X + X')/2 (in T)
  const unsigned tg = (p + 31) / 32;
  hipLaunchKernelGGL(k_transpose_lower, dim3(tg, tg), kBlock, 0, st, S.p, T.p, p);
  VK_HIP(hipMemsetAsync(X.p, 0, pp * sizeof(double), st));
  const unsigned slabs = (p + kSlab - 1) / kSlab;
  hipLaunchKernelGGL(k_forward_identity, slabs, kBlock, 0, st, S.p, X.p, p);
  hipLaunchKernelGGL(k_backward, slabs, kBlock, 0, st, T.p, X.p, p);
  hipLaunchKernelGGL(k_symmetrize, dim3(tg, tg), kBlock, 0, st, X.p, T.p, p);
  const double *G = T.p;
  // residual gate and condition bound
  const unsigned rt = (p + 63) / 64;
  DBuf<double> part, norms;
  part.alloc((uint64_t)rt * p);
  norms.alloc(3 * (uint64_t)p);
  hipLaunchKernelGGL(k_residual_colsq, dim3(rt, rt), kBlock, 0, st, Saug.p, G, p, part.p);
  hipLaunchKernelGGL(k_exact_norms, grid_for(p), kBlock, 0, st, part.p, rt, G, Saug.p, p, norms.p);
  // fit and the firm-column mean of D G D
  std::vector<double> v(p), z(p, 0.0);
  for (uint32_t i = 0; i < p; ++i) v[i] = q[i] * rhs[i];
  for (uint32_t i = W; i < p; ++i) z[i] = q[i];
  dv.upload(v);
  dout.alloc(p);
  hipLaunchKernelGGL(k_gemv, (p + 3) / 4, kBlock, 0, st, G, dv.p, p, dout.p);
  std::vector<double> gv(p), gz(p), nrm(3 * (uint64_t)p);
  VK_HIP(hipMemcpyAsync(gv.data(), dout.p, p * sizeof(double), hipMemcpyDeviceToHost, st));
  DBuf<double> dz, dgz;
  dz.upload(z);
  dgz.alloc(p);
  hipLaunchKernelGGL(k_gemv, (p + 3) / 4, kBlock, 0, st, G, dz.p, p, dgz.p);
  VK_HIP(hipMemcpyAsync(gz.data(), dgz.p, p * sizeof(double), hipMemcpyDeviceToHost, st));
  VK_HIP(hipMemcpyAsync(nrm.data(), norms.p, 3 * (uint64_t)p * sizeof(double), hipMemcpyDeviceToHost, st));
  VK_HIP(hipStreamSynchronize(st));
  double relres = 0.0, gfro = 0.0, smax = 0.0;
  for (uint32_t j = 0; j < p; ++j) {
    relres = std::max(relres, nrm[j]);
    gfro += nrm[p + j];
    smax = std::max(smax, nrm[2 * (uint64_t)p + j]);
  }
  if (!std::isfinite(relres) || relres > std::max(100.0 * opt.rank_tolerance, 1e-10))
    fail(kInverseResidual, "exact_estimator", "zero-sum quotient inverse failed its scaled residual gate");
  // lambda_min(S) >= 1/||S^-1||_F, lambda_max(S) <= max row sum: a certified
  // lower bound of the scaled quotient reciprocal condition number
  const double rcond = 1.0 / (std::sqrt(gfro) * smax);
  if (!(rcond > opt.rank_tolerance))
    fail(kSingularInformation, "exact_estimator", "full-firm zero-sum quotient is singular or numerically unidentified");
  std::vector<double> beta(p), mvec(p);
  for (uint32_t i = 0; i < p; ++i) {
    beta[i] = q[i] * gv[i];
    mvec[i] = q[i] * gz[i] / F;
  }
  double fmean = 0.0;
  for (uint32_t g = 0; g < F; ++g) fmean += beta[W + g];
  fmean /= F;
  for (uint32_t g = 0; g < F; ++g) beta[W + g] -= fmean;
  // complete fit residual (exact_estimator.rs:1871-1957): plain worker and
  // firm normal equations over the stored rows
  std::vector<double> nres(p, 0.0), nrhs(p, 0.0);
  double rss = 0.0;
  for (uint64_t r = 0; r < P.rows; ++r) {
    const uint32_t w = P.row_worker[r], f = W + P.row_firm[r];
    const double fr = (double)P.frequency[r];
    const double e = P.outcome[r] - beta[w] - beta[f];
    nrhs[w] += fr * P.outcome[r];
    nrhs[f] += fr * P.outcome[r];
    nres[w] += fr * e;
    nres[f] += fr * e;
    rss += fr * e * e;
  }
  double na = 0.0, nb = 0.0;
  for (uint32_t i = 0; i < p; ++i) {
    na += nres[i] * nres[i];
    nb += nrhs[i] * nrhs[i];
  }
  const double fit_rel = nb > 0.0 ? std::sqrt(na / nb) : std::sqrt(na);
  const double fit_tol = std::max(1e-11, 10.0 * opt.fit_tolerance);
  if (!std::isfinite(fit_rel) || fit_rel > fit_tol)
    fail(kFullResidual, "exact_estimator", "complete original full weighted fit normal-equation residual exceeds tolerance");
  SolveResult res;
  {
    std::vector<double> alpha(beta.begin(), beta.begin() + W), psi(beta.begin() + W, beta.end());
    plugin_components(P, alpha, psi, res.plugin);
  }
  const double factor_s = seconds_since(t0);
  // per-cell leverage and target bilinear forms
  dm.upload(mvec);
  dcell.alloc(4 * (uint64_t)C);
  const size_t lds = ((size_t)p + 24) * sizeof(double);
  const unsigned cgrid = std::min<uint32_t>(C, 4096);
  hipLaunchKernelGGL(k_exact_cell, cgrid, kBlock, lds, st, G, dq.p, dm.p, p, W, F, C, E.cell_worker.p,
                     E.cell_firm.p, dws.p, dfs.p, dcm.p, dcell.p);
  std::vector<double> cell(4 * (uint64_t)C);
  VK_HIP(hipMemcpyAsync(cell.data(), dcell.p, cell.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  VK_HIP(hipStreamSynchronize(st));
  Neumaier corr[3], corr_obs[3], corr_match[3];
  double max_leverage = 0.0, gap = std::numeric_limits<double>::infinity();
  // units [0, match_units) are match blocks; the observation units are the
  // last stored rows (all rows under observation deletion, the stayers in a
  // stayer hybrid, exact_estimator.rs:914-987 and its stayer plan)
  const uint64_t observation_units = P.units - P.match_units;
  uint64_t deleted_copies = 0;
  {
    // observation deletion: every physical copy of a stored row is its own
    // block, of leverage h_c and maker 1 - h_c
    for (uint64_t r = P.rows - observation_units; r < P.rows; ++r) {
      deleted_copies += P.frequency[r];
      const uint32_t c = P.row_cell[r];
      const double h = cell[4 * (uint64_t)c];
      const double maker = 1.0 - h;
      if (!std::isfinite(h) || h < -100.0 * opt.rank_tolerance)
        fail(kNonestimable, "exact_estimator", "physical observation leverage is invalid");
      if (maker <= opt.block_tolerance)
        fail(kNonestimable, "exact_estimator", "physical observation deletion has leverage at or above one");
      const double e = P.outcome[r] - beta[P.cell_worker[c]] - beta[W + P.cell_firm[c]];
      const double scale = (double)P.frequency[r] * P.outcome[r] * e / maker;
      for (int k = 0; k < 3; ++k) corr_obs[k].add(scale * cell[4 * (uint64_t)c + 1 + k]);
      for (int k = 0; k < 3; ++k) corr[k].add(scale * cell[4 * (uint64_t)c + 1 + k]);
      max_leverage = std::max(max_leverage, h);
      gap = std::min(gap, maker);
    }
  }
  // match blocks in closed form (exact_estimator.rs:1448-1696)
  for (uint64_t g = 0; g < P.match_units; ++g) {
    const uint32_t c = P.unit_cell[g];
    const double h = cell[4 * (uint64_t)c], n = (double)P.unit_count[g];
    const double maker = 1.0 - h * n;
    if (!std::isfinite(maker) || maker <= opt.block_tolerance)
      fail(kNonestimable, "exact_estimator", "a match residual block is singular");
    const double fy = P.unit_outcome[g];
    const double fe = fy - n * (beta[P.cell_worker[c]] + beta[W + P.cell_firm[c]]);
    const double scale = fy * fe / maker;
    for (int k = 0; k < 3; ++k) corr_match[k].add(scale * cell[4 * (uint64_t)c + 1 + k]);
    for (int k = 0; k < 3; ++k) corr[k].add(scale * cell[4 * (uint64_t)c + 1 + k]);
    max_leverage = std::max(max_leverage, h * n);
    gap = std::min(gap, maker);
  }
  for (int k = 0; k < 3; ++k) res.correction[k] = corr[k].value();
  for (int k = 0; k < 3; ++k) {
    res.exact.mover_correction[k] = corr_match[k].value();
    res.exact.stayer_correction[k] = corr_obs[k].value();
  }
  res.exact.mover_correction[3] =
      res.exact.mover_correction[0] + res.exact.mover_correction[1] + 2.0 * res.exact.mover_correction[2];
  res.exact.stayer_correction[3] =
      res.exact.stayer_correction[0] + res.exact.stayer_correction[1] + 2.0 * res.exact.stayer_correction[2];
  res.correction[3] = res.correction[0] + res.correction[1] + 2.0 * res.correction[2];
  if (P.stayers) {
    // exact_estimator.rs:1154-1158, 2334-2351: the stayer hybrid's correction
    // is add_components(mover, stayer), component by component and total by
    // total, so the mover/stayer split sums to it exactly
    for (int k = 0; k < 4; ++k) res.correction[k] = res.exact.mover_correction[k] + res.exact.stayer_correction[k];
  }
  for (int k = 0; k < 4; ++k) {
    if (!std::isfinite(res.correction[k]))
      fail(kCorrectionNonFinite, "exact_estimator", "exact correction is nonfinite");
    res.corrected[k] = res.plugin[k] - res.correction[k];
  }
  res.weighted_rss = rss;
  res.accounting_residual = std::max(component_identity(res.plugin), component_identity(res.correction));
  res.has_exact = true;
  ExactReceipt &x = res.exact;
  x.parameters = (uint64_t)W + F - 1;
  x.deletion_units = P.match_units + deleted_copies;
  x.cells = C;
  x.max_leverage = max_leverage;
  x.deletion_rank_gap = gap;
  x.information_rcond = rcond;
  x.inverse_relres = relres;
  x.fit_relres = fit_rel;
  x.fit_tolerance = fit_tol;
  {
    double zs = 0.0;
    for (uint32_t g = 0; g < F; ++g) zs += beta[W + g];
    x.firm_zero_sum_residual = std::fabs(zs);
  }
  x.seconds_factor = factor_s;
  x.seconds_cells = seconds_since(t0) - factor_s;
  res.seconds_total = seconds_since(t0);
  res.fitted_cell.resize(C);
  for (uint32_t c = 0; c < C; ++c) res.fitted_cell[c] = beta[P.cell_worker[c]] + beta[W + P.cell_firm[c]];
  return res;
}

}  // namespace vk
