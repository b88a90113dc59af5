// HIP kernels of the control path: the FWL residualisation of the controls
// and the outcome after their two-way solve, the weighted Gram reductions,
// the joint residual and coefficients, the match blocks' residual makers with
// the exact control part (generic_jla.rs:6053-6273, 9097-9310), and the
// control terms of the target contraction (statistical_batches.rs:259-346).
#pragma once

#include "kernels_engine.hpp"

namespace vk {
namespace dev {

constexpr int kMaxControls = 32;
constexpr int kGramTile = 64;         // rows per LDS tile of k_ctl_gram
constexpr int kGramColumns = 2 * kMaxControls + 1;
constexpr int kGramPerThread = 5;     // outputs per thread (<= 1280 per block)

// zry[r] = (z_r - (a_w + b_f)[0..q), y_r - (a_w + b_f)[q]): the FE residuals of
// the q controls and the outcome, from their K-lane two-way solution.
template <int K>
__global__ void k_ctl_residualize(uint64_t rows, uint32_t q, const uint32_t *__restrict__ row_cell,
                                  const uint32_t *__restrict__ cell_worker, const uint32_t *__restrict__ cell_firm,
                                  const double *__restrict__ worker, const double *__restrict__ firm,
                                  const double *__restrict__ z, const double *__restrict__ y,
                                  double *__restrict__ zry) {
  VK_GRID_LOOP(r, rows) {
    const uint32_t c = row_cell[r];
    const double *a = worker + (uint64_t)cell_worker[c] * K, *b = firm + (uint64_t)cell_firm[c] * K;
    double *out = zry + r * (q + 1);
    for (uint32_t j = 0; j < q; ++j) out[j] = z[r * q + j] - (a[j] + b[j]);
    out[q] = y[r] - (a[q] + b[q]);
  }
}

// Weighted cross products over the rows of x (rows x m, row-major), one
// partial per block: mode 0 the upper triangle (i <= j, row by row) of
// sum_r w_r x_ri x_rj (m <= 33); mode 1 the cross block sum_r w_r x_ri
// x_r(h + j), i, j < h = m / 2 (m <= 64).  w null means unit weights.
__device__ inline void gram_pair(uint32_t o, uint32_t m, int mode, int &i, int &j) {
  if (mode == 1) {
    const uint32_t h = m / 2;
    i = (int)(o / h);
    j = (int)(h + o % h);
    return;
  }
  uint32_t row = 0, start = 0;
  while (o >= start + (m - row)) {
    start += m - row;
    ++row;
  }
  i = (int)row;
  j = (int)(row + (o - start));
}

__global__ void k_ctl_gram(uint64_t rows, uint32_t m, int mode, const double *__restrict__ x,
                           const double *__restrict__ w, double *__restrict__ partial) {
  __shared__ double tile[kGramTile * kGramColumns];
  __shared__ double wt[kGramTile];
  const uint32_t outs = mode == 1 ? (m / 2) * (m / 2) : m * (m + 1) / 2;
  int pi[kGramPerThread], pj[kGramPerThread];
  double acc[kGramPerThread];
#pragma unroll
  for (int s = 0; s < kGramPerThread; ++s) {
    const uint32_t o = threadIdx.x + s * kBlock;
    pi[s] = -1;
    pj[s] = 0;
    acc[s] = 0.0;
    if (o < outs) gram_pair(o, m, mode, pi[s], pj[s]);
  }
  for (uint64_t base = (uint64_t)blockIdx.x * kGramTile; base < rows; base += (uint64_t)gridDim.x * kGramTile) {
    const uint32_t n = (uint32_t)min<uint64_t>(kGramTile, rows - base);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < n * m; k += kBlock) tile[(k / m) * kGramColumns + k % m] = x[base * m + k];
    for (uint32_t k = threadIdx.x; k < n; k += kBlock) wt[k] = w ? w[base + k] : 1.0;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kGramPerThread; ++s) {
      if (pi[s] < 0) continue;
      double a = 0.0;
      for (uint32_t k = 0; k < n; ++k) a += wt[k] * tile[k * kGramColumns + pi[s]] * tile[k * kGramColumns + pj[s]];
      acc[s] += a;
    }
  }
#pragma unroll
  for (int s = 0; s < kGramPerThread; ++s) {
    const uint32_t o = threadIdx.x + s * kBlock;
    if (o < outs) partial[(uint64_t)blockIdx.x * outs + o] = acc[s];
  }
}

// Column sums of a blocks x n partial table (compensated, blocks in order).
__global__ void k_ctl_finish(uint32_t blocks, uint32_t n, const double *__restrict__ partial,
                             double *__restrict__ out) {
  VK_GRID_LOOP(o, n) {
    double s = 0.0, c = 0.0;
    for (uint32_t b = 0; b < blocks; ++b) {
      const double v = partial[(uint64_t)b * n + o];
      const double t = s + v;
      c += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
      s = t;
    }
    out[o] = s + c;
  }
}

// The joint residual e_r = yr_r - zr_r' beta and per block the partial of
// sum_r f_r e_r^2.  beta: q doubles.
__global__ void k_ctl_resid(uint64_t rows, uint32_t q, const double *__restrict__ zry, const double *__restrict__ beta,
                            const double *__restrict__ f, double *__restrict__ e, double *__restrict__ partial) {
  __shared__ double red[kBlock];
  double acc = 0.0;
  VK_GRID_LOOP(r, rows) {
    const double *z = zry + r * (q + 1);
    double v = z[q];
    for (uint32_t j = 0; j < q; ++j) v -= z[j] * beta[j];
    e[r] = v;
    acc += f[r] * v * v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out[i] = sol[i K + q] - sum_j sol[i K + j] beta_j: the joint two-way
// coefficients from the K-lane solution of the controls and the outcome.
template <int K>
__global__ void k_ctl_combine(uint64_t n, uint32_t q, const double *__restrict__ sol, const double *__restrict__ beta,
                              double *__restrict__ out) {
  VK_GRID_LOOP(i, n) {
    const double *s = sol + i * K;
    double v = s[q];
    for (uint32_t j = 0; j < q; ++j) v -= s[j] * beta[j];
    out[i] = v;
  }
}

// u_r = sqrt(f_r) Phi' zr_r (Phi Phi' = S^-1, row-major q x q): the control
// columns of a row of the block factor L.
__global__ void k_ctl_whiten(uint64_t rows, uint32_t q, const double *__restrict__ zry, const double *__restrict__ f,
                             const double *__restrict__ phi, double *__restrict__ u) {
  __shared__ double sphi[kMaxControls * kMaxControls];
  for (uint32_t k = threadIdx.x; k < q * q; k += kBlock) sphi[k] = phi[k];
  __syncthreads();
  VK_GRID_LOOP(r, rows) {
    const double *z = zry + r * (q + 1);
    double zr[kMaxControls];
    for (uint32_t i = 0; i < q; ++i) zr[i] = z[i];
    const double root = sqrt(f[r]);
    for (uint32_t j = 0; j < q; ++j) {
      double v = 0.0;
      for (uint32_t i = 0; i < q; ++i) v += zr[i] * sphi[i * q + j];
      u[r * q + j] = root * v;
    }
  }
}

// Cyclic Jacobi eigendecomposition of the symmetric n x n matrix a (row
// stride N, overwritten: eigenvalues on the diagonal); v receives the
// eigenvectors as columns.
template <int N>
__device__ inline void jacobi_eigen(int n, double *a, double *v) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) v[i * N + j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < n; ++i) {
      diag += a[i * N + i] * a[i * N + i];
      for (int j = i + 1; j < n; ++j) off += a[i * N + j] * a[i * N + j];
    }
    if (!(off > 1e-32 * diag)) break;
    for (int p = 0; p < n; ++p)
      for (int r = p + 1; r < n; ++r) {
        const double apr = a[p * N + r];
        if (apr == 0.0) continue;
        const double theta = (a[r * N + r] - a[p * N + p]) / (2.0 * apr);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = a[k * N + p], akr = a[k * N + r];
          a[k * N + p] = c * akp - s * akr;
          a[k * N + r] = s * akp + c * akr;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = a[p * N + k], ark = a[r * N + k];
          a[p * N + k] = c * apk - s * ark;
          a[r * N + k] = s * apk + c * ark;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[k * N + p], vkr = v[k * N + r];
          v[k * N + p] = c * vkp - s * vkr;
          v[k * N + r] = s * vkp + c * vkr;
        }
      }
  }
}

struct BlockParams {
  double probes, block_tolerance, rank_tolerance;
};

// generic_jla.rs:6053-6273 with controls, one thread per match unit.  The
// block residual maker is M = I - L L' with the rows L_r = (sqrt(h) c_r, u_r),
// c_r = sqrt(f_r / F_g), h the unit's finite projection share; its inverse
// acts through the n x n maker when the unit has n <= 1 + q rows, else
// through the reduced (1 + q) x (1 + q) maker I - L'L (maker_actions,
// generic_jla.rs:9097-9310), either from a Jacobi eigendecomposition whose
// smallest eigenvalue must exceed the block tolerance (75).  The complete
// residual of both actions must pass max(100 rank_tolerance, 1e-10) (76).
// With x = M^-1 sqrt(f) e and m = M^-1 c, the adjusted deleted residual of
// each row is x + B m (c'x) - V m (c'm)(c'x); per unit the kernel writes
// contrib = A B with A = sum f y and B = sum sqrt(f) value, uab = (a, b)
// with a = sum f y zr and b = sum sqrt(f) value zr (q each), and umass =
// (sum f e, B).
template <int N>
__global__ void k_ctl_blocks(uint64_t units, uint32_t q, const uint64_t *__restrict__ unit_ptr,
                             const uint32_t *__restrict__ unit_rows, const double *__restrict__ adjust,
                             const double *__restrict__ f, const double *__restrict__ y,
                             const double *__restrict__ e, const double *__restrict__ zry,
                             const double *__restrict__ u, BlockParams prm, double *__restrict__ contrib,
                             double *__restrict__ uab, double *__restrict__ umass, StatusSink status) {
  VK_GRID_LOOP(g, units) {
    const uint64_t r0 = unit_ptr[g], width = unit_ptr[g + 1] - r0;
    const uint32_t rank = q + 1;
    const bool small = width <= rank;
    const int n = (int)(small ? width : rank);
    if (n > N) { report(status, g, 90); continue; }
    const double h = adjust[0 * units + g], bias = adjust[2 * units + g], var = adjust[3 * units + g];
    const double sh = sqrt(h);
    double total = 0.0;
    for (uint64_t k = 0; k < width; ++k) total += f[unit_rows[r0 + k]];
    const double rtot = 1.0 / sqrt(total);
    // L row k, column a (0: sqrt(h) c_k; 1 + j: u_kj)
    auto lrow = [&](uint64_t k, uint32_t col) -> double {
      const uint32_t r = unit_rows[r0 + k];
      return col == 0 ? sh * sqrt(f[r]) * rtot : u[(uint64_t)r * q + col - 1];
    };
    auto rhs = [&](uint64_t k, int col) -> double {
      const uint32_t r = unit_rows[r0 + k];
      return col == 0 ? sqrt(f[r]) * e[r] : sqrt(f[r]) * rtot;
    };
    double a[N * N], v[N * N];
    if (small) {
      for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
          double s = i == j ? 1.0 : 0.0;
          for (uint32_t c = 0; c < rank; ++c) s -= lrow(i, c) * lrow(j, c);
          a[i * N + j] = s;
          a[j * N + i] = s;
        }
    } else {
      for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
          double s = 0.0;
          for (uint64_t k = 0; k < width; ++k) s += lrow(k, i) * lrow(k, j);
          a[i * N + j] = (i == j ? 1.0 : 0.0) - s;
          a[j * N + i] = a[i * N + j];
        }
    }
    jacobi_eigen<N>(n, a, v);
    double lmin = INFINITY, lmax = -INFINITY;
    for (int i = 0; i < n; ++i) {
      lmin = fmin(lmin, a[i * N + i]);
      lmax = fmax(lmax, a[i * N + i]);
    }
    if (!isfinite(lmax) || !(lmin > prm.block_tolerance)) { report(status, g, 75); continue; }
    // solve: t (n x 2) = inverse applied to the width-form rhs or to L'rhs
    double t[2][N], sol[2][N];
    for (int col = 0; col < 2; ++col) {
      for (int i = 0; i < n; ++i) {
        double s = 0.0;
        if (small) s = rhs(i, col);
        else
          for (uint64_t k = 0; k < width; ++k) s += lrow(k, i) * rhs(k, col);
        t[col][i] = s;
      }
      for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) {
          double w = 0.0;
          for (int l = 0; l < n; ++l) w += v[l * N + k] * t[col][l];
          s += v[i * N + k] * w / a[k * N + k];
        }
        sol[col][i] = s;
      }
    }
    // x_k (row k of column col) = sol (width form) or rhs + L_k sol (reduced)
    auto xval = [&](uint64_t k, int col) -> double {
      if (small) return sol[col][k];
      double s = rhs(k, col);
      for (int i = 0; i < n; ++i) s += lrow(k, i) * sol[col][i];
      return s;
    };
    // complete residual M x - rhs = x - L (L'x) - rhs
    double res2 = 0.0, rhs2 = 0.0;
    for (int col = 0; col < 2; ++col) {
      double ltx[kMaxControls + 1];
      for (uint32_t c = 0; c < rank; ++c) {
        double s = 0.0;
        for (uint64_t k = 0; k < width; ++k) s += lrow(k, c) * xval(k, col);
        ltx[c] = s;
      }
      for (uint64_t k = 0; k < width; ++k) {
        double s = xval(k, col) - rhs(k, col);
        for (uint32_t c = 0; c < rank; ++c) s -= lrow(k, c) * ltx[c];
        res2 += s * s;
        rhs2 += rhs(k, col) * rhs(k, col);
      }
    }
    const double relres = rhs2 == 0.0 ? sqrt(res2) : sqrt(res2 / rhs2);
    if (!isfinite(relres) || relres > fmax(100.0 * prm.rank_tolerance, 1e-10)) { report(status, g, 76); continue; }
    double ct = 0.0, ci = 0.0;
    for (uint64_t k = 0; k < width; ++k) {
      const double c = rhs(k, 1);
      ct += c * xval(k, 0);
      ci += c * xval(k, 1);
    }
    double A = 0.0, B = 0.0, fe = 0.0;
    double *ua = uab + g * 2 * q, *ub = ua + q;
    for (uint32_t j = 0; j < q; ++j) {
      ua[j] = 0.0;
      ub[j] = 0.0;
    }
    for (uint64_t k = 0; k < width; ++k) {
      const uint32_t r = unit_rows[r0 + k];
      const double value = xval(k, 0) + bias * xval(k, 1) * ct - var * xval(k, 1) * ci * ct;
      const double fy = f[r] * y[r], sv = sqrt(f[r]) * value;
      A += fy;
      B += sv;
      fe += f[r] * e[r];
      const double *z = zry + (uint64_t)r * (q + 1);
      for (uint32_t j = 0; j < q; ++j) {
        ua[j] += fy * z[j];
        ub[j] += sv * z[j];
      }
    }
    if (!isfinite(A) || !isfinite(B)) { report(status, g, 71); continue; }
    contrib[g] = A * B;
    umass[2 * g] = fe;
    umass[2 * g + 1] = B;
  }
}

// Per cell [Zc | hc] (C x 2q): the cell's frequency-weighted control sums and
// h_c = sum over its units of A_g b_g + B_g a_g (uab rows (a_g, b_g)).
__global__ void k_ctl_cells(uint64_t cells, uint32_t q, const uint64_t *__restrict__ cell_unit_ptr,
                            const uint32_t *__restrict__ cell_units, const double *__restrict__ unit_outcome,
                            const double *__restrict__ umass, const double *__restrict__ uab,
                            const double *__restrict__ cell_control, double *__restrict__ out) {
  VK_GRID_LOOP(c, cells) {
    double *o = out + c * 2 * q;
    for (uint32_t j = 0; j < q; ++j) {
      o[j] = cell_control[c * q + j];
      o[q + j] = 0.0;
    }
    for (uint64_t k = cell_unit_ptr[c]; k < cell_unit_ptr[c + 1]; ++k) {
      const uint32_t g = cell_units[k];
      const double A = unit_outcome[g], B = umass[2 * g + 1];
      const double *a = uab + (uint64_t)g * 2 * q, *b = a + q;
      for (uint32_t j = 0; j < q; ++j) o[q + j] += A * b[j] + B * a[j];
    }
  }
}

// Per target column j (K2 lanes): sum over cells of P_cj [Zc | hc]_c with
// P_cj = worker[w K2 + j] + firm[f K2 + j] (w, f the cell's worker and firm).
// With K2 = 64 a wavefront takes one cell at a time, its lanes the columns:
// the cell's indices and its [Zc | hc] row are wave-uniform (scalar loads)
// and only the two solution rows are vector loads; two cells per step keep
// both cells' loads in flight.  With fewer columns a wavefront takes 64 / K2
// cells.  Rows are read QM wide (cellmat carries QM doubles of padding at
// its end; columns m .. QM - 1 are never stored).  Per block the partial is
// (m x K2).
template <int K2, int QM>
__global__ void __launch_bounds__(256) k_ctl_target(uint64_t cells, uint32_t m, const uint32_t *__restrict__ cell_worker,
                                                    const uint32_t *__restrict__ cell_firm,
                                                    const double *__restrict__ worker, const double *__restrict__ firm,
                                                    const double *__restrict__ cellmat, double *__restrict__ partial) {
  constexpr int G = kBlock / K2;
  __shared__ double red[kBlock];
  const int j = threadIdx.x % K2;
  double acc[QM];
#pragma unroll
  for (int i = 0; i < QM; ++i) acc[i] = 0.0;
  if constexpr (K2 == 64) {
    const uint64_t wave = (uint64_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
    const uint64_t stride = (uint64_t)gridDim.x * G;
    uint64_t c = (uint64_t)blockIdx.x * G + wave;
    for (; c + stride < cells; c += 2 * stride) {
      const uint64_t d = c + stride;
      const double p = worker[(uint64_t)cell_worker[c] * K2 + j] + firm[(uint64_t)cell_firm[c] * K2 + j];
      const double p2 = worker[(uint64_t)cell_worker[d] * K2 + j] + firm[(uint64_t)cell_firm[d] * K2 + j];
      const double *row = cellmat + c * m, *row2 = cellmat + d * m;
#pragma unroll
      for (int i = 0; i < QM; ++i) acc[i] += p * row[i] + p2 * row2[i];
    }
    if (c < cells) {
      const double p = worker[(uint64_t)cell_worker[c] * K2 + j] + firm[(uint64_t)cell_firm[c] * K2 + j];
      const double *row = cellmat + c * m;
#pragma unroll
      for (int i = 0; i < QM; ++i)
        acc[i] += p * row[i];
    }
  } else {
    const int grp = threadIdx.x / K2;
    for (uint64_t c = (uint64_t)blockIdx.x * G + grp; c < cells; c += (uint64_t)gridDim.x * G) {
      const double p = worker[(uint64_t)cell_worker[c] * K2 + j] + firm[(uint64_t)cell_firm[c] * K2 + j];
      const double *row = cellmat + c * m;
#pragma unroll
      for (int i = 0; i < QM; ++i)
        acc[i] += p * row[i];
    }
  }
#pragma unroll
  for (int i = 0; i < QM; ++i) {
    if (i < (int)m) {  // m is uniform: every thread takes the same barriers
      __syncthreads();
      red[threadIdx.x] = acc[i];
      __syncthreads();
      if (threadIdx.x < K2) {
        double s = 0.0;
        for (int k = 0; k < G; ++k) s += red[k * K2 + threadIdx.x];
        partial[((uint64_t)blockIdx.x * m + i) * K2 + threadIdx.x] = s;
      }
    }
  }
}

// Column sums of a blocks x n partial table, one wavefront per column: each
// lane a compensated sum over every 64th block, then the lanes' sums.
__global__ void k_ctl_finish_wave(uint32_t blocks, uint32_t n, const double *__restrict__ partial,
                                  double *__restrict__ out) {
  const uint64_t o = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x % 64;
  if (o >= n) return;
  double s = 0.0, c = 0.0;
  for (uint32_t b = lane; b < blocks; b += 64) {
    const double v = partial[(uint64_t)b * n + o];
    const double t = s + v;
    c += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
    s = t;
  }
  s += c;
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) out[o] = s;
}

}  // namespace dev
}  // namespace vk
