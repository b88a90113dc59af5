// HIP kernels of the certified lockstep PCG and of the estimator passes
// around it (right-hand sides, extraction, moments, contraction).
#pragma once

#include "kernels.hpp"

namespace vk {
namespace dev {

#define VK_GRID_LOOP(t, count)                                                         \
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < (count);      \
       t += (uint64_t)gridDim.x * blockDim.x)

// Per-column state of the lockstep certified PCG (pcg.rs:1537-1909).
struct ColState {
  double rho, alpha, beta, rhs_norm, init_norm, tol, mean_x, mean_r, mean_z, xnorm, rnorm, fresh,
      orig, proj_norm;
  int32_t active, status, iters, restarts, need_fresh, restarted, first, pad;
};

struct PcgParams {
  uint32_t n;
  double rel, abs, bound;
  int32_t max_iterations, interval;
  // 0, or a cap c on the stop: ||r|| <= min(abs + rel (||b|| + bound ||x||), c ||b||)
  double cap;
};

// The certified stop of pcg.rs:2428-2436, abs + rel (||b|| + 2 maxdeg ||x||),
// optionally capped relative to ||b|| (the gate-aware first pass).
__device__ inline double stop_tolerance(const PcgParams &prm, double rhs_norm, double xnorm) {
  const double t = prm.abs + prm.rel * (rhs_norm + prm.bound * xnorm);
  return prm.cap > 0.0 ? fmin(t, prm.cap * rhs_norm) : t;
}

// Device flags of the lockstep PCG: [0] some lane is active (the gate of
// every iteration kernel), [1] some active lane needs a fresh residual,
// [2] constant 1 (always-open gate).
enum { kFlagActive = 0, kFlagFresh = 1, kFlagOne = 2 };

// The lane means the iteration kernels subtract (x's mean after its latest
// update, folded into the next update; the residual mean the V-cycle
// removes), zero for inactive lanes, and the flags.  Every pass that changes
// a lane's state (start, warm start, update, fresh residual, limit) writes
// them for that lane, so no separate pass recomputes them.
struct LanePublish {
  double *__restrict__ mean_x, *__restrict__ mean_r;
  int *__restrict__ flags;
  __device__ void put(int j, const ColState &s) const {
    mean_x[j] = s.active ? s.mean_x : 0.0;
    mean_r[j] = s.active ? s.mean_r : 0.0;
  }
};

#define VK_GATE(gate) \
  if ((gate) && *(gate) == 0) return

// ---- setup: compatibility check, norms (components.rs:669-744, pcg.rs:1596-1660)
template <int K>
struct PcgStats {
  static constexpr int kThreads = kBlock;
  uint32_t n;
  const double *__restrict__ b;
  double *__restrict__ partial;
  __device__ void operator()(Grid g) const {
    double acc[3] = {0.0, 0.0, 0.0};
    VK_GRID_STRIDE(g, t, (uint64_t)n * K) {
      const double v = b[t];
      acc[0] += v;
      acc[1] += fabs(v);
      acc[2] += v * v;
    }
    block_partial<K, 3>(acc, partial, g.block);
  }
};
template <int K>
__global__ void k_pcg_stats(PcgStats<K> f) {
  f(Grid::hw());
}
struct FinStats {
  __device__ bool enabled() const { return true; }
  ColState *__restrict__ st;
  const uint8_t *__restrict__ use;
  uint32_t n;
  __device__ void reduced(int j, const double (&tot)[3]) const {
    ColState &s = st[j];
    s = ColState{};
    s.mean_r = tot[0] / (double)n;
    s.rhs_norm = sqrt(tot[2]);
    s.active = use[j] ? 1 : 0;
    s.first = 1;
    if (s.active && fabs(tot[0]) > 1e-7 * fmax(tot[1], 1.0)) {
      s.status = 51;  // incompatible Laplacian right-hand side
      s.active = 0;
    }
  }
};

// proj = b - mean, r = proj, x = 0 (cold start; a warm start passes a
// scratch buffer as `x`).
template <int K>
struct PcgProject {
  static constexpr int kThreads = kBlock;
  uint32_t n;
  const double *__restrict__ b;
  double *__restrict__ proj, *__restrict__ r, *__restrict__ x;
  const ColState *__restrict__ st;
  __device__ void operator()(Grid g) const {
    VK_GRID_STRIDE(g, t, (uint64_t)n * K) {
      const int j = t % K;
      const double p = b[t] - st[j].mean_r;
      proj[t] = p;
      r[t] = p;
      x[t] = 0.0;
    }
  }
};
template <int K>
__global__ void k_pcg_project(PcgProject<K> f) {
  f(Grid::hw());
}

// Initial tolerance test with a zero start: initial residual = ||b||.
template <int K>
struct PcgStart {
  static constexpr int kThreads = 64;
  ColState *__restrict__ st;
  PcgParams prm;
  LanePublish pub;
  __device__ void operator()(Grid) const {
    const int j = threadIdx.x;
    bool active = false;
    if (j < K) {
      ColState &s = st[j];
      s.init_norm = s.rhs_norm;
      s.tol = stop_tolerance(prm, s.rhs_norm, 0.0);
      s.mean_r = 0.0;
      if (s.active && s.init_norm <= s.tol) {
        s.active = 0;
        s.iters = 0;
      }
      pub.put(j, s);
      active = s.active != 0;
    }
    const unsigned long long any = __ballot(active);
    if (j == 0) {
      pub.flags[kFlagActive] = any != 0ULL;
      pub.flags[kFlagFresh] = 0;
    }
  }
};

// The flags after a warm start (whose lanes published their means).
template <int K>
struct PcgFlags {
  static constexpr int kThreads = 64;
  const ColState *__restrict__ st;
  int *__restrict__ flags;
  __device__ void operator()(Grid) const {
    const int j = threadIdx.x;
    const unsigned long long any = __ballot(j < K && st[j].active);
    if (j == 0) {
      flags[kFlagActive] = any != 0ULL;
      flags[kFlagFresh] = 0;
    }
  }
};
template <int K>
__global__ void k_pcg_flags(PcgFlags<K> f) {
  f(Grid::hw());
}
template <int K>
__global__ void k_pcg_start(PcgStart<K> f) {
  f(Grid::hw());
}

// Warm start: r = proj - A x (x centred), initial residual of the original rhs.
template <int K>
struct OpWarm : CsrOp<double, K, OpWarm<K>> {
  const double *__restrict__ b, *__restrict__ proj, *__restrict__ x;
  double *__restrict__ r;
  ColState *__restrict__ st;
  PcgParams prm;
  __device__ double self(uint32_t i, int j, NoLane = {}) const { return x[(uint64_t)i * K + j]; }
  __device__ double nbr(uint32_t c, int j, NoLane = {}) const { return x[(uint64_t)c * K + j]; }
  template <int NA>
  __device__ void finish(uint32_t i, int j, double ax, double (&acc)[NA]) const {
    const uint64_t q = (uint64_t)i * K + j;
    const double xi = x[q], rr = proj[q] - ax;
    r[q] = rr;
    const double o = rr + (b[q] - proj[q]);
    acc[0] += o * o;
    acc[1] += xi * xi;
    acc[2] += rr;
    acc[3] += rr * rr;
  }
  LanePublish pub;
  template <int NV>
  __device__ void reduced(int j, const double (&tot)[NV]) const {
    ColState &s = st[j];
    s.init_norm = sqrt(tot[0]);
    s.tol = stop_tolerance(prm, s.rhs_norm, sqrt(tot[1]));
    s.mean_r = tot[2] / (double)prm.n;
    if (s.active && s.init_norm <= s.tol) {
      s.active = 0;
    } else if (s.active && tot[3] == 0.0) {
      s.status = 64;
      s.active = 0;
    }
    pub.put(j, s);
  }
};

template <int K>
struct CenterCols {
  static constexpr int kThreads = kBlock;
  uint64_t rows;
  double *__restrict__ v;
  const double *__restrict__ mean;
  const int *__restrict__ gate;
  __device__ void operator()(Grid g) const {
    VK_GATE(gate);
    static_assert(kBlock % K == 0, "a thread keeps its lane across the grid stride");
    const double m = mean[threadIdx.x % K];
    VK_GRID_STRIDE(g, t, rows * K) v[t] -= m;
  }
};
template <int K>
__global__ void k_center_cols(CenterCols<K> f) {
  f(Grid::hw());
}

// Zero the means of lanes that are not part of the current solve.
template <int K>
__global__ void k_mask_means(double *__restrict__ mean, const uint8_t *__restrict__ use) {
  const int j = threadIdx.x;
  if (j < K && !use[j]) mean[j] = 0.0;
}

// rho = r.z - mean_z * sum r with r already centred (pcg.rs:1912-1947).
template <int K, typename Z>
struct PcgRho {
  static constexpr int kThreads = kBlock;
  uint32_t n;
  const double *__restrict__ r, *__restrict__ mean_r;  // r centred on the fly (the V-cycle's input)
  const Z *__restrict__ z;
  double *__restrict__ partial;
  const int *__restrict__ gate;
  __device__ void operator()(Grid g) const {
    VK_GATE(gate);
    static_assert(kBlock % K == 0, "a thread keeps its lane across the grid stride");
    double acc[3] = {0.0, 0.0, 0.0};
    const double mr = mean_r[threadIdx.x % K];
    VK_GRID_STRIDE(g, t, (uint64_t)n * K) {
      const double zz = (double)z[t], rr = r[t] - mr;
      acc[0] += zz;
      acc[1] += rr * zz;
      acc[2] += rr;
    }
    block_partial<K, 3>(acc, partial, g.block);
  }
};
template <int K, typename Z>
__global__ void k_pcg_rho(PcgRho<K, Z> f) {
  f(Grid::hw());
}
struct FinRho {
  __device__ bool enabled() const { return true; }
  ColState *__restrict__ st;
  uint32_t n;
  __device__ void reduced(int j, const double (&tot)[3]) const {
    ColState &s = st[j];
    if (!s.active) return;
    s.mean_z = tot[0] / (double)n;
    const double rho = tot[1] - s.mean_z * tot[2];
    if (!(rho > 0.0) || !isfinite(rho)) {
      s.status = 61;  // preconditioner breakdown: r'Mr not positive
      s.active = 0;
      return;
    }
    s.beta = (s.first || s.restarted) ? 0.0 : rho / s.rho;
    s.rho = rho;
  }
};

// Direction and curvature in one pass (pcg.rs:1640-1690): p = z_c (first, or
// after a residual replacement) or z_c + beta p, A p, and the curvature
// p'Ap; alpha = rho / p'Ap.  A p is formed as A z + beta (A p)_prev: the
// Laplacian's rows sum to zero, so A z_c = A z, and each edge term
// w (z_i - z_c) takes z in the precision the cycle produced it (an FP32
// cycle output differs exactly in FP64, and is gathered at 4 bytes a lane).
template <int K, typename Z>
struct OpDirectionCurvature : CsrOp<double, K, OpDirectionCurvature<K, Z>> {
  const Z *__restrict__ z;
  double *__restrict__ p;
  double *__restrict__ ap;
  ColState *__restrict__ st;
  __device__ double self(uint32_t i, int j, NoLane = {}) const { return (double)z[(uint64_t)i * K + j]; }
  __device__ double nbr(uint32_t c, int j, NoLane = {}) const { return (double)z[(uint64_t)c * K + j]; }
  template <int NA>
  __device__ void finish(uint32_t i, int j, double az, double (&acc)[NA]) const {
    const ColState &s = st[j];
    if (!s.active) return;
    const uint64_t q = (uint64_t)i * K + j;
    const double zc = (double)z[q] - s.mean_z;
    const bool fresh = s.first || s.restarted;
    // p and A p stream through once per iteration: non-temporal accesses
    // leave the L2 to the neighbour gathers of z
    const double pn = fresh ? zc : zc + s.beta * __builtin_nontemporal_load(&p[q]);
    const double apn = fresh ? az : az + s.beta * __builtin_nontemporal_load(&ap[q]);
    __builtin_nontemporal_store(pn, &p[q]);
    __builtin_nontemporal_store(apn, &ap[q]);
    acc[0] += pn * apn;
  }
  template <int NV>
  __device__ void reduced(int j, const double (&tot)[NV]) const {
    ColState &s = st[j];
    s.first = 0;
    if (!s.active) return;
    const double curv = tot[0];
    if (!(curv > 0.0) || !isfinite(curv)) {
      s.status = 60;
      s.active = 0;
      return;
    }
    s.alpha = s.rho / curv;
  }
};

// x += alpha p, r = rc - alpha Ap (rc: the centred residual the V-cycle saw);
// norms of the centred x and of r; candidate and scheduled residual
// replacement flags (pcg.rs:1692-1760).
template <int K>
struct PcgUpdate {
  static constexpr int kThreads = kBlock;
  uint32_t n;
  const double *__restrict__ p, *__restrict__ ap;
  double *__restrict__ x;
  double *__restrict__ r, *__restrict__ partial;
  const ColState *__restrict__ st;
  // per lane, 0 for inactive lanes: x's mean after the previous update (its
  // centring, folded in here) and the residual mean the V-cycle removed
  const double *__restrict__ mean_x, *__restrict__ mean_r;
  const int *__restrict__ gate;
  __device__ void operator()(Grid g) const {
    VK_GATE(gate);
    static_assert(kBlock % K == 0, "a thread keeps its lane across the grid stride");
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const int j = (int)(threadIdx.x % K);
    const bool active = st[j].active != 0;
    const double alpha = st[j].alpha, mx = mean_x[j], mr = mean_r[j];
    VK_GRID_STRIDE(g, t, (uint64_t)n * K) {
      // x, p and A p are next touched an iteration later: non-temporal
      double xv = __builtin_nontemporal_load(&x[t]) - mx, rv = r[t] - mr;
      if (active) {
        xv += alpha * __builtin_nontemporal_load(&p[t]);
        rv -= alpha * __builtin_nontemporal_load(&ap[t]);
        __builtin_nontemporal_store(xv, &x[t]);
      }
      r[t] = rv;
      acc[0] += xv;
      acc[1] += xv * xv;
      acc[2] += rv;
      acc[3] += rv * rv;
    }
    block_partial<K, 4>(acc, partial, g.block);
  }
};
template <int K>
__global__ void k_pcg_update(PcgUpdate<K> f) {
  f(Grid::hw());
}
struct FinUpdate {
  __device__ bool enabled() const { return true; }
  ColState *__restrict__ st;
  PcgParams prm;
  LanePublish pub;
  __device__ void reduced(int j, const double (&tot)[4]) const {
    ColState &s = st[j];
    s.restarted = 0;
    s.need_fresh = 0;
    if (s.active) {
      const double n = (double)prm.n;
      s.iters += 1;
      s.mean_x = tot[0] / n;
      s.xnorm = sqrt(fmax(tot[1] - n * s.mean_x * s.mean_x, 0.0));
      s.mean_r = tot[2] / n;
      s.rnorm = sqrt(tot[3]);
      s.tol = stop_tolerance(prm, s.rhs_norm, s.xnorm);
      s.need_fresh = (s.rnorm <= s.tol) || (s.iters % prm.interval == 0);
      if (s.need_fresh) pub.flags[kFlagFresh] = 1;
    }
    pub.put(j, s);
  }
};

// Fresh residual for flagged lanes: r = proj - A x with x centred by the
// lane's pending mean, the certificate against the unprojected rhs
// (pcg.rs:1757-1817, 2209-2236).
template <int K>
struct OpFresh : CsrOp<double, K, OpFresh<K>> {
  const double *__restrict__ b, *__restrict__ proj, *__restrict__ x;
  double *__restrict__ r;
  ColState *__restrict__ st;
  PcgParams prm;
  const int *__restrict__ flags;
  LanePublish pub;  // its mean_x: x's mean after the latest update (0 for inactive lanes)
  __device__ bool enabled() const { return flags[kFlagFresh] != 0; }
  __device__ double lane(int j) const { return pub.mean_x[j]; }
  __device__ double self(uint32_t i, int j, double m) const { return x[(uint64_t)i * K + j] - m; }
  __device__ double nbr(uint32_t c, int j, double m) const { return x[(uint64_t)c * K + j] - m; }
  template <int NA>
  __device__ void finish(uint32_t i, int j, double ax, double (&acc)[NA]) const {
    if (!(st[j].active && st[j].need_fresh)) return;
    const uint64_t q = (uint64_t)i * K + j;
    const double rr = proj[q] - ax;
    r[q] = rr;
    const double o = rr + (b[q] - proj[q]);
    acc[0] += rr * rr;
    acc[1] += o * o;
    acc[2] += rr;
  }
  template <int NV>
  __device__ void reduced(int j, const double (&tot)[NV]) const {
    ColState &s = st[j];
    if (!(s.active && s.need_fresh)) return;
    s.restarted = 1;
    s.restarts += 1;
    s.fresh = sqrt(tot[0]);
    s.orig = sqrt(tot[1]);
    s.mean_r = tot[2] / (double)prm.n;
    if (s.fresh <= s.tol && s.orig <= s.tol) {
      s.active = 0;
    } else if (s.fresh <= s.tol && s.iters >= prm.max_iterations) {
      s.status = 64;
      s.active = 0;
    } else if (s.iters >= prm.max_iterations) {
      s.status = 63;
      s.active = 0;
    }
    pub.put(j, s);
  }
};

// Iteration limit and the active flag for the next iteration.
template <int K>
struct PcgLimit {
  static constexpr int kThreads = 64;
  ColState *__restrict__ st;
  PcgParams prm;
  LanePublish pub;
  __device__ void operator()(Grid) const {
    const int j = threadIdx.x;
    bool active = false;
    if (j < K) {
      ColState &s = st[j];
      if (s.active && !s.need_fresh && s.iters >= prm.max_iterations) {
        s.status = 63;
        s.active = 0;
      }
      pub.put(j, s);
      active = s.active != 0;
    }
    const unsigned long long any = __ballot(active);
    if (j == 0) {
      pub.flags[kFlagActive] = any != 0ULL;
      pub.flags[kFlagFresh] = 0;
    }
  }
};
template <int K>
__global__ void k_pcg_limit(PcgLimit<K> f) {
  f(Grid::hw());
}

// x -= the lane's pending mean (the centring of the last update of every
// lane, which the update of the next iteration would have folded in).
template <int K>
struct CenterState {
  static constexpr int kThreads = kBlock;
  uint64_t rows;
  double *__restrict__ x;
  const ColState *__restrict__ st;
  __device__ void operator()(Grid g) const {
    static_assert(kBlock % K == 0, "a thread keeps its lane across the grid stride");
    const double m = st[threadIdx.x % K].mean_x;
    VK_GRID_STRIDE(g, t, rows * K) x[t] -= m;
  }
};
template <int K>
__global__ void k_center_state(CenterState<K> f) {
  f(Grid::hw());
}

// Some lane of some sample active: the OR of the samples' any-lane-active
// flags (one block), the poll of a lockstep sample batch.
static __global__ void k_any_active(const int *const *__restrict__ flags, uint32_t count, int *__restrict__ out) {
  int any = 0;
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) any |= *flags[i] != 0;
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) *out = any;
}

// ---------------------------------------------------------------- estimator passes

// Column means over `rows` rows of a K-lane matrix.
template <int K>
struct LaneSum {
  static constexpr int kThreads = kBlock;
  uint64_t rows;
  const double *__restrict__ v;
  double *__restrict__ partial;
  __device__ void operator()(Grid g) const {
    double acc[1] = {0.0};
    VK_GRID_STRIDE(g, t, rows * K) acc[0] += v[t];
    block_partial<K, 1>(acc, partial, g.block);
  }
};
template <int K>
__global__ void k_lane_sum(LaneSum<K> f) {
  f(Grid::hw());
}
struct FinMean {
  __device__ bool enabled() const { return true; }
  double *__restrict__ mean;
  const uint8_t *__restrict__ use;  // optional: zero the means of unused lanes
  uint64_t rows;
  __device__ void reduced(int j, const double (&tot)[1]) const {
    mean[j] = (use && !use[j]) ? 0.0 : tot[0] / (double)rows;
  }
};

// Lane packing for the per-row gather passes: LP consecutive lanes per
// thread (16- and 32-byte loads) and TPR = K / LP threads per row, so a
// wavefront walks 64 / TPR rows at once (more loads in flight).
template <int K>
struct Pack {
  static constexpr int LP = K >= 4 ? 4 : K;
  static constexpr int TPR = K / LP;
};

// block_partial for packed threads: thread t holds lanes (t % TPR) * LP + q.
template <int K, int NV>
__device__ inline void block_partial_packed(const double (&val)[Pack<K>::LP][NV], double *__restrict__ out,
                                             uint32_t slot) {
  constexpr int LP = Pack<K>::LP, TPR = Pack<K>::TPR, R = kBlock / TPR;
  __shared__ double lds[NV][R * K];
  const int tid = threadIdx.x, row = tid / TPR, j0 = (tid % TPR) * LP;
#pragma unroll
  for (int q = 0; q < LP; ++q)
#pragma unroll
    for (int v = 0; v < NV; ++v) lds[v][row * K + j0 + q] = val[q][v];
  __syncthreads();
  for (int w = R / 2; w > 0; w >>= 1) {
    for (int e = tid; e < w * K; e += kBlock)
#pragma unroll
      for (int v = 0; v < NV; ++v) lds[v][e] += lds[v][e + w * K];
    __syncthreads();
  }
  if (tid < K) {
#pragma unroll
    for (int v = 0; v < NV; ++v) out[((uint64_t)v * K + tid) * kPartialStride + slot] = lds[v][tid];
  }
}

// Exact-integer leverage right-hand sides (engine.rs:2562-2619): per worker
// and per firm, the sum of the atoms of the units of its cells.
// With `cunit` (every cell one unit: the unit of each cell) the unit is read
// directly instead of through the cell's unit range.
template <int K>
struct UnitRhs {
  static constexpr int kThreads = kBlock;
  uint64_t rows;
  const uint64_t *__restrict__ ptr;
  const uint32_t *__restrict__ cells;
  const uint64_t *__restrict__ cu_ptr;
  const uint32_t *__restrict__ cu;
  const int32_t *__restrict__ atoms;
  uint64_t units;
  uint32_t kcols;
  double *__restrict__ out;
  const uint32_t *__restrict__ cunit;
  __device__ void operator()(Grid g) const {
    constexpr int LP = Pack<K>::LP, TPR = Pack<K>::TPR;
    VK_GRID_STRIDE(g, t, rows * TPR) {
      const uint64_t i = t / TPR;
      const int j0 = (int)(t % TPR) * LP;
      double s[LP];
  #pragma unroll
      for (int q = 0; q < LP; ++q) s[q] = 0.0;
      for (uint64_t k = ptr[i]; k < ptr[i + 1]; ++k) {
        const uint32_t c = cells ? cells[k] : (uint32_t)k;
        if (cunit) {
          const int32_t *a = atoms + (uint64_t)cunit[c] * K + j0;
  #pragma unroll
          for (int q = 0; q < LP; ++q) s[q] += (double)a[q];
          continue;
        }
        for (uint64_t u = cu_ptr[c]; u < cu_ptr[c + 1]; ++u) {
          const int32_t *a = atoms + (uint64_t)cu[u] * K + j0;
  #pragma unroll
          for (int q = 0; q < LP; ++q) s[q] += (double)a[q];
        }
      }
  #pragma unroll
      for (int q = 0; q < LP; ++q) out[i * K + j0 + q] = j0 + q < (int)kcols ? s[q] : 0.0;
    }
  }
};
template <int K>
__global__ void k_unit_rhs(UnitRhs<K> f) {
  f(Grid::hw());
}

// wsum[w] = sum_c w_c v[f_c] / d_w  (operator.rs:210-262 first half)
template <int K>
struct WorkerGather {
  static constexpr int kThreads = kBlock;
  uint64_t workers;
  const uint64_t *__restrict__ wptr;
  const uint32_t *__restrict__ cell_firm;
  const double *__restrict__ cell_weight, *__restrict__ wdiag, *__restrict__ v;
  double *__restrict__ out;
  __device__ void operator()(Grid g) const {
    constexpr int LP = Pack<K>::LP, TPR = Pack<K>::TPR;
    VK_GRID_STRIDE(g, t, workers * TPR) {
      const uint64_t w = t / TPR;
      const int j0 = (int)(t % TPR) * LP;
      double s[LP];
#pragma unroll
      for (int q = 0; q < LP; ++q) s[q] = 0.0;
      for (uint64_t c = wptr[w]; c < wptr[w + 1]; ++c) {
        const double cw = cell_weight[c];
        const double *f = v + (uint64_t)cell_firm[c] * K + j0;
#pragma unroll
        for (int q = 0; q < LP; ++q) s[q] += cw * f[q];
      }
      const double d = wdiag[w];
#pragma unroll
      for (int q = 0; q < LP; ++q) out[w * K + j0 + q] = s[q] / d;
    }
  }
};
template <int K>
__global__ void k_worker_gather(WorkerGather<K> f) {
  f(Grid::hw());
}

// reduced relative residual ||rhs_F - (out - mean)|| / ||rhs_F||  (full_cmg.rs:1425-1443)
template <int K>
struct ReducedResidual {
  static constexpr int kThreads = kBlock;
  uint32_t firms;
  const double *__restrict__ rhs, *__restrict__ out, *__restrict__ mean;
  double *__restrict__ partial;
  __device__ void operator()(Grid g) const {
    double acc[2] = {0.0, 0.0};
    VK_GRID_STRIDE(g, t, (uint64_t)firms * K) {
      const int j = t % K;
      const double a = rhs[t] - (out[t] - mean[j]);
      acc[0] += a * a;
      acc[1] += rhs[t] * rhs[t];
    }
    block_partial<K, 2>(acc, partial, g.block);
  }
};
template <int K>
__global__ void k_reduced_residual(ReducedResidual<K> f) {
  f(Grid::hw());
}
struct FinRelative {
  __device__ bool enabled() const { return true; }
  double *__restrict__ result;
  __device__ void reduced(int j, const double (&tot)[2]) const {
    result[j] = tot[1] == 0.0 ? sqrt(tot[0]) : sqrt(tot[0]) / sqrt(tot[1]);
  }
};

// worker = (wrhs - sum_c w_c firm[f_c]) / d_w (operator.rs:311-343) and the
// worker block of the complete residual (operator.rs:386-458),
// r_w = (wrhs - sum_c w_c firm[f_c]) - d_w worker with the computed worker.
template <int K>
struct WorkerSolution {
  static constexpr int kThreads = kBlock;
  uint64_t workers;
  const uint64_t *__restrict__ wptr;
  const uint32_t *__restrict__ cell_firm;
  const double *__restrict__ cell_weight, *__restrict__ wdiag, *__restrict__ wrhs, *__restrict__ firm;
  double *__restrict__ worker, *__restrict__ partial;
  __device__ void operator()(Grid g) const {
    constexpr int LP = Pack<K>::LP, TPR = Pack<K>::TPR;
    double acc[LP][2];
#pragma unroll
    for (int q = 0; q < LP; ++q) acc[q][0] = acc[q][1] = 0.0;
    VK_GRID_STRIDE(g, t, workers * TPR) {
      const uint64_t w = t / TPR;
      const int j0 = (int)(t % TPR) * LP;
      double b[LP], s[LP];
#pragma unroll
      for (int q = 0; q < LP; ++q) s[q] = b[q] = wrhs[w * K + j0 + q];
      for (uint64_t c = wptr[w]; c < wptr[w + 1]; ++c) {
        const double cw = cell_weight[c];
        const double *f = firm + (uint64_t)cell_firm[c] * K + j0;
#pragma unroll
        for (int q = 0; q < LP; ++q) s[q] -= cw * f[q];
      }
      const double d = wdiag[w];
#pragma unroll
      for (int q = 0; q < LP; ++q) {
        const double x = s[q] / d;
        worker[w * K + j0 + q] = x;
        const double rw = s[q] - d * x;
        acc[q][0] += rw * rw;
        acc[q][1] += b[q] * b[q];
      }
    }
    block_partial_packed<K, 2>(acc, partial, g.block);
  }
};
template <int K>
__global__ void k_worker_solution(WorkerSolution<K> f) {
  f(Grid::hw());
}
template <int K>
struct FinSums {
  __device__ bool enabled() const { return true; }
  double *__restrict__ sums;
  __device__ void reduced(int j, const double (&tot)[2]) const {
    sums[j] = tot[0];
    sums[K + j] = tot[1];
  }
};

// ---------------------------------------------------------------- firm passes
// Firm rows walk their cells (firm_cells order); Pareto-sized firms own
// hundreds of thousands of cells, so they run as chunked segments.

// Chunk of a firm's cells: sum over cells c of w_c * term(worker of c).
template <int K, class D>
struct FirmCellOp {
  const uint32_t *__restrict__ fcells, *__restrict__ cell_worker;
  const double *__restrict__ cell_weight;
  // the worker and weight of each cell in fcells order: a wavefront that
  // owns a chunk loads them with one coalesced load and keeps 16 worker-row
  // gathers in flight, summing in the cell-by-cell walk's order
  const uint32_t *__restrict__ fworker = nullptr;
  const double *__restrict__ fweight = nullptr;
  template <int STEP = Shape<K>::V>
  __device__ double partial(const uint4 &d, int j, int s) const {
    const D &op = static_cast<const D &>(*this);
    if constexpr (Shape<K>::G == 64 && Shape<K>::V == 1 && STEP == 1) {
      if (fworker) {
        constexpr int kStage = 16;
        const uint32_t lane = threadIdx.x & 63u;
        double acc = 0.0;
        for (uint32_t base = d.y; base < d.z; base += 64) {
          const uint32_t cnt = d.z - base < 64 ? d.z - base : 64;  // wave-uniform
          uint32_t mw = 0;
          double mv = 0.0;
          if (lane < cnt) {
            mw = fworker[base + lane];
            mv = fweight[base + lane];
          }
          for (uint32_t r = 0; r < cnt; r += kStage) {
            double g[kStage], wt[kStage];
#pragma unroll
            for (int u = 0; u < kStage; ++u) {
              const uint32_t idx = r + u < cnt ? r + u : 0u;
              wt[u] = lane_value<true>(mv, idx);
              g[u] = op.term(lane_value<true>(mw, idx), j);
            }
#pragma unroll
            for (int u = 0; u < kStage; ++u)
              if (r + u < cnt) acc += wt[u] * g[u];
          }
        }
        return acc;
      }
    }
    double acc = 0.0;
#pragma unroll 4
    for (uint32_t k = d.y + s; k < d.z; k += STEP) {
      const uint32_t c = fcells[k];
      acc += cell_weight[c] * op.term(cell_worker[c], j);
    }
    return acc;
  }
  __device__ bool enabled() const { return true; }
  template <int NV>
  __device__ void reduced(int, const double (&)[NV]) const {}
};

// Exact-integer leverage right-hand side of a firm (engine.rs:2562-2619):
// sum of the atoms of the units of its cells.
template <int K>
struct OpFirmUnits {
  const uint32_t *__restrict__ fcells;
  const uint64_t *__restrict__ cu_ptr;
  const uint32_t *__restrict__ cu;
  const int32_t *__restrict__ atoms;
  uint64_t units;
  uint32_t kcols;
  double *__restrict__ out;
  // when every cell is one unit: the unit of each firm-ordered cell, so a
  // wavefront that owns a chunk loads its units with one coalesced load and
  // keeps 16 atom-row gathers in flight (the same integer sums in the same
  // order as the cell-by-cell walk)
  const uint32_t *__restrict__ funits = nullptr;
  template <int STEP = Shape<K>::V>
  __device__ double partial(const uint4 &d, int j, int s) const {
    if constexpr (Shape<K>::G == 64 && Shape<K>::V == 1 && STEP == 1) {
      if (funits) {
        constexpr int kStage = 16;
        const uint32_t lane = threadIdx.x & 63u;
        double acc = 0.0;
        for (uint32_t base = d.y; base < d.z; base += 64) {
          const uint32_t cnt = d.z - base < 64 ? d.z - base : 64;  // wave-uniform
          const uint32_t mu = lane < cnt ? funits[base + lane] : 0u;
          for (uint32_t r = 0; r < cnt; r += kStage) {
            int32_t g[kStage];
#pragma unroll
            for (int u = 0; u < kStage; ++u) {
              const uint32_t idx = r + u < cnt ? r + u : 0u;
              const uint32_t unit = (uint32_t)__builtin_amdgcn_readlane((int)mu, (int)idx);
              g[u] = atoms[(uint64_t)unit * K + j];
            }
#pragma unroll
            for (int u = 0; u < kStage; ++u)
              if (r + u < cnt) acc += (double)g[u];
          }
        }
        return j < (int)kcols ? acc : 0.0;
      }
    }
    double acc = 0.0;
    if (j >= (int)kcols) return acc;
    for (uint32_t k = d.y + s; k < d.z; k += STEP) {
      const uint32_t c = fcells[k];
      for (uint64_t q = cu_ptr[c]; q < cu_ptr[c + 1]; ++q) acc += (double)atoms[(uint64_t)cu[q] * K + j];
    }
    return acc;
  }
  __device__ bool enabled() const { return true; }
  template <int NA>
  __device__ void finish(uint32_t f, int j, double v, double (&)[NA]) const { out[(uint64_t)f * K + j] = v; }
  template <int NV>
  __device__ void reduced(int, const double (&)[NV]) const {}
};

// Schur right-hand side b_F = firm_rhs - sum_c w_c worker_rhs[w_c] / d_w
// (operator.rs:269-304), uncentred, plus its column mean.
template <int K>
struct OpSchurRhs : FirmCellOp<K, OpSchurRhs<K>> {
  const double *__restrict__ wdiag, *__restrict__ wrhs, *__restrict__ frhs;
  double *__restrict__ out, *__restrict__ mean;
  uint32_t firms;
  __device__ double term(uint32_t w, int j) const { return wrhs[(uint64_t)w * K + j] / wdiag[w]; }
  template <int NA>
  __device__ void finish(uint32_t f, int j, double sum, double (&acc)[NA]) const {
    const double v = frhs[(uint64_t)f * K + j] - sum;
    out[(uint64_t)f * K + j] = v;
    acc[0] += v;
  }
  template <int NV>
  __device__ void reduced(int j, const double (&tot)[NV]) const { mean[j] = tot[0] / (double)firms; }
};

// out[f] = d_f red[f] - sum_c w_c wsum[w_c]  and its column mean.
template <int K>
struct OpSchurApply : FirmCellOp<K, OpSchurApply<K>> {
  const double *__restrict__ fdiag, *__restrict__ red, *__restrict__ wsum;
  double *__restrict__ out, *__restrict__ mean;
  uint32_t firms;
  __device__ double term(uint32_t w, int j) const { return wsum[(uint64_t)w * K + j]; }
  template <int NA>
  __device__ void finish(uint32_t f, int j, double sum, double (&acc)[NA]) const {
    const uint64_t q = (uint64_t)f * K + j;
    const double v = fdiag[f] * red[q] - sum;
    out[q] = v;
    acc[0] += v;
  }
  template <int NV>
  __device__ void reduced(int j, const double (&tot)[NV]) const { mean[j] = tot[0] / (double)firms; }
};

// Firm block of the complete two-way residual (operator.rs:386-458); the
// result is the complete relative residual with the worker block sums.
template <int K>
struct OpFirmResidual : FirmCellOp<K, OpFirmResidual<K>> {
  const double *__restrict__ fdiag, *__restrict__ frhs, *__restrict__ firm, *__restrict__ worker,
      *__restrict__ wsums;
  double *__restrict__ result, *__restrict__ rhs_norm;
  __device__ double term(uint32_t w, int j) const { return worker[(uint64_t)w * K + j]; }
  template <int NA>
  __device__ void finish(uint32_t f, int j, double sum, double (&acc)[NA]) const {
    const uint64_t q = (uint64_t)f * K + j;
    const double b = frhs[q];
    const double rf = b - fdiag[f] * firm[q] - sum;
    acc[0] += rf * rf;
    acc[1] += b * b;
  }
  template <int NV>
  __device__ void reduced(int j, const double (&tot)[NV]) const {
    const double absn = sqrt(fmax(wsums[j] + tot[0], 0.0));
    const double rhsn = sqrt(fmax(wsums[K + j] + tot[1], 0.0));
    result[j] = rhsn == 0.0 ? absn : absn / rhsn;
    rhs_norm[j] = rhsn;
  }
};

// Lane blocks between original and renumbered vertex order:
// to_renumbered (forward 1): out[p K + j] = in[perm[p] K + j];
// from_renumbered (forward 0): out[perm[p] K + j] = in[p K + j].
static __global__ void k_permute_lanes(uint64_t n, uint32_t K, const uint32_t *__restrict__ perm,
                                       const double *__restrict__ in, double *__restrict__ out, int forward) {
  VK_GRID_LOOP(t, n * K) {
    const uint64_t p = t / K, j = t % K, o = (uint64_t)perm[p] * K + j;
    if (forward)
      out[t] = in[o];
    else
      out[o] = in[t];
  }
}

// Failure report of the per-unit estimator gates.  The reference stops at the
// first failing unit of its serial loops, batches in probe order
// (engine.rs:1190-1228, 2269-2410); here every failing thread offers the key
// launch << 40 | unit << 8 | code and the smallest key wins, so the reported
// code is that of the first failing unit of the first failing launch,
// independent of thread scheduling.  The host numbers launches in issue order.
struct StatusSink {
  unsigned long long *word;
  uint64_t launch;
};
__device__ inline void report(StatusSink s, uint64_t unit, int code) {
  atomicMin(s.word, (unsigned long long)((s.launch << 40) | ((unit & 0xffffffffull) << 8) | (uint64_t)(code & 0xff)));
}

// FiveMoments accumulation per deletion unit (engine.rs:1190-1228, 354-374):
// SoA state m[10][units] = 5 x (sum, correction), one compensated addition
// per batch of probes.
__device__ inline void neumaier_add(double &s, double &c, double v) {
  const double next = s + v;
  c += (fabs(s) >= fabs(v)) ? ((s - next) + v) : ((v - next) + s);
  s = next;
}

// With unit_cell null, cell_worker and cell_firm are indexed by unit (the
// unit's worker and firm, precomputed for match units).
template <int K>
struct Moments {
  static constexpr int kThreads = kBlock;
  uint64_t units;
  uint32_t kcols;
  const uint32_t *__restrict__ unit_cell;
  const uint64_t *__restrict__ unit_count;
  const uint32_t *__restrict__ cell_worker;
  const uint32_t *__restrict__ cell_firm;
  const double *__restrict__ worker;
  const double *__restrict__ firm;
  const int32_t *__restrict__ atoms;
  double *__restrict__ m;
  StatusSink status;
  __device__ void operator()(Grid grid) const {
    // TPR threads per unit, LP lanes each (packed loads); the batch's five sums
    // are combined in-thread, then by a fixed xor butterfly over the unit's
    // threads, and added to the unit's running compensated sums by its first
    // thread, batches in probe order.  A thread takes two units per step, their
    // loads issued together.
    constexpr int LP = Pack<K>::LP, TPR = Pack<K>::TPR, U = 2;
    const uint64_t stride = (uint64_t)grid.blocks * blockDim.x, total = units * TPR;
    for (uint64_t t0 = grid.block * (uint64_t)blockDim.x + threadIdx.x; t0 < total; t0 += U * stride) {
      const int j0 = (int)(t0 % TPR) * LP;  // the same for both units (stride is a multiple of TPR)
      uint64_t g[U];
      bool on[U];
      double wv[U][LP], fv[U][LP], av[U][LP], root[U];
  #pragma unroll
      for (int q = 0; q < U; ++q) {
        const uint64_t t = t0 + q * stride;
        on[q] = t < total;
        g[q] = on[q] ? t / TPR : t0 / TPR;
        const uint64_t c = unit_cell ? unit_cell[g[q]] : g[q];
        const uint64_t w = cell_worker[c], f = cell_firm[c];
        root[q] = sqrt((double)unit_count[g[q]]);
  #pragma unroll
        for (int l = 0; l < LP; ++l) {
          const int j = j0 + l;
          wv[q][l] = worker[w * K + j];
          fv[q][l] = firm[f * K + j];
          av[q][l] = (double)atoms[g[q] * K + j];
        }
      }
  #pragma unroll
      for (int q = 0; q < U; ++q) {
        double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  #pragma unroll
        for (int l = 0; l < LP; ++l) {
          if (j0 + l < (int)kcols) {
            const double proj = root[q] * (wv[q][l] + fv[q][l]);
            const double res = av[q][l] / root[q] - proj;
            if (on[q] && (!isfinite(proj) || !isfinite(res))) report(status, g[q], 73);
            const double ps = proj * proj, rs = res * res;
            v[0] += ps;
            v[1] += rs;
            v[2] += ps * ps;
            v[3] += rs * rs;
            v[4] += ps * rs;
          }
        }
  #pragma unroll
        for (int off = 1; off < TPR; off <<= 1) {
  #pragma unroll
          for (int r = 0; r < 5; ++r) v[r] += __shfl_xor(v[r], off);
        }
        if (j0 == 0 && on[q]) {
  #pragma unroll
          for (int r = 0; r < 5; ++r) {
            double s = m[(uint64_t)(2 * r) * units + g[q]], e = m[(uint64_t)(2 * r + 1) * units + g[q]];
            neumaier_add(s, e, v[r]);
            m[(uint64_t)(2 * r) * units + g[q]] = s;
            m[(uint64_t)(2 * r + 1) * units + g[q]] = e;
          }
        }
      }
    }
  }
};
template <int K>
__global__ void k_moments(Moments<K> f) {
  f(Grid::hw());
}

// Leverage adjustment per unit (engine.rs:2269-2410), outcome mean 0.
struct AdjustParams {
  double probes, block_tolerance, rank_tolerance;
};
static __global__ void k_adjust(uint64_t units, const double *__restrict__ m, const uint32_t *__restrict__ unit_cell,
                         const uint64_t *__restrict__ unit_count, const double *__restrict__ unit_outcome,
                         const double *__restrict__ fitted_cell, AdjustParams prm,
                         double *__restrict__ out /* 7 x units */, StatusSink status) {
  VK_GRID_LOOP(g, units) {
    const double pf = m[0 * units + g] + m[1 * units + g];
    const double mf = m[2 * units + g] + m[3 * units + g];
    const double total = (pf + mf) / prm.probes;
    if (!isfinite(total) || total <= prm.block_tolerance) { report(status, g, 74); continue; }
    const double proj = (pf / prm.probes) / total;
    const double res = (mf / prm.probes) / total;
    const double p2 = (m[4 * units + g] + m[5 * units + g]) / prm.probes;
    const double m2 = (m[6 * units + g] + m[7 * units + g]) / prm.probes;
    const double mixed = (m[8 * units + g] + m[9 * units + g]) / prm.probes;
    const double bias = (res * p2 - proj * m2 + (res - proj) * mixed) / prm.probes;
    double var = (res * res * p2 + proj * proj * m2 - 2.0 * proj * res * mixed) / prm.probes;
    if (!isfinite(proj) || !isfinite(res) || !isfinite(bias) || !isfinite(var)) { report(status, g, 73); continue; }
    if (var < -100.0 * prm.rank_tolerance) { report(status, g, 73); continue; }
    var = fmax(var, 0.0);
    if (1.0 - proj <= prm.block_tolerance) { report(status, g, 75); continue; }
    const double recip = 1.0 / res;
    const double rr = fabs(res * recip - 1.0);
    if (!isfinite(recip) || !isfinite(rr) || rr > fmax(100.0 * prm.rank_tolerance, 1e-10)) { report(status, g, 76); continue; }
    const uint32_t c = unit_cell[g];
    const double resid = unit_outcome[g] - (double)unit_count[g] * fitted_cell[c];
    const double mult = recip + bias * (recip * recip) - var * (recip * recip * recip);
    const double deleted = resid * mult;
    if (!isfinite(resid) || !isfinite(mult) || !isfinite(deleted)) { report(status, g, 71); continue; }
    out[0 * units + g] = proj;
    out[1 * units + g] = res;
    out[2 * units + g] = bias;
    out[3 * units + g] = var;
    out[4 * units + g] = resid;
    out[5 * units + g] = deleted;
    out[6 * units + g] = unit_outcome[g] * deleted;
  }
}

// ------------------------------------------------ observation deletion
// Generic-engine observation deletion without controls.  One unit per stored
// row, whose atom sums the signs of the row's own physical copies (k_atoms
// with copy offsets).

// ObservationMoments and ObservationCorrelations (statistical_batches.rs:79-206):
// per stored row the probe sums of p^2 and p^4, per physical copy those of
// s p and s p^3, p = alpha_w + psi_f of the row's cell and s the copy's sign,
// bit (copy % 64) of Counter-V1 word (entity, probe, copy / 64)
// (generic_jla.rs:10321-10329).  m: 4 x rows (p^2 sum, correction, p^4 sum,
// correction); corr: 4 x physical copies (s p sum, correction, s p^3 sum,
// correction); one compensated addition per batch, batches in probe order.
template <int K>
struct ObsMoments {
  static constexpr int kThreads = kBlock;
  uint64_t rows;
  uint32_t kcols;
  uint64_t first_probe;
  uint64_t seed;
  uint64_t tag;
  const uint32_t *__restrict__ unit_cell;
  const uint64_t *__restrict__ unit_count;
  const uint64_t *__restrict__ unit_rank;
  const uint64_t *__restrict__ unit_offset;
  const uint64_t *__restrict__ row_phys;
  uint64_t physical;
  const uint32_t *__restrict__ cell_worker;
  const uint32_t *__restrict__ cell_firm;
  const double *__restrict__ worker;
  const double *__restrict__ firm;
  double *__restrict__ m;
  double *__restrict__ corr;
  StatusSink status;
  __device__ void operator()(Grid grid) const {
    constexpr int LP = Pack<K>::LP, TPR = Pack<K>::TPR;
    const uint64_t k0 = splitmix64(seed), k1 = splitmix64(seed ^ 0xd1b54a32d192ed03ULL);
    VK_GRID_STRIDE(grid, t, rows * TPR) {
      const uint64_t g = t / TPR;
      const int j0 = (int)(t % TPR) * LP;
      const uint32_t c = unit_cell[g];
      const uint64_t w = cell_worker[c], f = cell_firm[c];
      double p[LP], v[2] = {0.0, 0.0};
  #pragma unroll
      for (int q = 0; q < LP; ++q) {
        const int j = j0 + q;
        p[q] = 0.0;
        if (j < (int)kcols) {
          p[q] = worker[w * K + j] + firm[f * K + j];
          if (!isfinite(p[q])) report(status, g, 73);
          const double p2 = p[q] * p[q];
          v[0] += p2;
          v[1] += p2 * p2;
        }
      }
  #pragma unroll
      for (int off = 1; off < TPR; off <<= 1) {
        v[0] += __shfl_xor(v[0], off);
        v[1] += __shfl_xor(v[1], off);
      }
      if (j0 == 0) {
  #pragma unroll
        for (int q = 0; q < 2; ++q) {
          double s = m[(uint64_t)(2 * q) * rows + g], e = m[(uint64_t)(2 * q + 1) * rows + g];
          neumaier_add(s, e, v[q]);
          m[(uint64_t)(2 * q) * rows + g] = s;
          m[(uint64_t)(2 * q + 1) * rows + g] = e;
        }
      }
      // every thread of the row runs the same copy loops, so the butterflies
      // below see the whole group
      const uint64_t lo = unit_offset[g], hi = lo + unit_count[g], entity = unit_rank[g], base = row_phys[g];
      for (uint64_t word = lo / 64; word * 64 < hi; ++word) {
        uint64_t bits[LP], blk[4] = {0, 0, 0, 0}, cached = ~0ULL;
  #pragma unroll
        for (int q = 0; q < LP; ++q) {
          const uint64_t probe = first_probe + (uint64_t)(j0 + q), pb = probe / 4;
          if (pb != cached) {
            blk[0] = entity; blk[1] = pb; blk[2] = word; blk[3] = tag;
            philox(k0, k1, blk);
            cached = pb;
          }
          bits[q] = blk[probe % 4];
        }
        const uint64_t c0 = lo > word * 64 ? lo : word * 64, c1 = hi < word * 64 + 64 ? hi : word * 64 + 64;
        for (uint64_t copy = c0; copy < c1; ++copy) {
          double a[2] = {0.0, 0.0};
  #pragma unroll
          for (int q = 0; q < LP; ++q) {
            if (j0 + q < (int)kcols) {
              const double sgn = ((bits[q] >> (copy % 64)) & 1) ? 1.0 : -1.0;
              a[0] += sgn * p[q];
              a[1] += sgn * (p[q] * p[q] * p[q]);
            }
          }
  #pragma unroll
          for (int off = 1; off < TPR; off <<= 1) {
            a[0] += __shfl_xor(a[0], off);
            a[1] += __shfl_xor(a[1], off);
          }
          if (j0 == 0) {
            const uint64_t x = base + (copy - lo);
  #pragma unroll
            for (int q = 0; q < 2; ++q) {
              double s = corr[(uint64_t)(2 * q) * physical + x], e = corr[(uint64_t)(2 * q + 1) * physical + x];
              neumaier_add(s, e, a[q]);
              corr[(uint64_t)(2 * q) * physical + x] = s;
              corr[(uint64_t)(2 * q + 1) * physical + x] = e;
            }
          }
        }
      }
    }
  }
};
template <int K>
__global__ void k_obs_moments(ObsMoments<K> f) {
  f(Grid::hw());
}

// observation_deleted_adjustment (generic_jla.rs:5893-6017) with
// FiveMoments::from_totals and finite (:441-513), no control leverage: per
// copy the finite-projection inverse 1/m + b/m^2 - v/m^3, d = e (sum over
// the row's copies) / f.  out (7 x rows): means over the copies of the
// projection share, residual share, bias and variance; e; d; and f y d, the
// row's share of its cell's correction weight.
static __global__ void k_obs_adjust(uint64_t rows, const double *__restrict__ m, const double *__restrict__ corr,
                                    uint64_t physical, const uint64_t *__restrict__ row_phys,
                                    const uint64_t *__restrict__ unit_count, const uint32_t *__restrict__ unit_cell,
                                    const double *__restrict__ row_outcome, const double *__restrict__ fitted_cell,
                                    AdjustParams prm, double *__restrict__ out, StatusSink status) {
  VK_GRID_LOOP(g, rows) {
    const double pf = m[0 * rows + g] + m[1 * rows + g];
    const double ps = m[2 * rows + g] + m[3 * rows + g];
    double is = 0.0, ic = 0.0, ls = 0.0, lc = 0.0, rs = 0.0, rc = 0.0, bs = 0.0, bc = 0.0, vs = 0.0, vc = 0.0;
    int fail = 0;
    for (uint64_t x = row_phys[g]; x < row_phys[g + 1]; ++x) {
      const double first = corr[0 * physical + x] + corr[1 * physical + x];
      const double third = corr[2 * physical + x] + corr[3 * physical + x];
      const double t0 = pf, t1 = prm.probes + pf - 2.0 * first, t2 = ps;
      const double t3 = prm.probes + 6.0 * pf + ps - 4.0 * first - 4.0 * third, t4 = pf + ps - 2.0 * third;
      const double total = (t0 + t1) / prm.probes;
      if (!isfinite(total) || total <= prm.block_tolerance) { fail = 74; break; }
      const double proj = (t0 / prm.probes) / total, res = (t1 / prm.probes) / total;
      const double p4 = t2 / prm.probes, m4 = t3 / prm.probes, mixed = t4 / prm.probes;
      const double bias = (res * p4 - proj * m4 + (res - proj) * mixed) / prm.probes;
      double var = (res * res * p4 + proj * proj * m4 - 2.0 * proj * res * mixed) / prm.probes;
      if (!isfinite(proj) || !isfinite(res) || !isfinite(bias) || !isfinite(var) ||
          var < -100.0 * prm.rank_tolerance) { fail = 73; break; }
      var = fmax(var, 0.0);
      if (!isfinite(res) || res <= prm.block_tolerance) { fail = 75; break; }
      const double r = 1.0 / res;
      const double inv = r + bias * (r * r) - var * (r * r * r);
      if (!isfinite(inv) || inv <= 0.0) { fail = 76; break; }
      neumaier_add(is, ic, inv);
      neumaier_add(ls, lc, proj);
      neumaier_add(rs, rc, res);
      neumaier_add(bs, bc, bias);
      neumaier_add(vs, vc, var);
    }
    if (fail) { report(status, g, fail); continue; }
    const double f = (double)unit_count[g];
    const double y = row_outcome[g], e = y - fitted_cell[unit_cell[g]];
    const double deleted = e * ((is + ic) / f);
    if (!isfinite(deleted)) { report(status, g, 71); continue; }
    out[0 * rows + g] = (ls + lc) / f;
    out[1 * rows + g] = (rs + rc) / f;
    out[2 * rows + g] = (bs + bc) / f;
    out[3 * rows + g] = (vs + vc) / f;
    out[4 * rows + g] = e;
    out[5 * rows + g] = deleted;
    out[6 * rows + g] = f * y * deleted;
  }
}

// ------------------------------------------------ numerical MCSE attachment
// engine/numerical.rs and numerical_mc.rs: a local linearisation of every
// unit's leverage adjustment in its five moment means, fold sums of the
// target solutions, and replayed leverage probes scored against both.

// FiniteDerivative::new and observation(0) (numerical_mc.rs:54-125): the
// gradient of 1/m + bias/m^2 - variance/m^3 in the five means.  Returns 0
// (defined), 1 (no finite derivative), 5 (nonsmooth: variance < 0) or 2 (no
// finite gradient); h, m and the variance are set unless it returns 1.
__host__ __device__ inline int unit_linearisation(const double (&u)[5], double r, double (&grad)[5], double &h_out,
                                                  double &m_out, double &var_out) {
  const double p = u[0], mm = u[1], a = u[2], b = u[3], c = u[4];
  const double s = p + mm;
  if (!(s > 0.0)) return 1;
#pragma unroll
  for (int j = 0; j < 5; ++j)
    if (!__builtin_isfinite(u[j])) return 1;
  const double h = p / s, m = mm / s;
  const double bias = (m * a - h * b + (m - h) * c) / r;
  const double var = (m * m * a + h * h * b - 2.0 * h * m * c) / r;
  double dh[5] = {m / s, -h / s, 0.0, 0.0, 0.0}, db[5], dv[5];
  bool finite = __builtin_isfinite(h) && __builtin_isfinite(m) && __builtin_isfinite(bias) && __builtin_isfinite(var);
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const double da = j == 2, dbb = j == 3, dc = j == 4;
    db[j] = (-(a + b + 2.0 * c) * dh[j] + m * da - h * dbb + (m - h) * dc) / r;
    dv[j] = ((-2.0 * m * a + 2.0 * h * b - 2.0 * c * (m - h)) * dh[j] + m * m * da + h * h * dbb -
             2.0 * h * m * dc) / r;
    finite = finite && __builtin_isfinite(dh[j]) && __builtin_isfinite(db[j]) && __builtin_isfinite(dv[j]);
  }
  if (!finite) return 1;
  h_out = h;
  m_out = m;
  var_out = var;
  if (var < 0.0) return 5;
  if (m <= 0.0) return 2;
  const double inv = 1.0 / m, inv2 = inv * inv, inv3 = inv2 * inv, inv4 = inv2 * inv2;
  const double factor = -inv2 - 2.0 * bias * inv3 + 3.0 * var * inv4;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    grad[j] = -factor * dh[j] + inv2 * db[j] - inv3 * dv[j];
    if (!__builtin_isfinite(grad[j])) return 2;
  }
  return 0;
}

// Per unit: gradient (5 x units), offset gradient . u (Neumaier), and block
// partials for the attachment's diagnostics: [min u0 + u1, min m, max
// sqrt(variance)/m, last unit with an event (unit * 8 + code, or -1), any
// clipped adjustment variance (engine.rs:2327-2338)].
static __global__ void __launch_bounds__(kBlock) k_numerical_state(uint64_t units, const double *__restrict__ m,
                                                                   double r, double *__restrict__ grad,
                                                                   double *__restrict__ offs,
                                                                   double *__restrict__ part) {
  __shared__ double lds[3][kBlock];
  __shared__ long long lev[kBlock], lclip[kBlock];
  double mn_c = INFINITY, mn_m = INFINITY, mx_s = 0.0;
  long long last = -1, clipped = 0;
  VK_GRID_LOOP(g, units) {
    double u[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) u[q] = (m[(2 * q) * units + g] + m[(2 * q + 1) * units + g]) / r;
    // the adjustment's own variance (engine.rs:2299-2338) marks clipping
    {
      const double pf = m[0 * units + g] + m[1 * units + g], mf = m[2 * units + g] + m[3 * units + g];
      const double total = (pf + mf) / r;
      const double proj = (pf / r) / total, res = (mf / r) / total;
      const double var = (res * res * u[2] + proj * proj * u[3] - 2.0 * proj * res * u[4]) / r;
      if (var < 0.0) clipped = 1;
    }
    double gr[5] = {0, 0, 0, 0, 0}, h = 0.0, mm = 0.0, var = 0.0;
    const int code = unit_linearisation(u, r, gr, h, mm, var);
    if (code != 1) {  // numerical.rs:95-100: after a finite derivative exists
      mn_c = fmin(mn_c, u[0] + u[1]);
      mn_m = fmin(mn_m, mm);
      mx_s = fmax(mx_s, sqrt(fmax(var, 0.0)) / mm);
    }
    // event status: 5 nonsmooth adjustment, 6 nonfinite derivative
    if (code) last = (long long)g * 8 + (code == 5 ? 5 : 6);
    double s = 0.0, e = 0.0;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      grad[(uint64_t)q * units + g] = code ? 0.0 : gr[q];
      if (!code) neumaier_add(s, e, gr[q] * u[q]);
    }
    offs[g] = s + e;
  }
  lds[0][threadIdx.x] = mn_c;
  lds[1][threadIdx.x] = mn_m;
  lds[2][threadIdx.x] = mx_s;
  lev[threadIdx.x] = last;
  lclip[threadIdx.x] = clipped;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      lds[0][threadIdx.x] = fmin(lds[0][threadIdx.x], lds[0][threadIdx.x + w]);
      lds[1][threadIdx.x] = fmin(lds[1][threadIdx.x], lds[1][threadIdx.x + w]);
      lds[2][threadIdx.x] = fmax(lds[2][threadIdx.x], lds[2][threadIdx.x + w]);
      lev[threadIdx.x] = lev[threadIdx.x] > lev[threadIdx.x + w] ? lev[threadIdx.x] : lev[threadIdx.x + w];
      lclip[threadIdx.x] |= lclip[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double *o = part + 5 * (uint64_t)blockIdx.x;
    o[0] = lds[0][0];
    o[1] = lds[1][0];
    o[2] = lds[2][0];
    o[3] = (double)lev[0];
    o[4] = (double)lclip[0];
  }
}

// Target folds (engine/numerical.rs:124-157): for every probe of a target
// batch in probe order, unit g adds mass * (sw^2, sf^2, sw sf) to its fold's
// compensated sums, mass = outcome_g * residual_mass_g / fold size; sw and sf
// are the worker- and firm-score solutions (columns 2q and 2q + 1) at the
// unit's cell.  folds: sums then errors, [(fold * 3 + j) * units + g].
template <int K2>
__global__ void __launch_bounds__(kBlock) k_numerical_folds(uint64_t units, uint32_t probes, uint64_t first,
                                                            uint64_t R, const uint32_t *__restrict__ unit_cell,
                                                            const uint32_t *__restrict__ cell_worker,
                                                            const uint32_t *__restrict__ cell_firm,
                                                            const double *__restrict__ unit_outcome,
                                                            const double *__restrict__ residual_mass,
                                                            const double *__restrict__ worker,
                                                            const double *__restrict__ firm,
                                                            double *__restrict__ folds) {
  VK_GRID_LOOP(g, units) {
    const uint32_t c = unit_cell[g];
    const uint64_t w = cell_worker[c], f = cell_firm[c];
    const double base = unit_outcome[g] * residual_mass[g];
    for (uint32_t q = 0; q < probes; ++q) {
      const uint64_t probe = first + q;
      const int fold = (int)(probe % 2);
      const double count = fold == 0 ? (double)((R + 1) / 2) : (double)(R / 2);
      const double sw = worker[w * K2 + 2 * q] + firm[f * K2 + 2 * q];
      const double sf = worker[w * K2 + 2 * q + 1] + firm[f * K2 + 2 * q + 1];
      const double mass = base / count;
      const double kappa[3] = {sw * sw, sf * sf, sw * sf};
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double &s = folds[(uint64_t)(fold * 3 + j) * units + g];
        double &e = folds[(uint64_t)(6 + fold * 3 + j) * units + g];
        neumaier_add(s, e, mass * kappa[j]);
      }
    }
  }
}

// Replay scores (engine/numerical.rs:246-291): per replayed probe (lane) the
// six sums over units of -fold[g][f][j] * alpha_g, alpha_g = gradient_g . u_g
// - offset_g with u_g the unit's (p^2, m^2, p^4, m^4, p^2 m^2).
template <int K>
__global__ void __launch_bounds__(kBlock) k_numerical_scores(uint64_t units, uint32_t kcols,
                                                             const uint32_t *__restrict__ unit_cell,
                                                             const uint64_t *__restrict__ unit_count,
                                                             const uint32_t *__restrict__ cell_worker,
                                                             const uint32_t *__restrict__ cell_firm,
                                                             const double *__restrict__ worker,
                                                             const double *__restrict__ firm,
                                                             const int32_t *__restrict__ atoms,
                                                             const double *__restrict__ grad,
                                                             const double *__restrict__ offs,
                                                             const double *__restrict__ folds,
                                                             double *__restrict__ partial) {
  double acc[6] = {0, 0, 0, 0, 0, 0};
  VK_GRID_LOOP(t, units * K) {
    const uint64_t g = t / K;
    const int j = t % K;
    if (j >= (int)kcols) continue;
    const uint32_t c = unit_cell[g];
    const uint64_t w = cell_worker[c], f = cell_firm[c];
    const double root = sqrt((double)unit_count[g]);
    const double p = root * (worker[w * K + j] + firm[f * K + j]);
    const double mres = (double)atoms[g * K + j] / root - p;
    const double p2 = p * p, m2 = mres * mres;
    const double u[5] = {p2, m2, p2 * p2, m2 * m2, p2 * m2};
    double s = 0.0, e = 0.0;
    neumaier_add(s, e, -offs[g]);
#pragma unroll
    for (int q = 0; q < 5; ++q) neumaier_add(s, e, grad[(uint64_t)q * units + g] * u[q]);
    const double alpha = s + e;
#pragma unroll
    for (int v = 0; v < 6; ++v) acc[v] -= (folds[(uint64_t)v * units + g] + folds[(uint64_t)(6 + v) * units + g]) * alpha;
  }
  block_partial<K, 6>(acc, partial);
}
// Replay scores of observation units (generic_jla/numerical.rs:617-641): per
// replayed probe (lane) the six sums over units of -fold[g][f][j] * alpha_g,
// alpha_g = g2 p^2 + g4 p^4 - offset + (sum_copies q1 s) p + (sum_copies q3 s) p^3,
// with the copies' signs s regenerated from the leverage Counter-V1 words.
// grad: 3 x units (g2, g4, offset); copy: 2 x copies (q1, q3).
template <int K>
__global__ void __launch_bounds__(kBlock) k_obs_scores(uint64_t units, uint32_t kcols, uint64_t first_probe,
                                                       uint64_t seed, uint64_t tag,
                                                       const uint32_t *__restrict__ unit_cell,
                                                       const uint64_t *__restrict__ unit_count,
                                                       const uint64_t *__restrict__ unit_rank,
                                                       const uint64_t *__restrict__ unit_offset,
                                                       const uint64_t *__restrict__ row_phys, uint64_t physical,
                                                       const uint32_t *__restrict__ cell_worker,
                                                       const uint32_t *__restrict__ cell_firm,
                                                       const double *__restrict__ worker,
                                                       const double *__restrict__ firm,
                                                       const double *__restrict__ grad,
                                                       const double *__restrict__ copy,
                                                       const double *__restrict__ folds,
                                                       double *__restrict__ partial) {
  const uint64_t k0 = splitmix64(seed), k1 = splitmix64(seed ^ 0xd1b54a32d192ed03ULL);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  VK_GRID_LOOP(t, units * K) {
    const uint64_t g = t / K;
    const int j = t % K;
    if (j >= (int)kcols) continue;
    const uint32_t c = unit_cell[g];
    const uint64_t w = cell_worker[c], f = cell_firm[c];
    const double p = worker[w * K + j] + firm[f * K + j];
    const uint64_t probe = first_probe + (uint64_t)j;
    const uint64_t lo = unit_offset[g], hi = lo + unit_count[g], base = row_phys[g];
    double w1 = 0.0, w1e = 0.0, w3 = 0.0, w3e = 0.0;
    for (uint64_t word = lo / 64; word * 64 < hi; ++word) {
      uint64_t blk[4] = {unit_rank[g], probe / 4, word, tag};
      philox(k0, k1, blk);
      const uint64_t bits = blk[probe % 4];
      const uint64_t c0 = lo > word * 64 ? lo : word * 64, c1 = hi < word * 64 + 64 ? hi : word * 64 + 64;
      for (uint64_t cp = c0; cp < c1; ++cp) {
        const double s = ((bits >> (cp % 64)) & 1) ? 1.0 : -1.0;
        const uint64_t x = base + (cp - lo);
        neumaier_add(w1, w1e, copy[x] * s);
        neumaier_add(w3, w3e, copy[physical + x] * s);
      }
    }
    double a = 0.0, ae = 0.0;
    neumaier_add(a, ae, grad[g] * p * p);
    neumaier_add(a, ae, grad[units + g] * (p * p * p * p));
    neumaier_add(a, ae, -grad[2 * units + g]);
    neumaier_add(a, ae, (w1 + w1e) * p);
    neumaier_add(a, ae, (w3 + w3e) * (p * p * p));
    const double alpha = a + ae;
#pragma unroll
    for (int v = 0; v < 6; ++v) acc[v] -= (folds[(uint64_t)v * units + g] + folds[(uint64_t)(6 + v) * units + g]) * alpha;
  }
  block_partial<K, 6>(acc, partial);
}

struct FinScores {
  double *__restrict__ out;  // [lane][fold * 3 + j]
  __device__ bool enabled() const { return true; }
  __device__ void reduced(int j, const double (&tot)[6]) const {
#pragma unroll
    for (int v = 0; v < 6; ++v) out[j * 6 + v] = tot[v];
  }
};

// cell_correction_weight[c] = compensated sum over the cell's units ascending.
static __global__ void k_cell_weight(uint64_t cells, const uint64_t *__restrict__ cu_ptr, const uint32_t *__restrict__ cu,
                              const double *__restrict__ contrib, double *__restrict__ out) {
  VK_GRID_LOOP(c, cells) {
    double s = 0.0, k = 0.0;
    for (uint64_t q = cu_ptr[c]; q < cu_ptr[c + 1]; ++q) neumaier_add(s, k, contrib[cu[q]]);
    out[c] = s + k;
  }
}

// Target direction per cell: compensated sum over its strata of
// sqrt(per_copy / T) atom (engine.rs:2621-2703), totals for centring.
// atoms: strata x KT, interleaved.
template <int KT>
struct TargetDirection {
  static constexpr int kThreads = kBlock;
  uint64_t cells;
  const uint64_t *__restrict__ csp;
  const double *__restrict__ per_copy;
  double target_total;
  const int32_t *__restrict__ atoms;
  double *__restrict__ dir;
  double *__restrict__ partial;
  __device__ void operator()(Grid g) const {
    double acc[2] = {0.0, 0.0};
    VK_GRID_STRIDE(g, t, cells * KT) {
      const uint64_t c = t / KT;
      const int j = t % KT;
      double s = 0.0, k = 0.0;
      for (uint64_t q = csp[c]; q < csp[c + 1]; ++q)
        neumaier_add(s, k, sqrt(per_copy[q] / target_total) * (double)atoms[q * KT + j]);
      const double d = s + k;
      dir[t] = d;
      acc[0] += d;
      acc[1] += fabs(d);
    }
    block_partial<KT, 2>(acc, partial, g.block);
  }
};
template <int KT>
__global__ void k_target_direction(TargetDirection<KT> f) {
  f(Grid::hw());
}
template <int KT>
struct FinTotals {
  __device__ bool enabled() const { return true; }
  double *__restrict__ totals;
  __device__ void reduced(int j, const double (&tot)[2]) const {
    totals[j] = tot[0];
    totals[KT + j] = tot[1] + fabs(tot[0]);  // reference scale
  }
};

// Centre, then transpose to worker/firm scores in the 2-column layout:
// column 2j = (score_w, 0), column 2j+1 = (0, score_f)  (engine.rs:2705-2971).
template <int KT>
struct TargetScores {
  static constexpr int kThreads = kBlock;
  uint64_t rows;
  const uint64_t *__restrict__ ptr;
  const uint32_t *__restrict__ cells;
  const double *__restrict__ dir;
  const double *__restrict__ cell_target;
  double target_total;
  const double *__restrict__ totals;
  int parity;
  double *__restrict__ out;
  __device__ void operator()(Grid g) const {
    constexpr int K2 = 2 * KT;
    VK_GRID_STRIDE(g, t, rows * K2) {
      const uint64_t i = t / K2;
      const int l = t % K2;
      const int j = l / 2;
      double s = 0.0;
      if ((l & 1) == parity) {
        for (uint64_t k = ptr[i]; k < ptr[i + 1]; ++k) {
          const uint64_t c = cells ? cells[k] : k;
          s += dir[c * KT + j] - cell_target[c] / target_total * totals[j];
        }
      }
      out[t] = s;
    }
  }
};
template <int KT>
__global__ void k_target_scores(TargetScores<KT> f) {
  f(Grid::hw());
}

// Firm scores of the target directions in the 2-column layout: column
// 2j + 1 = sum over the firm's cells of the centred direction, column 2j = 0
// (engine.rs:2705-2971).  Chunked like the other firm passes.
template <int KT>
struct OpFirmTargetScores {
  const uint32_t *__restrict__ fcells;
  const double *__restrict__ dir, *__restrict__ cell_target, *__restrict__ totals;
  double target_total;
  double *__restrict__ out;
  template <int STEP = Shape<2 * KT>::V>
  __device__ double partial(const uint4 &d, int l, int s) const {
    double acc = 0.0;
    if ((l & 1) == 0) return acc;
    const int j = l / 2;
    for (uint32_t k = d.y + s; k < d.z; k += STEP) {
      const uint32_t c = fcells[k];
      acc += dir[(uint64_t)c * KT + j] - cell_target[c] / target_total * totals[j];
    }
    return acc;
  }
  __device__ bool enabled() const { return true; }
  template <int NA>
  __device__ void finish(uint32_t f, int l, double v, double (&)[NA]) const { out[(uint64_t)f * 2 * KT + l] = v; }
  template <int NV>
  __device__ void reduced(int, const double (&)[NV]) const {}
};

// Compatibility repair of the last score entry (engine.rs:2842-2881).
template <int K2>
struct BalanceSum {
  static constexpr int kThreads = kBlock;
  uint64_t rows;
  const double *__restrict__ v;
  double *__restrict__ partial;
  __device__ void operator()(Grid g) const {
    double acc[1] = {0.0};
    VK_GRID_STRIDE(g, t, (rows - 1) * K2) acc[0] += v[t];
    block_partial<K2, 1>(acc, partial, g.block);
  }
};
template <int K2>
__global__ void k_balance_sum(BalanceSum<K2> f) {
  f(Grid::hw());
}
template <int K2>
struct FinBalance {
  __device__ bool enabled() const { return true; }
  double *__restrict__ v;
  uint64_t rows;
  int parity;
  const double *__restrict__ totals;
  StatusSink status;
  __device__ void reduced(int l, const double (&tot)[1]) const {
    if ((l & 1) != parity) return;
    const double ref = totals[K2 / 2 + l / 2];
    const double last = -tot[0];
    const double original = v[(rows - 1) * K2 + l];
    const double gate = 4096.0 * 2.220446049250313e-16;
    if ((ref == 0.0 && last != original) || (ref > 0.0 && fabs(last - original) > gate * ref))
      report(status, (uint64_t)l, 77);
    v[(rows - 1) * K2 + l] = last;
  }
};

// Target contraction per probe (engine.rs:3003-3100).
template <int KT>
struct Contract {
  static constexpr int kThreads = kBlock;
  uint64_t cells;
  const uint32_t *__restrict__ cell_worker;
  const uint32_t *__restrict__ cell_firm;
  const double *__restrict__ weight;
  const double *__restrict__ worker;
  const double *__restrict__ firm;
  double *__restrict__ partial;
  __device__ void operator()(Grid g) const {
    constexpr int K2 = 2 * KT;
    double acc[3] = {0.0, 0.0, 0.0};
    VK_GRID_STRIDE(g, t, cells * KT) {
      const uint64_t c = t / KT;
      const int j = t % KT;
      const uint64_t w = cell_worker[c], f = cell_firm[c];
      const double pw = worker[w * K2 + 2 * j] + firm[f * K2 + 2 * j];
      const double pf = worker[w * K2 + 2 * j + 1] + firm[f * K2 + 2 * j + 1];
      const double wt = weight[c];
      acc[0] += wt * pw * pw;
      acc[1] += wt * pf * pf;
      acc[2] += wt * pw * pf;
    }
    block_partial<KT, 3>(acc, partial, g.block);
  }
};
template <int KT>
__global__ void k_contract(Contract<KT> f) {
  f(Grid::hw());
}
struct FinDraws {
  __device__ bool enabled() const { return true; }
  double *__restrict__ draws;
  __device__ void reduced(int j, const double (&tot)[3]) const {
    double *d = draws + 4 * j;
    d[0] = tot[0];
    d[1] = tot[1];
    d[2] = tot[2];
    d[3] = tot[0] + tot[1] + 2.0 * tot[2];
  }
};

// Harmonic lift of auxiliary vertices for a warm start (cmg_impl.rs.in:459-500):
// x at the renumbered position of aux vertex a = weighted mean of its firms'
// normalised values `firm`.
template <int K>
__global__ void k_lift(uint32_t firms, uint32_t aux, const uint64_t *__restrict__ aptr,
                       const uint32_t *__restrict__ anbr, const double *__restrict__ aw,
                       const uint32_t *__restrict__ pos, const double *__restrict__ firm,
                       double *__restrict__ x) {
  VK_GRID_LOOP(t, (uint64_t)aux * K) {
    const uint32_t a = t / K;
    const int j = t % K;
    double num = 0.0, den = 0.0;
    for (uint64_t k = aptr[a]; k < aptr[a + 1]; ++k) {
      num += aw[k] * firm[(uint64_t)anbr[k] * K + j];
      den += aw[k];
    }
    x[(uint64_t)pos[firms + a] * K + j] = num / den;
  }
}

// Hybrid right-hand side in the renumbered order: firm entries from the
// centred Schur right-hand side, auxiliary entries zero (full_cmg.rs:842-895).
template <int K>
struct HybridRhs {
  static constexpr int kThreads = kBlock;
  uint64_t rows;
  uint32_t firms;
  const uint32_t *__restrict__ perm;
  const double *__restrict__ sch;
  double *__restrict__ b;
  __device__ void operator()(Grid g) const {
    VK_GRID_STRIDE(g, t, rows * K) {
      const uint32_t v = perm[t / K];
      b[t] = v < firms ? sch[(uint64_t)v * K + t % K] : 0.0;
    }
  }
};
template <int K>
__global__ void k_hybrid_rhs(HybridRhs<K> f) {
  f(Grid::hw());
}

// out = in over `count` doubles (a device copy that can join a multi-sample
// launch).
struct CopyDoubles {
  static constexpr int kThreads = kBlock;
  uint64_t count;
  const double *__restrict__ in;
  double *__restrict__ out;
  __device__ void operator()(Grid g) const {
    VK_GRID_STRIDE(g, t, count) out[t] = in[t];
  }
};
static __global__ void k_copy_doubles(CopyDoubles f) { f(Grid::hw()); }

}  // namespace dev
}  // namespace vk
