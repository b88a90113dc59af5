// HIP kernels of the leave-out engine (gfx950).
//
// Vectors of a batch of K right-hand sides are stored vertex-major and
// interleaved, v[row * K + lane], so one wavefront reads K contiguous lanes of
// a row and the CSR structure is read once per batch.  K is a compile-time
// power of two in [1, 64].  Reductions are deterministic: every block writes a
// per-lane partial and the last block to finish (atomic ticket) sums the
// partials in block order and runs the per-lane scalar logic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vk {
namespace dev {

constexpr int kBlock = 256;
constexpr int kMaxReduceBlocks = 1024;
constexpr int kMaxReduceValues = 4;

// ---------------------------------------------------------------- RNG
// Counter-V1 Philox4x64-10 with SplitMix64 keys (rng.rs:14-99, 309-327).
__device__ inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

__device__ inline void philox(uint64_t k0, uint64_t k1, uint64_t c[4]) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    uint64_t lo0 = 0xd2b74407b1ce6e93ULL * c[0], hi0 = __umul64hi(0xd2b74407b1ce6e93ULL, c[0]);
    uint64_t lo1 = 0xca5a826395121157ULL * c[2], hi1 = __umul64hi(0xca5a826395121157ULL, c[2]);
    uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    if (round != 9) {
      k0 += 0x9e3779b97f4a7c15ULL;
      k1 += 0xbb67ae8584caa73bULL;
    }
  }
}

// Rademacher atoms 2*popcount - n for `columns` probes starting at
// `first_probe`; out[col * n + row] as exact binary64 integers.
__global__ void k_atoms(uint64_t n, const uint64_t *__restrict__ entity,
                        const uint64_t *__restrict__ trials, uint64_t seed, uint64_t tag,
                        uint64_t first_probe, uint32_t columns, double *__restrict__ out) {
  const uint64_t first_block = first_probe / 4, last_block = (first_probe + columns - 1) / 4;
  const uint64_t blocks = last_block - first_block + 1;
  const uint64_t k0 = splitmix64(seed), k1 = splitmix64(seed ^ 0xd1b54a32d192ed03ULL);
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n * blocks;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t row = t % n, pb = first_block + t / n;
    const uint64_t tr = trials[row], complete = tr / 64, rem = tr % 64;
    uint64_t pos[4] = {0, 0, 0, 0};
    for (uint64_t w = 0; w <= complete; ++w) {
      if (w == complete && rem == 0) break;
      uint64_t c[4] = {entity[row], pb, w, tag};
      philox(k0, k1, c);
      const uint64_t mask = (w == complete) ? ((1ULL << rem) - 1) : ~0ULL;
#pragma unroll
      for (int l = 0; l < 4; ++l) pos[l] += __popcll(c[l] & mask);
    }
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const uint64_t probe = pb * 4 + l;
      if (probe < first_probe || probe >= first_probe + columns) continue;
      out[(probe - first_probe) * n + row] = (double)((int64_t)(2 * pos[l]) - (int64_t)tr);
    }
  }
}

// ---------------------------------------------------------------- reductions

// Block-reduce NV values per lane (K lanes, kBlock/K rows per block), write
// block partials and let the last block sum all partials in block order.
// Returns true in the last block; totals[v * K + lane] then holds the sums for
// threads with threadIdx.x < K.
template <int K, int NV>
__device__ inline bool reduce_lanes(double (&val)[NV], double *__restrict__ partial,
                                    unsigned *__restrict__ counter, double (&total)[NV]) {
  __shared__ double lds[NV][kBlock];
  __shared__ bool last;
  constexpr int R = kBlock / K;
  const int tid = threadIdx.x, row = tid / K;
#pragma unroll
  for (int v = 0; v < NV; ++v) lds[v][tid] = val[v];
  __syncthreads();
  for (int s = R / 2; s > 0; s >>= 1) {
    if (row < s) {
#pragma unroll
      for (int v = 0; v < NV; ++v) lds[v][tid] += lds[v][tid + s * K];
    }
    __syncthreads();
  }
  if (tid < K) {
#pragma unroll
    for (int v = 0; v < NV; ++v) partial[((uint64_t)blockIdx.x * NV + v) * K + tid] = lds[v][tid];
  }
  __threadfence();
  __syncthreads();
  if (tid == 0) last = atomicAdd(counter, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last) return false;
  __threadfence();
  const volatile double *vp = partial;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double s = 0.0;
    for (unsigned b = row; b < gridDim.x; b += R) s += vp[((uint64_t)b * NV + v) * K + (tid % K)];
    lds[v][tid] = s;
  }
  __syncthreads();
  for (int s = R / 2; s > 0; s >>= 1) {
    if (row < s) {
#pragma unroll
      for (int v = 0; v < NV; ++v) lds[v][tid] += lds[v][tid + s * K];
    }
    __syncthreads();
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) total[v] = lds[v][tid];
  if (tid == 0) atomicExch(counter, 0u);  // device-scope, visible to the next launch
  return true;
}

// ---------------------------------------------------------------- CSR rows
// y_i = sum_j w_ij (x_i - x_j), neighbours ascending: the same sum order as
// the reference's canonical edge scatter (vendor/cmg/src/graph.rs:664-705).
template <typename T, int K>
__device__ inline T row_laplacian(const uint32_t *__restrict__ rp, const uint32_t *__restrict__ col,
                                  const T *__restrict__ val, const T *__restrict__ x, uint32_t i,
                                  int j, T xi) {
  T acc = T(0);
  const uint32_t e = rp[i + 1];
  for (uint32_t k = rp[i]; k < e; ++k) acc += val[k] * (xi - x[(uint64_t)col[k] * K + j]);
  return acc;
}

// Same row with x replaced by x + P e (prolongation folded into the read).
template <typename T, int K>
__device__ inline T row_laplacian_prolonged(const uint32_t *__restrict__ rp,
                                            const uint32_t *__restrict__ col,
                                            const T *__restrict__ val, const T *__restrict__ x,
                                            const uint32_t *__restrict__ lab,
                                            const T *__restrict__ e, uint32_t i, int j, T xi) {
  T acc = T(0);
  const uint32_t end = rp[i + 1];
  for (uint32_t k = rp[i]; k < end; ++k) {
    const uint32_t c = col[k];
    acc += val[k] * (xi - (x[(uint64_t)c * K + j] + e[(uint64_t)lab[c] * K + j]));
  }
  return acc;
}

// ---------------------------------------------------------------- V-cycle
// cycle.rs:342-437: output = D^-1 rhs; a non-null mean first centres rhs in
// place (coarse component centring, cycle.rs:459-478).
template <typename T, int K>
__global__ void k_scale(uint32_t n, const T *__restrict__ inv, T *__restrict__ b,
                        T *__restrict__ x, const double *__restrict__ mean) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < (uint64_t)n * K;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = t / K;
    const int j = t % K;
    T bi = b[t];
    if (mean) {
      bi = bi - (T)mean[j];
      b[t] = bi;
    }
    x[t] = inv[i] * bi;
  }
}

// y = x + D^-1 (b - A x)   (cycle.rs:380-437 after the level matvec)
template <typename T, int K>
__global__ void k_smooth(uint32_t n, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ col,
                         const T *__restrict__ val, const T *__restrict__ inv, const T *__restrict__ b,
                         const T *__restrict__ x, T *__restrict__ y) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < (uint64_t)n * K;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = t / K;
    const int j = t % K;
    const T xi = x[t];
    const T ax = row_laplacian<T, K>(rp, col, val, x, i, j, xi);
    y[t] = xi + inv[i] * (b[t] - ax);
  }
}

// y = x' + D^-1 (b - A x') with x' = x + P e  (prolong-add then smoothing)
template <typename T, int K>
__global__ void k_prolong_smooth(uint32_t n, const uint32_t *__restrict__ rp,
                                 const uint32_t *__restrict__ col, const T *__restrict__ val,
                                 const T *__restrict__ inv, const uint32_t *__restrict__ lab,
                                 const T *__restrict__ e, const T *__restrict__ b,
                                 const T *__restrict__ x, T *__restrict__ y) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < (uint64_t)n * K;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = t / K;
    const int j = t % K;
    const T xi = x[t] + e[(uint64_t)lab[i] * K + j];
    const T ax = row_laplacian_prolonged<T, K>(rp, col, val, x, lab, e, i, j, xi);
    y[t] = xi + inv[i] * (b[t] - ax);
  }
}

// coarse[c] = sum over members v ascending of (b_v - (A x)_v)  (coarsen.rs:476-503)
// plus the coarse mean per lane for the centring of the next level.
template <typename T, int K>
__global__ void k_restrict(uint32_t nc, const uint32_t *__restrict__ aptr,
                           const uint32_t *__restrict__ amem, const uint32_t *__restrict__ rp,
                           const uint32_t *__restrict__ col, const T *__restrict__ val,
                           const T *__restrict__ b, const T *__restrict__ x, T *__restrict__ bc,
                           double *__restrict__ partial, unsigned *__restrict__ counter,
                           double *__restrict__ mean) {
  double acc[1] = {0.0};
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < (uint64_t)nc * K;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t c = t / K;
    const int j = t % K;
    T s = T(0);
    const uint32_t end = aptr[c + 1];
    for (uint32_t m = aptr[c]; m < end; ++m) {
      const uint32_t v = amem[m];
      const T xv = x[(uint64_t)v * K + j];
      s += b[(uint64_t)v * K + j] - row_laplacian<T, K>(rp, col, val, x, v, j, xv);
    }
    bc[t] = s;
    acc[0] += (double)s;
  }
  double tot[1];
  if (reduce_lanes<K, 1>(acc, partial, counter, tot) && threadIdx.x < K)
    mean[threadIdx.x] = tot[0] / (double)nc;
}

// Direct terminal: x = G (b - mean), G the grounded inverse (ldl.rs:892-997).
template <typename T, int K>
__global__ void k_terminal_direct(uint32_t n, const double *__restrict__ g, const T *__restrict__ b,
                                  const double *__restrict__ mean, T *__restrict__ x) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < (uint64_t)n * K;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = t / K;
    const int j = t % K;
    const double mu = mean ? mean[j] : 0.0;
    double s = 0.0;
    for (uint32_t k = 0; k < n; ++k) s += g[(uint64_t)i * n + k] * ((double)b[(uint64_t)k * K + j] - mu);
    x[t] = (T)s;
  }
}

// Plain level product y = A x (diagnostics and tests).
template <typename T, int K>
__global__ void k_spmv(uint32_t n, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ col,
                       const T *__restrict__ val, const T *__restrict__ x, T *__restrict__ y) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < (uint64_t)n * K;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = t / K;
    const int j = t % K;
    y[t] = row_laplacian<T, K>(rp, col, val, x, i, j, x[t]);
  }
}

// ---------------------------------------------------------------- precision casts
template <int K>
__global__ void k_to_float(uint64_t count, const double *__restrict__ a, float *__restrict__ b) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < count;
       t += (uint64_t)gridDim.x * blockDim.x)
    b[t] = (float)a[t];
}
template <int K>
__global__ void k_to_double(uint64_t count, const float *__restrict__ a, double *__restrict__ b) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < count;
       t += (uint64_t)gridDim.x * blockDim.x)
    b[t] = (double)a[t];
}

}  // namespace dev
}  // namespace vk
