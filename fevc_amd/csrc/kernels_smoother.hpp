// Learned CMG smoother weights: the message-passing network (an input layer
// and three message-passing layers of width 64) that predicts a per-node
// Jacobi weight omega_i in (0, 0.95) for every smoothed level, its backward
// pass, and the training V-cycle with its exact adjoint.
//
// No counterpart in the reference: the reference cycle smooths with the fixed
// inverse diagonal 0.5 / d_i (vendor/cmg/src/preconditioner/cycle.rs:136-311).
// The learned cycle uses omega_i / d_i in the same places.  Every omega_i < 1
// keeps 2 D / omega - A positive definite, so the symmetric V-cycle remains a
// valid PCG preconditioner.
//
// Layouts: node-major rows of H = 64 hidden channels, h[i * 64 + c]; the 8
// input features X[i * 8 + f]; training vectors interleave 32 error columns,
// v[i * 32 + lane].  Training runs in FP64 (its gradient is compared with the
// CPU oracle); the solve path's inference runs in BF16 with FP32 accumulation
// on the matrix cores (k_mpnn_layer_bf16).
#pragma once

#include "kernels.hpp"

namespace vk {
namespace dev {
namespace mpnn {

constexpr int F = 8;
constexpr int H = 64;
constexpr int LAYERS = 3;  // message-passing layers after the input layer
constexpr double kOmegaMax = 0.95;
// flat parameter layout: W0[H][F], b0[H], then per layer t = 1..LAYERS
// U_t[H][H], V_t[H][H], c_t[H]; then a[H], a0
constexpr int kW0 = 0, kB0 = kW0 + H * F, kLayer0 = kB0 + H, kLayerSize = 2 * H * H + H;
constexpr int kU(int t) { return kLayer0 + (t - 1) * kLayerSize; }
constexpr int kV(int t) { return kU(t) + H * H; }
constexpr int kC(int t) { return kV(t) + H * H; }
constexpr int kA = kLayer0 + LAYERS * kLayerSize, kA0 = kA + H;
constexpr int P = kA0 + 1;
constexpr int kNodesPerBlock = kBlock / H;  // 4

}  // namespace mpnn

// h0 = relu(W0 x + b0)
static __global__ void __launch_bounds__(kBlock) k_mpnn_in(uint32_t n, const double *__restrict__ X,
                                                    const double *__restrict__ th, double *__restrict__ h0) {
  using namespace mpnn;
  VK_GRID_LOOP(t, (uint64_t)n * H) {
    const uint64_t i = t / H;
    const int c = t % H;
    double s = th[kB0 + c];
#pragma unroll
    for (int f = 0; f < F; ++f) s += th[kW0 + c * F + f] * X[i * F + f];
    h0[t] = s > 0.0 ? s : 0.0;
  }
}

// h = relu(U hp + V m + c)
static __global__ void __launch_bounds__(kBlock) k_mpnn_layer(uint32_t n, const double *__restrict__ hp,
                                                       const double *__restrict__ m, const double *__restrict__ th,
                                                       int u_off, int v_off, int c_off, double *__restrict__ h) {
  using namespace mpnn;
  VK_GRID_LOOP(t, (uint64_t)n * H) {
    const uint64_t i = t / H;
    const int c = t % H;
    double s = th[c_off + c];
    const double *hr = hp + i * H;
    const double *mr = m + i * H;
#pragma unroll 8
    for (int k = 0; k < H; ++k) s += th[u_off + c * H + k] * hr[k] + th[v_off + c * H + k] * mr[k];
    h[t] = s > 0.0 ? s : 0.0;
  }
}

// omega = 0.95 sigmoid(a . h_L + a0); the smoother inverse diagonal is
// omega / d = 2 omega inv_plain (inv_plain = 0.5 / d).
static __global__ void __launch_bounds__(kBlock) k_mpnn_out(uint32_t n, const double *__restrict__ hl,
                                                     const double *__restrict__ th,
                                                     const double *__restrict__ inv_plain, double *__restrict__ omega,
                                                     double *__restrict__ inv64, float *__restrict__ inv32) {
  using namespace mpnn;
  VK_GRID_LOOP(i, (uint64_t)n) {
    double z = th[kA0];
    for (int c = 0; c < H; ++c) z += th[kA + c] * hl[i * H + c];
    const double w = kOmegaMax / (1.0 + exp(-z));
    if (omega) omega[i] = w;
    const double v = 2.0 * w * inv_plain[i];
    if (inv64) inv64[i] = v;
    if (inv32) inv32[i] = (float)v;
  }
}

// ------------------------------------------------ BF16 inference (solve path)
// The solve path evaluates the network with BF16 activations and weights and
// FP32 accumulation: the input layer and messages on the vector units, the
// 64 x 128 layer products on the matrix cores (v_mfma_f32_16x16x16_bf16).
// BF16 values are the upper half of an FP32, rounded to nearest even.
__device__ inline uint16_t bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf16_value(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Layer weights of every message-passing layer as BF16 rows
// wcat[t - 1][c][k] = U_t[c][k] (k < H) and V_t[c][k - H] (k >= H).
static __global__ void __launch_bounds__(kBlock) k_mpnn_weights_bf16(const double *__restrict__ th,
                                                                     uint16_t *__restrict__ wcat) {
  using namespace mpnn;
  VK_GRID_LOOP(e, (uint64_t)LAYERS * H * 2 * H) {
    const int t = (int)(e / (H * 2 * H)) + 1, c = (int)(e / (2 * H) % H), k = (int)(e % (2 * H));
    const double w = k < H ? th[kU(t) + c * H + k] : th[kV(t) + c * H + (k - H)];
    wcat[e] = bf16_bits((float)w);
  }
}

// h0 = relu(W0 x + b0) in FP32, stored as BF16.
static __global__ void __launch_bounds__(kBlock) k_mpnn_in_bf16(uint32_t n, const double *__restrict__ X,
                                                                const double *__restrict__ th,
                                                                uint16_t *__restrict__ h0) {
  using namespace mpnn;
  VK_GRID_LOOP(t, (uint64_t)n * H) {
    const uint64_t i = t / H;
    const int c = t % H;
    float s = (float)th[kB0 + c];
#pragma unroll
    for (int f = 0; f < F; ++f) s += (float)th[kW0 + c * F + f] * (float)X[i * F + f];
    h0[t] = bf16_bits(s > 0.0f ? s : 0.0f);
  }
}

// Messages m_i = rs_i sum_k w_ik h_k over the level's chunked CSR rows (64
// channels per row, FP32 sums in entry order), stored as BF16.
template <int K>
struct OpMessageBf16 {
  const uint32_t *__restrict__ col;
  const double *__restrict__ val;
  const uint16_t *__restrict__ h;
  const double *__restrict__ rs;
  uint16_t *__restrict__ m;
  template <int STEP = Shape<K>::V>
  __device__ float partial(const uint4 &d, int j, int s) const {
    float acc = 0.0f;
    for (uint32_t k = d.y + s; k < d.z; k += STEP) acc += (float)val[k] * bf16_value(h[(uint64_t)col[k] * K + j]);
    return acc;
  }
  __device__ bool enabled() const { return true; }
  template <int NA>
  __device__ void finish(uint32_t i, int j, float sum, double (&)[NA]) const {
    m[(uint64_t)i * K + j] = bf16_bits((float)rs[i] * sum);
  }
  template <int NV>
  __device__ void reduced(int, const double (&)[NV]) const {}
};

// One message-passing layer on the matrix cores: out = relu([h m] W_t^T + c_t)
// for a 16-node tile per wavefront (four tiles per block).  The MFMA operand
// layouts of v_mfma_f32_16x16x16_bf16: lane l holds A[l % 16][4 (l / 16) + i]
// and B[4 (l / 16) + i][l % 16], i = 0..3, and receives D[4 (l / 16) + i][l % 16].
// The layer output is BF16 (out16) or, for the last layer, FP32 (out32).
typedef short mfma_bf16x4 __attribute__((ext_vector_type(4)));
typedef float mfma_f32x4 __attribute__((ext_vector_type(4)));
static __global__ void __launch_bounds__(kBlock) k_mpnn_layer_bf16(uint32_t n, const uint16_t *__restrict__ h,
                                                                   const uint16_t *__restrict__ m,
                                                                   const uint16_t *__restrict__ wcat,
                                                                   const double *__restrict__ th, int c_off,
                                                                   uint16_t *__restrict__ out16,
                                                                   float *__restrict__ out32) {
  using namespace mpnn;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lo = lane & 15, quad = lane >> 4;
  const uint64_t tiles = ((uint64_t)n + 15) / 16;
  for (uint64_t tile = (uint64_t)blockIdx.x * (kBlock / 64) + wave; tile < tiles;
       tile += (uint64_t)gridDim.x * (kBlock / 64)) {
    const uint64_t row = tile * 16 + lo;
    mfma_f32x4 acc[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[cb] = mfma_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2 * H / 16; ++kb) {
      const int k = kb * 16 + quad * 4;  // the lane's four K indices, k .. k + 3
      mfma_bf16x4 a = {0, 0, 0, 0};
      if (row < n) {
        const uint16_t *src = k < H ? h + row * H + k : m + row * H + (k - H);
        a = *reinterpret_cast<const mfma_bf16x4 *>(src);
      }
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        const mfma_bf16x4 b = *reinterpret_cast<const mfma_bf16x4 *>(wcat + (uint64_t)(cb * 16 + lo) * 2 * H + k);
        acc[cb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[cb], 0, 0, 0);
      }
    }
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const int c = cb * 16 + lo;
      const float bias = (float)th[c_off + c];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint64_t node = tile * 16 + quad * 4 + i;
        if (node >= n) continue;
        float v = acc[cb][i] + bias;
        v = v > 0.0f ? v : 0.0f;
        if (out32) out32[node * H + c] = v;
        else out16[node * H + c] = bf16_bits(v);
      }
    }
  }
}

// omega = 0.95 sigmoid(a . h_L + a0) from the FP32 last layer (FP32 sums).
static __global__ void __launch_bounds__(kBlock) k_mpnn_out_f32(uint32_t n, const float *__restrict__ hl,
                                                                const double *__restrict__ th,
                                                                const double *__restrict__ inv_plain,
                                                                double *__restrict__ omega, double *__restrict__ inv64,
                                                                float *__restrict__ inv32) {
  using namespace mpnn;
  VK_GRID_LOOP(i, (uint64_t)n) {
    float z = (float)th[kA0];
    for (int c = 0; c < H; ++c) z += (float)th[kA + c] * hl[i * H + c];
    const double w = kOmegaMax / (1.0 + exp(-(double)z));
    if (omega) omega[i] = w;
    const double v = 2.0 * w * inv_plain[i];
    if (inv64) inv64[i] = v;
    if (inv32) inv32[i] = (float)v;
  }
}

// Constant smoother weight omega on every node (host-forced fallback tests):
// the inverse diagonals omega / d from the plain 0.5 / d.
static __global__ void __launch_bounds__(kBlock) k_constant_omega(uint32_t n, double omega,
                                                           const double *__restrict__ inv_plain,
                                                           double *__restrict__ inv64, float *__restrict__ inv32) {
  VK_GRID_LOOP(i, (uint64_t)n) {
    const double v = 2.0 * omega * inv_plain[i];
    inv64[i] = v;
    inv32[i] = (float)v;
  }
}

// Output layer backward: dz = g_omega * 0.95 s (1 - s), dh_L = dz a.
static __global__ void __launch_bounds__(kBlock) k_mpnn_out_back(uint32_t n, const double *__restrict__ omega,
                                                          const double *__restrict__ g_omega,
                                                          const double *__restrict__ th, double *__restrict__ dz,
                                                          double *__restrict__ dhl) {
  using namespace mpnn;
  VK_GRID_LOOP(t, (uint64_t)n * H) {
    const uint64_t i = t / H;
    const int c = t % H;
    const double s = omega[i] / kOmegaMax;
    const double d = g_omega[i] * kOmegaMax * s * (1.0 - s);
    dhl[t] = d * th[kA + c];
    if (c == 0) dz[i] = d;
  }
}

// Hidden layer backward for nodes in groups of kNodesPerBlock per block:
//   dpre = dh * [h > 0], dhp = U^T dpre, dms = (V^T dpre) / s_i
// (dms is then scattered through the adjacency: dhp += W dms).
static __global__ void __launch_bounds__(kBlock) k_mpnn_layer_back(uint32_t n, const double *__restrict__ dh,
                                                            const double *__restrict__ h,
                                                            const double *__restrict__ th, int u_off, int v_off,
                                                            const double *__restrict__ rs, double *__restrict__ dpre,
                                                            double *__restrict__ dhp, double *__restrict__ dms) {
  using namespace mpnn;
  __shared__ double lds[kBlock];
  const int local = threadIdx.x / H, k = threadIdx.x % H;
  for (uint64_t base = (uint64_t)blockIdx.x * kNodesPerBlock; base < n;
       base += (uint64_t)gridDim.x * kNodesPerBlock) {
    const uint64_t i = base + local;
    const bool live = i < n;
    double d = 0.0;
    if (live) {
      const uint64_t q = i * H + k;
      d = h[q] > 0.0 ? dh[q] : 0.0;
      dpre[q] = d;
    }
    lds[threadIdx.x] = d;
    __syncthreads();
    if (live) {
      double a = 0.0, b = 0.0;
      const double *row = lds + local * H;
#pragma unroll 8
      for (int c = 0; c < H; ++c) {
        a += th[u_off + c * H + k] * row[c];
        b += th[v_off + c * H + k] * row[c];
      }
      dhp[i * H + k] = a;
      dms[i * H + k] = b * rs[i];
    }
    __syncthreads();
  }
}

// dpre0 = dh0 * [h0 > 0]
static __global__ void __launch_bounds__(kBlock) k_relu_back(uint64_t count, const double *__restrict__ dh,
                                                      const double *__restrict__ h, double *__restrict__ dpre) {
  VK_GRID_LOOP(t, count) dpre[t] = h[t] > 0.0 ? dh[t] : 0.0;
}

// Per-block partial of G[c][k] = sum_i A[i][c] B[i][k] (c < KA, k < KB) and,
// with bias, of sum_i A[i][c]; the block's nodes are the contiguous range
// [blockIdx.x * n / gridDim.x, ...).  Partials are summed by k_partial_sum in
// block order.
template <int KA, int KB, bool Bias>
__global__ void __launch_bounds__(kBlock) k_outer_partial(uint32_t n, const double *__restrict__ A,
                                                          const double *__restrict__ B,
                                                          double *__restrict__ partial) {
  constexpr int E = KA * KB + (Bias ? KA : 0);
  constexpr int PER = (E + kBlock - 1) / kBlock;
  constexpr int TILE = 8;
  __shared__ double la[TILE * KA], lb[TILE * KB];
  double acc[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) acc[e] = 0.0;
  const uint64_t lo = (uint64_t)n * blockIdx.x / gridDim.x, hi = (uint64_t)n * (blockIdx.x + 1) / gridDim.x;
  for (uint64_t base = lo; base < hi; base += TILE) {
    const int rows = (int)(hi - base < (uint64_t)TILE ? hi - base : (uint64_t)TILE);
    for (int t = threadIdx.x; t < TILE * KA; t += kBlock) la[t] = t / KA < rows ? A[base * KA + t] : 0.0;
    for (int t = threadIdx.x; t < TILE * KB; t += kBlock) lb[t] = t / KB < rows ? B[base * KB + t] : 0.0;
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int idx = threadIdx.x + e * kBlock;
      if (idx < KA * KB) {
        const int c = idx / KB, k = idx % KB;
        for (int r = 0; r < TILE; ++r) acc[e] += la[r * KA + c] * lb[r * KB + k];
      } else if (Bias && idx < E) {
        const int c = idx - KA * KB;
        for (int r = 0; r < TILE; ++r) acc[e] += la[r * KA + c];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = threadIdx.x + e * kBlock;
    if (idx < E) partial[(uint64_t)blockIdx.x * E + idx] = acc[e];
  }
}

// out[e] += sum_b partial[b][e] in block order (deterministic).
static __global__ void __launch_bounds__(kBlock) k_partial_sum(const double *__restrict__ partial, uint32_t blocks,
                                                        uint32_t count, double *__restrict__ out) {
  VK_GRID_LOOP(e, (uint64_t)count) {
    double s = 0.0;
    for (uint32_t b = 0; b < blocks; ++b) s += partial[(uint64_t)b * count + e];
    out[e] += s;
  }
}

// Adam (Kingma & Ba 2015) with bias correction.
static __global__ void __launch_bounds__(kBlock) k_adam(uint32_t count, double *__restrict__ th,
                                                 const double *__restrict__ g, double *__restrict__ m,
                                                 double *__restrict__ v, double lr, double b1, double b2, double eps,
                                                 double c1, double c2) {
  VK_GRID_LOOP(t, (uint64_t)count) {
    m[t] = b1 * m[t] + (1.0 - b1) * g[t];
    v[t] = b2 * v[t] + (1.0 - b2) * g[t] * g[t];
    th[t] -= lr * (m[t] / c1) / (sqrt(v[t] / c2) + eps);
  }
}

// ---------------------------------------------------------------- messages

// out[i] = scale_i * sum_k val_k x[col_k]  (accumulate: out[i] += ...), K
// lanes per row: the adjacency product of the message passing.
template <int K>
struct OpAdjacency {
  const uint32_t *__restrict__ col;
  const double *__restrict__ val;
  const double *__restrict__ x;
  const double *__restrict__ scale;
  double *__restrict__ out;
  bool accumulate;
  template <int STEP = Shape<K>::V>
  __device__ double partial(const uint4 &d, int j, int s) const {
    double acc = 0.0;
    for (uint32_t k = d.y + s; k < d.z; k += STEP) acc += val[k] * x[(uint64_t)col[k] * K + j];
    return acc;
  }
  __device__ bool enabled() const { return true; }
  template <int NA>
  __device__ void finish(uint32_t i, int j, double sum, double (&)[NA]) const {
    const uint64_t q = (uint64_t)i * K + j;
    const double v = scale ? scale[i] * sum : sum;
    out[q] = accumulate ? out[q] + v : v;
  }
  template <int NV>
  __device__ void reduced(int, const double (&)[NV]) const {}
};

// ------------------------------------------------------ training V-cycle

// y = w (b + s) + e[lab]: post-smoothing output (x1 = w b, y = x1 + P e + w s)
template <int K>
__global__ void __launch_bounds__(kBlock) k_train_post(uint32_t n, const double *__restrict__ w,
                                                       const double *__restrict__ b, const double *__restrict__ s,
                                                       const double *__restrict__ e, const uint32_t *__restrict__ lab,
                                                       double *__restrict__ y) {
  VK_GRID_LOOP(t, (uint64_t)n * K) {
    const uint64_t i = t / K;
    const int j = t % K;
    y[t] = w[i] * b[t] + e[(uint64_t)lab[i] * K + j] + w[i] * s[t];
  }
}

// x = w * b (row scaling)
template <int K>
__global__ void __launch_bounds__(kBlock) k_row_scale(uint32_t n, const double *__restrict__ w,
                                                      const double *__restrict__ b, double *__restrict__ x) {
  VK_GRID_LOOP(t, (uint64_t)n * K) x[t] = w[t / K] * b[t];
}

// out = add + e[lab] (prolongation, optionally added to `add`)
template <int K>
__global__ void __launch_bounds__(kBlock) k_prolong(uint32_t n, const uint32_t *__restrict__ lab,
                                                    const double *__restrict__ e, const double *__restrict__ add,
                                                    double *__restrict__ out) {
  VK_GRID_LOOP(t, (uint64_t)n * K) {
    const uint64_t i = t / K;
    const int j = t % K;
    const double v = e[(uint64_t)lab[i] * K + j];
    out[t] = add ? add[t] + v : v;
  }
}

// out = v - mean (per lane); mean may be null (copy)
template <int K>
__global__ void __launch_bounds__(kBlock) k_lane_centre(uint32_t n, const double *__restrict__ v,
                                                        const double *__restrict__ mean, double *__restrict__ out) {
  VK_GRID_LOOP(t, (uint64_t)n * K) out[t] = mean ? v[t] - mean[t % K] : v[t];
}

// out = G^T v for the dense terminal (the adjoint of k_terminal_direct)
template <int K>
__global__ void __launch_bounds__(kBlock) k_dense_transpose(uint32_t n, const double *__restrict__ g,
                                                            const double *__restrict__ v, double *__restrict__ out) {
  VK_GRID_LOOP(t, (uint64_t)n * K) {
    const uint64_t i = t / K;
    const int j = t % K;
    double s = 0.0;
    for (uint32_t k = 0; k < n; ++k) s += g[(uint64_t)k * n + i] * v[(uint64_t)k * K + j];
    out[t] = s;
  }
}

// Adjoint of the smoothing steps of one level (one thread per node):
//   g_omega_i = (sum_j ly s + lx1 b) / d_i = 2 inv_plain_i * (...)
//   lb = lr1 + w lx1
template <int K>
__global__ void __launch_bounds__(kBlock) k_train_smooth_adjoint(
    uint32_t n, const double *__restrict__ w, const double *__restrict__ inv_plain, const double *__restrict__ ly,
    const double *__restrict__ s, const double *__restrict__ lx1, const double *__restrict__ b,
    const double *__restrict__ lr1, double *__restrict__ lb, double *__restrict__ g_omega) {
  VK_GRID_LOOP(i, (uint64_t)n) {
    double g = 0.0;
    for (int j = 0; j < K; ++j) {
      const uint64_t q = i * K + j;
      g += ly[q] * s[q];
    }
    double g2 = 0.0;
    for (int j = 0; j < K; ++j) {
      const uint64_t q = i * K + j;
      g2 += lx1[q] * b[q];
      lb[q] = lr1[q] + w[i] * lx1[q];
    }
    g_omega[i] = 2.0 * inv_plain[i] * (g + g2);
  }
}

// Per-lane dot product partials: block_partial of sum_i a[i][j] b[i][j].
template <int K>
__global__ void __launch_bounds__(kBlock) k_lane_dot(uint32_t n, const double *__restrict__ a,
                                                     const double *__restrict__ b, double *__restrict__ partial) {
  double acc[1] = {0.0};
  VK_GRID_LOOP(t, (uint64_t)n * K) acc[0] += a[t] * b[t];
  block_partial<K, 1>(acc, partial);
}

struct FinStore {
  double *__restrict__ out;
  __device__ bool enabled() const { return true; }
  __device__ void reduced(int j, const double (&tot)[1]) const { out[j] = tot[0]; }
};

// lam = -2 A u / (K q_j): the loss gradient with respect to the cycle output
template <int K>
__global__ void __launch_bounds__(kBlock) k_loss_seed(uint32_t n, const double *__restrict__ au,
                                                      const double *__restrict__ q, double *__restrict__ lam) {
  VK_GRID_LOOP(t, (uint64_t)n * K) lam[t] = -2.0 * au[t] / ((double)K * q[t % K]);
}

// e -= y (the error after one cycle: e' = e - M A e)
static __global__ void __launch_bounds__(kBlock) k_sub(uint64_t count, const double *__restrict__ y, double *__restrict__ e) {
  VK_GRID_LOOP(t, count) e[t] -= y[t];
}

// Uniform(-1, 1) training errors from a counter hash (SplitMix64 finaliser
// over seed and index).
static __global__ void __launch_bounds__(kBlock) k_random_errors(uint64_t count, uint64_t seed, double *__restrict__ e) {
  VK_GRID_LOOP(t, count) {
    uint64_t z = seed + 0x9e3779b97f4a7c15ULL * (t + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    e[t] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

}  // namespace dev
}  // namespace vk
