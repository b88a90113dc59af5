// HIP kernels of the leave-out engine (gfx950): RNG, deterministic
// reductions, and the chunked segment kernel behind every sparse pass of the
// CMG V-cycle and of PCG.
//
// Vectors of a batch of K right-hand sides are stored vertex-major and
// interleaved, v[row * K + lane], so K neighbouring threads read one row's K
// lanes in one contiguous access and the CSR structure is read once per
// batch.  K is a compile-time power of two in [1, 64].
//
// Sparse passes walk "segments": CSR rows (y_i = sum_j w_ij (x_i - x_j)) or
// aggregates (coarse_c = sum of the contiguous members' residuals).  The host
// cuts segments into chunks of at most 64 entries.  A chunk is processed by a
// group of G = K * V threads, V slices per lane striding over its entries and
// combined by a fixed butterfly.  A segment with several chunks stores chunk
// partials that a second, small kernel sums in chunk order; the kernel
// boundary publishes them, so the hot pass needs no device-scope fence (on
// gfx950 such a fence writes back and invalidates the L2 of the XCD).
// Every order is fixed, so results do not depend on scheduling.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include <cstdint>

namespace vk {
namespace dev {

// Build-time tuning knobs (measured defaults): unroll of the CSR walk, grid
// cap of streaming launches, short segments per grouped chunk.
#ifndef VK_SEG_UNROLL
#define VK_SEG_UNROLL 4
#endif
// the same for the FP64 lanes of the Krylov passes (level-0 products)
#ifndef VK_SEG_UNROLL_F64
#define VK_SEG_UNROLL_F64 VK_SEG_UNROLL
#endif
#ifndef VK_GRID_CAP
#define VK_GRID_CAP 8192
#endif
#ifndef VK_SEG_GROUP
#define VK_SEG_GROUP 4
#endif

constexpr int kBlock = 256;
constexpr int kMaxReduceBlocks = 4096;
// block cap of the reductions of the PCG iteration (update, rho, direction
// and A p, fresh residual): fewer, longer blocks measured faster there (the
// finishes read fewer partials), while the per-batch passes keep the larger cap
constexpr int kKrylovReduceBlocks = 2048;
constexpr int kMaxReduceValues = 6;
// Block partials: value v of lane j from block slot b sits at
// [(v * K + j) * kPartialStride + b], so the final per-lane sum reads
// consecutive slots (pass-1 blocks, then combine blocks).
constexpr uint32_t kPartialStride = 2 * kMaxReduceBlocks;
constexpr uint32_t kSingle = 0xffffffffu;
constexpr uint32_t kGroupBase = 0xffffff00u;

// Block coordinates of a pass: the hardware grid of a plain launch, or one
// sample's share of a multi-sample launch (k_multi).  Kernels that can run
// inside a multi-sample launch are written as functors over a Grid; the
// plain launch is a named kernel that calls the functor with Grid::hw().
struct Grid {
  uint32_t block, blocks;
  __device__ static Grid hw() { return {blockIdx.x, gridDim.x}; }
};
#define VK_GRID_STRIDE(g, t, count)                                                     \
  for (uint64_t t = (g).block * (uint64_t)blockDim.x + threadIdx.x; t < (count);       \
       t += (uint64_t)(g).blocks * blockDim.x)

// One launch over several independent samples' instances of one pass: sample
// s runs its own functor fs[s] on blocks [ends[s - 1], ends[s]) (ends[-1] =
// 0), as that functor would run alone on a grid of ends[s] - ends[s - 1]
// blocks.  At most 64 samples per launch: one wavefront reads the block ends
// and the sample is the number of ends at or below the block (wave-uniform).
constexpr uint32_t kMultiMax = 64;
template <class F>
__global__ void __launch_bounds__(kBlock) k_multi(const F *__restrict__ fs, const uint32_t *__restrict__ ends,
                                               uint32_t count) {
  const uint32_t b = blockIdx.x, lane = threadIdx.x & 63u;
  const uint32_t e = lane < count ? ends[lane] : 0xffffffffu;
  const uint32_t s = (uint32_t)__popcll(__ballot(e <= b));
  const uint32_t begin = s ? (uint32_t)__builtin_amdgcn_readlane((int)e, (int)s - 1) : 0u;
  const uint32_t end = (uint32_t)__builtin_amdgcn_readlane((int)e, (int)s);
  const F f = fs[s];
  f(Grid{b - begin, end - begin});
}

// Threads per chunk: V slices of K lanes, G = K * V <= 64 (one wavefront).
template <int K>
struct Shape {
  static constexpr int V = K >= 64 ? 1 : (K >= 4 ? 64 / K : 16);
  static constexpr int G = K * V;
};

// Lane vectors: a segment op runs on K "lanes" of type T, where T may pack
// four FP32 probe columns (float4, 16-byte loads); values that are one per
// row (matrix values, inverse diagonals) stay scalar.
template <typename T>
struct ScalarOf {
  using type = T;
};
template <>
struct ScalarOf<float4> {
  using type = float;
};
template <typename T>
using scalar_t = typename ScalarOf<T>::type;

__device__ inline float shfl_xor(float v, int off) { return __shfl_xor(v, off); }
__device__ inline double shfl_xor(double v, int off) { return __shfl_xor(v, off); }
__device__ inline float4 shfl_xor(float4 v, int off) {
  return make_float4(__shfl_xor(v.x, off), __shfl_xor(v.y, off), __shfl_xor(v.z, off), __shfl_xor(v.w, off));
}

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
// `first_probe`, as exact integers, interleaved: out[row * ld + col].  The
// probe block varies fastest across threads so that stores are contiguous.
// Row r sums the signs of copies [offset[r], offset[r] + trials[r]) of its
// entity, copy j being bit j % 64 of Counter-V1 word j / 64 (rng.rs:121-179;
// generic_jla.rs:10321-10329 with a copy offset); offset may be null (all
// zero).  Callers guarantee trials < 2^31.
struct Atoms {
  static constexpr int kThreads = kBlock;
  uint64_t n;
  const uint64_t *__restrict__ entity;
  const uint64_t *__restrict__ trials;
  const uint64_t *__restrict__ offset;
  uint64_t seed;
  uint64_t tag;
  uint64_t first_probe;
  uint32_t columns;
  uint32_t ld;
  int32_t *__restrict__ out;
  __device__ void operator()(Grid g) const {
    const uint64_t first_block = first_probe / 4, last_block = (first_probe + columns - 1) / 4;
    const uint64_t blocks = last_block - first_block + 1;
    const uint64_t k0 = splitmix64(seed), k1 = splitmix64(seed ^ 0xd1b54a32d192ed03ULL);
    for (uint64_t t = g.block * (uint64_t)blockDim.x + threadIdx.x; t < n * blocks;
         t += (uint64_t)g.blocks * blockDim.x) {
      const uint64_t row = t / blocks, pb = first_block + t % blocks;
      const uint64_t tr = trials[row], lo = offset ? offset[row] : 0, hi = lo + tr;
      uint64_t pos[4] = {0, 0, 0, 0};
      for (uint64_t w = lo / 64; w * 64 < hi; ++w) {
        const uint64_t b0 = lo > w * 64 ? lo - w * 64 : 0, b1 = hi - w * 64 < 64 ? hi - w * 64 : 64;
        const uint64_t mask = (b1 == 64 ? ~0ULL : ((1ULL << b1) - 1)) & ~((1ULL << b0) - 1);
        uint64_t c[4] = {entity[row], pb, w, tag};
        philox(k0, k1, c);
  #pragma unroll
        for (int l = 0; l < 4; ++l) pos[l] += __popcll(c[l] & mask);
      }
  #pragma unroll
      for (int l = 0; l < 4; ++l) {
        const uint64_t probe = pb * 4 + l;
        if (probe < first_probe || probe >= first_probe + columns) continue;
        out[row * ld + (probe - first_probe)] = (int32_t)((int64_t)(2 * pos[l]) - (int64_t)tr);
      }
    }
  }
};
static __global__ void k_atoms(Atoms f) {
  f(Grid::hw());
}

// ---------------------------------------------------------------- reductions
// Two-kernel deterministic reductions without device-scope fences (on
// gfx950 a device-scope fence writes back and invalidates the L2 of the
// XCD): every block of a pass stores one partial per lane (block_partial),
// then k_finish runs one block per lane that sums the partials of all blocks
// in a fixed order and hands the totals to Fin::reduced(lane, totals).

template <int K, int NV>
__device__ inline void block_partial(const double (&val)[NV], double *__restrict__ out, uint32_t slot) {
  __shared__ double lds[NV][kBlock];
  constexpr int R = kBlock / K;
  const int tid = threadIdx.x, row = tid / K;
#pragma unroll
  for (int v = 0; v < NV; ++v) lds[v][tid] = val[v];
  __syncthreads();
  for (int w = R / 2; w > 0; w >>= 1) {
    if (row < w) {
#pragma unroll
      for (int v = 0; v < NV; ++v) lds[v][tid] += lds[v][tid + w * K];
    }
    __syncthreads();
  }
  if (tid < K) {
#pragma unroll
    for (int v = 0; v < NV; ++v) out[((uint64_t)v * K + tid) * kPartialStride + slot] = lds[v][tid];
  }
}
template <int K, int NV>
__device__ inline void block_partial(const double (&val)[NV], double *__restrict__ out) {
  block_partial<K, NV>(val, out, blockIdx.x);
}

// Lane g.block of `items` block partials: thread t adds items t, t + kBlock,
// ... in order, then a fixed LDS tree; thread 0 calls fin.reduced.
template <int K, int NV, class Fin>
struct Finish {
  static constexpr int kThreads = kBlock;
  const double *__restrict__ partial;
  uint32_t items;
  Fin fin;
  const int *__restrict__ gate;
  __device__ void operator()(Grid g) const {
    if (gate && *gate == 0) return;
    if (!fin.enabled()) return;
    __shared__ double lds[NV][kBlock];
    const int j = g.block, tid = threadIdx.x;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double s = 0.0;
      const double *row = partial + ((uint64_t)v * K + j) * kPartialStride;
      uint32_t b = tid;
      for (; b + 3 * kBlock < items; b += 4 * kBlock) {
        const double a0 = row[b];
        const double a1 = row[b + kBlock];
        const double a2 = row[b + 2 * kBlock];
        const double a3 = row[b + 3 * kBlock];
        s += a0;
        s += a1;
        s += a2;
        s += a3;
      }
      for (; b < items; b += kBlock) s += row[b];
      lds[v][tid] = s;
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
      if (tid < w) {
#pragma unroll
        for (int v = 0; v < NV; ++v) lds[v][tid] += lds[v][tid + w];
      }
      __syncthreads();
    }
    if (tid == 0) {
      double tot[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) tot[v] = lds[v][0];
      fin.reduced(j, tot);
    }
  }
};
template <int K, int NV, class Fin>
__global__ void __launch_bounds__(kBlock) k_finish(Finish<K, NV, Fin> f) {
  f(Grid::hw());
}

// ---------------------------------------------------------------- segments

// A chunk descriptor is (segment, begin, end, tag): tag kSingle for a whole
// segment, a multi index for one chunk of a long segment, or kGroupBase + c
// for a group of c consecutive short segments starting at `segment` (their
// entry ranges come from the segment pointers `ptr`).
struct Chunks {
  const uint4 *__restrict__ desc;
  const uint32_t *__restrict__ mrange;  // first and end chunk of every multi segment
  const uint32_t *__restrict__ ptr;     // segment pointers (tables with groups)
  uint32_t count, n_multi;
};

// Shared scratch of the segment kernels (one stream, kernels in sequence).
struct SegScratch {
  double *partial;      // chunk partials of multi-chunk segments, chunks x K (as T)
  double *bpartial;     // block partials of the reduction: pass 1, then pass 2 after `blocks1`
  const int *gate;      // kernels return at once when *gate == 0
};

template <typename T, int K>
__device__ inline T slice_sum(T v) {
#pragma unroll
  for (int off = K; off < Shape<K>::G; off <<= 1) v += shfl_xor(v, off);
  return v;
}

// Pass 1 of a segment operation.  Op supplies
//   T partial<STEP = V>(const uint4 &d, int j, int s)
//                                            sum over the chunk's entries
//                                            d.y + s, d.y + s + STEP, ... < d.z
//   void finish(uint32_t seg, int j, T total, double (&acc)[NA])
//   void reduced(int j, const double (&tot)[NV])   (NV > 0 only)
//   bool enabled()                           uniform early exit
// Single-chunk segments and groups of short segments are finished here (a
// group puts one short segment on each slice, so a wavefront walks several
// rows at once); multi-chunk segments leave their
// chunk partials for pass 2 (the kernel boundary publishes them, so no
// device-scope fence is needed).  With NV > 0 every block stores its
// reduction partial (no ticket).
template <typename T, int K, int NV, class Op>
struct SegPass {
  static constexpr int kThreads = kBlock;
  Chunks ch;
  Op op;
  SegScratch sc;
  __device__ void operator()(Grid g) const {
    if (sc.gate && *sc.gate == 0) return;
    if (!op.enabled()) return;
    constexpr int G = Shape<K>::G, NA = NV > 0 ? NV : 1;
    const int j = threadIdx.x % K, s = (threadIdx.x % G) / K;
    double acc[NA];
#pragma unroll
    for (int v = 0; v < NA; ++v) acc[v] = 0.0;
    T *partial = reinterpret_cast<T *>(sc.partial);
    const uint64_t total = (uint64_t)ch.count * G;
    for (uint64_t t = g.block * (uint64_t)kBlock + threadIdx.x; t < total;
         t += (uint64_t)g.blocks * kBlock) {
      uint32_t u = (uint32_t)(t / G);
      if constexpr (G == 64) u = __builtin_amdgcn_readfirstlane(u);  // one chunk per wavefront
      uint4 d = ch.desc[u];
      if constexpr (G == 64) {
        // wave-uniform chunk: the CSR walk below then uses scalar loads
        d.x = __builtin_amdgcn_readfirstlane(d.x);
        d.y = __builtin_amdgcn_readfirstlane(d.y);
        d.z = __builtin_amdgcn_readfirstlane(d.z);
        d.w = __builtin_amdgcn_readfirstlane(d.w);
      }
      if (d.w > kGroupBase && d.w != kSingle) {
        // a group of short segments: slice s finishes segments s, s + V, ...
        // of the group, each walked by its own K threads
        for (uint32_t q = s; q < d.w - kGroupBase; q += Shape<K>::V) {
          const uint32_t seg = d.x + q;
          const uint4 e = make_uint4(seg, ch.ptr[seg], ch.ptr[seg + 1], kSingle);
          op.finish(seg, j, op.template partial<1>(e, j, 0), acc);
        }
        continue;
      }
      const T part = slice_sum<T, K>(op.partial(d, j, s));
      if (s != 0) continue;
      if (d.w == kSingle)
        op.finish(d.x, j, part, acc);
      else
        partial[(uint64_t)u * K + j] = part;
    }
    if constexpr (NV > 0) block_partial<K, NA>(acc, sc.bpartial, g.block);
  }
};
template <typename T, int K, int NV, class Op>
__global__ void __launch_bounds__(kBlock) k_segments(SegPass<T, K, NV, Op> f) {
  f(Grid::hw());
}

// Pass 2: every multi-chunk segment is summed by one block (kBlock / K
// slices per lane, a fixed LDS tree) in chunk order and finished; with NV > 0
// each block stores its reduction partial after the `blocks1` pass-1
// partials, and k_finish over both passes runs Op::reduced.
template <typename T, int K, int NV, class Op>
struct SegCombine {
  static constexpr int kThreads = kBlock;
  Chunks ch;
  Op op;
  SegScratch sc;
  uint32_t blocks1;
  __device__ void operator()(Grid g) const {
    if (sc.gate && *sc.gate == 0) return;
    if (!op.enabled()) return;
    constexpr int S = kBlock / K, NA = NV > 0 ? NV : 1;
    __shared__ T tree[kBlock];
    const int j = threadIdx.x % K, s = threadIdx.x / K;
    double acc[NA];
#pragma unroll
    for (int v = 0; v < NA; ++v) acc[v] = 0.0;
    const T *partial = reinterpret_cast<const T *>(sc.partial);
    for (uint32_t m = g.block; m < ch.n_multi; m += g.blocks) {
      const uint32_t first = ch.mrange[2 * m], end = ch.mrange[2 * m + 1];
      T sum = T(0);
      uint32_t q = first + s;
      for (; q + 3 * S < end; q += 4 * S) {
        const T a0 = partial[(uint64_t)q * K + j], a1 = partial[(uint64_t)(q + S) * K + j];
        const T a2 = partial[(uint64_t)(q + 2 * S) * K + j], a3 = partial[(uint64_t)(q + 3 * S) * K + j];
        sum += a0;
        sum += a1;
        sum += a2;
        sum += a3;
      }
      for (; q < end; q += S) sum += partial[(uint64_t)q * K + j];
      tree[threadIdx.x] = sum;
      __syncthreads();
      for (int w = S / 2; w > 0; w >>= 1) {
        if (s < w) tree[threadIdx.x] += tree[threadIdx.x + w * K];
        __syncthreads();
      }
      if (s == 0) op.finish(ch.desc[first].x, j, tree[j], acc);
      __syncthreads();
    }
    if constexpr (NV > 0) block_partial<K, NA>(acc, sc.bpartial, blocks1 + g.block);
  }
};
template <typename T, int K, int NV, class Op>
__global__ void __launch_bounds__(kBlock) k_segments_combine(SegCombine<T, K, NV, Op> f) {
  f(Grid::hw());
}

// Value of lane `src` of a wavefront.  With `uniform` (the index is the same
// on every lane) a scalar read-lane, else a cross-lane permute.
template <bool uniform>
__device__ inline uint32_t lane_value(uint32_t v, uint32_t src) {
  if constexpr (uniform) return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)src);
  else return (uint32_t)__shfl((int)v, (int)src, 64);
}
template <bool uniform>
__device__ inline float lane_value(float v, uint32_t src) {
  return __uint_as_float(lane_value<uniform>(__float_as_uint(v), src));
}
template <bool uniform>
__device__ inline double lane_value(double v, uint32_t src) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const uint32_t lo = lane_value<uniform>((uint32_t)b, src), hi = lane_value<uniform>((uint32_t)(b >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Row chunk of y = A x with the neighbour values supplied by D:
// sum over entries of w_ik (x_i - x_k), the reference's per-edge form
// (vendor/cmg/src/graph.rs:664-705).
// An operator may hoist per-lane values out of the walk with lane(j); the
// result is handed to self/nbr (NoLane when nothing is hoisted).
//
// When a wavefront owns the chunk (Shape<K>::G == 64) the walk is staged:
// the wavefront loads up to 64 entries' columns and weights with one
// coalesced load, then every slice issues kStage neighbour gathers at once
// (columns and weights handed across lanes) before it accumulates them.
// Each slice still sums its entries s, s + V, ... in increasing order, so the
// result is the plain walk's bit for bit; only the number of gathers in
// flight changes (a latency-bound walk of dependent loads otherwise).
struct NoLane {};

template <typename T, int K, class D>
struct CsrOp {
  const uint32_t *__restrict__ col;
  const scalar_t<T> *__restrict__ val;
  __device__ NoLane lane(int) const { return {}; }
  template <int STEP = Shape<K>::V>
  __device__ T partial(const uint4 &d, int j, int s) const {
    const D &op = static_cast<const D &>(*this);
    const auto ln = op.lane(j);
    const T xi = op.self(d.x, j, ln);
    T acc = T(0);
    if constexpr (Shape<K>::G == 64 && STEP == Shape<K>::V) {
      constexpr int V = Shape<K>::V;
      constexpr int kStage = V == 1 ? 16 : 8;
      const uint32_t lane = threadIdx.x & 63u;
      for (uint32_t base = d.y; base < d.z; base += 64) {
        const uint32_t cnt = d.z - base < 64 ? d.z - base : 64;  // wave-uniform
        uint32_t mc = 0;
        scalar_t<T> mv = scalar_t<T>(0);
        if (lane < cnt) {
          mc = col[base + lane];
          mv = val[base + lane];
        }
        const uint32_t rounds = (cnt + V - 1) / V;
        for (uint32_t r = 0; r < rounds; r += kStage) {
          T g[kStage];
          scalar_t<T> w[kStage];
          // With one slice the rounds are wave-uniform: every gather is
          // issued unconditionally so none waits on a branch, a round past
          // the end re-reading the column of the block's first entry (a
          // valid index of the gathered vector, whose space may differ from
          // the rows', as for precomposed coarse columns).  With several
          // slices validity differs across lanes and the gathers are masked.
#pragma unroll
          for (int u = 0; u < kStage; ++u) {
            const uint32_t idx = (r + u) * V + (uint32_t)s;
            const bool ok = r + u < rounds && idx < cnt;
            const uint32_t src = ok ? idx : 0u;
            w[u] = lane_value<V == 1>(mv, src);
            const uint32_t c = lane_value<V == 1>(mc, src);
            if constexpr (V == 1) g[u] = op.nbr(c, j, ln);
            else g[u] = ok ? op.nbr(c, j, ln) : xi;
          }
#pragma unroll
          for (int u = 0; u < kStage; ++u) {
            const uint32_t idx = (r + u) * V + (uint32_t)s;
            if constexpr (V == 1) {
              const T next = acc + w[u] * (xi - g[u]);
              acc = r + u < rounds ? next : acc;
            } else if (r + u < rounds && idx < cnt) {
              acc += w[u] * (xi - g[u]);
            }
          }
        }
      }
      return acc;
    }
    if constexpr (std::is_same_v<T, double>) {
#pragma unroll VK_SEG_UNROLL_F64
      for (uint32_t k = d.y + s; k < d.z; k += STEP) acc += val[k] * (xi - op.nbr(col[k], j, ln));
    } else {
#pragma unroll VK_SEG_UNROLL
      for (uint32_t k = d.y + s; k < d.z; k += STEP) acc += val[k] * (xi - op.nbr(col[k], j, ln));
    }
    return acc;
  }
  __device__ bool enabled() const { return true; }
  template <int NV>
  __device__ void reduced(int, const double (&)[NV]) const {}
};

// ---------------------------------------------------------------- V-cycle ops
// cycle.rs:136-437.  Level vectors: braw (the parent's restricted residual),
// b (centred), x, y, r.

// First visit: b = braw - mean, x = D^-1 b, r = b - A x  (it == 0 scale,
// matvec; coarse centring cycle.rs:459-478 folded in).
template <typename T, int K>
struct OpScaleResidual : CsrOp<T, K, OpScaleResidual<T, K>> {
  const T *__restrict__ braw;
  const double *__restrict__ mean;
  const scalar_t<T> *__restrict__ inv;
  T *__restrict__ b, *__restrict__ x, *__restrict__ r;
  __device__ T lane(int j) const {  // the lane's coarse mean
    if (!mean) return T(0);
    if constexpr (std::is_same_v<T, float4>)
      return make_float4((float)mean[4 * j], (float)mean[4 * j + 1], (float)mean[4 * j + 2], (float)mean[4 * j + 3]);
    else
      return (T)mean[j];
  }
  __device__ T self(uint32_t i, int j, T mu) const { return inv[i] * (braw[(uint64_t)i * K + j] - mu); }
  __device__ T nbr(uint32_t c, int j, T mu) const { return inv[c] * (braw[(uint64_t)c * K + j] - mu); }
  template <int NA>
  __device__ void finish(uint32_t i, int j, T ax, double (&)[NA]) const {
    const uint64_t q = (uint64_t)i * K + j;
    const T bi = braw[q] - lane(j);
    b[q] = bi;
    x[q] = inv[i] * bi;
    r[q] = bi - ax;
  }
};

// First visit of an uncentred level whose x0 = D^-1 braw was written by the
// producer of braw (the restriction, or the FP64 -> FP32 split on level 0):
// r = braw - A x0.  b is braw itself and x is x0, so neither is rewritten,
// and a neighbour costs one gather instead of two.
template <typename T, int K>
struct OpFirstResidual : CsrOp<T, K, OpFirstResidual<T, K>> {
  const T *__restrict__ braw, *__restrict__ x0;
  T *__restrict__ r;
  __device__ T self(uint32_t i, int j, NoLane = {}) const { return x0[(uint64_t)i * K + j]; }
  __device__ T nbr(uint32_t c, int j, NoLane = {}) const { return x0[(uint64_t)c * K + j]; }
  template <int NA>
  __device__ void finish(uint32_t i, int j, T ax, double (&)[NA]) const {
    const uint64_t q = (uint64_t)i * K + j;
    if constexpr (std::is_same_v<T, float4>) {
      // b is read for the last time in the cycle: non-temporal
      typedef float v4 __attribute__((ext_vector_type(4)));
      const v4 bn = __builtin_nontemporal_load(reinterpret_cast<const v4 *>(braw + q));
      r[q] = T{bn.x, bn.y, bn.z, bn.w} - ax;
    } else {
      r[q] = braw[q] - ax;
    }
  }
};

// Later visits: y = x + D^-1 (b - A x)
template <typename T, int K>
struct OpSmooth : CsrOp<T, K, OpSmooth<T, K>> {
  const T *__restrict__ b, *__restrict__ x;
  const scalar_t<T> *__restrict__ inv;
  T *__restrict__ y;
  __device__ T self(uint32_t i, int j, NoLane = {}) const { return x[(uint64_t)i * K + j]; }
  __device__ T nbr(uint32_t c, int j, NoLane = {}) const { return x[(uint64_t)c * K + j]; }
  template <int NA>
  __device__ void finish(uint32_t i, int j, T ax, double (&)[NA]) const {
    const uint64_t q = (uint64_t)i * K + j;
    y[q] = x[q] + inv[i] * (b[q] - ax);
  }
};

// r = b - A x
template <typename T, int K>
struct OpResidual : CsrOp<T, K, OpResidual<T, K>> {
  const T *__restrict__ b, *__restrict__ x;
  T *__restrict__ r;
  __device__ T self(uint32_t i, int j, NoLane = {}) const { return x[(uint64_t)i * K + j]; }
  __device__ T nbr(uint32_t c, int j, NoLane = {}) const { return x[(uint64_t)c * K + j]; }
  template <int NA>
  __device__ void finish(uint32_t i, int j, T ax, double (&)[NA]) const {
    const uint64_t q = (uint64_t)i * K + j;
    r[q] = b[q] - ax;
  }
};

// y = x' + D^-1 (b - A x'), x' = x + P e  (prolong-add, matvec, Jacobi).
// With r = b - A x kept from the restriction, A x' = (b - r) + A P e, so
// y = x' + D^-1 (r - A P e): the walk gathers only the coarse correction e
// (small, cache resident), never the fine iterate.  The CSR columns handed to
// this operator are already coarse (lab[col[k]], composed once per layout),
// so a neighbour costs one dependent load, not two.  `Finest` only names the
// finest-level instance (its own kernel symbol for profiling).
template <typename T, int K, bool Finest = false>
struct OpProlongSmooth : CsrOp<T, K, OpProlongSmooth<T, K, Finest>> {
  const T *__restrict__ r, *__restrict__ x, *__restrict__ e;
  const scalar_t<T> *__restrict__ inv;
  const uint32_t *__restrict__ lab;
  T *__restrict__ y;
  __device__ T self(uint32_t i, int j, NoLane = {}) const { return e[(uint64_t)lab[i] * K + j]; }
  __device__ T nbr(uint32_t c, int j, NoLane = {}) const { return e[(uint64_t)c * K + j]; }
  template <int NA>
  __device__ void finish(uint32_t i, int j, T ape, double (&)[NA]) const {
    const uint64_t q = (uint64_t)i * K + j;
    if constexpr (Finest && std::is_same_v<T, float4>) {
      // the finest x and r end their use here: non-temporal, so the cycle's
      // output (gathered next by the Krylov passes) keeps the cache
      typedef float v4 __attribute__((ext_vector_type(4)));
      const v4 xn = __builtin_nontemporal_load(reinterpret_cast<const v4 *>(x + q));
      const v4 rn = __builtin_nontemporal_load(reinterpret_cast<const v4 *>(r + q));
      const T xv{xn.x, xn.y, xn.z, xn.w}, rv{rn.x, rn.y, rn.z, rn.w};
      y[q] = xv + self(i, j) + inv[i] * (rv - ape);
    } else {
      y[q] = x[q] + self(i, j) + inv[i] * (r[q] - ape);
    }
  }
};

// y = A x (diagnostics)
template <typename T, int K>
struct OpProduct : CsrOp<T, K, OpProduct<T, K>> {
  const T *__restrict__ x;
  T *__restrict__ y;
  __device__ T self(uint32_t i, int j, NoLane = {}) const { return x[(uint64_t)i * K + j]; }
  __device__ T nbr(uint32_t c, int j, NoLane = {}) const { return x[(uint64_t)c * K + j]; }
  template <int NA>
  __device__ void finish(uint32_t i, int j, T ax, double (&)[NA]) const {
    y[(uint64_t)i * K + j] = ax;
  }
};

// Restriction: coarse[c] = sum over the aggregate's contiguous members of r
// (coarsen.rs:476-503), plus the per-lane coarse mean for the centring.  With
// `x0`, also the coarse level's first iterate x0 = D_c^-1 coarse (when the
// coarse level is not centred).
template <typename T, int K>
struct OpAggregate {
  const T *__restrict__ r;
  T *__restrict__ out;
  double *__restrict__ mean;
  uint32_t coarse;
  const scalar_t<T> *__restrict__ inv = nullptr;
  T *__restrict__ x0 = nullptr;
  template <int STEP = Shape<K>::V>
  __device__ T partial(const uint4 &d, int j, int s) const {
    T acc = T(0);
    for (uint32_t m = d.y + s; m < d.z; m += STEP) acc += r[(uint64_t)m * K + j];
    return acc;
  }
  __device__ bool enabled() const { return true; }
  template <int NA>
  __device__ void finish(uint32_t c, int j, T sum, double (&acc)[NA]) const {
    out[(uint64_t)c * K + j] = sum;
    if (x0) x0[(uint64_t)c * K + j] = inv[c] * sum;
    if constexpr (!std::is_same_v<T, float4>) acc[0] += (double)sum;
  }
  template <int NV>
  __device__ void reduced(int j, const double (&tot)[NV]) const {
    mean[j] = tot[0] / (double)coarse;
  }
};

// Direct terminal: b = braw - mean, x = G b with G the grounded inverse
// (ldl.rs:892-997).  G is dense (n <= the direct threshold, 700 by default),
// so this is a small dense product of an n x n matrix with n x K lanes.  A
// block owns a tile of Terminal<K>::kRows rows for all K lanes and walks the
// inner index in chunks of kChunk: the chunk's centred right-hand side
// (kChunk x K) and the tile's slice of G (rows x kChunk) are loaded with
// coalesced reads into registers one chunk ahead, staged through LDS, and
// every thread accumulates its outputs over the chunk.  Each output is still
// the serial sum over k = 0 .. n - 1 in increasing order, so the result is
// that of a plain row-by-row dot product bit for bit.
template <int K>
struct Terminal {
  static constexpr int kGroups = kBlock / K;            // row groups of K lanes
  static constexpr int kPerThread = K >= 32 ? 2 : 1;    // rows per thread
  static constexpr int kRows = kGroups * kPerThread;    // rows per tile
  static constexpr int kChunk = kRows >= 128 ? 16 : 64; // inner indices per stage
  static constexpr int kLoadB = (kChunk * K + kBlock - 1) / kBlock;
  static constexpr int kLoadG = (kRows * kChunk + kBlock - 1) / kBlock;
};
template <int K>
inline uint64_t terminal_threads(uint32_t n) {
  return (uint64_t)((n + Terminal<K>::kRows - 1) / Terminal<K>::kRows) * kBlock;
}

template <typename T, int K>
struct TerminalDirect {
  static constexpr int kThreads = kBlock;
  uint32_t n;
  const double *__restrict__ ginv;
  const T *__restrict__ braw;
  const double *__restrict__ mean;
  T *__restrict__ b, *__restrict__ x;
  const int *__restrict__ gate;
  __device__ void operator()(Grid g) const {
    if (gate && *gate == 0) return;
    using TT = Terminal<K>;
    constexpr int R = TT::kRows, KC = TT::kChunk, NB = TT::kLoadB, NG = TT::kLoadG, RG = TT::kGroups;
    __shared__ double bt[KC][K];
    __shared__ double gt[R][KC + 1];
    const int tid = threadIdx.x, j = tid % K, row = tid / K;
    for (uint32_t i0 = g.block * (uint32_t)R; i0 < n; i0 += g.blocks * (uint32_t)R) {
      double rb[NB], rg[NG];
      auto load = [&](uint32_t k0) {
#pragma unroll
        for (int e = 0; e < NB; ++e) {
          const int idx = tid + e * kBlock, kk = idx / K, jj = idx % K;
          rb[e] = 0.0;
          if (idx < KC * K && k0 + kk < n)
            rb[e] = (double)braw[(uint64_t)(k0 + kk) * K + jj] - (mean ? mean[jj] : 0.0);
        }
#pragma unroll
        for (int e = 0; e < NG; ++e) {
          const int idx = tid + e * kBlock, rr = idx / KC, kk = idx % KC;
          rg[e] = 0.0;
          if (idx < R * KC && i0 + rr < n && k0 + kk < n) rg[e] = ginv[(uint64_t)(i0 + rr) * n + k0 + kk];
        }
      };
      double s[TT::kPerThread];
#pragma unroll
      for (int q = 0; q < TT::kPerThread; ++q) s[q] = 0.0;
      load(0);
      for (uint32_t k0 = 0; k0 < n; k0 += KC) {
#pragma unroll
        for (int e = 0; e < NB; ++e) {
          const int idx = tid + e * kBlock;
          if (idx < KC * K) bt[idx / K][idx % K] = rb[e];
        }
#pragma unroll
        for (int e = 0; e < NG; ++e) {
          const int idx = tid + e * kBlock;
          if (idx < R * KC) gt[idx / KC][idx % KC] = rg[e];
        }
        __syncthreads();
        if (k0 + KC < n) load(k0 + KC);  // the next chunk's loads fly during this chunk's products
        const uint32_t kc = n - k0 < (uint32_t)KC ? n - k0 : (uint32_t)KC;
        for (uint32_t kk = 0; kk < kc; ++kk) {
          const double bv = bt[kk][j];
#pragma unroll
          for (int q = 0; q < TT::kPerThread; ++q) s[q] += gt[row + RG * q][kk] * bv;
        }
        __syncthreads();
      }
      const double mu = mean ? mean[j] : 0.0;
#pragma unroll
      for (int q = 0; q < TT::kPerThread; ++q) {
        const uint32_t i = i0 + row + RG * q;
        if (i < n) {
          const uint64_t t = (uint64_t)i * K + j;
          x[t] = (T)s[q];
          b[t] = (T)((double)braw[t] - mu);
        }
      }
    }
  }
};
template <typename T, int K>
__global__ void __launch_bounds__(kBlock) k_terminal_direct(TerminalDirect<T, K> f) {
  f(Grid::hw());
}

// Iterative terminal: b = braw - mean, x = D^-1 b  (cycle.rs:342-379)
template <typename T, int K>
struct TerminalScale {
  static constexpr int kThreads = kBlock;
  uint32_t n;
  const T *__restrict__ inv, *__restrict__ braw;
  const double *__restrict__ mean;
  T *__restrict__ b, *__restrict__ x;
  const int *__restrict__ gate;
  __device__ void operator()(Grid g) const {
    if (gate && *gate == 0) return;
    VK_GRID_STRIDE(g, t, (uint64_t)n * K) {
      const uint32_t i = t / K;
      const int j = t % K;
      const T bi = mean ? braw[t] - (T)mean[j] : braw[t];
      b[t] = bi;
      x[t] = inv[i] * bi;
    }
  }
};
template <typename T, int K>
__global__ void k_terminal_scale(TerminalScale<T, K> f) {
  f(Grid::hw());
}

// Plain stream copy (HBM traffic calibration of the profiler counters).
static __global__ void k_stream_copy(uint64_t count, const double4 *__restrict__ a, double4 *__restrict__ b) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < count;
       t += (uint64_t)gridDim.x * blockDim.x)
    b[t] = a[t];
}

// ---------------------------------------------------------------- precision casts
// rc = r - mean (FP64, kept for the Krylov recurrences) and its FP32 copy for
// the V-cycle; with `inv`, also the cycle's first iterate x0 = D^-1 rf.
template <int K>
struct CentreSplit {
  static constexpr int kThreads = kBlock;
  uint64_t rows;
  const double *__restrict__ r, *__restrict__ mean;
  double *__restrict__ rc;
  float *__restrict__ rf;
  const float *__restrict__ inv;
  float *__restrict__ x0;
  const int *__restrict__ gate;
  __device__ void operator()(Grid g) const {
    if (gate && *gate == 0) return;
    static_assert(kBlock % K == 0, "a thread keeps its lane across the grid stride");
    const double m = mean ? mean[threadIdx.x % K] : 0.0;
    VK_GRID_STRIDE(g, t, rows * K) {
      const double v = mean ? r[t] - m : r[t];
      if (rc) rc[t] = v;
      rf[t] = (float)v;
      if (inv) x0[t] = inv[t / K] * (float)v;
    }
  }
};
template <int K>
__global__ void k_centre_split(CentreSplit<K> f) {
  f(Grid::hw());
}
template <int K>
__global__ void k_to_double(uint64_t count, const float *__restrict__ a, double *__restrict__ b,
                            const int *__restrict__ gate) {
  if (gate && *gate == 0) return;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < count;
       t += (uint64_t)gridDim.x * blockDim.x)
    b[t] = (double)a[t];
}

// Index composition out[k] = map[a[k]] (coarse column of every CSR entry).
static __global__ void k_compose(uint64_t count, const uint32_t *__restrict__ a, const uint32_t *__restrict__ map,
                                 uint32_t *__restrict__ out) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < count;
       t += (uint64_t)gridDim.x * blockDim.x)
    out[t] = map[a[t]];
}

// Row permutation of K-lane vectors: out[p] = in[perm[p]] (gather).
template <int K>
struct GatherRows {
  static constexpr int kThreads = kBlock;
  uint64_t rows;
  const uint32_t *__restrict__ idx;
  const double *__restrict__ in;
  double *__restrict__ out;
  __device__ void operator()(Grid g) const {
    VK_GRID_STRIDE(g, t, rows * K) out[t] = in[(uint64_t)idx[t / K] * K + t % K];
  }
};
template <int K>
__global__ void k_gather_rows(GatherRows<K> f) {
  f(Grid::hw());
}

}  // namespace dev
}  // namespace vk
