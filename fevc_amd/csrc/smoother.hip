// Learned CMG smoother weights on the device: inference into the V-cycle's
// inverse diagonals, the training loss and its exact gradient, Adam training,
// and the plain-cycle fallback around the solve path.
//
// No counterpart in the reference (whose cycle smooths with 0.5 / d,
// vendor/cmg/src/preconditioner/cycle.rs:136-311).  The kernels are in
// kernels_smoother.hpp.  Training objective, for 32 error columns e at level 0:
//   L = mean_j |e_j - M A e_j|_A^2 / |e_j|_A^2
// with M the training cycle (one visit per level; level 0 uncentred, coarse
// right-hand sides centred by their lane means).  Error columns are uniform
// random vectors passed through `passes` plain-cycle error steps
// e <- e - M_plain A e, so they hold the modes the plain cycle leaves.
#include "device_impl.hpp"
#include "kernels_smoother.hpp"

#include <cmath>

namespace vk {
namespace {

namespace M = dev::mpnn;
static_assert(M::P == (int)kSmootherParams && M::LAYERS == kSmootherLayers,
              "network layout and the declared parameter count differ");
constexpr int KT = 32;  // training columns
constexpr unsigned kOuterBlocks = 512;

// ------------------------------------------------------------- host inputs

// Features of one smoothed level, rows in the device (renumbered) order:
//   1, log d - mean log d, log(1 + degree) / 4, max w / s, intra-aggregate
//   weight / s, log aggregate size / 4, depth / 8, sum w (log d_j - log d_i) / s
// with s the weighted degree; rs = 1 / s (0 for isolated vertices).
void level_features(const Level &L, const std::vector<uint32_t> &perm, uint32_t depth, std::vector<double> &X,
                    std::vector<double> &rs) {
  const Laplacian &g = L.graph;
  const uint64_t n = g.n;
  std::vector<double> s(n, 0.0), wmax(n, 0.0), intra(n, 0.0), contrast(n, 0.0), ld(n, 0.0);
  std::vector<uint64_t> cnt(n, 0), asz(std::max<uint64_t>(L.coarse, 1), 0);
  double lsum = 0.0;
  uint64_t lcount = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (g.diag[i] > 0.0) {
      ld[i] = std::log(g.diag[i]);
      lsum += ld[i];
      ++lcount;
    }
  const double mu = lcount ? lsum / (double)lcount : 0.0;
  for (size_t e = 0; e < g.u.size(); ++e) {
    const uint32_t a = g.u[e], b = g.v[e];
    const double w = g.w[e];
    s[a] += w;
    s[b] += w;
    cnt[a]++;
    cnt[b]++;
    wmax[a] = std::max(wmax[a], w);
    wmax[b] = std::max(wmax[b], w);
    if (L.labels[a] == L.labels[b]) {
      intra[a] += w;
      intra[b] += w;
    }
    contrast[a] += w * (ld[b] - ld[a]);
    contrast[b] += w * (ld[a] - ld[b]);
  }
  for (uint64_t i = 0; i < n; ++i) asz[L.labels[i]]++;
  X.assign(n * M::F, 0.0);
  rs.assign(n, 0.0);
  for (uint64_t r = 0; r < n; ++r) {
    const uint32_t i = perm[r];
    const double inv_s = s[i] > 0.0 ? 1.0 / s[i] : 0.0;
    double *x = &X[r * M::F];
    x[0] = 1.0;
    x[1] = g.diag[i] > 0.0 ? ld[i] - mu : 0.0;
    x[2] = std::log1p((double)cnt[i]) / 4.0;
    x[3] = wmax[i] * inv_s;
    x[4] = intra[i] * inv_s;
    x[5] = std::log((double)asz[L.labels[i]]) / 4.0;
    x[6] = depth / 8.0;
    x[7] = contrast[i] * inv_s;
    rs[r] = inv_s;
  }
}

void ensure_features(Impl &E) {
  SmootherState &S = E.sm;
  // the network's segment passes run H lanes: the segment scratch must hold
  // H lanes per chunk whatever width the solves use
  E.ensure(M::H);
  if (S.features) return;
  const Hierarchy &h = *E.H;
  S.L.clear();
  size_t maxn = 0;
  for (size_t l = 0; l < h.levels.size(); ++l) {
    if (h.levels[l].terminal) continue;
    S.L.emplace_back();
    SmootherLevel &Z = S.L.back();
    Z.level = (uint32_t)l;
    Z.n = E.levels[l].n;
    std::vector<double> X, rs;
    level_features(h.levels[l], E.levels[l].perm, (uint32_t)l, X, rs);
    Z.X.upload(X);
    Z.rs.upload(rs);
    const size_t m = (size_t)Z.n * M::H;
    for (auto &v : Z.h) v.alloc(m);
    for (auto &v : Z.m) v.alloc(m);
    for (auto &v : Z.hb) v.alloc(m);
    Z.mb.alloc(m);
    Z.h32.alloc(m);
    Z.omega.alloc(Z.n);
    Z.gomega.alloc(Z.n);
    E.levels[l].linv.alloc(Z.n);
    E.levels[l].linv32.alloc(Z.n);
    maxn = std::max<size_t>(maxn, Z.n);
  }
  const size_t m = std::max<size_t>(maxn, 1) * M::H;
  for (auto *v : {&S.dh, &S.dpre, &S.dms, &S.dtmp}) v->alloc(m);
  S.dz.alloc(std::max<size_t>(maxn, 1));
  S.wcat.alloc((size_t)M::LAYERS * M::H * 2 * M::H);
  S.theta.alloc(M::P);
  S.grad.alloc(M::P);
  S.am.alloc(M::P);
  S.av.alloc(M::P);
  S.partial.alloc((size_t)kOuterBlocks * (M::H * M::H + M::H));
  S.features = true;
}

// ------------------------------------------------------------ the network

void adjacency(Impl &E, const DevLevel &D, const double *x, const double *scale, double *out, bool accumulate) {
  segments<double, M::H, 0>(E, D.rows, dev::OpAdjacency<M::H>{D.col.p, D.val.p, x, scale, out, accumulate});
}

// omega (and the learned inverse diagonals) of every smoothed level
void mpnn_forward(Impl &E) {
  SmootherState &S = E.sm;
  const double *th = S.theta.p;
  for (SmootherLevel &Z : S.L) {
    DevLevel &D = E.levels[Z.level];
    const uint64_t m = (uint64_t)Z.n * M::H;
    hipLaunchKernelGGL(dev::k_mpnn_in, grid_for(m), kBlock, 0, E.stream, Z.n, Z.X.p, th, Z.h[0].p);
    for (int t = 1; t <= M::LAYERS; ++t) {
      adjacency(E, D, Z.h[t - 1].p, Z.rs.p, Z.m[t - 1].p, false);
      hipLaunchKernelGGL(dev::k_mpnn_layer, grid_for(m), kBlock, 0, E.stream, Z.n, Z.h[t - 1].p, Z.m[t - 1].p, th,
                         M::kU(t), M::kV(t), M::kC(t), Z.h[t].p);
    }
    hipLaunchKernelGGL(dev::k_mpnn_out, grid_for(Z.n), kBlock, 0, E.stream, Z.n, Z.h[M::LAYERS].p, th, D.inv.p,
                       Z.omega.p, D.linv.p, D.linv32.p);
  }
}

// The solve path's inference of the same network: BF16 weights and
// activations, FP32 accumulation, the layer products on the matrix cores.
void mpnn_infer(Impl &E) {
  SmootherState &S = E.sm;
  const double *th = S.theta.p;
  const uint64_t wc = (uint64_t)M::LAYERS * M::H * 2 * M::H;
  hipLaunchKernelGGL(dev::k_mpnn_weights_bf16, grid_for(wc), kBlock, 0, E.stream, th, S.wcat.p);
  for (SmootherLevel &Z : S.L) {
    DevLevel &D = E.levels[Z.level];
    const uint64_t m = (uint64_t)Z.n * M::H;
    hipLaunchKernelGGL(dev::k_mpnn_in_bf16, grid_for(m), kBlock, 0, E.stream, Z.n, Z.X.p, th, Z.hb[0].p);
    const uint64_t tiles = ((uint64_t)Z.n + 15) / 16;
    const unsigned blocks = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((tiles + 3) / 4, 16384));
    int cur = 0;
    for (int t = 1; t <= M::LAYERS; ++t) {
      segments<float, M::H, 0>(E, D.rows, dev::OpMessageBf16<M::H>{D.col.p, D.val.p, Z.hb[cur].p, Z.rs.p, Z.mb.p});
      const bool last = t == M::LAYERS;
      hipLaunchKernelGGL(dev::k_mpnn_layer_bf16, blocks, kBlock, 0, E.stream, Z.n, Z.hb[cur].p, Z.mb.p,
                         S.wcat.p + (uint64_t)(t - 1) * M::H * 2 * M::H, th, M::kC(t),
                         last ? nullptr : Z.hb[1 - cur].p, last ? Z.h32.p : nullptr);
      cur = 1 - cur;
    }
    hipLaunchKernelGGL(dev::k_mpnn_out_f32, grid_for(Z.n), kBlock, 0, E.stream, Z.n, Z.h32.p, th, D.inv.p,
                       Z.omega.p, D.linv.p, D.linv32.p);
  }
}

template <int KA, int KB, bool Bias>
void outer(Impl &E, uint32_t n, const double *A, const double *B, double *grad) {
  constexpr uint32_t count = KA * KB + (Bias ? KA : 0);
  const unsigned blocks = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(kOuterBlocks, (n + 63) / 64));
  hipLaunchKernelGGL((dev::k_outer_partial<KA, KB, Bias>), blocks, kBlock, 0, E.stream, n, A, B, E.sm.partial.p);
  hipLaunchKernelGGL(dev::k_partial_sum, grid_for(count), kBlock, 0, E.stream, E.sm.partial.p, blocks, count, grad);
}

// dL/dtheta += backward of one level from its g_omega
void mpnn_backward(Impl &E, SmootherLevel &Z) {
  SmootherState &S = E.sm;
  DevLevel &D = E.levels[Z.level];
  const double *th = S.theta.p;
  double *g = S.grad.p;
  const uint32_t n = Z.n;
  const uint64_t m = (uint64_t)n * M::H;
  const unsigned lb = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(8192, (n + M::kNodesPerBlock - 1) /
                                                                                   M::kNodesPerBlock));
  hipLaunchKernelGGL(dev::k_mpnn_out_back, grid_for(m), kBlock, 0, E.stream, n, Z.omega.p, Z.gomega.p, th, S.dz.p,
                     S.dh.p);
  outer<1, M::H, true>(E, n, S.dz.p, Z.h[M::LAYERS].p, g + M::kA);
  // layers L..1: dh (= dh_t) -> dtmp (= dh_{t-1}), the buffers then swap
  double *dh = S.dh.p, *dnext = S.dtmp.p;
  for (int t = M::LAYERS; t >= 1; --t) {
    hipLaunchKernelGGL(dev::k_mpnn_layer_back, lb, kBlock, 0, E.stream, n, dh, Z.h[t].p, th, M::kU(t), M::kV(t),
                       Z.rs.p, S.dpre.p, dnext, S.dms.p);
    outer<M::H, M::H, false>(E, n, S.dpre.p, Z.h[t - 1].p, g + M::kU(t));
    outer<M::H, M::H, true>(E, n, S.dpre.p, Z.m[t - 1].p, g + M::kV(t));
    adjacency(E, D, S.dms.p, nullptr, dnext, true);
    std::swap(dh, dnext);
  }
  // input layer
  hipLaunchKernelGGL(dev::k_relu_back, grid_for(m), kBlock, 0, E.stream, m, dh, Z.h[0].p, S.dpre.p);
  outer<M::H, M::F, true>(E, n, S.dpre.p, Z.X.p, g + M::kW0);
}

// ------------------------------------------------------- training V-cycle

void ensure_train(Impl &E) {
  SmootherState &S = E.sm;
  if (S.train_buffers) return;
  E.ensure(KT);
  S.T.resize(E.levels.size());
  for (size_t l = 0; l < E.levels.size(); ++l) {
    const size_t m = (size_t)E.levels[l].n * KT;
    TrainLevel &T = S.T[l];
    for (auto *v : {&T.braw, &T.b, &T.s, &T.t1, &T.t2, &T.t3, &T.y, &T.ly}) v->alloc(m);
    T.mean.alloc(KT);
  }
  const size_t m0 = (size_t)E.n0 * KT;
  for (auto *v : {&S.err, &S.ae, &S.u, &S.au, &S.lam}) v->alloc(m0);
  S.q.alloc(KT);
  S.uau.alloc(KT);
  S.train_buffers = true;
}

// smoother inverse diagonal of a level for the training cycle
const double *train_w(Impl &E, size_t l, bool learned) {
  return learned ? E.levels[l].linv.p : E.levels[l].inv.p;
}

void lane_mean(Impl &E, uint32_t n, const double *v, double *mean) {
  const unsigned g = reduce_grid((uint64_t)n * KT);
  run(E, dev::k_lane_sum<KT>, g, dev::LaneSum<KT>{(uint64_t)n, v, E.rpartial.p});
  finish<KT, 1>(E, E.rpartial.p, g, dev::FinMean{mean, nullptr, (uint64_t)n});
}

void lane_dot(Impl &E, uint32_t n, const double *a, const double *b, double *out) {
  const unsigned g = reduce_grid((uint64_t)n * KT);
  hipLaunchKernelGGL((dev::k_lane_dot<KT>), g, kBlock, 0, E.stream, n, a, b, E.rpartial.p);
  finish<KT, 1>(E, E.rpartial.p, g, dev::FinStore{out});
}

// y = M braw at level l (braw's lane means are in T[l].mean for l > 0)
const double *train_forward(Impl &E, size_t l, const double *braw, bool learned) {
  DevLevel &D = E.levels[l];
  TrainLevel &T = E.sm.T[l];
  const uint32_t n = D.n;
  const uint64_t m = (uint64_t)n * KT;
  const double *mean = l ? T.mean.p : nullptr;
  if (D.terminal) {
    if (D.terminal == kDirect)
      run(E, dev::k_terminal_direct<double, KT>, grid_for(dev::terminal_threads<KT>(n)),
          dev::TerminalDirect<double, KT>{n, E.ginv.p, braw, mean, T.b.p, T.y.p, nullptr});
    else
      run(E, dev::k_terminal_scale<double, KT>, grid_for(m),
          dev::TerminalScale<double, KT>{n, D.inv.p, braw, mean, T.b.p, T.y.p, nullptr});
    return T.y.p;
  }
  const double *w = train_w(E, l, learned);
  hipLaunchKernelGGL((dev::k_lane_centre<KT>), grid_for(m), kBlock, 0, E.stream, n, braw, mean, T.b.p);
  hipLaunchKernelGGL((dev::k_row_scale<KT>), grid_for(m), kBlock, 0, E.stream, n, w, T.b.p, T.t1.p);
  segments<double, KT, 0>(E, D.rows, dev::OpResidual<double, KT>{{D.col.p, D.val.p}, T.b.p, T.t1.p, T.t2.p});
  DevLevel &Dc = E.levels[l + 1];
  TrainLevel &Tc = E.sm.T[l + 1];
  segments<double, KT, 1>(E, D.aggs, dev::OpAggregate<double, KT>{T.t2.p, Tc.braw.p, Tc.mean.p, Dc.n});
  const double *ec = train_forward(E, l + 1, Tc.braw.p, learned);
  hipLaunchKernelGGL((dev::k_prolong<KT>), grid_for(m), kBlock, 0, E.stream, n, D.lab.p, ec, (const double *)nullptr,
                     T.t3.p);
  segments<double, KT, 0>(E, D.rows, dev::OpResidual<double, KT>{{D.col.p, D.val.p}, T.t2.p, T.t3.p, T.s.p});
  hipLaunchKernelGGL((dev::k_train_post<KT>), grid_for(m), kBlock, 0, E.stream, n, w, T.b.p, T.s.p, ec, D.lab.p,
                     T.y.p);
  return T.y.p;
}

// Adjoint of train_forward: returns dL/dbraw; writes g_omega of every
// smoothed level into its SmootherLevel.
const double *train_adjoint(Impl &E, size_t l, const double *ly, const std::vector<SmootherLevel *> &by_level) {
  DevLevel &D = E.levels[l];
  TrainLevel &T = E.sm.T[l];
  const uint32_t n = D.n;
  const uint64_t m = (uint64_t)n * KT;
  const double *lb;
  if (D.terminal) {
    if (D.terminal == kDirect)
      hipLaunchKernelGGL((dev::k_dense_transpose<KT>), grid_for(m), kBlock, 0, E.stream, n, E.ginv.p, ly, T.t1.p);
    else
      hipLaunchKernelGGL((dev::k_row_scale<KT>), grid_for(m), kBlock, 0, E.stream, n, D.inv.p, ly, T.t1.p);
    lb = T.t1.p;
  } else {
    const double *w = D.linv.p;
    // ls = w ly (t1); lpe = ly - A ls (t2); lec = P^T lpe (coarse ly)
    hipLaunchKernelGGL((dev::k_row_scale<KT>), grid_for(m), kBlock, 0, E.stream, n, w, ly, T.t1.p);
    segments<double, KT, 0>(E, D.rows, dev::OpResidual<double, KT>{{D.col.p, D.val.p}, ly, T.t1.p, T.t2.p});
    DevLevel &Dc = E.levels[l + 1];
    TrainLevel &Tc = E.sm.T[l + 1];
    segments<double, KT, 1>(E, D.aggs, dev::OpAggregate<double, KT>{T.t2.p, Tc.ly.p, Tc.mean.p, Dc.n});
    const double *lc = train_adjoint(E, l + 1, Tc.ly.p, by_level);
    // lr1 = ls + P lc (t3); lx1 = ly - A lr1 (t2)
    hipLaunchKernelGGL((dev::k_prolong<KT>), grid_for(m), kBlock, 0, E.stream, n, D.lab.p, lc, T.t1.p, T.t3.p);
    segments<double, KT, 0>(E, D.rows, dev::OpResidual<double, KT>{{D.col.p, D.val.p}, ly, T.t3.p, T.t2.p});
    SmootherLevel *Z = by_level[l];
    hipLaunchKernelGGL((dev::k_train_smooth_adjoint<KT>), grid_for(n), kBlock, 0, E.stream, n, w, D.inv.p, ly, T.s.p,
                       T.t2.p, T.b.p, T.t3.p, T.t1.p, Z->gomega.p);
    lb = T.t1.p;
  }
  if (l == 0) return lb;
  lane_mean(E, n, lb, T.mean.p);
  double *out = D.terminal ? T.t2.p : T.t3.p;
  hipLaunchKernelGGL((dev::k_lane_centre<KT>), grid_for(m), kBlock, 0, E.stream, n, lb, T.mean.p, out);
  return out;
}

// Loss and gradient (into S.grad, zeroed here) for the error columns in S.err.
double loss_and_gradient(Impl &E) {
  SmootherState &S = E.sm;
  DevLevel &D0 = E.levels[0];
  const uint32_t n = D0.n;
  const uint64_t m = (uint64_t)n * KT;
  mpnn_forward(E);
  segments<double, KT, 0>(E, D0.rows, dev::OpProduct<double, KT>{{D0.col.p, D0.val.p}, S.err.p, S.ae.p});
  lane_dot(E, n, S.err.p, S.ae.p, S.q.p);
  const double *y = train_forward(E, 0, S.ae.p, true);
  VK_HIP(hipMemcpyAsync(S.u.p, S.err.p, m * sizeof(double), hipMemcpyDeviceToDevice, E.stream));
  hipLaunchKernelGGL(dev::k_sub, grid_for(m), kBlock, 0, E.stream, m, y, S.u.p);
  segments<double, KT, 0>(E, D0.rows, dev::OpProduct<double, KT>{{D0.col.p, D0.val.p}, S.u.p, S.au.p});
  lane_dot(E, n, S.u.p, S.au.p, S.uau.p);
  hipLaunchKernelGGL((dev::k_loss_seed<KT>), grid_for(m), kBlock, 0, E.stream, n, S.au.p, S.q.p, S.lam.p);
  std::vector<SmootherLevel *> by_level(E.levels.size(), nullptr);
  for (SmootherLevel &Z : S.L) by_level[Z.level] = &Z;
  (void)train_adjoint(E, 0, S.lam.p, by_level);
  VK_HIP(hipMemsetAsync(S.grad.p, 0, M::P * sizeof(double), E.stream));
  for (SmootherLevel &Z : S.L) mpnn_backward(E, Z);
  std::vector<double> q, uau;
  S.q.download(q, KT, E.stream);
  S.uau.download(uau, KT, E.stream);
  double L = 0.0;
  for (int j = 0; j < KT; ++j) L += uau[j] / q[j];
  return L / KT;
}

// Training errors: uniform random columns, centred, then `passes` plain-cycle
// error steps e <- e - M_plain A e.
void make_errors(Impl &E, uint64_t seed, uint32_t passes) {
  SmootherState &S = E.sm;
  DevLevel &D0 = E.levels[0];
  const uint32_t n = D0.n;
  const uint64_t m = (uint64_t)n * KT;
  hipLaunchKernelGGL(dev::k_random_errors, grid_for(m), kBlock, 0, E.stream, m, seed, S.u.p);
  lane_mean(E, n, S.u.p, S.q.p);
  hipLaunchKernelGGL((dev::k_lane_centre<KT>), grid_for(m), kBlock, 0, E.stream, n, S.u.p, S.q.p, S.err.p);
  for (uint32_t p = 0; p < passes; ++p) {
    segments<double, KT, 0>(E, D0.rows, dev::OpProduct<double, KT>{{D0.col.p, D0.val.p}, S.err.p, S.ae.p});
    const double *y = train_forward(E, 0, S.ae.p, false);
    hipLaunchKernelGGL(dev::k_sub, grid_for(m), kBlock, 0, E.stream, m, y, S.err.p);
  }
}

void check_params(const std::vector<double> &theta) {
  if (theta.size() != (size_t)M::P)
    fail(kInvalidInput, "smoother", "expected " + std::to_string(M::P) + " network parameters");
  for (double v : theta)
    if (!std::isfinite(v)) fail(kInvalidInput, "smoother", "network parameters must be finite");
}

void upload_theta(Impl &E, const std::vector<double> &theta) {
  VK_HIP(hipMemcpy(E.sm.theta.p, theta.data(), M::P * sizeof(double), hipMemcpyHostToDevice));
}

bool has_smoothed_levels(const Impl &E) {
  for (const DevLevel &D : E.levels)
    if (!D.terminal) return true;
  return false;
}

std::vector<uint32_t> tuned_shape(const Impl &E);
std::vector<uint32_t> full_shape(const Impl &E);
double iteration_work(const Impl &E, const std::vector<uint32_t> &shape);

}  // namespace

void DeviceEngine::smoother_omega(const std::vector<double> &theta, std::vector<double> &omega, bool bf16) {
  Impl &E = *impl_;
  check_params(theta);
  E.gate = E.always();
  E.ensure(M::H);
  ensure_features(E);
  upload_theta(E, theta);
  if (bf16) mpnn_infer(E);
  else mpnn_forward(E);
  omega.clear();
  for (SmootherLevel &Z : E.sm.L) {
    std::vector<double> w;
    Z.omega.download(w, Z.n, E.stream);
    const std::vector<double> orig = from_renumbered(w, E.levels[Z.level].perm, 1);
    omega.insert(omega.end(), orig.begin(), orig.end());
  }
  if (E.sm.active) {  // restore the inverse diagonals of the weights in use
    upload_theta(E, E.sm.params);
    mpnn_infer(E);
    VK_HIP(hipStreamSynchronize(E.stream));
  }
}

double DeviceEngine::smoother_gradient(const std::vector<double> &theta, const std::vector<double> &errors,
                                       std::vector<double> &grad) {
  Impl &E = *impl_;
  check_params(theta);
  if (!has_smoothed_levels(E)) fail(kInvalidInput, "smoother", "the hierarchy has no smoothed level");
  if (errors.size() != (size_t)E.n0 * KT) fail(kInvalidInput, "smoother", "errors must be n x 32");
  PlainCycle keep(E);
  E.gate = E.always();
  ensure_features(E);
  ensure_train(E);
  upload_theta(E, theta);
  const std::vector<double> rn = to_renumbered(errors, E.levels[0].perm, KT);
  VK_HIP(hipMemcpy(E.sm.err.p, rn.data(), rn.size() * sizeof(double), hipMemcpyHostToDevice));
  const double L = loss_and_gradient(E);
  E.sm.grad.download(grad, M::P, E.stream);
  if (keep.saved) {
    upload_theta(E, E.sm.params);
    mpnn_infer(E);
    VK_HIP(hipStreamSynchronize(E.stream));
  }
  return L;
}

void DeviceEngine::smoother_train(std::vector<double> &state, uint64_t step0, uint32_t steps, double lr, uint64_t seed,
                                  uint32_t passes, std::vector<double> &losses) {
  Impl &E = *impl_;
  if (state.size() != 3 * (size_t)M::P) fail(kInvalidInput, "smoother", "training state must hold 3 x parameters");
  for (double v : state)
    if (!std::isfinite(v)) fail(kInvalidInput, "smoother", "training state must be finite");
  if (!(lr > 0.0) || !std::isfinite(lr)) fail(kInvalidInput, "smoother", "learning rate must be positive");
  if (!has_smoothed_levels(E)) fail(kInvalidInput, "smoother", "the hierarchy has no smoothed level");
  PlainCycle keep(E);
  E.gate = E.always();
  ensure_features(E);
  ensure_train(E);
  SmootherState &S = E.sm;
  VK_HIP(hipMemcpy(S.theta.p, state.data(), M::P * sizeof(double), hipMemcpyHostToDevice));
  VK_HIP(hipMemcpy(S.am.p, state.data() + M::P, M::P * sizeof(double), hipMemcpyHostToDevice));
  VK_HIP(hipMemcpy(S.av.p, state.data() + 2 * M::P, M::P * sizeof(double), hipMemcpyHostToDevice));
  const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
  losses.assign(steps, 0.0);
  for (uint32_t k = 0; k < steps; ++k) {
    const uint64_t t = step0 + k + 1;
    make_errors(E, seed ^ (0x9e3779b97f4a7c15ULL * t), passes);
    losses[k] = loss_and_gradient(E);
    if (!std::isfinite(losses[k])) fail(kInvalidInput, "smoother", "training loss is not finite");
    hipLaunchKernelGGL(dev::k_adam, grid_for(M::P), kBlock, 0, E.stream, (uint32_t)M::P, S.theta.p, S.grad.p, S.am.p,
                       S.av.p, lr, b1, b2, eps, 1.0 - std::pow(b1, (double)t), 1.0 - std::pow(b2, (double)t));
  }
  std::vector<double> th, mm, vv;
  S.theta.download(th, M::P, E.stream);
  S.am.download(mm, M::P, E.stream);
  S.av.download(vv, M::P, E.stream);
  std::copy(th.begin(), th.end(), state.begin());
  std::copy(mm.begin(), mm.end(), state.begin() + M::P);
  std::copy(vv.begin(), vv.end(), state.begin() + 2 * M::P);
  if (keep.saved) {
    upload_theta(E, S.params);
    mpnn_infer(E);
    VK_HIP(hipStreamSynchronize(E.stream));
  }
}

namespace {

// Levels the tuned cycle visits: level 0, every level that at least halves
// the last visited one, and the terminal.  Levels that coarsen less add
// smoothing work but (measured) no convergence.
std::vector<uint32_t> tuned_shape(const Impl &E) {
  const uint32_t T = (uint32_t)E.levels.size() - 1;
  std::vector<uint32_t> s{0};
  for (uint32_t l = 1; l < T; ++l)
    if (2ull * E.levels[l].n <= E.levels[s.back()].n) s.push_back(l);
  if (T > 0) s.push_back(T);
  return s;
}

std::vector<uint32_t> full_shape(const Impl &E) {
  std::vector<uint32_t> s(E.levels.size());
  for (uint32_t l = 0; l < s.size(); ++l) s[l] = l;
  return s;
}

// Work of one PCG iteration with a cycle shape, in CSR-entry units: the
// Krylov SpMV and vector passes at level 0 plus two level products and the
// vector passes of every visited smoothed level.
double iteration_work(const Impl &E, const std::vector<uint32_t> &shape) {
  double w = 3.0 * (double)E.levels[0].nnz + 8.0 * E.levels[0].n;
  for (uint32_t l : shape)
    if (!E.levels[l].terminal) w += 2.0 * (double)E.levels[l].nnz + 4.0 * E.levels[l].n;
  return w;
}

}  // namespace

void DeviceEngine::smoother_mode(int mode, const std::vector<double> &theta, int precision) {
  Impl &E = *impl_;
  if (mode != 0 && mode != 1) fail(kInvalidInput, "smoother", "mode must be 0 (plain) or 1 (learned)");
  SmootherState &S = E.sm;
  S.active = false;
  S.shaped = false;
  S.calibrated = false;
  S.breakdown = false;
  S.plain_iterations = S.learned_iterations = 0;
  S.calibration.clear();
  S.mode = mode;
  E.precision_hint = precision;
  if (mode == 0) return;
  check_params(theta);
  S.params = theta;
  if (!has_smoothed_levels(E)) {  // a direct terminal alone: nothing to tune
    S.calibrated = true;
    return;
  }
  E.gate = E.always();
  E.ensure(M::H);
  ensure_features(E);
  upload_theta(E, theta);
  mpnn_infer(E);
  VK_HIP(hipStreamSynchronize(E.stream));
  // Calibration: one leverage batch of 16 probes to rel.res 1e-8 with the
  // reference cycle, then with the learned weights and/or the tuned shape.
  // The configuration of least work (iterations x iteration work) among
  // those needing no more iterations than the reference cycle is kept.
  SolveOptions o;
  o.leverage_width = 16;
  o.precision = precision;
  o.probe_tolerance = 1e-8;
  const std::vector<uint32_t> full = full_shape(E), tuned = tuned_shape(E);
  const bool try_shape = tuned.size() < full.size();
  set_cycle_shape(E, tuned);
  auto iterations = [&](bool learned, bool shaped) -> uint64_t {
    S.active = learned;
    S.shaped = shaped;
    uint64_t solves = 0, it = 0;
    try {
      (void)time_leverage_batches(o, 1, solves, it);
    } catch (const Error &e) {
      if (!learned || (e.code != kPcgCurvature && e.code != kPcgPreconditioner)) throw;
      S.breakdown = true;
      it = UINT64_MAX;
    }
    return it;
  };
  struct Option { bool learned, shaped; uint64_t it; double work; };
  std::vector<Option> opts;
  opts.push_back({false, false, iterations(false, false), iteration_work(E, full)});
  opts.push_back({true, false, iterations(true, false), iteration_work(E, full)});
  if (try_shape) {
    opts.push_back({false, true, iterations(false, true), iteration_work(E, tuned)});
    opts.push_back({true, true, iterations(true, true), iteration_work(E, tuned)});
  }
  // learned weights only where they cut iterations against the plain weights
  // on the same shape; a shape only within 10% of the reference iterations
  const uint64_t base = opts[0].it;
  size_t best = 0;
  for (size_t k = 1; k < opts.size(); ++k) {
    if (opts[k].learned && opts[k].it >= opts[k - 1].it) continue;
    if ((double)opts[k].it > 1.1 * (double)base) continue;
    const double a = (double)opts[k].it * opts[k].work, b = (double)opts[best].it * opts[best].work;
    if (a < b || (a == b && opts[k].it < opts[best].it)) best = k;
  }
  S.calibration.assign(4, UINT32_MAX);
  for (size_t k = 0; k < opts.size(); ++k) S.calibration[k] = (uint32_t)std::min<uint64_t>(opts[k].it, UINT32_MAX);
  S.active = opts[best].learned;
  S.shaped = opts[best].shaped;
  S.plain_iterations = (uint32_t)std::min<uint64_t>(base, UINT32_MAX);
  S.learned_iterations = (uint32_t)std::min<uint64_t>(opts[best].it, UINT32_MAX);
  S.calibrated = true;
}

void DeviceEngine::smoother_default(const std::vector<double> &theta) {
  SmootherState &S = impl_->sm;
  check_params(theta);
  S.mode = 1;
  S.params = theta;
  S.active = S.calibrated = S.breakdown = false;
  S.plain_iterations = S.learned_iterations = 0;
}

void DeviceEngine::smoother_force(double omega) {
  Impl &E = *impl_;
  SmootherState &S = E.sm;
  if (!std::isfinite(omega) || omega <= 0.0) fail(kInvalidInput, "smoother", "forced omega must be positive");
  if (!has_smoothed_levels(E)) fail(kInvalidInput, "smoother", "the hierarchy has no smoothed level");
  E.ensure(M::H);
  ensure_features(E);
  for (SmootherLevel &Z : S.L) {
    DevLevel &D = E.levels[Z.level];
    hipLaunchKernelGGL(dev::k_constant_omega, grid_for(Z.n), kBlock, 0, E.stream, Z.n, omega, D.inv.p, D.linv.p,
                       D.linv32.p);
  }
  VK_HIP(hipStreamSynchronize(E.stream));
  S.mode = 1;
  S.active = true;
  S.shaped = false;
  S.calibrated = true;
  S.breakdown = false;
  S.plain_iterations = S.learned_iterations = 0;
  S.calibration.clear();
}

DeviceEngine::SmootherStatus DeviceEngine::smoother_status() const {
  const SmootherState &S = impl_->sm;
  SmootherStatus s;
  s.mode = S.mode;
  s.active = S.active;
  s.breakdown = S.breakdown;
  s.plain_iterations = S.plain_iterations;
  s.learned_iterations = S.learned_iterations;
  s.shaped = S.shaped;
  if (S.shaped) s.shape = S.shape;
  s.calibration = S.calibration;
  return s;
}

// The solve path with the smoother policy: the learned weights are used when
// calibrated to help; a solve that reports loss of positive definiteness with
// them (curvature or preconditioner breakdown) is solved again from the start
// with the plain cycle, which then stays in use.
SolveResult DeviceEngine::run(const SolveOptions &opt) {
  if (opt.exact) return run_exact(opt);
  Impl &E = *impl_;
  SmootherState &S = E.sm;
  if (S.mode == 1 && !S.calibrated) smoother_mode(1, S.params, opt.precision);
  if (!S.active && !S.shaped) return run_once(opt);
  try {
    return run_once(opt);
  } catch (const Error &e) {
    if (e.code != kPcgCurvature && e.code != kPcgPreconditioner) throw;
    S.active = false;
    S.shaped = false;
    S.breakdown = true;
    return run_once(opt);
  }
}

}  // namespace vk
