// Hybrid Laplacian and CMG hierarchy construction (host, exact).
//
// Restates the deterministic setup of the reference route CMG_FULL_V2:
//   degree-4 hybrid graph            cmg_impl.rs.in:78-363
//   canonical Laplacian              vendor/cmg/src/graph.rs:279-414
//   heavy forest, split, low degree  vendor/cmg/src/forest.rs:94-633
//   Galerkin contraction             vendor/cmg/src/coarsen.rs:359-401
//   level loop and stopping rules    vendor/cmg/src/hierarchy.rs:437-690
//   grounded LDL^T terminal          vendor/cmg/src/ldl.rs:677-1041
//   penultimate repeat recalibration vendor/cmg/src/preconditioner.rs:874-957
// The same operation order as the reference makes every label, weight and
// repeat count identical, so the device runs the reference's hierarchy.
#include "host_cmg.hpp"

#include <algorithm>
#include <numeric>

#include "common.hpp"

namespace vk {
namespace {

struct KeyWeight {
  uint64_t key;
  double w;
};

uint64_t repeat_from_nonzeros(uint64_t fine, uint64_t coarse) {
  if (coarse == 0) return 1;
  uint64_t q = fine / coarse;
  q = q > 0 ? q - 1 : 0;
  return std::max<uint64_t>(q, 1);
}

uint32_t uf_find(std::vector<uint64_t> &p, uint64_t x) {
  uint64_t r = x;
  while (p[r] != r) r = p[r];
  while (p[x] != x) {
    uint64_t nx = p[x];
    p[x] = r;
    x = nx;
  }
  return static_cast<uint32_t>(r);
}

// Labels in first-occurrence order of union-find roots that always keep the
// smaller index (forest.rs:550-607, components.rs:492-567).
uint64_t first_occurrence_labels(uint64_t n, const std::vector<std::pair<uint64_t, uint64_t>> &pairs,
                                 std::vector<uint32_t> &labels) {
  std::vector<uint64_t> p(n);
  std::iota(p.begin(), p.end(), 0);
  for (const auto &e : pairs) {
    uint64_t a = uf_find(p, e.first), b = uf_find(p, e.second);
    if (a != b) p[std::max(a, b)] = std::min(a, b);
  }
  std::vector<uint32_t> root_label(n, UINT32_MAX);
  labels.resize(n);
  uint32_t count = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t r = uf_find(p, i);
    if (root_label[r] == UINT32_MAX) root_label[r] = count++;
    labels[i] = root_label[r];
  }
  return count;
}

uint64_t component_count(const Laplacian &g) {
  std::vector<std::pair<uint64_t, uint64_t>> pairs(g.u.size());
  for (size_t e = 0; e < g.u.size(); ++e) pairs[e] = {g.u[e], g.v[e]};
  std::vector<uint32_t> labels;
  return first_occurrence_labels(g.n, pairs, labels);
}

// forest.rs:137-152 + 325-337: heaviest incident edge, ties to lowest neighbour.
void heavy_forest(const Laplacian &g, std::vector<uint64_t> &parent, std::vector<double> &weight) {
  parent.resize(g.n);
  std::iota(parent.begin(), parent.end(), 0);
  weight.assign(g.n, 0.0);
  for (size_t e = 0; e < g.u.size(); ++e) {
    uint64_t a = g.u[e], b = g.v[e];
    double w = g.w[e];
    if (w > weight[a] || (w == weight[a] && b < parent[a])) { weight[a] = w; parent[a] = b; }
    if (w > weight[b] || (w == weight[b] && a < parent[b])) { weight[b] = w; parent[b] = a; }
  }
}

// forest.rs:431-547: diameter cuts (walks longer than 5), then conductance cuts.
void split_forest(std::vector<uint64_t> &forest) {
  const uint64_t n = forest.size();
  std::vector<int64_t> anc(n, 0);
  std::vector<uint64_t> indeg(n, 0);
  std::vector<uint8_t> visited(n, 0);
  std::vector<uint64_t> walk;
  std::vector<int64_t> fresh;
  walk.reserve(64);
  fresh.reserve(64);
  for (uint64_t i = 0; i < n; ++i) indeg[forest[i]]++;
  for (uint64_t start = 0; start < n; ++start) {
    uint64_t cur = start;
    bool cont = true;
    while (cont && indeg[cur] == 0 && !visited[cur]) {
      cont = false;
      int64_t path = 0;
      walk.assign(1, cur);
      fresh.assign(1, 0);
      uint64_t k = 0;
      while (k <= 5 || visited[cur]) {
        cur = forest[cur];
        if (cur == walk[k] || (k > 0 && cur == walk[k - 1])) break;
        ++k;
        walk.push_back(cur);
        path += visited[cur] ? 0 : 1;
        fresh.push_back(path);
      }
      if (k > 5) {
        uint64_t mid = k / 2;
        forest[walk[mid]] = walk[mid];
        uint64_t next = walk[mid + 1];
        indeg[next]--;
        int64_t removed = anc[walk[mid]];
        for (uint64_t i = mid + 1; i <= k; ++i) anc[walk[i]] -= removed;
        for (uint64_t i = 0; i <= mid; ++i) {
          visited[walk[i]] = 1;
          anc[walk[i]] += fresh[i];
        }
        cur = next;
        cont = true;
      }
      if (!cont)
        for (uint64_t i = 0; i <= k; ++i) {
          anc[walk[i]] += fresh[i];
          visited[walk[i]] = 1;
        }
    }
  }
  for (uint64_t start = 0; start < n; ++start) {
    uint64_t cur = start;
    bool again = true;
    while (again && indeg[cur] == 0) {
      again = false;
      uint64_t prev = cur;
      for (;;) {
        uint64_t next = forest[cur];
        if (next == cur || next == prev) break;
        if (anc[cur] > 2 && anc[next] - anc[cur] > 2) {
          forest[cur] = cur;
          indeg[next]--;
          int64_t removed = anc[cur];
          uint64_t ap = cur, ac = next;
          for (;;) {
            anc[ac] -= removed;
            uint64_t an = forest[ac];
            if (an == ac || an == ap) break;
            ap = ac;
            ac = an;
          }
          cur = next;
          again = true;
          break;
        }
        prev = cur;
        cur = next;
      }
    }
  }
}

// forest.rs:339-375
void low_degree_correction(const Laplacian &g, double threshold, const std::vector<double> &selected,
                           std::vector<uint64_t> &parent) {
  bool any = false;
  for (uint64_t i = 0; i < g.n && !any; ++i)
    any = g.diag[i] > 0.0 && selected[i] / g.diag[i] < threshold;
  if (!any) return;
  std::vector<double> incident(g.n, 0.0);
  for (uint64_t i = 0; i < g.n; ++i)
    if (parent[i] != i) {
      incident[i] += selected[i];
      incident[parent[i]] += selected[i];
    }
  for (uint64_t i = 0; i < g.n; ++i)
    if (g.diag[i] > 0.0 && incident[i] / g.diag[i] < threshold) parent[i] = i;
}

uint64_t aggregate(const Laplacian &g, double threshold, std::vector<uint32_t> &labels) {
  std::vector<uint64_t> parent;
  std::vector<double> weight;
  heavy_forest(g, parent, weight);
  split_forest(parent);
  low_degree_correction(g, threshold, weight, parent);
  std::vector<std::pair<uint64_t, uint64_t>> pairs(g.n);
  for (uint64_t i = 0; i < g.n; ++i) pairs[i] = {i, parent[i]};
  return first_occurrence_labels(g.n, pairs, labels);
}

Laplacian contract(const Laplacian &g, const std::vector<uint32_t> &labels, uint64_t coarse) {
  std::vector<uint64_t> keys;
  std::vector<double> ws;
  keys.reserve(g.u.size());
  ws.reserve(g.u.size());
  for (size_t e = 0; e < g.u.size(); ++e) {
    uint32_t a = labels[g.u[e]], b = labels[g.v[e]];
    if (a == b) continue;
    keys.push_back((static_cast<uint64_t>(std::min(a, b)) << 32) | std::max(a, b));
    ws.push_back(g.w[e]);
  }
  return Laplacian::canonical(coarse, keys, ws);
}

void fill_layout(Level &L) {
  const Laplacian &g = L.graph;
  L.row_ptr.assign(g.n + 1, 0);
  for (size_t e = 0; e < g.u.size(); ++e) {
    L.row_ptr[g.u[e] + 1]++;
    L.row_ptr[g.v[e] + 1]++;
  }
  for (uint64_t i = 0; i < g.n; ++i) L.row_ptr[i + 1] += L.row_ptr[i];
  if (g.u.size() * 2 > UINT32_MAX) fail(kResourceLimit, "cmg_layout", "level exceeds 32-bit CSR offsets");
  L.col.resize(2 * g.u.size());
  L.val.resize(2 * g.u.size());
  std::vector<uint32_t> cur(L.row_ptr.begin(), L.row_ptr.end() - 1);
  // canonical edge order appends each row's neighbours in ascending order
  for (size_t e = 0; e < g.u.size(); ++e) {
    uint32_t a = cur[g.u[e]]++, b = cur[g.v[e]]++;
    L.col[a] = g.v[e]; L.val[a] = g.w[e];
    L.col[b] = g.u[e]; L.val[b] = g.w[e];
  }
  L.inv_diag.resize(g.n);
  for (uint64_t i = 0; i < g.n; ++i) L.inv_diag[i] = g.diag[i] > 0.0 ? 0.5 / g.diag[i] : 0.0;
  if (!L.labels.empty()) {
    L.agg_ptr.assign(L.coarse + 1, 0);
    for (uint32_t l : L.labels) L.agg_ptr[l + 1]++;
    for (uint64_t c = 0; c < L.coarse; ++c) L.agg_ptr[c + 1] += L.agg_ptr[c];
    L.agg_members.resize(g.n);
    std::vector<uint32_t> ac(L.agg_ptr.begin(), L.agg_ptr.end() - 1);
    for (uint64_t i = 0; i < g.n; ++i) L.agg_members[ac[L.labels[i]]++] = static_cast<uint32_t>(i);
  }
}

// ldl.rs:677-803 + 999-1041: ground the last vertex of every component,
// order by static pattern count then index, dense LDL^T.
void factor_terminal(const Laplacian &g, Hierarchy &H) {
  std::vector<std::pair<uint64_t, uint64_t>> pairs(g.u.size());
  for (size_t e = 0; e < g.u.size(); ++e) pairs[e] = {g.u[e], g.v[e]};
  std::vector<uint32_t> comp;
  uint64_t count = first_occurrence_labels(g.n, pairs, comp);
  H.anchors.assign(count, 0);
  for (uint64_t i = 0; i < g.n; ++i) H.anchors[comp[i]] = i;
  std::vector<uint8_t> anchor(g.n, 0);
  for (uint64_t a : H.anchors) anchor[a] = 1;
  std::vector<uint64_t> pattern(g.n, 0);
  for (uint64_t i = 0; i < g.n; ++i)
    if (!anchor[i]) pattern[i] = 1;
  for (size_t e = 0; e < g.u.size(); ++e)
    if (!anchor[g.u[e]] && !anchor[g.v[e]]) { pattern[g.u[e]]++; pattern[g.v[e]]++; }
  H.perm.clear();
  for (uint64_t i = 0; i < g.n; ++i)
    if (!anchor[i]) H.perm.push_back(i);
  std::sort(H.perm.begin(), H.perm.end(), [&](uint64_t a, uint64_t b) {
    return pattern[a] != pattern[b] ? pattern[a] < pattern[b] : a < b;
  });
  const uint64_t dim = H.perm.size();
  H.ldl_dim = dim;
  std::vector<double> A(dim * dim, 0.0), Lw(dim * dim, 0.0);
  std::vector<uint64_t> fidx(g.n, UINT64_MAX);
  for (uint64_t k = 0; k < dim; ++k) {
    fidx[H.perm[k]] = k;
    A[k * dim + k] = g.diag[H.perm[k]];
  }
  for (size_t e = 0; e < g.u.size(); ++e) {
    uint64_t x = fidx[g.u[e]], y = fidx[g.v[e]];
    if (x == UINT64_MAX || y == UINT64_MAX) continue;
    A[x * dim + y] -= g.w[e];
    A[y * dim + x] -= g.w[e];
  }
  H.ldl_diag.assign(dim, 0.0);
  for (uint64_t r = 0; r < dim; ++r) Lw[r * dim + r] = 1.0;
  for (uint64_t c = 0; c < dim; ++c) {
    double pivot = A[c * dim + c];
    for (uint64_t p = 0; p < c; ++p) {
      double v = Lw[c * dim + p];
      pivot -= v * v * H.ldl_diag[p];
    }
    if (!std::isfinite(pivot) || pivot <= 0.0)
      fail(kCmgSetup, "cmg_terminal", "nonpositive pivot in the grounded terminal factor");
    H.ldl_diag[c] = pivot;
    for (uint64_t r = c + 1; r < dim; ++r) {
      double v = A[r * dim + c];
      for (uint64_t p = 0; p < c; ++p) v -= Lw[r * dim + p] * Lw[c * dim + p] * H.ldl_diag[p];
      Lw[r * dim + c] = v / pivot;
    }
  }
  uint64_t strict = 0;
  for (uint64_t r = 0; r < dim; ++r)
    for (uint64_t c = 0; c < r; ++c) strict += Lw[r * dim + c] != 0.0;
  H.factor_nnz = dim + strict;
  H.ldl_lower = Lw;
  // Grounded inverse: column k of (L D L^T)^{-1}, scattered to graph indices.
  H.terminal_inverse.assign(g.n * g.n, 0.0);
  std::vector<double> y(dim), x(dim);
  for (uint64_t k = 0; k < dim; ++k) {
    for (uint64_t r = 0; r < dim; ++r) {
      double s = (r == k) ? 1.0 : 0.0;
      for (uint64_t c = 0; c < r; ++c) s -= Lw[r * dim + c] * y[c];
      y[r] = s;
    }
    for (uint64_t r = 0; r < dim; ++r) y[r] /= H.ldl_diag[r];
    for (uint64_t r = dim; r-- > 0;) {
      double s = y[r];
      for (uint64_t c = r + 1; c < dim; ++c) s -= Lw[c * dim + r] * x[c];
      x[r] = s;
    }
    for (uint64_t r = 0; r < dim; ++r) H.terminal_inverse[H.perm[r] * g.n + H.perm[k]] = x[r];
  }
}

}  // namespace

Laplacian Laplacian::canonical(uint64_t n, std::vector<uint64_t> &keys, std::vector<double> &weights) {
  const size_t m = keys.size();
  std::vector<KeyWeight> kw(m);
  for (size_t i = 0; i < m; ++i) kw[i] = {keys[i], weights[i]};
  std::sort(kw.begin(), kw.end(), [](const KeyWeight &a, const KeyWeight &b) {
    return a.key != b.key ? a.key < b.key : total_key(a.w) < total_key(b.w);
  });
  Laplacian g;
  g.n = n;
  g.diag.assign(n, 0.0);
  for (size_t i = 0; i < m;) {
    uint64_t key = kw[i].key;
    Neumaier s;
    while (i < m && kw[i].key == key) s.add(kw[i++].w);
    double w = s.value();
    if (!std::isfinite(w) || w <= 0.0) fail(kCmgSetup, "cmg_graph", "merged edge weight is invalid");
    uint32_t a = static_cast<uint32_t>(key >> 32), b = static_cast<uint32_t>(key);
    g.u.push_back(a);
    g.v.push_back(b);
    g.w.push_back(w);
    g.diag[a] += w;
    g.diag[b] += w;
  }
  uint64_t dnnz = 0;
  double maxdeg = 0.0;
  for (double d : g.diag) {
    dnnz += d != 0.0;
    maxdeg = std::max(maxdeg, d);
  }
  g.nnz = dnnz + 2 * g.u.size();
  g.norm_bound = 2.0 * maxdeg;
  return g;
}

Laplacian laplacian_from_edges(uint64_t n, uint64_t m, const uint32_t *u, const uint32_t *v, const double *w) {
  std::vector<uint64_t> keys(m);
  std::vector<double> ws(w, w + m);
  for (uint64_t e = 0; e < m; ++e) {
    if (u[e] >= n || v[e] >= n || u[e] == v[e] || !std::isfinite(w[e]) || w[e] <= 0.0)
      fail(kCmgSetup, "cmg_graph", "invalid edge " + std::to_string(e));
    keys[e] = (static_cast<uint64_t>(std::min(u[e], v[e])) << 32) | std::max(u[e], v[e]);
  }
  return Laplacian::canonical(n, keys, ws);
}

HybridSystem build_hybrid(const Prepared &p) {
  HybridSystem hs;
  const uint64_t F = p.firms, W = p.workers;
  hs.firms = F;
  struct Contribution { uint64_t key; uint64_t order; double w; };
  std::vector<Contribution> raw;
  raw.reserve(p.cells * 2);
  uint32_t aux = 0;
  for (uint64_t w = 0; w < W; ++w) {
    uint64_t b = p.worker_ptr[w], e = p.worker_ptr[w + 1], deg = e - b;
    double mass = 0.0;
    for (uint64_t c = b; c < e; ++c) mass += p.cell_weight[c];
    if (!std::isfinite(mass) || mass <= 0.0) fail(kInvalidWeight, "cmg_graph", "worker information mass is invalid");
    if (deg <= 1) continue;
    if (deg <= 4) {
      for (uint64_t l = b; l < e; ++l)
        for (uint64_t r = l + 1; r < e; ++r) {
          uint32_t a = p.cell_firm[l], c = p.cell_firm[r];
          raw.push_back({(static_cast<uint64_t>(std::min(a, c)) << 32) | std::max(a, c), raw.size(),
                         p.cell_weight[l] * p.cell_weight[r] / mass});
        }
    } else {
      uint32_t av = static_cast<uint32_t>(F + aux++);
      for (uint64_t c = b; c < e; ++c)
        raw.push_back({(static_cast<uint64_t>(p.cell_firm[c]) << 32) | av, raw.size(), p.cell_weight[c]});
    }
  }
  hs.aux = aux;
  // BTreeMap accumulation (cmg_impl.rs.in:634-665): key order, insertion-order sums
  std::sort(raw.begin(), raw.end(), [](const Contribution &a, const Contribution &b) {
    return a.key != b.key ? a.key < b.key : a.order < b.order;
  });
  std::vector<uint64_t> keys;
  std::vector<double> ws;
  for (size_t i = 0; i < raw.size();) {
    uint64_t key = raw[i].key;
    double s = raw[i++].w;
    while (i < raw.size() && raw[i].key == key) s = s + raw[i++].w;
    keys.push_back(key);
    ws.push_back(s);
  }
  const uint64_t n = F + aux;
  hs.graph = Laplacian::canonical(n, keys, ws);
  // auxiliary incidence sorted by (weight desc, neighbour structural key, edge)
  std::vector<uint64_t> primary(F, 0);
  {
    std::vector<uint32_t> fo(F);
    std::iota(fo.begin(), fo.end(), 0u);
    auto sig_cmp = [&](uint32_t x, uint32_t y) {
      uint64_t i = p.firm_ptr[x], ie = p.firm_ptr[x + 1], j = p.firm_ptr[y], je = p.firm_ptr[y + 1];
      for (; i < ie && j < je; ++i, ++j) {
        uint32_t ci = p.firm_cells[i], cj = p.firm_cells[j];
        if (p.cell_worker[ci] != p.cell_worker[cj]) return p.cell_worker[ci] < p.cell_worker[cj] ? -1 : 1;
        uint64_t bi = bits_of(p.cell_weight[ci]), bj = bits_of(p.cell_weight[cj]);
        if (bi != bj) return bi < bj ? -1 : 1;
      }
      if (i < ie) return 1;
      if (j < je) return -1;
      return 0;
    };
    std::sort(fo.begin(), fo.end(), [&](uint32_t a, uint32_t b) { return sig_cmp(a, b) < 0; });
    uint64_t rank = 0;
    for (uint64_t i = 0; i < F; ++i) {
      if (i == 0 || sig_cmp(fo[i - 1], fo[i]) != 0) ++rank;
      primary[fo[i]] = rank;
    }
  }
  hs.aux_ptr.assign(aux + 1, 0);
  const Laplacian &g = hs.graph;
  for (size_t e = 0; e < g.u.size(); ++e)
    if (g.v[e] >= F) hs.aux_ptr[g.v[e] - F + 1]++;
  for (uint64_t a = 0; a < aux; ++a) hs.aux_ptr[a + 1] += hs.aux_ptr[a];
  struct Arc { uint32_t nbr, edge; double w; };
  std::vector<Arc> arcs(hs.aux_ptr[aux]);
  std::vector<uint64_t> cur(hs.aux_ptr.begin(), hs.aux_ptr.end() - 1);
  for (size_t e = 0; e < g.u.size(); ++e)
    if (g.v[e] >= F) arcs[cur[g.v[e] - F]++] = {g.u[e], static_cast<uint32_t>(e), g.w[e]};
  for (uint64_t a = 0; a < aux; ++a)
    std::sort(arcs.begin() + hs.aux_ptr[a], arcs.begin() + hs.aux_ptr[a + 1], [&](const Arc &x, const Arc &y) {
      if (total_key(x.w) != total_key(y.w)) return total_key(x.w) > total_key(y.w);
      if (primary[x.nbr] != primary[y.nbr]) return primary[x.nbr] < primary[y.nbr];
      return x.edge < y.edge;
    });
  for (const Arc &a : arcs) {
    hs.aux_nbr.push_back(a.nbr);
    hs.aux_w.push_back(a.w);
  }
  return hs;
}

Hierarchy build_hierarchy(const Laplacian &graph, const CmgOptions &opt) {
  Hierarchy H;
  const uint64_t initial_nnz = graph.nnz;
  uint64_t cumulative = 0;
  Laplacian current = graph;
  for (;;) {
    const uint64_t n = current.n;
    Level L;
    if (n == 0 || n <= 1 || n < opt.direct_threshold) {
      H.terminal_reason = kDirect;
      L.graph = std::move(current);
      L.terminal = kDirect;
      H.levels.push_back(std::move(L));
      break;
    }
    std::vector<uint32_t> labels;
    uint64_t coarse = aggregate(current, opt.low_degree_threshold, labels);
    int terminal = kNotTerminal;
    if (coarse == 1) {
      terminal = kFullContraction;
    } else {
      cumulative += current.nnz;
      if (coarse >= (n > 0 ? n - 1 : 0)) terminal = kStagnated;
      else if (static_cast<double>(cumulative) > opt.max_hierarchy_nnz_factor * static_cast<double>(initial_nnz))
        terminal = kFill;
      else if (H.levels.size() + 1 >= opt.max_levels)
        terminal = kMaxLevels;
    }
    if (terminal != kNotTerminal) {
      H.terminal_reason = terminal;
      L.graph = std::move(current);
      L.labels = std::move(labels);
      L.coarse = coarse;
      L.terminal = terminal;
      H.levels.push_back(std::move(L));
      break;
    }
    Laplacian next = contract(current, labels, coarse);
    L.repeat = repeat_from_nonzeros(current.nnz, next.nnz);
    L.graph = std::move(current);
    L.labels = std::move(labels);
    L.coarse = coarse;
    H.levels.push_back(std::move(L));
    current = std::move(next);
  }
  H.cumulative_nnz = cumulative;
  if (H.terminal_reason == kDirect) {
    factor_terminal(H.levels.back().graph, H);
    if (H.levels.size() >= 2) {
      Level &pen = H.levels[H.levels.size() - 2];
      pen.repeat = repeat_from_nonzeros(pen.graph.nnz, H.factor_nnz);
    }
  }
  for (Level &L : H.levels) fill_layout(L);
  H.finest_components = component_count(H.levels[0].graph);
  return H;
}

ChunkTable chunk_segments(const std::vector<uint32_t> &ptr, uint32_t chunk, uint32_t group) {
  if (chunk == 0) fail(kInvalidInput, "device_layout", "chunk size must be positive");
  if (group == 0 || group >= kSingleChunk - kGroupChunk) fail(kInvalidInput, "device_layout", "bad group size");
  ChunkTable t;
  const uint64_t segments = ptr.empty() ? 0 : ptr.size() - 1;
  const uint64_t short_len = chunk / 4;
  if (group > 1) t.ptr = ptr;
  for (uint64_t s = 0; s < segments; ++s) {
    const uint64_t begin = ptr[s], len = ptr[s + 1] - begin;
    if (group > 1 && len <= short_len) {
      uint64_t e = s + 1;
      while (e < segments && e - s < group && ptr[e + 1] - ptr[e] <= short_len) ++e;
      if (e - s > 1) {
        t.desc.push_back(static_cast<uint32_t>(s));
        t.desc.push_back(ptr[s]);
        t.desc.push_back(ptr[e]);
        t.desc.push_back(kGroupChunk + static_cast<uint32_t>(e - s));
        s = e - 1;
        continue;
      }
    }
    const uint64_t parts = std::max<uint64_t>(1, (len + chunk - 1) / chunk);
    const uint32_t first = static_cast<uint32_t>(t.chunks());
    const uint32_t multi = parts > 1 ? t.multi() : kSingleChunk;
    for (uint64_t q = 0; q < parts; ++q) {
      t.desc.push_back(static_cast<uint32_t>(s));
      t.desc.push_back(static_cast<uint32_t>(begin + len * q / parts));
      t.desc.push_back(static_cast<uint32_t>(begin + len * (q + 1) / parts));
      t.desc.push_back(multi);
    }
    if (parts > 1) {
      if (t.multi() >= kGroupChunk) fail(kInvalidInput, "device_layout", "too many multi-chunk segments");
      t.mrange.push_back(first);
      t.mrange.push_back(static_cast<uint32_t>(t.chunks()));
    }
  }
  return t;
}

std::vector<LevelLayout> device_layout(const Hierarchy &H, uint32_t chunk) {
  const size_t L = H.levels.size();
  std::vector<LevelLayout> out(L);
  for (size_t li = L; li-- > 0;) {
    const Level &lv = H.levels[li];
    LevelLayout &d = out[li];
    const uint32_t n = static_cast<uint32_t>(lv.graph.n);
    d.perm.resize(n);
    d.pos.resize(n);
    const bool has_child = li + 1 < L;
    if (has_child) {
      // stable counting sort of the vertices by the renumbered coarse index
      const std::vector<uint32_t> &cpos = out[li + 1].pos;
      const uint32_t nc = static_cast<uint32_t>(cpos.size());
      if (lv.labels.size() != n || lv.coarse != nc) fail(kInvariant, "device_layout", "labels do not match the coarse level");
      d.agg_ptr.assign(nc + 1, 0);
      for (uint32_t v = 0; v < n; ++v) d.agg_ptr[cpos[lv.labels[v]] + 1]++;
      for (uint32_t c = 0; c < nc; ++c) d.agg_ptr[c + 1] += d.agg_ptr[c];
      std::vector<uint32_t> cur(d.agg_ptr.begin(), d.agg_ptr.end() - 1);
      for (uint32_t v = 0; v < n; ++v) d.perm[cur[cpos[lv.labels[v]]]++] = v;
    } else {
      std::iota(d.perm.begin(), d.perm.end(), 0u);
    }
    for (uint32_t p = 0; p < n; ++p) d.pos[d.perm[p]] = p;
    d.row_ptr.assign(n + 1, 0);
    for (uint32_t p = 0; p < n; ++p) {
      const uint32_t v = d.perm[p];
      d.row_ptr[p + 1] = d.row_ptr[p] + (lv.row_ptr[v + 1] - lv.row_ptr[v]);
    }
    d.col.resize(lv.col.size());
    d.val.resize(lv.val.size());
    d.inv.resize(n);
    for (uint32_t p = 0; p < n; ++p) {
      const uint32_t v = d.perm[p];
      uint32_t k = d.row_ptr[p];
      for (uint32_t e = lv.row_ptr[v]; e < lv.row_ptr[v + 1]; ++e, ++k) {
        d.col[k] = d.pos[lv.col[e]];
        d.val[k] = lv.val[e];
      }
      d.inv[p] = lv.inv_diag[v];
    }
    if (has_child) {
      const std::vector<uint32_t> &cpos = out[li + 1].pos;
      d.lab.resize(n);
      for (uint32_t p = 0; p < n; ++p) d.lab[p] = cpos[lv.labels[d.perm[p]]];
      d.aggs = chunk_segments(d.agg_ptr, chunk);
    }
    d.rows = chunk_segments(d.row_ptr, chunk);
  }
  return out;
}

}  // namespace vk
