// Host preparation for the MI355X leave-out engine.
//
// Restates, in the reference's exact selection and ordering rules:
//   validation                types.rs:185-263
//   canonical identifiers     problem.rs:102-172, 696-746 (redense)
//   topology checksum         problem.rs:1213-1248
//   leave-out connected set   graph.rs:262-512 (largest component by
//                             (firms, mass), movers, fixed point of
//                             one-unit workers / worker articulations /
//                             bridge deletion units)
//   compression               problem.rs:206-606
//   JLA plan                  jla.rs:48-105, 413-790
// The selection works on precomputed coordinate ids so each fixed-point pass
// is a linear scan plus one low-link DFS.
#include "host_prepare.hpp"

#include <algorithm>
#include <chrono>
#include <numeric>

#include "common.hpp"

namespace vk {
namespace {

// Dense ids in sorted distinct order; returns level count.
std::vector<uint32_t> redense(const uint64_t *v, uint64_t n, uint64_t &levels) {
  std::vector<uint64_t> sorted(v, v + n);
  std::sort(sorted.begin(), sorted.end());
  sorted.erase(std::unique(sorted.begin(), sorted.end()), sorted.end());
  if (sorted.size() > UINT32_MAX) fail(kResourceLimit, "canonicalize", "cardinality exceeds u32");
  std::vector<uint32_t> out(n);
  for (uint64_t i = 0; i < n; ++i)
    out[i] = static_cast<uint32_t>(std::lower_bound(sorted.begin(), sorted.end(), v[i]) - sorted.begin());
  levels = sorted.size();
  return out;
}

uint64_t fnv_checksum(uint64_t n, const uint32_t *w, const uint32_t *f, const uint32_t *d,
                      const uint64_t *q) {
  struct Rec { uint32_t w, f, d; uint64_t q; };
  std::vector<Rec> r(n);
  for (uint64_t i = 0; i < n; ++i) r[i] = {w[i], f[i], d[i], q[i]};
  std::sort(r.begin(), r.end(), [](const Rec &a, const Rec &b) {
    if (a.w != b.w) return a.w < b.w;
    if (a.f != b.f) return a.f < b.f;
    if (a.d != b.d) return a.d < b.d;
    return a.q < b.q;
  });
  uint64_t h = 0xcbf29ce484222325ULL;
  auto word = [&](uint64_t x) {
    for (int i = 0; i < 8; ++i) {
      h ^= (x >> (8 * i)) & 0xffu;
      h *= 0x00000100000001b3ULL;
    }
  };
  for (const Rec &x : r) {
    word(x.w); word(x.f); word(x.d); word(x.q);
  }
  return h;
}

struct UnionFind {
  std::vector<uint32_t> p;
  explicit UnionFind(uint64_t n) : p(n) { std::iota(p.begin(), p.end(), 0u); }
  uint32_t find(uint32_t x) {
    uint32_t r = x;
    while (p[r] != r) r = p[r];
    while (p[x] != r) { uint32_t nx = p[x]; p[x] = r; x = nx; }
    return r;
  }
  void unite(uint32_t a, uint32_t b) {
    a = find(a); b = find(b);
    if (a != b) p[std::max(a, b)] = std::min(a, b);
  }
};

// Coordinate-level view of the canonical input used by graph selection.
struct Graph {
  uint64_t rows, W, F, D, K;  // K distinct coordinates (worker, firm)
  const uint32_t *w, *f, *d;
  const uint64_t *freq;
  std::vector<uint32_t> row_coord;       // coordinate of each row
  std::vector<uint32_t> coord_w, coord_f;
  std::vector<uint32_t> del_coord;       // coordinate of each deletion id
};

// Iterative low-link DFS over an undirected multigraph in CSR form.
// mode 0 marks articulation vertices, mode 1 marks bridge edges.
void lowlink(uint64_t nodes, const std::vector<uint64_t> &off, const std::vector<uint32_t> &to,
             const std::vector<uint32_t> &eid, int mode, std::vector<uint8_t> &flag) {
  std::vector<uint32_t> disc(nodes, 0), low(nodes, 0), parent(nodes, UINT32_MAX), kids(nodes, 0);
  struct Frame { uint32_t node, pedge; uint64_t next; };
  std::vector<Frame> st;
  uint32_t clock = 0;
  for (uint64_t root = 0; root < nodes; ++root) {
    if (off[root] == off[root + 1] || disc[root]) continue;
    disc[root] = low[root] = ++clock;
    st.push_back({static_cast<uint32_t>(root), UINT32_MAX, off[root]});
    while (!st.empty()) {
      Frame &top = st.back();
      uint32_t node = top.node;
      if (top.next < off[node + 1]) {
        uint64_t k = top.next++;
        uint32_t t = to[k], e = eid[k];
        if (e == top.pedge) continue;
        if (!disc[t]) {
          parent[t] = node;
          kids[node]++;
          disc[t] = low[t] = ++clock;
          st.push_back({t, e, off[t]});
        } else {
          low[node] = std::min(low[node], disc[t]);
        }
      } else {
        uint32_t pe = top.pedge, par = parent[node];
        st.pop_back();
        if (par != UINT32_MAX) {
          if (mode == 1 && low[node] > disc[par]) flag[pe] = 1;
          low[par] = std::min(low[par], low[node]);
          if (mode == 0 && parent[par] != UINT32_MAX && low[node] >= disc[par]) flag[par] = 1;
        }
      }
    }
    if (mode == 0) flag[root] = kids[root] > 1;
  }
}

void build_csr(uint64_t nodes, const std::vector<uint32_t> &eu, const std::vector<uint32_t> &ev,
               std::vector<uint64_t> &off, std::vector<uint32_t> &to, std::vector<uint32_t> &eid) {
  uint64_t m = eu.size();
  off.assign(nodes + 1, 0);
  for (uint64_t e = 0; e < m; ++e) { off[eu[e] + 1]++; off[ev[e] + 1]++; }
  for (uint64_t i = 0; i < nodes; ++i) off[i + 1] += off[i];
  to.resize(2 * m);
  eid.resize(2 * m);
  std::vector<uint64_t> cur(off.begin(), off.end() - 1);
  for (uint64_t e = 0; e < m; ++e) {
    uint64_t a = cur[eu[e]]++, b = cur[ev[e]]++;
    to[a] = ev[e]; eid[a] = static_cast<uint32_t>(e);
    to[b] = eu[e]; eid[b] = static_cast<uint32_t>(e);
  }
}

// graph.rs:625-731: keep the unique largest component by (firms, mass).
uint64_t largest_component(const Graph &g, std::vector<uint8_t> &active) {
  uint64_t nodes = g.W + g.F;
  std::vector<uint8_t> cact(g.K, 0);
  for (uint64_t r = 0; r < g.rows; ++r)
    if (active[r]) cact[g.row_coord[r]] = 1;
  UnionFind uf(nodes);
  std::vector<uint8_t> touched(nodes, 0);
  for (uint64_t c = 0; c < g.K; ++c)
    if (cact[c]) {
      uf.unite(g.coord_w[c], static_cast<uint32_t>(g.W + g.coord_f[c]));
      touched[g.coord_w[c]] = touched[g.W + g.coord_f[c]] = 1;
    }
  std::vector<uint32_t> label(nodes, UINT32_MAX);
  uint32_t count = 0;
  for (uint64_t i = 0; i < nodes; ++i) {
    if (!touched[i]) continue;
    uint32_t r = uf.find(static_cast<uint32_t>(i));
    if (label[r] == UINT32_MAX) label[r] = count++;
    label[i] = label[r];
  }
  if (!count) fail(kGraphEmpty, "graph", "no connected worker-firm component remains");
  std::vector<uint64_t> firms(count, 0), mass(count, 0);
  for (uint64_t f = 0; f < g.F; ++f)
    if (label[g.W + f] != UINT32_MAX) firms[label[g.W + f]]++;
  for (uint64_t r = 0; r < g.rows; ++r)
    if (active[r]) mass[label[g.w[r]]] += g.freq[r];
  uint32_t best = 0;
  for (uint32_t c = 1; c < count; ++c)
    if (std::make_pair(firms[c], mass[c]) > std::make_pair(firms[best], mass[best])) best = c;
  uint32_t winners = 0;
  for (uint32_t c = 0; c < count; ++c) winners += firms[c] == firms[best] && mass[c] == mass[best];
  if (winners != 1)
    fail(kGraphUnidentified, "graph", "multiple connected components tie for largest-component selection");
  for (uint64_t r = 0; r < g.rows; ++r)
    if (active[r] && label[g.w[r]] != best) active[r] = 0;
  return count;
}

// graph.rs:793-808: distinct active deletion units per worker.
std::vector<uint32_t> units_per_worker(const Graph &g, const std::vector<uint8_t> &active) {
  std::vector<uint8_t> seen(g.D, 0);
  std::vector<uint32_t> cnt(g.W, 0);
  for (uint64_t r = 0; r < g.rows; ++r)
    if (active[r] && !seen[g.d[r]]) {
      seen[g.d[r]] = 1;
      cnt[g.w[r]]++;
    }
  return cnt;
}

// graph.rs:872-886: articulation workers of the coordinate graph.
std::vector<uint8_t> articulation_workers(const Graph &g, const std::vector<uint8_t> &active) {
  std::vector<uint8_t> cact(g.K, 0);
  for (uint64_t r = 0; r < g.rows; ++r)
    if (active[r]) cact[g.row_coord[r]] = 1;
  std::vector<uint32_t> eu, ev;
  for (uint64_t c = 0; c < g.K; ++c)
    if (cact[c]) {
      eu.push_back(g.coord_w[c]);
      ev.push_back(static_cast<uint32_t>(g.W + g.coord_f[c]));
    }
  uint64_t nodes = g.W + g.F;
  std::vector<uint64_t> off;
  std::vector<uint32_t> to, eid;
  build_csr(nodes, eu, ev, off, to, eid);
  std::vector<uint8_t> flag(nodes, 0);
  lowlink(nodes, off, to, eid, 0, flag);
  flag.resize(g.W);
  return flag;
}

// graph.rs:938-1021: deletion units that are bridges of the unit multigraph.
uint64_t bridge_units(const Graph &g, const std::vector<uint8_t> &active, std::vector<uint8_t> &is_bridge) {
  std::vector<uint8_t> seen(g.D, 0);
  for (uint64_t r = 0; r < g.rows; ++r)
    if (active[r]) seen[g.d[r]] = 1;
  std::vector<uint32_t> eu, ev, ed;
  for (uint64_t d = 0; d < g.D; ++d)
    if (seen[d]) {
      uint32_t c = g.del_coord[d];
      eu.push_back(g.coord_w[c]);
      ev.push_back(static_cast<uint32_t>(g.W + g.coord_f[c]));
      ed.push_back(static_cast<uint32_t>(d));
    }
  is_bridge.assign(g.D, 0);
  if (eu.empty()) return 0;
  uint64_t nodes = g.W + g.F;
  std::vector<uint64_t> off;
  std::vector<uint32_t> to, eid;
  build_csr(nodes, eu, ev, off, to, eid);
  std::vector<uint8_t> flag(eu.size(), 0);
  lowlink(nodes, off, to, eid, 1, flag);
  uint64_t bridges = 0;
  for (uint64_t k = 0; k < eu.size(); ++k)
    if (flag[k]) { is_bridge[ed[k]] = 1; ++bridges; }
  return bridges;
}

uint64_t count_active(const std::vector<uint8_t> &a) {
  uint64_t n = 0;
  for (uint8_t x : a) n += x;
  return n;
}

// graph.rs:262-512
void select(const Graph &g, std::vector<uint8_t> &active, SelectionReceipt &rec) {
  active.assign(g.rows, 1);
  rec.input_rows = g.rows;
  uint64_t comps = largest_component(g, active);
  rec.initial_components = rec.maximum_components = comps;
  rec.initial_component_rows = count_active(active);
  auto upw = units_per_worker(g, active);
  for (uint64_t r = 0; r < g.rows; ++r)
    if (active[r]) active[r] = upw[g.w[r]] > 1;
  rec.mover_input_rows = count_active(active);
  if (!rec.mover_input_rows)
    fail(kGraphEmpty, "graph", "no observations remain in the mover target population");
  rec.maximum_components = std::max<uint64_t>(rec.maximum_components, largest_component(g, active));
  {
    std::vector<uint8_t> seen(g.D, 0);
    for (uint64_t r = 0; r < g.rows; ++r)
      if (active[r] && !seen[g.d[r]]) { seen[g.d[r]] = 1; rec.initial_deletion_edges++; }
  }
  uint64_t bound = g.W + rec.initial_deletion_edges + 1;
  std::vector<uint8_t> is_bridge;
  for (;;) {
    if (rec.fixed_point_iterations > bound)
      fail(kGraphCertificate, "graph", "sample fixed-point pruning exceeded its finite iteration bound");
    rec.maximum_components = std::max<uint64_t>(rec.maximum_components, largest_component(g, active));
    if (!count_active(active)) fail(kGraphEmpty, "graph", "no component remains after fixed-point pruning");
    upw = units_per_worker(g, active);
    uint64_t removed = 0;
    for (uint64_t w = 0; w < g.W; ++w) removed += upw[w] == 1;
    if (removed) {
      for (uint64_t r = 0; r < g.rows; ++r)
        if (active[r] && upw[g.w[r]] == 1) active[r] = 0;
      rec.insufficient_workers_removed += removed;
      rec.degree_iterations++;
      rec.fixed_point_iterations++;
      continue;
    }
    auto art = articulation_workers(g, active);
    uint64_t arts = 0;
    for (uint8_t x : art) arts += x;
    if (arts) {
      for (uint64_t r = 0; r < g.rows; ++r)
        if (active[r] && art[g.w[r]]) active[r] = 0;
      rec.articulation_workers_removed += arts;
      rec.articulation_iterations++;
      rec.fixed_point_iterations++;
      continue;
    }
    uint64_t bridges = bridge_units(g, active, is_bridge);
    if (!bridges) break;
    uint64_t removed_rows = 0;
    for (uint64_t r = 0; r < g.rows; ++r)
      if (active[r] && is_bridge[g.d[r]]) { active[r] = 0; ++removed_rows; }
    rec.bridge_units_removed += bridges;
    rec.bridge_rows_removed += removed_rows;
    rec.bridge_iterations++;
    rec.fixed_point_iterations++;
  }
  rec.maximum_components = std::max<uint64_t>(rec.maximum_components, largest_component(g, active));
  if (bridge_units(g, active, is_bridge))
    fail(kGraphCertificate, "graph", "final retained deletion-unit multigraph still contains a bridge");
  {
    auto art = articulation_workers(g, active);
    for (uint8_t x : art)
      if (x) fail(kGraphCertificate, "graph", "final retained graph still contains a worker articulation");
    auto u = units_per_worker(g, active);
    for (uint32_t x : u)
      if (x == 1) fail(kGraphCertificate, "graph", "final retained graph still contains a one-deletion-unit worker");
  }
  rec.retained_rows = count_active(active);
  if (!rec.retained_rows) fail(kGraphEmpty, "graph", "no component remains after fixed-point graph pruning");
  rec.retained_physical_mass = 0;
  std::vector<uint8_t> seen(g.D, 0);
  for (uint64_t r = 0; r < g.rows; ++r)
    if (active[r]) {
      rec.retained_physical_mass += g.freq[r];
      if (!seen[g.d[r]]) { seen[g.d[r]] = 1; rec.retained_deletion_edges++; }
    }
}

// graph.rs:82-242 (select_observation_deletion_graph): largest component,
// then to a fixed point drop workers with one physical observation, else the
// articulation workers; stayers with several copies stay.
void select_observation(const Graph &g, std::vector<uint8_t> &active, SelectionReceipt &rec) {
  active.assign(g.rows, 1);
  rec.input_rows = g.rows;
  uint64_t comps = largest_component(g, active);
  rec.initial_components = rec.maximum_components = comps;
  rec.initial_component_rows = count_active(active);
  rec.mover_input_rows = count_active(active);
  {
    std::vector<uint8_t> seen(g.K, 0);
    for (uint64_t r = 0; r < g.rows; ++r)
      if (active[r] && !seen[g.row_coord[r]]) { seen[g.row_coord[r]] = 1; rec.initial_deletion_edges++; }
  }
  auto physical = [&]() {
    std::vector<uint64_t> p(g.W, 0);
    for (uint64_t r = 0; r < g.rows; ++r)
      if (active[r]) p[g.w[r]] += g.freq[r];
    return p;
  };
  const uint64_t bound = g.W + 1;
  for (;;) {
    if (rec.fixed_point_iterations > bound)
      fail(kGraphCertificate, "graph", "observation-deletion pruning exceeded its finite iteration bound");
    rec.maximum_components = std::max<uint64_t>(rec.maximum_components, largest_component(g, active));
    if (!count_active(active)) fail(kGraphEmpty, "graph", "no component remains after observation-deletion pruning");
    const auto phys = physical();
    uint64_t removed = 0;
    for (uint64_t w = 0; w < g.W; ++w) removed += phys[w] == 1;
    if (removed) {
      for (uint64_t r = 0; r < g.rows; ++r)
        if (active[r] && phys[g.w[r]] == 1) active[r] = 0;
      rec.insufficient_workers_removed += removed;
      rec.degree_iterations++;
      rec.fixed_point_iterations++;
      continue;
    }
    auto art = articulation_workers(g, active);
    uint64_t arts = 0;
    for (uint8_t x : art) arts += x;
    if (!arts) break;
    for (uint64_t r = 0; r < g.rows; ++r)
      if (active[r] && art[g.w[r]]) active[r] = 0;
    rec.articulation_workers_removed += arts;
    rec.articulation_iterations++;
    rec.fixed_point_iterations++;
  }
  rec.maximum_components = std::max<uint64_t>(rec.maximum_components, largest_component(g, active));
  for (uint64_t p : physical())
    if (p == 1) fail(kGraphCertificate, "graph", "final observation-deletion graph still contains a physical-singleton worker");
  {
    auto art = articulation_workers(g, active);
    for (uint8_t x : art)
      if (x) fail(kGraphCertificate, "graph", "final observation-deletion graph still contains a worker articulation");
  }
  rec.retained_rows = count_active(active);
  if (!rec.retained_rows) fail(kGraphEmpty, "graph", "no component remains after observation-deletion graph pruning");
  rec.retained_physical_mass = 0;
  std::vector<uint8_t> seen(g.K, 0);
  for (uint64_t r = 0; r < g.rows; ++r)
    if (active[r]) {
      rec.retained_physical_mass += g.freq[r];
      if (!seen[g.row_coord[r]]) { seen[g.row_coord[r]] = 1; rec.retained_deletion_edges++; }
    }
}

int cmp_total(double a, double b) {
  int64_t x = total_key(a), y = total_key(b);
  return x < y ? -1 : (x > y ? 1 : 0);
}

// Observation semantic ranks (generic_jla.rs:4267-4334, no controls): dense
// 1-based ranks over (worker, firm, per-copy target mass, outcome); `order`
// returns the rows sorted by (rank, row).
std::vector<uint64_t> observation_ranks(const Prepared &P, std::vector<uint32_t> &order) {
  const uint64_t R = P.rows;
  auto cmp = [&](uint32_t a, uint32_t b) {
    if (P.row_worker[a] != P.row_worker[b]) return P.row_worker[a] < P.row_worker[b] ? -1 : 1;
    if (P.row_firm[a] != P.row_firm[b]) return P.row_firm[a] < P.row_firm[b] ? -1 : 1;
    int c = cmp_total(P.per_copy[a], P.per_copy[b]);
    if (c) return c;
    return cmp_total(P.outcome[a], P.outcome[b]);
  };
  order.resize(R);
  std::iota(order.begin(), order.end(), 0u);
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    int c = cmp(a, b);
    return c ? c < 0 : a < b;
  });
  std::vector<uint64_t> rank(R, 0);
  uint64_t cur = 0;
  for (uint64_t i = 0; i < R; ++i) {
    if (i == 0 || cmp(order[i - 1], order[i]) != 0) ++cur;
    rank[order[i]] = cur;
  }
  return rank;
}

// Observation deletion plan (generic_jla.rs:4158-4215, statistical_batches.rs:15-53):
// classes of equal observation rank (rows ascending), and one unit per stored
// row holding the class rank and the row's copy offset in its class, so that
// the row's atom is the sum of its own physical copies' Rademacher signs.
void observation_plan(Prepared &P, std::vector<uint32_t> &order) {
  const uint64_t R = P.rows;
  P.row_rank = observation_ranks(P, order);
  P.unit_offset.assign(R, 0);
  uint64_t offset = 0;
  for (uint64_t i = 0; i < R; ++i) {
    if (i == 0 || P.row_rank[order[i - 1]] != P.row_rank[order[i]]) offset = 0;
    P.unit_offset[order[i]] = offset;  // rows of a class ascend in `order`
    offset += P.frequency[order[i]];
  }
  P.units = R;
  P.match_units = 0;
  P.unit_ptr.resize(R + 1);
  std::iota(P.unit_ptr.begin(), P.unit_ptr.end(), 0ull);
  P.unit_rows.resize(R);
  std::iota(P.unit_rows.begin(), P.unit_rows.end(), 0u);
  P.unit_cell = P.row_cell;
  P.unit_count = P.frequency;
  P.unit_rank = P.row_rank;
  P.unit_outcome.resize(R);
  for (uint64_t i = 0; i < R; ++i) P.unit_outcome[i] = static_cast<double>(P.frequency[i]) * P.outcome[i];
  P.unit_target = P.target;
}

}  // namespace

Prepared prepare(const Columns &cols, const uint8_t *mask, bool observation) {
  using clock = std::chrono::steady_clock;
  auto mark = clock::now();
  auto lap = [&mark]() {
    const auto now = clock::now();
    const uint64_t ns = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(now - mark).count();
    mark = now;
    return ns;
  };
  const uint64_t n = cols.rows;
  if (n == 0) fail(kInvalidInput, "ingest", "input sample is empty");
  Prepared P;
  P.input_rows = n;
  uint64_t physical = 0;
  double target_total = 0.0;
  for (uint64_t r = 0; r < n; ++r) {
    if (!std::isfinite(cols.outcome[r]))
      fail(kInvalidInput, "ingest", "outcome is nonfinite at zero-based row " + std::to_string(r));
    if (cols.frequency[r] == 0)
      fail(kInvalidWeight, "ingest", "frequency is zero at zero-based row " + std::to_string(r));
    if (!std::isfinite(cols.target[r]) || cols.target[r] < 0.0)
      fail(kInvalidTargetWeight, "ingest", "target weight is invalid at zero-based row " + std::to_string(r));
    physical += cols.frequency[r];
    if (physical > (1ULL << 53))
      fail(kResourceLimit, "ingest", "literal physical-frequency total exceeds the exact binary64 integer range");
    target_total += cols.target[r];
  }
  if (!std::isfinite(target_total) || target_total <= 0.0)
    fail(kInvalidTargetWeight, "ingest", "target weights must have positive finite total mass");
  P.input_physical = physical;
  P.input_target_total = target_total;

  uint64_t W0, F0, D0;
  std::vector<uint32_t> cw = redense(cols.worker, n, W0);
  std::vector<uint32_t> cf = redense(cols.firm, n, F0);
  std::vector<uint32_t> cd = redense(cols.deletion, n, D0);
  Graph g{n, W0, F0, D0, 0, cw.data(), cf.data(), cd.data(), cols.frequency, {}, {}, {}, {}};
  // distinct coordinates; every deletion id must stay inside one coordinate
  {
    std::vector<uint64_t> key(n);
    for (uint64_t r = 0; r < n; ++r) key[r] = static_cast<uint64_t>(cw[r]) * F0 + cf[r];
    std::vector<uint64_t> uniq(key);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    g.K = uniq.size();
    g.row_coord.resize(n);
    g.coord_w.resize(g.K);
    g.coord_f.resize(g.K);
    for (uint64_t c = 0; c < g.K; ++c) {
      g.coord_w[c] = static_cast<uint32_t>(uniq[c] / F0);
      g.coord_f[c] = static_cast<uint32_t>(uniq[c] % F0);
    }
    for (uint64_t r = 0; r < n; ++r)
      g.row_coord[r] = static_cast<uint32_t>(std::lower_bound(uniq.begin(), uniq.end(), key[r]) - uniq.begin());
    g.del_coord.assign(D0, UINT32_MAX);
    for (uint64_t r = 0; r < n; ++r) {
      uint32_t &dc = g.del_coord[cd[r]];
      if (dc == UINT32_MAX) dc = g.row_coord[r];
      else if (dc != g.row_coord[r])
        fail(kInvalidIdentifier, "canonicalize",
             "each deletion identifier must stay within one worker-firm coordinate");
    }
  }
  P.input_topology_checksum = fnv_checksum(n, cw.data(), cf.data(), cd.data(), cols.frequency);
  P.observation = observation;
  P.canonicalize_ns = lap();
  if (mask) {
    P.mask.assign(mask, mask + n);
  } else if (observation) {
    select_observation(g, P.mask, P.selection);
  } else {
    select(g, P.mask, P.selection);
  }

  P.graph_ns = lap();
  // ---- compression (problem.rs:206-606)
  std::vector<uint64_t> kept;
  for (uint64_t r = 0; r < n; ++r)
    if (P.mask[r]) kept.push_back(r);
  const uint64_t R = kept.size();
  if (!R) fail(kGraphEmpty, "compression", "no rows remain after graph selection");
  P.row_input = kept;
  auto relabel = [&](const std::vector<uint32_t> &ids, uint64_t levels, uint64_t &count) {
    std::vector<uint8_t> present(levels, 0);
    for (uint64_t r : kept) present[ids[r]] = 1;
    std::vector<uint32_t> map(levels, UINT32_MAX);
    uint32_t next = 0;
    for (uint64_t l = 0; l < levels; ++l)
      if (present[l]) map[l] = next++;
    count = next;
    return map;
  };
  uint64_t W, F, G;
  auto wmap = relabel(cw, W0, W), fmap = relabel(cf, F0, F), dmap = relabel(cd, D0, G);
  if (F < 2) fail(kGraphUnidentified, "compression", "retained graph must contain at least two firms");
  P.rows = R; P.workers = W; P.firms = F; P.units = G;
  P.row_worker.resize(R); P.row_firm.resize(R); P.row_deletion.resize(R); P.row_cell.resize(R);
  P.outcome.resize(R); P.frequency.resize(R); P.target.resize(R);
  for (uint64_t i = 0; i < R; ++i) {
    uint64_t r = kept[i];
    P.row_worker[i] = wmap[cw[r]];
    P.row_firm[i] = fmap[cf[r]];
    P.row_deletion[i] = dmap[cd[r]];
    P.outcome[i] = cols.outcome[r];
    P.frequency[i] = cols.frequency[r];
    P.target[i] = cols.target[r];
  }
  std::vector<uint32_t> order(R);
  std::iota(order.begin(), order.end(), 0u);
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    if (P.row_worker[a] != P.row_worker[b]) return P.row_worker[a] < P.row_worker[b];
    if (P.row_firm[a] != P.row_firm[b]) return P.row_firm[a] < P.row_firm[b];
    return a < b;
  });
  for (uint64_t i = 0; i < R;) {
    uint64_t b = i;
    uint32_t w = P.row_worker[order[i]], f = P.row_firm[order[i]];
    while (i < R && P.row_worker[order[i]] == w && P.row_firm[order[i]] == f) ++i;
    uint32_t cell = static_cast<uint32_t>(P.cell_worker.size());
    Kahan kw, ko, kt;
    for (uint64_t j = b; j < i; ++j) {
      uint32_t row = order[j];
      double weight = static_cast<double>(P.frequency[row]);
      kw.add(weight);
      ko.add(weight * P.outcome[row]);
      kt.add(P.target[row]);
      P.row_cell[row] = cell;
    }
    P.cell_worker.push_back(w);
    P.cell_firm.push_back(f);
    P.cell_weight.push_back(kw.sum);
    P.cell_outcome.push_back(ko.sum);
    P.cell_target.push_back(kt.sum);
  }
  const uint64_t C = P.cell_worker.size();
  P.cells = C;
  P.worker_ptr.assign(W + 1, 0);
  for (uint64_t c = 0; c < C; ++c) P.worker_ptr[P.cell_worker[c] + 1]++;
  for (uint64_t w = 0; w < W; ++w) P.worker_ptr[w + 1] += P.worker_ptr[w];
  P.firm_cells.resize(C);
  std::iota(P.firm_cells.begin(), P.firm_cells.end(), 0u);
  std::sort(P.firm_cells.begin(), P.firm_cells.end(), [&](uint32_t a, uint32_t b) {
    if (P.cell_firm[a] != P.cell_firm[b]) return P.cell_firm[a] < P.cell_firm[b];
    if (P.cell_worker[a] != P.cell_worker[b]) return P.cell_worker[a] < P.cell_worker[b];
    return a < b;
  });
  P.firm_ptr.assign(F + 1, 0);
  for (uint64_t c = 0; c < C; ++c) P.firm_ptr[P.cell_firm[c] + 1]++;
  for (uint64_t f = 0; f < F; ++f) P.firm_ptr[f + 1] += P.firm_ptr[f];
  // exact target strata (problem.rs:959-1035): count only
  {
    auto key_cmp = [&](uint32_t a, uint32_t b) {
      if (P.row_worker[a] != P.row_worker[b]) return P.row_worker[a] < P.row_worker[b] ? -1 : 1;
      if (P.row_firm[a] != P.row_firm[b]) return P.row_firm[a] < P.row_firm[b] ? -1 : 1;
      if (P.row_deletion[a] != P.row_deletion[b]) return P.row_deletion[a] < P.row_deletion[b] ? -1 : 1;
      int c = cmp_total(P.outcome[a], P.outcome[b]);
      if (c) return c;
      if (P.frequency[a] != P.frequency[b]) return P.frequency[a] < P.frequency[b] ? -1 : 1;
      return cmp_total(P.target[a], P.target[b]);
    };
    std::iota(order.begin(), order.end(), 0u);
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
      int c = key_cmp(a, b);
      return c ? c < 0 : a < b;
    });
    P.exact_strata = 0;
    for (uint64_t i = 0; i < R; ++i)
      if (i == 0 || key_cmp(order[i - 1], order[i]) != 0) P.exact_strata++;
  }
  P.unit_ptr.assign(G + 1, 0);
  for (uint64_t i = 0; i < R; ++i) P.unit_ptr[P.row_deletion[i] + 1]++;
  for (uint64_t u = 0; u < G; ++u) P.unit_ptr[u + 1] += P.unit_ptr[u];
  P.unit_rows.resize(R);
  {
    std::vector<uint64_t> cur(P.unit_ptr.begin(), P.unit_ptr.end() - 1);
    for (uint64_t i = 0; i < R; ++i) P.unit_rows[cur[P.row_deletion[i]]++] = static_cast<uint32_t>(i);
  }
  Kahan tt;
  for (uint64_t i = 0; i < R; ++i) {
    P.physical_total += P.frequency[i];
    tt.add(P.target[i]);
  }
  P.target_total = tt.sum;
  if (!std::isfinite(P.target_total) || P.target_total <= 0.0)
    fail(kInvalidTargetWeight, "compression", "retained target mass is not positive and finite");
  P.topology_checksum = fnv_checksum(R, P.row_worker.data(), P.row_firm.data(), P.row_deletion.data(),
                                     P.frequency.data());

  P.compress_ns = lap();
  // ---- JLA plan (jla.rs)
  P.per_copy.resize(R);
  for (uint64_t i = 0; i < R; ++i) {
    double v = canonical_zero(P.target[i] / static_cast<double>(P.frequency[i]));
    if (!std::isfinite(v) || v < 0.0)
      fail(kInvalidTargetWeight, "jla_plan", "per-copy target mass is invalid at zero-based row " + std::to_string(i));
    P.per_copy[i] = v;
  }
  if (P.observation) {
    observation_plan(P, order);
  } else {
  auto semantic_cmp = [&](uint32_t a, uint32_t b) {
    if (P.row_worker[a] != P.row_worker[b]) return P.row_worker[a] < P.row_worker[b] ? -1 : 1;
    if (P.row_firm[a] != P.row_firm[b]) return P.row_firm[a] < P.row_firm[b] ? -1 : 1;
    if (P.row_deletion[a] != P.row_deletion[b]) return P.row_deletion[a] < P.row_deletion[b] ? -1 : 1;
    int c = cmp_total(canonical_zero(P.per_copy[a]), canonical_zero(P.per_copy[b]));
    if (c) return c;
    return cmp_total(canonical_zero(P.outcome[a]), canonical_zero(P.outcome[b]));
  };
  std::iota(order.begin(), order.end(), 0u);
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    int c = semantic_cmp(a, b);
    return c ? c < 0 : a < b;
  });
  P.row_rank.resize(R);
  {
    uint64_t cur = 0;
    for (uint64_t i = 0; i < R; ++i) {
      if (i == 0 || semantic_cmp(order[i - 1], order[i]) != 0) ++cur;
      P.row_rank[order[i]] = cur;
    }
  }
  P.unit_cell.resize(G); P.unit_count.resize(G); P.unit_rank.resize(G);
  P.unit_outcome.resize(G); P.unit_target.resize(G);
  for (uint64_t u = 0; u < G; ++u) {
    uint64_t phys = 0, minr = UINT64_MAX;
    uint32_t cell = UINT32_MAX;
    Neumaier os, ts;
    for (uint64_t k = P.unit_ptr[u]; k < P.unit_ptr[u + 1]; ++k) {
      uint32_t row = P.unit_rows[k];
      phys += P.frequency[row];
      if (cell == UINT32_MAX) cell = P.row_cell[row];
      else if (cell != P.row_cell[row])
        fail(kInvalidIdentifier, "jla_plan", "one deletion unit crosses coefficient cells");
      os.add(static_cast<double>(P.frequency[row]) * P.outcome[row]);
      ts.add(P.target[row]);
      minr = std::min(minr, P.row_rank[row]);
    }
    P.unit_cell[u] = cell;
    P.unit_count[u] = phys;
    P.unit_outcome[u] = os.value();
    P.unit_target[u] = ts.value();
    P.unit_rank[u] = minr;
  }
  }
  std::iota(order.begin(), order.end(), 0u);
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    if (P.row_cell[a] != P.row_cell[b]) return P.row_cell[a] < P.row_cell[b];
    int c = cmp_total(P.per_copy[a], P.per_copy[b]);
    if (c) return c < 0;
    if (P.row_rank[a] != P.row_rank[b]) return P.row_rank[a] < P.row_rank[b];
    return a < b;
  });
  for (uint64_t i = 0; i < R;) {
    uint64_t b = i;
    uint32_t cell = P.row_cell[order[i]];
    uint64_t bits = bits_of(P.per_copy[order[i]]);
    while (i < R && P.row_cell[order[i]] == cell && bits_of(P.per_copy[order[i]]) == bits) ++i;
    uint64_t phys = 0, minr = UINT64_MAX;
    Neumaier ts;
    for (uint64_t j = b; j < i; ++j) {
      uint32_t row = order[j];
      phys += P.frequency[row];
      ts.add(P.target[row]);
      minr = std::min(minr, P.row_rank[row]);
    }
    P.stratum_cell.push_back(cell);
    P.stratum_per_copy.push_back(P.per_copy[order[b]]);
    P.stratum_count.push_back(phys);
    P.stratum_target.push_back(ts.value());
    P.stratum_rank.push_back(minr);
  }
  P.strata = P.stratum_cell.size();
  // gather indices for the device: units and strata of each cell
  // (P.units: the deletion units, or the stored rows under observation deletion)
  P.cell_unit_ptr.assign(C + 1, 0);
  for (uint64_t u = 0; u < P.units; ++u) P.cell_unit_ptr[P.unit_cell[u] + 1]++;
  for (uint64_t c = 0; c < C; ++c) P.cell_unit_ptr[c + 1] += P.cell_unit_ptr[c];
  P.cell_units.resize(P.units);
  {
    std::vector<uint64_t> cur(P.cell_unit_ptr.begin(), P.cell_unit_ptr.end() - 1);
    for (uint64_t u = 0; u < P.units; ++u) P.cell_units[cur[P.unit_cell[u]]++] = static_cast<uint32_t>(u);
  }
  P.cell_stratum_ptr.assign(C + 1, 0);
  for (uint64_t s = 0; s < P.strata; ++s) P.cell_stratum_ptr[P.stratum_cell[s] + 1]++;
  for (uint64_t c = 0; c < C; ++c) P.cell_stratum_ptr[c + 1] += P.cell_stratum_ptr[c];
  if (!P.observation) P.match_units = P.units;
  P.plan_ns = lap();
  return P;
}

Prepared augment_stayers(const Prepared &M, const StayerColumns &s, StayerReceipt &rec) {
  if (M.observation)
    fail(kUnsupported, "session_stayer_augmentation", "the stayer hybrid requires match-deletion mover preparation");
  if (M.stayers)
    fail(kContextPoisoned, "session_stayer_augmentation", "the prepared generation already owns a stayer augmentation");
  const uint64_t S = s.rows;
  uint64_t SW = 0;
  for (uint64_t r = 0; r < S; ++r) SW = std::max<uint64_t>(SW, s.worker[r]);
  if ((S == 0) != (SW == 0))
    fail(kInvalidIdentifier, "stayer_augmentation",
         "eligible stayer workers must be a nonempty dense map when rows are supplied");
  std::vector<uint64_t> worker_firm(SW, 0), worker_physical(SW, 0);
  uint64_t stayer_physical = 0;
  Neumaier stayer_target;
  for (uint64_t r = 0; r < S; ++r) {
    const uint64_t f = s.firm[r], w = s.worker[r];
    if (f == 0 || f > M.firms)
      fail(kInvalidIdentifier, "stayer_augmentation",
           "stayer row " + std::to_string(r) + " is not attached to a retained mover firm");
    if (w == 0 || w > SW)
      fail(kInvalidIdentifier, "stayer_augmentation", "stayer row " + std::to_string(r) + " has a non-dense worker identifier");
    if (worker_firm[w - 1] && worker_firm[w - 1] != f)
      fail(kInvalidIdentifier, "stayer_augmentation",
           "eligible stayer worker " + std::to_string(w) + " spans more than one firm");
    worker_firm[w - 1] = f;
    worker_physical[w - 1] += s.frequency[r];
    stayer_physical += s.frequency[r];
    stayer_target.add(s.target[r]);
  }
  for (uint64_t w = 0; w < SW; ++w)
    if (!worker_firm[w]) fail(kInvalidIdentifier, "stayer_augmentation", "eligible stayer worker identifiers are not dense");
  for (uint64_t w = 0; w < SW; ++w)
    if (worker_physical[w] < 2)
      fail(kInvalidInput, "stayer_augmentation",
           "every eligible stayer worker must represent at least two physical observations");
  // combined columns: the mover problem's dense identifiers, then the stayers
  // with worker M.workers + w and one deletion unit per stored row
  const uint64_t RM = M.rows, R = RM + S;
  std::vector<uint64_t> cw(R), cf(R), cd(R), cq(R);
  std::vector<double> cy(R), ct(R);
  for (uint64_t r = 0; r < RM; ++r) {
    cw[r] = M.row_worker[r] + 1ull; cf[r] = M.row_firm[r] + 1ull; cd[r] = M.row_deletion[r] + 1ull;
    cy[r] = M.outcome[r]; cq[r] = M.frequency[r]; ct[r] = M.target[r];
  }
  for (uint64_t r = 0; r < S; ++r) {
    cw[RM + r] = M.workers + s.worker[r]; cf[RM + r] = s.firm[r]; cd[RM + r] = M.units + r + 1;
    cy[RM + r] = s.outcome[r]; cq[RM + r] = s.frequency[r]; ct[RM + r] = s.target[r];
  }
  std::vector<uint8_t> all(R, 1);
  Columns cols{R, cw.data(), cf.data(), cd.data(), cy.data(), cq.data(), ct.data()};
  Prepared P = prepare(cols, all.data(), false);
  if (P.rows != R || P.workers != M.workers + SW || P.firms != M.firms || P.units != M.units + S)
    fail(kInvariant, "stayer_augmentation", "combined exact problem dimensions do not reconcile");
  for (uint64_t r = 0; r < RM; ++r)
    if (P.row_worker[r] != M.row_worker[r] || P.row_firm[r] != M.row_firm[r] ||
        P.row_deletion[r] != M.row_deletion[r] || P.frequency[r] != M.frequency[r])
      fail(kInvariant, "stayer_augmentation", "combined preparation changed a retained mover row");
  // keep the mover generation's identity (input rows, mask, selection)
  P.input_rows = M.input_rows; P.input_physical = M.input_physical; P.input_target_total = M.input_target_total;
  P.input_topology_checksum = M.input_topology_checksum; P.mask = M.mask; P.selection = M.selection;
  // the hybrid plan (generic_jla.rs:2170-2213): mover match groups keep their
  // semantic entities; each stayer row is an observation unit whose entity is
  // its observation class (ranked over all rows, generic_jla.rs:4217-4265) and
  // whose copies continue those of the class's earlier stayer rows
  P.stayers = true;
  P.match_units = M.units;
  P.unit_offset.assign(P.units, 0);
  if (S) {
    std::vector<uint32_t> order;
    const std::vector<uint64_t> orank = observation_ranks(P, order);
    uint64_t offset = 0, last = 0;
    for (uint64_t i = 0; i < R; ++i) {
      const uint32_t row = order[i];
      if (row < RM) continue;
      if (orank[row] != last) offset = 0;
      last = orank[row];
      const uint64_t u = M.units + (row - RM);
      P.unit_rank[u] = orank[row];
      P.unit_offset[u] = offset;
      offset += P.frequency[row];
    }
  }
  rec = StayerReceipt{};
  rec.mover_stored_rows = RM; rec.stayer_stored_rows = S; rec.combined_stored_rows = R;
  rec.mover_physical_mass = M.physical_total; rec.stayer_physical_mass = stayer_physical;
  rec.combined_physical_mass = P.physical_total;
  rec.mover_workers = M.workers; rec.stayer_workers = SW; rec.combined_workers = P.workers; rec.firms = P.firms;
  rec.mover_deletion_units = M.units; rec.stayer_deletion_units = stayer_physical;
  rec.combined_deletion_units = M.units + stayer_physical;
  rec.mover_target_mass = M.target_total; rec.stayer_target_mass = stayer_target.value();
  rec.combined_target_mass = P.target_total; rec.topology_checksum = P.topology_checksum;
  if (!std::isfinite(rec.stayer_target_mass) || rec.stayer_target_mass < 0.0)
    fail(kInvalidTargetWeight, "stayer_augmentation", "stayer target mass is not finite and nonnegative");
  return P;
}

void attach_controls(Prepared &P, const double *const *columns, uint32_t q) {
  if (q > 32) fail(kInvalidInput, "engine_prepare", "at most 32 controls");
  if (q && !columns) fail(kInvalidInput, "engine_ingest", "control-pointer array is null");
  if (!q && columns) fail(kInvalidInput, "engine_ingest", "zero controls require a null control-pointer array");
  for (uint32_t j = 0; j < q; ++j) {
    if (!columns[j]) fail(kInvalidInput, "engine_ingest", "control column " + std::to_string(j) + " is null");
    for (uint64_t r = 0; r < P.input_rows; ++r)
      if (!std::isfinite(columns[j][r]))
        fail(kInvalidInput, "engine_ingest", "control column " + std::to_string(j) + " has a nonfinite value");
  }
  P.controls = q;
  P.control.assign(P.rows * q, 0.0);
  for (uint64_t i = 0; i < P.rows; ++i)
    for (uint32_t j = 0; j < q; ++j) P.control[i * q + j] = columns[j][P.row_input[i]];
  std::vector<Kahan> sums((size_t)P.cells * q);
  for (uint64_t i = 0; i < P.rows; ++i) {
    const double f = static_cast<double>(P.frequency[i]);
    for (uint32_t j = 0; j < q; ++j) sums[(size_t)P.row_cell[i] * q + j].add(f * P.control[i * q + j]);
  }
  P.cell_control.resize(sums.size());
  for (size_t k = 0; k < sums.size(); ++k) P.cell_control[k] = sums[k].sum;
}

void set_outcome(Prepared &P, const std::vector<double> &y) {
  if (y.size() != P.rows) fail(kInvariant, "working_outcome", "working outcome has the wrong length");
  P.outcome = y;
  // cell sums in (cell, ascending row) order, as the compression forms them
  std::vector<Kahan> cell(P.cells);
  for (uint64_t i = 0; i < P.rows; ++i) cell[P.row_cell[i]].add(static_cast<double>(P.frequency[i]) * y[i]);
  for (uint64_t c = 0; c < P.cells; ++c) P.cell_outcome[c] = cell[c].sum;
  for (uint64_t u = 0; u < P.units; ++u) {
    Neumaier os;
    for (uint64_t k = P.unit_ptr[u]; k < P.unit_ptr[u + 1]; ++k) {
      const uint32_t row = P.unit_rows[k];
      os.add(static_cast<double>(P.frequency[row]) * y[row]);
    }
    P.unit_outcome[u] = os.value();
  }
}

}  // namespace vk
