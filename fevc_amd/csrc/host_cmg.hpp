// Host construction of the hybrid firm-Schur Laplacian and its CMG hierarchy,
// laid out for the device V-cycle.
#pragma once

#include <cstdint>
#include <vector>

#include "host_prepare.hpp"

namespace vk {

// Canonical weighted Laplacian: edges u < v sorted by (u, v), duplicates
// merged with a compensated sum (vendor/cmg/src/graph.rs:279-414).
struct Laplacian {
  uint64_t n = 0;
  std::vector<uint32_t> u, v;
  std::vector<double> w;
  std::vector<double> diag;
  uint64_t nnz = 0;
  double norm_bound = 0.0;
  static Laplacian canonical(uint64_t n, std::vector<uint64_t> &keys, std::vector<double> &weights);
};

enum Terminal : int { kNotTerminal = 0, kDirect = 1, kFullContraction = 2, kStagnated = 3, kFill = 4, kMaxLevels = 5 };

struct Level {
  Laplacian graph;
  std::vector<double> inv_diag;  // 0.5 / degree
  std::vector<uint32_t> labels;  // aggregation labels (empty at a direct terminal)
  uint64_t coarse = 0;
  uint64_t repeat = 0;           // effective recursive repeat count
  int terminal = kNotTerminal;
  // device gather layout
  std::vector<uint32_t> row_ptr, col;  // CSR of off-diagonal entries, columns ascending
  std::vector<double> val;             // edge weights (positive)
  std::vector<uint32_t> agg_ptr, agg_members;  // fine members of each aggregate, ascending
};

struct Hierarchy {
  std::vector<Level> levels;
  int terminal_reason = kDirect;
  uint64_t cumulative_nnz = 0;
  // grounded LDL^T of a direct terminal (ldl.rs:677-803)
  uint64_t ldl_dim = 0, factor_nnz = 0;
  std::vector<uint64_t> anchors, perm;
  std::vector<double> ldl_lower;  // dense strict lower, row-major dim x dim
  std::vector<double> ldl_diag;
  std::vector<double> terminal_inverse;  // n_T x n_T grounded inverse, row-major
  uint64_t finest_components = 1;
};

struct CmgOptions {
  uint64_t direct_threshold = 700;
  uint64_t max_levels = 128;
  double max_hierarchy_nnz_factor = 5.0;
  double low_degree_threshold = 0.125;
};

struct HybridSystem {
  uint64_t firms = 0, aux = 0;
  Laplacian graph;
  // auxiliary incidence for the harmonic lift (cmg_impl.rs.in:459-500)
  std::vector<uint64_t> aux_ptr;
  std::vector<uint32_t> aux_nbr;
  std::vector<double> aux_w;
};

HybridSystem build_hybrid(const Prepared &p);
Laplacian laplacian_from_edges(uint64_t n, uint64_t m, const uint32_t *u, const uint32_t *v, const double *w);
Hierarchy build_hierarchy(const Laplacian &g, const CmgOptions &opt);

// ------------------------------------------------------------------ device layout
// Segments (CSR rows, or aggregates over their contiguous members) are cut
// into chunks of at most `chunk` entries so that no GPU thread group walks a
// long row or a large aggregate alone.  A segment with several chunks gets a
// "multi" index: its chunk partials are combined in a second pass.  With
// `group` > 1, runs of up to `group` consecutive short segments (at most
// chunk / 4 entries each) share one chunk tagged kGroupChunk + count, so that
// the slices of one wavefront walk several short rows at once.
constexpr uint32_t kSingleChunk = 0xffffffffu;
constexpr uint32_t kGroupChunk = 0xffffff00u;

struct ChunkTable {
  std::vector<uint32_t> desc;    // 4 per chunk: segment, begin, end, tag
  std::vector<uint32_t> mrange;  // 2 per multi-chunk segment: first chunk, end chunk
  std::vector<uint32_t> ptr;     // segment pointers (grouped tables only)
  uint64_t chunks() const { return desc.size() / 4; }
  uint32_t multi() const { return static_cast<uint32_t>(mrange.size() / 2); }
};
ChunkTable chunk_segments(const std::vector<uint32_t> &ptr, uint32_t chunk, uint32_t group = 1);

// One level renumbered so that every aggregate is a contiguous run of
// vertices, runs ordered like the coarse level's own renumbering; the last
// level keeps its numbering.  Row entries keep the reference order (ascending
// original neighbour index).
struct LevelLayout {
  std::vector<uint32_t> perm, pos;  // new -> old, old -> new
  std::vector<uint32_t> row_ptr, col;
  std::vector<double> val, inv;
  std::vector<uint32_t> lab;        // coarse vertex (renumbered) of every vertex; monotone
  std::vector<uint32_t> agg_ptr;    // member range of every coarse vertex
  ChunkTable rows, aggs;
};
std::vector<LevelLayout> device_layout(const Hierarchy &H, uint32_t chunk);

}  // namespace vk
