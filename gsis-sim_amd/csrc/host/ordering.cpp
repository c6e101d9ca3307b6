// Cell orderings for the LU-SGS sweeps: reverse Cuthill-McKee over the whole mesh
// (cell_ordering = "rcm") and recursive coordinate bisection into sweep blocks with a reverse
// Cuthill-McKee numbering inside each block (cell_ordering = "rcb", the default).
//
// The numbering decides which cells share a sweep block and the order of the in-block
// dependency chain, so it fixes every multi-block iterate; it must be the reference's
// (src/mesh/ordering.cpp:30-120 and src/kinetic/kinetic_setup.cpp:46-178). Every choice is a
// total order, so the result does not depend on container order:
//   * "lightest cell" = smallest (degree, cell index), degree counted inside the cell set being
//     ordered (the whole mesh, or one block);
//   * Cuthill-McKee appends, for each cell taken from the queue in turn, its not yet numbered
//     neighbours in lightest-first order; components are started from the lightest unnumbered cell;
//   * the start of a component is found from its seed through the last level of a breadth-first
//     layering: over the whole mesh the lightest cell of the last level replaces the start while
//     that strictly deepens the layering (at most 16 times); inside a block it replaces the seed
//     once, unconditionally;
//   * bisection splits a set of P blocks into P/2 and P - P/2 along the axis of the largest extent
//     of the cell-centre bounding box (the first such axis on ties), giving the first part the
//     round(n (P/2) / P) cells smallest by (centre coordinate, cell index); blocks are numbered in
//     depth-first order, first part first.

#include <algorithm>
#include <cmath>
#include <numeric>

#include "mesh.hpp"

namespace gsk {

namespace {

// The cells of one set (the whole mesh or one block) and their adjacency restricted to the set.
class CellSet {
 public:
  CellSet(const CellAdjacency& adj, const std::vector<int>* set_of, int set)
      : adj_(adj), set_of_(set_of), set_(set) {}

  bool contains(Index c) const { return set_of_ == nullptr || (*set_of_)[static_cast<std::size_t>(c)] == set_; }

  template <class Fn>
  void for_each_neighbour(Index c, Fn&& fn) const {
    const auto lo = adj_.offsets[static_cast<std::size_t>(c)], hi = adj_.offsets[static_cast<std::size_t>(c) + 1];
    for (Offset k = lo; k < hi; ++k) {
      const Index nb = adj_.cells[static_cast<std::size_t>(k)];
      if (contains(nb)) fn(nb);
    }
  }

  Index degree(Index c) const {
    if (set_of_ == nullptr) return adj_.degree(c);
    Index d = 0;
    for_each_neighbour(c, [&](Index) { ++d; });
    return d;
  }

 private:
  const CellAdjacency& adj_;
  const std::vector<int>* set_of_;
  int set_;
};

struct Lighter {
  const std::vector<Index>& degree;  // indexed by cell
  bool operator()(Index a, Index b) const {
    const Index da = degree[static_cast<std::size_t>(a)], db = degree[static_cast<std::size_t>(b)];
    return da != db ? da < db : a < b;
  }
};

// Breadth-first layering from root inside the set: returns the number of levels and leaves the
// cells of the deepest level in `last`. `visit` holds per-cell stamps; `stamp` is advanced.
int layering(const CellSet& set, Index root, std::vector<long long>& visit, long long& stamp,
             std::vector<Index>& last) {
  const long long mark = ++stamp;
  std::vector<Index> level{root}, next;
  visit[static_cast<std::size_t>(root)] = mark;
  int depth = 1;
  for (;;) {
    next.clear();
    for (const Index c : level)
      set.for_each_neighbour(c, [&](Index nb) {
        if (visit[static_cast<std::size_t>(nb)] != mark) {
          visit[static_cast<std::size_t>(nb)] = mark;
          next.push_back(nb);
        }
      });
    if (next.empty()) break;
    level.swap(next);
    ++depth;
  }
  last = level;
  return depth;
}

// Reverse Cuthill-McKee numbering of the cells of one set (listed in `cells`), appended to out.
// numbered: per-cell flags shared across sets. deepen: the whole-mesh start search (see top).
void reverse_cuthill_mckee(const CellSet& set, const std::vector<Index>& cells, bool deepen,
                           std::vector<Index>& degree, std::vector<char>& numbered, std::vector<long long>& visit,
                           long long& stamp, std::vector<Index>& out) {
  for (const Index c : cells) degree[static_cast<std::size_t>(c)] = set.degree(c);
  const Lighter lighter{degree};
  std::vector<Index> seeds = cells;
  std::sort(seeds.begin(), seeds.end(), lighter);
  const std::size_t base = out.size();
  std::vector<Index> last, fresh;
  for (const Index seed : seeds) {
    if (numbered[static_cast<std::size_t>(seed)]) continue;
    Index start = seed;
    int depth = layering(set, seed, visit, stamp, last);
    for (int round = 0; round < (deepen ? 16 : 1); ++round) {
      const Index cand = *std::min_element(last.begin(), last.end(), lighter);
      if (!deepen) {
        start = cand;
        break;
      }
      std::vector<Index> cand_last;
      const int cand_depth = layering(set, cand, visit, stamp, cand_last);
      if (cand_depth <= depth) break;
      start = cand;
      depth = cand_depth;
      last.swap(cand_last);
    }
    // Cuthill-McKee from start: the output itself is the queue
    std::size_t head = out.size();
    out.push_back(start);
    numbered[static_cast<std::size_t>(start)] = 1;
    for (; head < out.size(); ++head) {
      fresh.clear();
      set.for_each_neighbour(out[head], [&](Index nb) {
        if (!numbered[static_cast<std::size_t>(nb)]) {
          numbered[static_cast<std::size_t>(nb)] = 1;
          fresh.push_back(nb);
        }
      });
      std::sort(fresh.begin(), fresh.end(), lighter);
      out.insert(out.end(), fresh.begin(), fresh.end());
    }
  }
  std::reverse(out.begin() + static_cast<std::ptrdiff_t>(base), out.end());
}

// Recursive coordinate bisection of all cells into P blocks (see top); blocks in depth-first order.
std::vector<std::vector<Index>> bisect_blocks(const std::vector<Vec3>& centre, Index n_cells, int P) {
  struct Task {
    std::vector<Index> cells;
    int parts;
  };
  std::vector<std::vector<Index>> blocks;
  std::vector<Task> stack;
  stack.push_back({std::vector<Index>(static_cast<std::size_t>(n_cells)), P});
  std::iota(stack.back().cells.begin(), stack.back().cells.end(), 0);
  while (!stack.empty()) {
    Task t = std::move(stack.back());
    stack.pop_back();
    if (t.parts <= 1 || t.cells.size() <= 1) {
      blocks.push_back(std::move(t.cells));
      continue;
    }
    Vec3 lo = centre[static_cast<std::size_t>(t.cells[0])], hi = lo;
    for (const Index c : t.cells)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], centre[static_cast<std::size_t>(c)][a]);
        hi[a] = std::max(hi[a], centre[static_cast<std::size_t>(c)][a]);
      }
    int axis = 0;
    if (hi[1] - lo[1] > hi[axis] - lo[axis]) axis = 1;
    if (hi[2] - lo[2] > hi[axis] - lo[axis]) axis = 2;
    const int first_parts = t.parts / 2;
    const auto n_first = static_cast<std::size_t>(std::llround(static_cast<double>(t.cells.size()) *
                                                               static_cast<double>(first_parts) /
                                                               static_cast<double>(t.parts)));
    std::sort(t.cells.begin(), t.cells.end(), [&](Index a, Index b) {
      const double xa = centre[static_cast<std::size_t>(a)][axis], xb = centre[static_cast<std::size_t>(b)][axis];
      return xa != xb ? xa < xb : a < b;
    });
    Task first{std::vector<Index>(t.cells.begin(), t.cells.begin() + static_cast<std::ptrdiff_t>(n_first)), first_parts};
    Task second{std::vector<Index>(t.cells.begin() + static_cast<std::ptrdiff_t>(n_first), t.cells.end()),
                t.parts - first_parts};
    stack.push_back(std::move(second));  // popped after the first part's whole subtree
    stack.push_back(std::move(first));
  }
  return blocks;
}

}  // namespace

std::vector<Index> rcm_ordering(const PolyMesh& m) {
  const CellAdjacency adj = build_cell_adjacency(m);
  const auto n = static_cast<std::size_t>(m.n_cells);
  std::vector<Index> cells(n), degree(n), out;
  std::iota(cells.begin(), cells.end(), 0);
  std::vector<char> numbered(n, 0);
  std::vector<long long> visit(n, -1);
  long long stamp = 0;
  out.reserve(n);
  reverse_cuthill_mckee(CellSet(adj, nullptr, 0), cells, true, degree, numbered, visit, stamp, out);
  return out;
}

std::vector<Index> rcb_ordering(const PolyMesh& m, const MeshGeometry& g, int P, std::vector<Index>& block_offsets) {
  const auto n = static_cast<std::size_t>(m.n_cells);
  const std::vector<std::vector<Index>> blocks = bisect_blocks(g.cell_center, m.n_cells, P);
  std::vector<int> block_of(n, -1);
  for (std::size_t b = 0; b < blocks.size(); ++b)
    for (const Index c : blocks[b]) block_of[static_cast<std::size_t>(c)] = static_cast<int>(b);
  const CellAdjacency adj = build_cell_adjacency(m);
  std::vector<Index> degree(n), out;
  std::vector<char> numbered(n, 0);
  std::vector<long long> visit(n, -1);
  long long stamp = 0;
  out.reserve(n);
  block_offsets.assign(1, 0);
  for (std::size_t b = 0; b < blocks.size(); ++b) {
    reverse_cuthill_mckee(CellSet(adj, &block_of, static_cast<int>(b)), blocks[b], false, degree, numbered, visit,
                          stamp, out);
    GSK_CHECK(out.size() == static_cast<std::size_t>(block_offsets.back()) + blocks[b].size(),
              "internal error: a block ordering lost cells");
    block_offsets.push_back(static_cast<Index>(out.size()));
  }
  return out;
}

}  // namespace gsk
