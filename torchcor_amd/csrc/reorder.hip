// Reverse Cuthill-McKee node order of a symmetric sparsity pattern (host).
//
// The reference offers RCM as torchcor.core.reorder.RCM
// (torchcor/core/reorder.py:7-48, through scipy's reverse_cuthill_mckee) but
// does not use it in Monodomain.  Here it is applied inside the simulator
// when it narrows the pattern: on a mesh with scattered node numbers (a
// randomly permuted unstructured mesh) every neighbour gather of the SELL
// passes would otherwise touch its own cache line.
//
// Method: Cuthill-McKee breadth-first numbering (neighbours of each node
// appended in increasing degree, ties by index), started per connected
// component from a pseudo-peripheral node found with the George-Liu
// iteration (start from a minimum-degree node; repeatedly take the
// minimum-degree node of the deepest level of the rooted level structure
// while that increases the depth), and the whole sequence reversed.
#include <algorithm>
#include <new>
#include <vector>

#include "tc_internal.h"

namespace tc {
namespace {

struct Graph {
  int64_t n;
  const int32_t* rp;
  const int32_t* col;
  std::vector<int32_t> deg;  // off-diagonal entries of each row
};

// Rooted level structure from `root` over nodes not yet numbered: returns its
// depth (number of levels) and leaves the last level's nodes in `last`.
// `seen` holds the epoch of the search that reached each node.
int level_structure(const Graph& g, int32_t root, const std::vector<uint8_t>& numbered,
                    std::vector<int32_t>& seen, int32_t epoch, std::vector<int32_t>& queue,
                    std::vector<int32_t>& last) {
  queue.clear();
  queue.push_back(root);
  seen[root] = epoch;
  size_t level_begin = 0;
  int depth = 0;
  while (level_begin < queue.size()) {
    const size_t level_end = queue.size();
    ++depth;
    for (size_t q = level_begin; q < level_end; ++q) {
      const int32_t v = queue[q];
      for (int32_t e = g.rp[v]; e < g.rp[v + 1]; ++e) {
        const int32_t u = g.col[e];
        if (u == v || numbered[u] || seen[u] == epoch) continue;
        seen[u] = epoch;
        queue.push_back(u);
      }
    }
    if (queue.size() == level_end) {
      last.assign(queue.begin() + level_begin, queue.begin() + level_end);
      break;
    }
    level_begin = level_end;
  }
  return depth;
}

int32_t min_degree(const Graph& g, const std::vector<int32_t>& nodes) {
  int32_t best = nodes[0];
  for (int32_t v : nodes)
    if (g.deg[v] < g.deg[best] || (g.deg[v] == g.deg[best] && v < best)) best = v;
  return best;
}

int64_t bandwidth(const Graph& g, const int32_t* pos) {
  int64_t bw = 0;
  for (int64_t i = 0; i < g.n; ++i)
    for (int32_t e = g.rp[i]; e < g.rp[i + 1]; ++e) {
      const int64_t d = pos ? (int64_t)pos[i] - pos[g.col[e]] : i - (int64_t)g.col[e];
      bw = std::max(bw, d < 0 ? -d : d);
    }
  return bw;
}

int rcm_order(int64_t n, const int32_t* row_ptr_host, const int32_t* col_host,
              int32_t* order_host, int64_t* bw_in_host, int64_t* bw_out_host) {
  TC_CHECK(n >= 0 && n < (int64_t)INT32_MAX, TC_ERR_ARG, "tc_rcm_order: bad n %lld", (long long)n);
  TC_CHECK(n == 0 || (row_ptr_host && col_host && order_host), TC_ERR_ARG,
           "tc_rcm_order: null array");
  if (n == 0) {
    if (bw_in_host) *bw_in_host = 0;
    if (bw_out_host) *bw_out_host = 0;
    return TC_OK;
  }
  TC_CHECK(row_ptr_host[0] == 0, TC_ERR_ARG, "tc_rcm_order: row_ptr[0] != 0");
  Graph g{n, row_ptr_host, col_host, std::vector<int32_t>(n)};
  for (int64_t i = 0; i < n; ++i) {
    TC_CHECK(row_ptr_host[i + 1] >= row_ptr_host[i], TC_ERR_ARG,
             "tc_rcm_order: row_ptr decreases at row %lld", (long long)i);
    int32_t d = 0;
    for (int32_t e = row_ptr_host[i]; e < row_ptr_host[i + 1]; ++e) {
      const int32_t c = col_host[e];
      TC_CHECK(c >= 0 && c < n, TC_ERR_ARG, "tc_rcm_order: column %d out of range", c);
      d += (c != i);
    }
    g.deg[i] = d;
  }
  if (bw_in_host) *bw_in_host = bandwidth(g, nullptr);

  // candidate seeds: nodes by (degree, index)
  std::vector<int32_t> by_deg(n);
  for (int64_t i = 0; i < n; ++i) by_deg[i] = (int32_t)i;
  std::stable_sort(by_deg.begin(), by_deg.end(),
                   [&](int32_t a, int32_t b) { return g.deg[a] < g.deg[b]; });

  std::vector<uint8_t> numbered(n, 0);
  std::vector<int32_t> seen(n, -1), queue, last, nb;
  queue.reserve(n);
  int32_t epoch = 0;
  int64_t count = 0;
  size_t seed_at = 0;
  while (count < n) {
    while (numbered[by_deg[seed_at]]) ++seed_at;
    int32_t root = by_deg[seed_at];
    int depth = level_structure(g, root, numbered, seen, epoch++, queue, last);
    for (int it = 0; it < 32; ++it) {
      const int32_t x = min_degree(g, last);
      std::vector<int32_t> last_x;
      const int dx = level_structure(g, x, numbered, seen, epoch++, queue, last_x);
      if (dx <= depth) break;
      root = x;
      depth = dx;
      last.swap(last_x);
    }
    // Cuthill-McKee numbering of the component from root
    int64_t head = count;
    order_host[count++] = root;
    numbered[root] = 1;
    while (head < count) {
      const int32_t v = order_host[head++];
      nb.clear();
      for (int32_t e = g.rp[v]; e < g.rp[v + 1]; ++e) {
        const int32_t u = g.col[e];
        if (numbered[u]) continue;
        numbered[u] = 1;
        nb.push_back(u);
      }
      std::sort(nb.begin(), nb.end(), [&](int32_t a, int32_t b) {
        return g.deg[a] < g.deg[b] || (g.deg[a] == g.deg[b] && a < b);
      });
      for (int32_t u : nb) order_host[count++] = u;
    }
  }
  std::reverse(order_host, order_host + n);
  if (bw_out_host) {
    std::vector<int32_t> pos(n);
    for (int64_t i = 0; i < n; ++i) pos[order_host[i]] = (int32_t)i;
    *bw_out_host = bandwidth(g, pos.data());
  }
  return TC_OK;
}

}  // namespace
}  // namespace tc

extern "C" int tc_rcm_order(int64_t n, const int32_t* row_ptr_host, const int32_t* col_host,
                            int32_t* order_host, int64_t* bw_in_host, int64_t* bw_out_host) {
  try {
    return tc::rcm_order(n, row_ptr_host, col_host, order_host, bw_in_host, bw_out_host);
  } catch (const std::bad_alloc&) {
    tc::set_error("tc_rcm_order: out of host memory (n = %lld)", (long long)n);
    return TC_ERR_SIZE;
  }
}
