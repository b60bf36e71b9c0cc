// fot_partition.cpp -- partitioners and orderings of the loop engine:
// the reference's heuristic block partitioner, a k-way partitioner standing in
// for METIS and bandwidth-reducing orderings standing in for Scotch's GPS.
#include <algorithm>
#include <bitset>
#include <map>
#include <numeric>
#include <queue>
#include <vector>

#include "fot_internal.hpp"

using fot::fail;

namespace {

using Bits = std::bitset<128>;
constexpr uint32_t kUndef = 0xffffffffu;

// The candidate ranking of HeuristicalPartition::priority_t
// (details/heuristical_partition.hpp:29-48).  It reads the colours in use by
// the block being grown, a value that changes while candidates sit in the
// ordered container; the reference keeps it in a static that the comparator
// reads, so the container below does the same to rank candidates
// identically.
Bits g_block_colours;

struct Candidate {
  Bits occupied;
  uint32_t reuse = 0;
  uint32_t added = 0;
  bool operator<(const Candidate &o) const {
    const int a = (~occupied & g_block_colours).any();
    const int b = (~o.occupied & g_block_colours).any();
    if (reuse == 0 || o.reuse == 0) return reuse < o.reuse;
    if (a != b) return a < b;
    if (reuse != o.reuse) return reuse < o.reuse;
    return added < o.added;
  }
};

// Index-addressable max-queue (priority_queue.hpp:7-58): an ordered
// multimap of candidates plus a per-vertex handle; the maximum is the last
// element, so equal candidates pop most-recently-inserted first.
class CandidateQueue {
public:
  using Map = std::multimap<Candidate, uint32_t>;
  explicit CandidateQueue(uint32_t n) : handle_(n), present_(n, 0) {}
  void push(uint32_t v, const Candidate &c) {
    handle_[v] = map_.emplace(c, v);
    present_[v] = 1;
  }
  bool empty() const { return map_.empty(); }
  uint32_t pop_max(Candidate &c) {
    auto it = std::prev(map_.end());
    const uint32_t v = it->second;
    c = it->first;
    map_.erase(it);
    present_[v] = 0;
    return v;
  }
  bool present(uint32_t v) const { return present_[v] != 0; }
  const Candidate &get(uint32_t v) const { return handle_[v]->first; }
  void update(uint32_t v, const Candidate &c) {
    map_.erase(handle_[v]);
    handle_[v] = map_.emplace(c, v);
  }

private:
  Map map_;
  std::vector<Map::iterator> handle_;
  std::vector<uint8_t> present_;
};

} // namespace

static fot_status fot_heuristic_partition_impl(const uint32_t *ci,
                                              const uint32_t *ce, uint32_t n,
                                              const uint32_t *coarse_in,
                                              uint32_t block_size,
                                              uint32_t *result_out,
                                              uint32_t *num_blocks) {
  if (!ci || !coarse_in || !result_out || !num_blocks || n == 0 || block_size == 0)
    return fail(FOT_ERR_INVALID_ARGUMENT, "fot_heuristic_partition: bad arguments");
  // reorderCSRToPartition (partition.cu:232-243): vertices stably sorted by
  // coarse part, graph relabelled.
  std::vector<uint32_t> inv(n), pos(n);
  std::iota(inv.begin(), inv.end(), 0u);
  std::stable_sort(inv.begin(), inv.end(),
                   [&](uint32_t a, uint32_t b) { return coarse_in[a] < coarse_in[b]; });
  for (uint32_t i = 0; i < n; ++i) pos[inv[i]] = i;
  std::vector<uint32_t> coarse(n), off(size_t(n) + 1, 0), adj;
  adj.reserve(ci[n]);
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t o = inv[i];
    coarse[i] = coarse_in[o];
    for (uint32_t r = ci[o]; r < ci[o + 1]; ++r) {
      if (ce[r] >= n) return fail(FOT_ERR_INVALID_INPUT, "neighbour out of range");
      adj.push_back(pos[ce[r]]);
    }
    off[i + 1] = uint32_t(adj.size());
  }
  std::vector<uint32_t> result(n, kUndef);
  uint32_t block_id = 0;
  CandidateQueue queue(n);
  // growBlock (partition.cu:176-230)
  auto grow = [&](uint32_t start, uint32_t to, uint32_t size, uint32_t id) -> bool {
    g_block_colours.reset();
    for (uint32_t i = start; i < to; ++i)
      if (result[i] == kUndef) queue.push(i, Candidate{});
    for (uint32_t k = 0; k < size; ++k) {
      if (queue.empty()) return false;
      Candidate cur;
      const uint32_t v = queue.pop_max(cur);
      result[v] = id;
      Bits avail = ~cur.occupied & g_block_colours;
      if (avail.none()) {
        g_block_colours <<= 1;
        g_block_colours.set(0);
        avail = ~cur.occupied & g_block_colours;
      }
      unsigned colour = 0;
      while (colour < 128 && !avail[colour]) ++colour;
      for (uint32_t r = off[v]; r < off[v + 1]; ++r) {
        const uint32_t w = adj[r];
        if (result[w] != kUndef || !queue.present(w)) continue;
        Candidate c = queue.get(w);
        c.added = k + 1;
        ++c.reuse;
        c.occupied.set(colour);
        queue.update(w, c);
      }
    }
    // the reference drops the queue after each block
    while (!queue.empty()) {
      Candidate c;
      queue.pop_max(c);
    }
    return true;
  };
  // partitionWithin (partition.cu:146-174) per coarse part
  auto within = [&](uint32_t from, uint32_t to) -> bool {
    const uint32_t cbs = to - from;
    const uint32_t nb = (cbs + block_size - 1) / block_size;
    const uint32_t small = cbs / nb, large = (cbs + nb - 1) / nb, nlarge = cbs % nb;
    uint32_t sp = from;
    for (uint32_t b = 0; b < nb; ++b) {
      while (result[sp] != kUndef) ++sp;
      if (!grow(sp, to, b < nlarge ? large : small, block_id++)) return false;
    }
    return true;
  };
  uint32_t from = 0;
  for (uint32_t to = 1; to <= n; ++to) {
    if (to == n || coarse[to] != coarse[to - 1]) {
      if (!within(from, to)) return fail(FOT_ERR_INVALID_INPUT, "partition ran dry");
      from = to;
    }
  }
  for (uint32_t i = 0; i < n; ++i) result_out[inv[i]] = result[i];
  *num_blocks = block_id;
  return FOT_OK;
}
extern "C" fot_status fot_heuristic_partition(const uint32_t *ci,
                                              const uint32_t *ce, uint32_t n,
                                              const uint32_t *coarse_in,
                                              uint32_t block_size,
                                              uint32_t *result_out,
                                              uint32_t *num_blocks) {
  return fot::guarded("fot_heuristic_partition", [&] {
    return fot_heuristic_partition_impl(ci, ce, n, coarse_in, block_size, result_out, num_blocks);
  });
}

namespace {

// Breadth-first level structure from `root`; returns the last level's
// vertices and the depth.  Only vertices with comp[v] == c are visited.
uint32_t bfs_levels(const uint32_t *ci, const uint32_t *ce, uint32_t root,
                    std::vector<uint32_t> &level, std::vector<uint32_t> &queue) {
  queue.clear();
  queue.push_back(root);
  level[root] = 0;
  uint32_t depth = 0;
  for (size_t h = 0; h < queue.size(); ++h) {
    const uint32_t v = queue[h];
    for (uint32_t r = ci[v]; r < ci[v + 1]; ++r) {
      const uint32_t w = ce[r];
      if (level[w] == kUndef) {
        level[w] = level[v] + 1;
        depth = std::max(depth, level[w]);
        queue.push_back(w);
      }
    }
  }
  return depth;
}

// George-Liu pseudo-peripheral vertex of the component of `start`.
uint32_t pseudo_peripheral(const uint32_t *ci, const uint32_t *ce, uint32_t start,
                           std::vector<uint32_t> &level, std::vector<uint32_t> &queue) {
  uint32_t root = start, depth = 0;
  for (int iter = 0; iter < 16; ++iter) {
    for (uint32_t v : queue) level[v] = kUndef;
    const uint32_t d = bfs_levels(ci, ce, root, level, queue);
    if (iter > 0 && d <= depth) break;
    depth = d;
    // last level: pick the vertex of minimum degree (lowest id on ties)
    uint32_t best = root, bestdeg = kUndef;
    for (uint32_t v : queue)
      if (level[v] == d) {
        const uint32_t dg = ci[v + 1] - ci[v];
        if (dg < bestdeg || (dg == bestdeg && v < best)) {
          bestdeg = dg;
          best = v;
        }
      }
    if (best == root) break;
    root = best;
  }
  for (uint32_t v : queue) level[v] = kUndef;
  return root;
}

// Cuthill-McKee numbering from pseudo-peripheral roots, component by
// component (lowest unvisited vertex first); neighbours by ascending degree,
// then id.  `order` receives vertices in numbering order.
void cuthill_mckee(const uint32_t *ci, const uint32_t *ce, uint32_t n,
                   std::vector<uint32_t> &order) {
  std::vector<uint32_t> level(n, kUndef), queue, nb;
  std::vector<uint8_t> seen(n, 0);
  order.clear();
  order.reserve(n);
  for (uint32_t s = 0; s < n; ++s) {
    if (seen[s]) continue;
    queue.clear();
    queue.push_back(s);
    const uint32_t root = pseudo_peripheral(ci, ce, s, level, queue);
    size_t head = order.size();
    order.push_back(root);
    seen[root] = 1;
    for (; head < order.size(); ++head) {
      const uint32_t v = order[head];
      nb.clear();
      for (uint32_t r = ci[v]; r < ci[v + 1]; ++r)
        if (!seen[ce[r]]) {
          seen[ce[r]] = 1;
          nb.push_back(ce[r]);
        }
      std::sort(nb.begin(), nb.end(), [&](uint32_t a, uint32_t b) {
        const uint32_t da = ci[a + 1] - ci[a], db = ci[b + 1] - ci[b];
        return da != db ? da < db : a < b;
      });
      order.insert(order.end(), nb.begin(), nb.end());
    }
  }
}

} // namespace

// Reverse Cuthill-McKee as scipy.sparse.csgraph.reverse_cuthill_mckee
// computes it (scipy 1.15, the routine scripts/rcm.py:20 calls), restated
// from its published algorithm:
//  * the pattern is that of A + A^T (rows ascending, no duplicates);
//  * a vertex's degree is its row length, plus one when the row holds the
//    vertex itself;
//  * components are seeded in a degree order (`seed_order`: scipy takes
//    numpy's argsort of the degrees, whose order among equal degrees is the
//    host numpy's; NULL here means a stable sort, ties to the lower id);
//  * breadth first from each seed: each vertex of the current level appends
//    its unvisited neighbours in row order, then those are insertion-sorted
//    by degree (stable);
//  * the visiting order, reversed, is the new numbering.
static fot_status fot_rcm_order_impl(const uint32_t *ci, const uint32_t *ce, uint32_t n,
                                    const uint32_t *seed_order, uint32_t *perm) {
  if (!ci || !perm || (n && !ce && ci[n])) return fail(FOT_ERR_INVALID_ARGUMENT, "null");
  // the symmetric pattern, rows sorted and unique
  std::vector<uint32_t> off(size_t(n) + 1, 0);
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t r = ci[i]; r < ci[i + 1]; ++r) {
      if (ce[r] >= n) return fail(FOT_ERR_INVALID_ARGUMENT, "endpoint %u >= %u", ce[r], n);
      ++off[i + 1];
      ++off[ce[r] + 1];
    }
  for (uint32_t i = 0; i < n; ++i) off[i + 1] += off[i];
  std::vector<uint32_t> adj(off[n]), fill(off.begin(), off.end() - 1);
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t r = ci[i]; r < ci[i + 1]; ++r) {
      adj[fill[i]++] = ce[r];
      adj[fill[ce[r]]++] = i;
    }
  std::vector<uint32_t> row(size_t(n) + 1, 0), deg(n);
  {
    uint32_t w = 0;
    for (uint32_t i = 0; i < n; ++i) {
      auto b = adj.begin() + off[i], e = adj.begin() + off[i + 1];
      std::sort(b, e);
      e = std::unique(b, e);
      row[i] = w;
      const bool diag = std::binary_search(b, e, i);
      for (auto it = b; it != e; ++it) adj[w++] = *it;
      deg[i] = uint32_t(e - b) + (diag ? 1u : 0u);
    }
    row[n] = w;
  }
  std::vector<uint32_t> seeds(n);
  if (seed_order) {
    std::vector<uint8_t> hit(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
      if (seed_order[i] >= n || hit[seed_order[i]])
        return fail(FOT_ERR_INVALID_ARGUMENT, "seed_order is not a permutation");
      hit[seed_order[i]] = 1;
      seeds[i] = seed_order[i];
    }
  } else {
    for (uint32_t i = 0; i < n; ++i) seeds[i] = i;
    std::stable_sort(seeds.begin(), seeds.end(),
                     [&](uint32_t a, uint32_t b) { return deg[a] < deg[b]; });
  }
  std::vector<uint32_t> order;
  order.reserve(n);
  std::vector<uint8_t> seen(n, 0);
  for (uint32_t z = 0; z < n && order.size() < n; ++z) {
    const uint32_t seed = seeds[z];
    if (seen[seed]) continue;
    seen[seed] = 1;
    order.push_back(seed);
    size_t level_start = order.size() - 1, level_end = order.size();
    while (level_start < level_end) {
      for (size_t ii = level_start; ii < level_end; ++ii) {
        const uint32_t v = order[ii];
        const size_t first = order.size();
        for (uint32_t r = row[v]; r < row[v + 1]; ++r)
          if (!seen[adj[r]]) {
            seen[adj[r]] = 1;
            order.push_back(adj[r]);
          }
        // insertion sort of the appended vertices by degree, stable
        for (size_t k = first + 1; k < order.size(); ++k) {
          const uint32_t t = order[k];
          size_t l = k;
          while (l > first && deg[t] < deg[order[l - 1]]) {
            order[l] = order[l - 1];
            --l;
          }
          order[l] = t;
        }
      }
      level_start = level_end;
      level_end = order.size();
    }
  }
  for (uint32_t i = 0; i < n; ++i) perm[order[i]] = n - 1 - i; // reversed
  return FOT_OK;
}
extern "C" fot_status fot_rcm_order(const uint32_t *ci, const uint32_t *ce, uint32_t n,
                                    const uint32_t *seed_order, uint32_t *perm) {
  return fot::guarded("fot_rcm_order",
                      [&] { return fot_rcm_order_impl(ci, ce, n, seed_order, perm); });
}

static fot_status fot_gps_order_impl(const uint32_t *ci, const uint32_t *ce,
                                    uint32_t n, uint32_t *perm) {
  if (!ci || !perm || (n && !ce && ci[n])) return fail(FOT_ERR_INVALID_ARGUMENT, "null");
  // Level-structure numbering from a pseudo-diameter endpoint (the first
  // phase of Gibbs-Poole-Stockmeyer), not reversed.
  std::vector<uint32_t> order;
  cuthill_mckee(ci, ce, n, order);
  for (uint32_t i = 0; i < n; ++i) perm[order[i]] = i;
  return FOT_OK;
}
extern "C" fot_status fot_gps_order(const uint32_t *ci, const uint32_t *ce,
                                    uint32_t n, uint32_t *perm) {
  return fot::guarded("fot_gps_order", [&] { return fot_gps_order_impl(ci, ce, n, perm); });
}

static fot_status fot_kway_partition_impl(const uint32_t *ci, const uint32_t *ce,
                                         uint32_t n, uint32_t max_part,
                                         uint32_t *part, uint32_t *num_parts) {
  if (!ci || !part || !num_parts || max_part == 0)
    return fail(FOT_ERR_INVALID_ARGUMENT, "fot_kway_partition: bad arguments");
  // Greedy graph growing: seeds are taken in Cuthill-McKee order (so parts
  // follow a band through the graph), each part grows by repeatedly adding
  // the frontier vertex with the most neighbours already in the part (ties:
  // earliest in the seed order) until it holds `target` vertices.
  std::vector<uint32_t> order;
  cuthill_mckee(ci, ce, n, order);
  std::vector<uint32_t> rank(n);
  for (uint32_t i = 0; i < n; ++i) rank[order[i]] = i;
  const uint32_t nparts = (n + max_part - 1) / max_part;
  const uint32_t target = nparts ? (n + nparts - 1) / nparts : 0;
  std::fill(part, part + n, kUndef);
  std::vector<uint32_t> gain(n, 0), touched;
  // max-heap of (gain, -rank) with lazy deletion
  using Entry = std::pair<uint64_t, uint32_t>;
  std::priority_queue<Entry> heap;
  uint32_t pid = 0, cursor = 0;
  while (cursor < n) {
    while (cursor < n && part[order[cursor]] != kUndef) ++cursor;
    if (cursor >= n) break;
    uint32_t size = 0;
    heap = std::priority_queue<Entry>();
    touched.clear();
    auto key = [&](uint32_t v) { return (uint64_t(gain[v]) << 32) | (0xffffffffu - rank[v]); };
    heap.emplace(key(order[cursor]), order[cursor]);
    touched.push_back(order[cursor]);
    while (size < target && !heap.empty()) {
      const Entry e = heap.top();
      heap.pop();
      const uint32_t v = e.second;
      if (part[v] != kUndef || e.first != key(v)) continue;
      part[v] = pid;
      ++size;
      for (uint32_t r = ci[v]; r < ci[v + 1]; ++r) {
        const uint32_t w = ce[r];
        if (part[w] != kUndef) continue;
        if (gain[w] == 0) touched.push_back(w);
        ++gain[w];
        heap.emplace(key(w), w);
      }
    }
    for (uint32_t v : touched) gain[v] = 0;
    ++pid;
  }
  *num_parts = pid;
  return FOT_OK;
}
extern "C" fot_status fot_kway_partition(const uint32_t *ci, const uint32_t *ce,
                                         uint32_t n, uint32_t max_part,
                                         uint32_t *part, uint32_t *num_parts) {
  return fot::guarded("fot_kway_partition", [&] {
    return fot_kway_partition_impl(ci, ce, n, max_part, part, num_parts);
  });
}

namespace {

// Recursive bisection state: verts[lo, hi) holds the vertices of one piece;
// mark[v] == stamp marks the piece being split.
struct Bisector {
  const uint32_t *ci, *ce;
  std::vector<uint32_t> verts, mark, order, tmp;
  std::vector<uint8_t> seen;
  uint32_t stamp = 0;
  uint32_t *part;

  // Breadth-first order of the piece from `root`, restricted to it; pieces
  // in several components continue from their lowest unvisited position.
  // Returns the last vertex reached from `root`'s component.
  uint32_t bfs(uint32_t lo, uint32_t hi, uint32_t root) {
    order.clear();
    uint32_t last = root;
    auto visit = [&](uint32_t s) {
      size_t head = order.size();
      order.push_back(s);
      seen[s] = 1;
      for (; head < order.size(); ++head) {
        const uint32_t v = order[head];
        for (uint32_t r = ci[v]; r < ci[v + 1]; ++r) {
          const uint32_t w = ce[r];
          if (mark[w] == stamp && !seen[w]) {
            seen[w] = 1;
            order.push_back(w);
          }
        }
      }
    };
    visit(root);
    last = order.back();
    for (uint32_t i = lo; i < hi; ++i)
      if (!seen[verts[i]]) visit(verts[i]);
    for (uint32_t v : order) seen[v] = 0;
    return last;
  }

  void split(uint32_t lo, uint32_t hi, uint32_t p, uint32_t first) {
    if (p <= 1 || hi - lo <= 1) {
      for (uint32_t i = lo; i < hi; ++i) part[verts[i]] = first;
      return;
    }
    ++stamp;
    for (uint32_t i = lo; i < hi; ++i) mark[verts[i]] = stamp;
    // two sweeps: the far end of a sweep from the piece's first vertex, then
    // the level order from that end (a pseudo-peripheral start)
    const uint32_t far = bfs(lo, hi, verts[lo]);
    bfs(lo, hi, far);
    std::copy(order.begin(), order.end(), verts.begin() + lo);
    const uint32_t p1 = p / 2;
    const uint32_t n1 = uint32_t((uint64_t(hi - lo) * p1 + p / 2) / p);
    split(lo, lo + n1, p1, first);
    split(lo + n1, hi, p - p1, first + p1);
  }
};

} // namespace

static fot_status fot_bisect_partition_impl(const uint32_t *ci, const uint32_t *ce, uint32_t n,
                                           uint32_t max_part, uint32_t *part,
                                           uint32_t *num_parts) {
  if (!ci || !part || !num_parts || max_part == 0 || (n && !ce && ci[n]))
    return fail(FOT_ERR_INVALID_ARGUMENT, "fot_bisect_partition: bad arguments");
  const uint32_t nparts = (n + max_part - 1) / max_part;
  Bisector b;
  b.ci = ci, b.ce = ce, b.part = part;
  b.verts.resize(n);
  for (uint32_t i = 0; i < n; ++i) b.verts[i] = i;
  b.mark.assign(n, 0);
  b.seen.assign(n, 0);
  b.split(0, n, nparts, 0);
  *num_parts = nparts;
  return FOT_OK;
}
extern "C" fot_status fot_bisect_partition(const uint32_t *ci, const uint32_t *ce, uint32_t n,
                                           uint32_t max_part, uint32_t *part,
                                           uint32_t *num_parts) {
  return fot::guarded("fot_bisect_partition", [&] {
    return fot_bisect_partition_impl(ci, ce, n, max_part, part, num_parts);
  });
}
