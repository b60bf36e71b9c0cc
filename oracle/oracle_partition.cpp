/*
 * oracle_partition.cpp -- CPU oracle for the heuristic block partitioner.
 *
 * TEST INFRASTRUCTURE ONLY (see fot_oracle.h).  The reference's
 * HeuristicalPartition (partition.cu:128-275) keeps its candidates in a
 * PriorityQueue built on std::multimap (priority_queue.hpp:7-58) whose
 * comparator reads a mutable static (details/heuristical_partition.hpp:29-48,
 * mutated at partition.cu:178,194-195).  Its result therefore depends on the
 * tree that libstdc++'s multimap builds, so this one oracle function is C++
 * and uses the same container with the same comparator and the same sequence
 * of operations.  The METIS coarse step (partition.cu:102-126) is not
 * available here; the coarse partition is an input.
 */
#include <algorithm>
#include <bitset>
#include <cstdint>
#include <map>
#include <numeric>
#include <vector>

#include "fot_oracle.h"

namespace {

using colourset = std::bitset<128>;
constexpr uint32_t UNDEF = 0xffffffffu;

/* the mutable static the comparator reads (partition.cu:98-99) */
colourset g_used_colours;

struct prio {
  colourset occupied;
  uint32_t reuse = 0;
  uint32_t added = 0;
  /* details/heuristical_partition.hpp:37-48 */
  bool operator<(const prio &rhs) const {
    const int d = (~occupied & g_used_colours).any();
    const int rd = (~rhs.occupied & g_used_colours).any();
    if (reuse == 0 || rhs.reuse == 0) return reuse < rhs.reuse;
    if (d != rd) return d < rd;
    if (reuse != rhs.reuse) return reuse < rhs.reuse;
    return added < rhs.added;
  }
};

/* priority_queue.hpp:7-58 */
template <class P> struct pqueue {
  std::map<uint32_t, typename std::multimap<P, uint32_t>::iterator> idx;
  std::multimap<P, uint32_t> pri;
  void push(uint32_t i, P p) {
    auto it = pri.insert(std::make_pair(p, i));
    idx[i] = it;
  }
  std::pair<uint32_t, P> pop_max() {
    auto r = pri.rbegin();
    auto it = (++r).base();
    auto res = std::make_pair(it->second, it->first);
    idx.erase(it->second);
    pri.erase(it);
    return res;
  }
  void modify(uint32_t i, P p) {
    auto it = idx.find(i);
    pri.erase(it->second);
    it->second = pri.insert(std::make_pair(p, i));
  }
  std::pair<bool, P> get(uint32_t i) const {
    auto it = idx.find(i);
    if (it == idx.end()) return std::make_pair(false, P{});
    return std::make_pair(true, it->second->first);
  }
};

struct heuristic {
  std::vector<uint32_t> ind, end; /* CSR after reorderInverse */
  std::vector<uint32_t> coarse;
  uint32_t block_size;
  std::vector<uint32_t> result, inverse;

  /* partition.cu:232-243 and reorder.hpp:177-202 */
  void reorder_to_coarse(const uint32_t *ci, const uint32_t *ce, uint32_t n) {
    inverse.resize(n);
    std::iota(inverse.begin(), inverse.end(), 0u);
    std::stable_sort(inverse.begin(), inverse.end(), [&](uint32_t a, uint32_t b) {
      return coarse[a] < coarse[b];
    });
    std::vector<uint32_t> c2(n), perm(n);
    for (uint32_t i = 0; i < n; ++i) c2[i] = coarse[inverse[i]];
    coarse.swap(c2);
    for (uint32_t i = 0; i < n; ++i) perm[inverse[i]] = i;
    ind.assign(n + 1, 0);
    end.clear();
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t o = inverse[i];
      for (uint32_t r = ci[o]; r < ci[o + 1]; ++r) end.push_back(perm[ce[r]]);
      ind[i + 1] = static_cast<uint32_t>(end.size());
    }
  }

  /* partition.cu:176-230 */
  void grow(uint32_t start, uint32_t to, uint32_t size, uint32_t block_id) {
    g_used_colours.reset();
    pqueue<prio> q;
    for (uint32_t i = start; i < to; ++i)
      if (result[i] == UNDEF) q.push(i, prio{});
    for (uint32_t i = 0; i < size; ++i) {
      auto mp = q.pop_max();
      const uint32_t cur = mp.first;
      const prio cp = mp.second;
      result[cur] = block_id;
      colourset avail = ~cp.occupied & g_used_colours;
      if (avail.none()) {
        g_used_colours <<= 1;
        g_used_colours.set(0);
        avail = ~cp.occupied & g_used_colours;
      }
      unsigned colour = 0;
      while (colour < 128 && !avail[colour]) ++colour;
      colourset assigned;
      assigned.set(colour);
      for (uint32_t j = ind[cur]; j < ind[cur + 1]; ++j) {
        const uint32_t nb = end[j];
        if (result[nb] != UNDEF) continue;
        auto op = q.get(nb);
        if (!op.first) continue;
        prio np = op.second;
        np.added = i + 1;
        ++np.reuse;
        np.occupied |= assigned;
        q.modify(nb, np);
      }
    }
  }

  /* partition.cu:146-174 */
  void within(uint32_t from, uint32_t to, uint32_t &block_id) {
    const uint32_t cbs = to - from;
    const uint32_t nb = (cbs + block_size - 1) / block_size;
    const uint32_t small = cbs / nb;
    const uint32_t large = (cbs + nb - 1) / nb;
    const uint32_t nlarge = cbs % nb;
    uint32_t sp = from;
    for (uint32_t b = 0; b < nb; ++b) {
      while (result[sp] != UNDEF) ++sp;
      grow(sp, to, b < nlarge ? large : small, block_id++);
    }
  }
};

} // namespace

extern "C" int oracle_heuristic_partition(const uint32_t *ci,
                                          const uint32_t *ce, uint32_t n,
                                          const uint32_t *coarse_partition,
                                          uint32_t block_size,
                                          uint32_t *result_out) {
  if (n == 0 || block_size == 0) return -1;
  heuristic h;
  h.coarse.assign(coarse_partition, coarse_partition + n);
  h.block_size = block_size;
  h.reorder_to_coarse(ci, ce, n);
  h.result.assign(n, UNDEF);
  uint32_t from = 0, block_id = 0;
  for (uint32_t to = 1; to < n; ++to) {
    if (h.coarse[to] != h.coarse[to - 1]) {
      h.within(from, to, block_id);
      from = to;
    }
  }
  h.within(from, n, block_id);
  /* reorderData<false>(result, 1, inverse) (partition.cu:142) */
  for (uint32_t i = 0; i < n; ++i) result_out[h.inverse[i]] = h.result[i];
  return static_cast<int>(block_id);
}

extern "C" int oracle_pq_run(const int *ops, uint32_t num_ops, int *out) {
  pqueue<int> q;
  int w = 0;
  for (uint32_t k = 0; k < num_ops; ++k) {
    const int op = ops[3 * k], a = ops[3 * k + 1], b = ops[3 * k + 2];
    if (op == 0) {
      q.push(static_cast<uint32_t>(a), b);
    } else if (op == 1) {
      auto r = q.pop_max();
      out[w++] = static_cast<int>(r.first);
      out[w++] = r.second;
    } else if (op == 2) {
      q.modify(static_cast<uint32_t>(a), b);
    } else {
      auto r = q.get(static_cast<uint32_t>(a));
      out[w++] = r.first ? 1 : 0;
      out[w++] = r.second;
    }
  }
  return w;
}
