// fot_plan.cpp -- host planning of the loop engine: global colouring, the
// hierarchical block / thread colouring plan, partition reordering, point
// renumbering and the cell graphs the partitioners consume.
//
// These are integer algorithms whose outputs must be identical to the
// reference's (colourings, plans, permutations are checked bit-exact), so the
// tie-breaking rules of the reference are kept; the data structures are not
// (sorts and flat arrays instead of std::map / std::set).
#include <algorithm>
#include <cstring>
#include <numeric>

#include "fot_internal.hpp"

using fot::Colours;
using fot::fail;

namespace {

// Mesh::getAvailableColour<true> (mesh.hpp:287-315): least-used available
// colour, first index on ties.
unsigned least_used(const Colours &avail, const uint32_t *set_sizes,
                    unsigned n) {
  unsigned colour = n;
  for (unsigned i = 0; i < n; ++i)
    if (avail.test(i) && (colour >= n || set_sizes[colour] > set_sizes[i]))
      colour = i;
  return colour;
}

// Point -> incident cells (ascending), CSR.
void point_cells(const uint32_t *map, uint32_t num_cells, uint32_t dim,
                 uint32_t num_points, std::vector<uint32_t> &off,
                 std::vector<uint32_t> &cells) {
  off.assign(size_t(num_points) + 1, 0);
  const size_t len = size_t(num_cells) * dim;
  for (size_t k = 0; k < len; ++k) ++off[map[k] + 1];
  for (uint32_t p = 0; p < num_points; ++p) off[p + 1] += off[p];
  cells.resize(len);
  std::vector<uint32_t> fill(off.begin(), off.end() - 1);
  for (size_t k = 0; k < len; ++k) cells[fill[map[k]]++] = uint32_t(k / dim);
}

} // namespace

namespace fot {

// In-block thread colouring, colourCellsWithOrdering (colouring.cu:12-102):
// cell graph with one entry per shared (slot, slot) pair, smallest-last
// order (minimum degree, the LAST such cell on ties), then first fit in
// reverse order.  The smallest-last selection uses degree buckets of cell
// bitsets instead of the reference's O(b^2) scan; it selects the same cell.
int colour_block(const uint32_t *local, uint32_t nc, uint32_t md,
                 uint8_t *colours) {
  if (nc == 0) return 0;
  uint32_t np = 0;
  for (uint32_t k = 0; k < nc * md; ++k) np = std::max(np, local[k] + 1);
  std::vector<uint32_t> poff, pcell;
  point_cells(local, nc, md, np, poff, pcell);
  std::vector<uint32_t> aoff(nc + 1, 0), adj;
  adj.reserve(size_t(nc) * md * 4);
  for (uint32_t i = 0; i < nc; ++i) {
    for (uint32_t o = 0; o < md; ++o) {
      const uint32_t p = local[i * md + o];
      for (uint32_t r = poff[p]; r < poff[p + 1]; ++r)
        if (pcell[r] != i) adj.push_back(pcell[r]);
    }
    aoff[i + 1] = uint32_t(adj.size());
  }
  std::vector<uint32_t> deg(nc);
  uint32_t maxdeg = 0;
  for (uint32_t i = 0; i < nc; ++i) {
    deg[i] = aoff[i + 1] - aoff[i];
    maxdeg = std::max(maxdeg, deg[i]);
  }
  const uint32_t words = (nc + 63) / 64;
  std::vector<uint64_t> bucket(size_t(maxdeg + 1) * words, 0);
  std::vector<uint32_t> bcount(maxdeg + 1, 0);
  std::vector<uint8_t> done(nc, 0);
  for (uint32_t i = 0; i < nc; ++i) {
    bucket[size_t(deg[i]) * words + i / 64] |= uint64_t(1) << (i % 64);
    ++bcount[deg[i]];
  }
  std::vector<uint32_t> order(nc);
  uint32_t dmin = 0;
  for (uint32_t it = 0; it < nc; ++it) {
    while (bcount[dmin] == 0) ++dmin;
    const uint64_t *b = &bucket[size_t(dmin) * words];
    uint32_t v = 0;
    for (uint32_t w = words; w-- > 0;)
      if (b[w]) {
        v = w * 64 + 63 - uint32_t(__builtin_clzll(b[w]));
        break;
      }
    bucket[size_t(dmin) * words + v / 64] &= ~(uint64_t(1) << (v % 64));
    --bcount[dmin];
    done[v] = 1;
    order[it] = v;
    for (uint32_t r = aoff[v]; r < aoff[v + 1]; ++r) {
      const uint32_t n = adj[r];
      if (done[n]) continue;
      bucket[size_t(deg[n]) * words + n / 64] &= ~(uint64_t(1) << (n % 64));
      --bcount[deg[n]];
      --deg[n];
      bucket[size_t(deg[n]) * words + n / 64] |= uint64_t(1) << (n % 64);
      ++bcount[deg[n]];
      if (deg[n] < dmin) dmin = deg[n];
    }
  }
  std::vector<Colours> cc(nc);
  Colours used;
  int ncol = 0;
  for (uint32_t i = nc; i-- > 0;) {
    const uint32_t v = order[i];
    Colours occ;
    for (uint32_t r = aoff[v]; r < aoff[v + 1]; ++r) occ |= cc[adj[r]];
    Colours avail = Colours::available(used, occ);
    if (avail.none()) {
      if (ncol >= 128) return -1;
      used.set(unsigned(ncol++));
      avail = Colours::available(used, occ);
    }
    const unsigned c = avail.first();
    colours[v] = uint8_t(c);
    cc[v].set(c);
  }
  return ncol;
}

// Block boundaries of the hierarchical plan (colouring.hpp:70-115): cells
// stably sorted by partition id, a new block where the id changes or the
// block reaches block_size; without a partition, consecutive chunks.
void split_blocks(uint32_t num_cells, const uint32_t *partition, uint32_t block_size,
                  std::vector<uint32_t> &order, std::vector<uint32_t> &blocks) {
  order.resize(num_cells);
  blocks.clear();
  if (partition) {
    std::vector<uint64_t> key(num_cells);
    for (uint32_t i = 0; i < num_cells; ++i) key[i] = (uint64_t(partition[i]) << 32) | i;
    std::sort(key.begin(), key.end());
    blocks.push_back(0);
    order[0] = uint32_t(key[0]);
    uint32_t tid = 0;
    for (uint32_t i = 1; i < num_cells; ++i) {
      order[i] = uint32_t(key[i]);
      if (++tid == block_size) {
        tid = 0;
        blocks.push_back(i);
        continue;
      }
      if ((key[i - 1] >> 32) != (key[i] >> 32)) {
        tid = 0;
        blocks.push_back(i);
      }
    }
  } else {
    std::iota(order.begin(), order.end(), 0u);
    for (uint32_t i = 0; i < num_cells; i += block_size) blocks.push_back(i);
  }
  blocks.push_back(num_cells);
}

} // namespace fot

namespace {
using fot::colour_block;
} // namespace

static fot_status fot_colour_cells_impl(const uint32_t *c2n, uint32_t num_cells,
                                       uint32_t mesh_dim, uint32_t num_points,
                                       uint8_t *colour_out,
                                       uint32_t *num_colours) {
  if ((!c2n && num_cells) || (!colour_out && num_cells) || !num_colours ||
      mesh_dim == 0)
    return fail(FOT_ERR_INVALID_ARGUMENT, "fot_colour_cells: bad arguments");
  // Mesh::colourCells (mesh.hpp:117-162)
  std::vector<Colours> pc(num_points);
  uint32_t set_sizes[64] = {0};
  Colours used;
  uint32_t ncol = 0;
  for (uint32_t i = 0; i < num_cells; ++i) {
    Colours occ;
    for (uint32_t j = 0; j < mesh_dim; ++j) {
      const uint32_t p = c2n[size_t(mesh_dim) * i + j];
      if (p >= num_points)
        return fail(FOT_ERR_INVALID_INPUT, "cell %u: point %u out of range", i, p);
      occ |= pc[p];
    }
    Colours avail = Colours::available(used, occ);
    if (avail.none()) {
      if (ncol >= 64)
        return fail(FOT_ERR_COLOUR_OVERFLOW,
                    "global colouring needs more than 64 colours");
      used.set(ncol++);
      avail = Colours::available(used, occ);
    }
    const unsigned c = least_used(avail, set_sizes, 64);
    colour_out[i] = uint8_t(c);
    for (uint32_t j = 0; j < mesh_dim; ++j) pc[c2n[size_t(mesh_dim) * i + j]].set(c);
    ++set_sizes[c];
  }
  *num_colours = ncol;
  return FOT_OK;
}
extern "C" fot_status fot_colour_cells(const uint32_t *c2n, uint32_t num_cells,
                                       uint32_t mesh_dim, uint32_t num_points,
                                       uint8_t *colour_out,
                                       uint32_t *num_colours) {
  return fot::guarded("fot_colour_cells", [&] {
    return fot_colour_cells_impl(c2n, num_cells, mesh_dim, num_points, colour_out, num_colours);
  });
}

static fot_status fot_colour_cells_with_ordering_impl(const uint32_t *mapping,
                                                     uint32_t num_cells,
                                                     uint32_t mapping_dim,
                                                     uint8_t *colours_out,
                                                     uint32_t *num_colours) {
  if ((!mapping && num_cells) || !num_colours || mapping_dim == 0)
    return fail(FOT_ERR_INVALID_ARGUMENT, "bad arguments");
  const int n = colour_block(mapping, num_cells, mapping_dim, colours_out);
  if (n < 0) return fail(FOT_ERR_COLOUR_OVERFLOW, "more than 128 thread colours");
  *num_colours = uint32_t(n);
  return FOT_OK;
}
extern "C" fot_status fot_colour_cells_with_ordering(const uint32_t *mapping,
                                                     uint32_t num_cells,
                                                     uint32_t mapping_dim,
                                                     uint8_t *colours_out,
                                                     uint32_t *num_colours) {
  return fot::guarded("fot_colour_cells_with_ordering", [&] {
    return fot_colour_cells_with_ordering_impl(mapping, num_cells, mapping_dim, colours_out,
                                               num_colours);
  });
}

namespace fot {

// HierarchicalColourMemory ctor (colouring.hpp:44-147) with colourBlock
// (:169-235), sortCellsByColours (:246-272) and cellListSOA (:300-308).
fot_status build_hplan(const uint32_t *const *c2n, const uint32_t *mesh_dim,
                       uint32_t num_mappings, uint32_t num_cells,
                       uint32_t num_points0, const uint32_t *partition,
                       uint32_t block_size, fot_hplan &out) {
  if (num_cells == 0 || block_size == 0 || num_mappings == 0 ||
      num_mappings > FOT_MAX_MAPPINGS)
    return fail(FOT_ERR_INVALID_ARGUMENT, "hplan: empty mesh or block size 0");
  const uint32_t md0 = mesh_dim[0];
  out = fot_hplan{};
  out.num_mappings = num_mappings;
  out.num_cells = num_cells;
  for (uint32_t m = 0; m < num_mappings; ++m) out.mesh_dim[m] = mesh_dim[m];
  // block boundaries over positions of `order`
  std::vector<uint32_t> order, blocks;
  split_blocks(num_cells, partition, block_size, order, blocks);
  for (uint32_t m = 0; m < num_mappings; ++m)
    for (size_t k = 0; k < size_t(num_cells) * mesh_dim[m]; ++k)
      if (m == 0 && c2n[0][k] >= num_points0)
        return fail(FOT_ERR_INVALID_INPUT, "mapping 0 entry %zu out of range", k);

  std::vector<Colours> point_colours(num_points0);
  Colours used;
  std::vector<uint32_t> set_sizes;
  std::vector<uint64_t> keys;
  std::vector<uint32_t> local, perm;
  std::vector<uint8_t> tcol;
  for (size_t b = 0; b + 1 < blocks.size(); ++b) {
    const uint32_t from = blocks[b], to = blocks[b + 1], bs = to - from;
    Colours occ;
    for (uint32_t i = from; i < to; ++i)
      for (uint32_t j = 0; j < md0; ++j)
        occ |= point_colours[c2n[0][size_t(md0) * order[i] + j]];
    Colours avail = Colours::available(used, occ);
    if (avail.none()) {
      if (set_sizes.size() >= 128)
        return fail(FOT_ERR_COLOUR_OVERFLOW, "more than 128 block colours");
      used.set(unsigned(set_sizes.size()));
      set_sizes.push_back(0);
      out.colours.emplace_back();
      for (uint32_t m = 0; m < num_mappings; ++m) out.colours.back().cell_list[m].clear();
      avail = Colours::available(used, occ);
    }
    const unsigned colour = least_used(avail, set_sizes.data(), unsigned(set_sizes.size()));
    ++set_sizes[colour];
    HostColour &hc = out.colours[colour];
    // colour the written points; local ids = rank in ascending point order
    keys.resize(size_t(bs) * md0);
    for (uint32_t i = 0; i < bs; ++i)
      for (uint32_t j = 0; j < md0; ++j) {
        const uint32_t p = c2n[0][size_t(md0) * order[from + i] + j];
        point_colours[p].set(colour);
        keys[size_t(i) * md0 + j] = (uint64_t(p) << 32) | (uint64_t(i) * md0 + j);
      }
    std::sort(keys.begin(), keys.end());
    local.resize(size_t(bs) * md0);
    const size_t cache_base = hc.points_to_be_cached.size();
    for (size_t r = 0; r < keys.size(); ++r) {
      const uint32_t p = uint32_t(keys[r] >> 32);
      if (r == 0 || p != uint32_t(keys[r - 1] >> 32)) hc.points_to_be_cached.push_back(p);
      local[uint32_t(keys[r])] = uint32_t(hc.points_to_be_cached.size() - cache_base - 1);
    }
    hc.points_to_be_cached_offsets.push_back(uint32_t(hc.points_to_be_cached.size()));
    hc.shared_size = std::max<uint32_t>(hc.shared_size,
                                        uint32_t(hc.points_to_be_cached.size() - cache_base));
    tcol.resize(bs);
    const int nt = colour_block(local.data(), bs, md0, tcol.data());
    if (nt < 0) return fail(FOT_ERR_COLOUR_OVERFLOW, "more than 128 thread colours");
    hc.num_cell_colours.push_back(uint8_t(nt));
    // sortCellsByColours: stable by thread colour
    perm.resize(bs);
    {
      uint32_t cnt[257] = {0};
      for (uint32_t i = 0; i < bs; ++i) ++cnt[tcol[i] + 1];
      for (int k = 0; k < 256; ++k) cnt[k + 1] += cnt[k];
      for (uint32_t i = 0; i < bs; ++i) perm[cnt[tcol[i]]++] = i;
    }
    for (uint32_t i = 0; i < bs; ++i) {
      const uint32_t src = perm[i];
      const uint32_t cell = order[from + src];
      hc.cell_colours.push_back(tcol[src]);
      hc.cell_inv_permutation.push_back(cell);
      for (uint32_t j = 0; j < md0; ++j) hc.cell_list[0].push_back(local[size_t(src) * md0 + j]);
      for (uint32_t m = 1; m < num_mappings; ++m)
        for (uint32_t j = 0; j < mesh_dim[m]; ++j)
          hc.cell_list[m].push_back(c2n[m][size_t(mesh_dim[m]) * cell + j]);
    }
    hc.block_offsets.push_back(hc.block_offsets.back() + bs);
  }
  // AOS -> SOA per colour (cellListSOA)
  std::vector<uint32_t> tmp;
  for (HostColour &hc : out.colours) {
    const size_t n = hc.cell_colours.size();
    for (uint32_t m = 0; m < num_mappings; ++m) {
      const uint32_t d = mesh_dim[m];
      tmp.resize(n * d);
      for (size_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < d; ++j) tmp[j * n + i] = hc.cell_list[m][i * d + j];
      hc.cell_list[m].swap(tmp);
    }
  }
  return FOT_OK;
}

} // namespace fot

static fot_status fot_hplan_build_impl(const uint32_t *const *c2n,
                                      const uint32_t *mesh_dim,
                                      uint32_t num_mappings, uint32_t num_cells,
                                      uint32_t num_points0,
                                      const uint32_t *partition,
                                      uint32_t block_size, fot_hplan **plan) {
  if (!c2n || !mesh_dim || !plan)
    return fail(FOT_ERR_INVALID_ARGUMENT, "fot_hplan_build: null argument");
  *plan = nullptr;
  fot_hplan *p = new (std::nothrow) fot_hplan;
  if (!p) return fail(FOT_ERR_OUT_OF_MEMORY, "fot_hplan_build");
  const fot_status s = fot::build_hplan(c2n, mesh_dim, num_mappings, num_cells,
                                        num_points0, partition, block_size, *p);
  if (s != FOT_OK) {
    delete p;
    return s;
  }
  *plan = p;
  return FOT_OK;
}
extern "C" fot_status fot_hplan_build(const uint32_t *const *c2n,
                                      const uint32_t *mesh_dim,
                                      uint32_t num_mappings, uint32_t num_cells,
                                      uint32_t num_points0,
                                      const uint32_t *partition,
                                      uint32_t block_size, fot_hplan **plan) {
  return fot::guarded("fot_hplan_build", [&] {
    return fot_hplan_build_impl(c2n, mesh_dim, num_mappings, num_cells, num_points0, partition,
                                block_size, plan);
  });
}

extern "C" void fot_hplan_free(fot_hplan *plan) { delete plan; }

static fot_status fot_hplan_info_get_impl(const fot_hplan *plan,
                                         fot_hplan_info *info) {
  if (!plan || !info) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  *info = fot_hplan_info{};
  info->num_colours = uint32_t(plan->colours.size());
  info->num_cells = plan->num_cells;
  double tc = 0;
  for (const auto &c : plan->colours) {
    info->num_blocks += uint32_t(c.num_cell_colours.size());
    info->num_cached += uint32_t(c.points_to_be_cached.size());
    info->max_cached = std::max(info->max_cached, c.shared_size);
    for (size_t b = 0; b + 1 < c.block_offsets.size(); ++b)
      info->max_block = std::max(info->max_block, c.block_offsets[b + 1] - c.block_offsets[b]);
    for (uint8_t n : c.num_cell_colours) {
      info->max_thread_colours = std::max<uint32_t>(info->max_thread_colours, n);
      tc += n;
    }
  }
  info->avg_thread_colours = info->num_blocks ? tc / info->num_blocks : 0.0;
  return FOT_OK;
}
extern "C" fot_status fot_hplan_info_get(const fot_hplan *plan,
                                         fot_hplan_info *info) {
  return fot::guarded("fot_hplan_info_get", [&] { return fot_hplan_info_get_impl(plan, info); });
}

static fot_status fot_hplan_colour_get_impl(const fot_hplan *plan, uint32_t colour,
                                           fot_hplan_colour *out) {
  if (!plan || !out || colour >= plan->colours.size())
    return fail(FOT_ERR_INVALID_ARGUMENT, "fot_hplan_colour_get: bad colour");
  const fot::HostColour &c = plan->colours[colour];
  *out = fot_hplan_colour{};
  out->num_blocks = uint32_t(c.num_cell_colours.size());
  out->num_cells = uint32_t(c.cell_colours.size());
  out->num_cached = uint32_t(c.points_to_be_cached.size());
  out->shared_size = c.shared_size;
  out->points_to_be_cached = c.points_to_be_cached.data();
  out->points_to_be_cached_offsets = c.points_to_be_cached_offsets.data();
  for (uint32_t m = 0; m < plan->num_mappings; ++m) out->cell_list[m] = c.cell_list[m].data();
  out->cell_colours = c.cell_colours.data();
  out->num_cell_colours = c.num_cell_colours.data();
  out->block_offsets = c.block_offsets.data();
  out->cell_inv_permutation = c.cell_inv_permutation.data();
  return FOT_OK;
}
extern "C" fot_status fot_hplan_colour_get(const fot_hplan *plan, uint32_t colour,
                                           fot_hplan_colour *out) {
  return fot::guarded("fot_hplan_colour_get", [&] {
    return fot_hplan_colour_get_impl(plan, colour, out);
  });
}

static fot_status fot_reorder_to_partition_impl(uint32_t *partition,
                                               uint32_t num_cells,
                                               uint32_t *inv_perm) {
  if ((!partition || !inv_perm) && num_cells)
    return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  // mesh.hpp:227-245: sort (partition, cell) pairs
  std::vector<uint64_t> key(num_cells);
  for (uint32_t i = 0; i < num_cells; ++i) key[i] = (uint64_t(partition[i]) << 32) | i;
  std::sort(key.begin(), key.end());
  for (uint32_t i = 0; i < num_cells; ++i) {
    partition[i] = uint32_t(key[i] >> 32);
    inv_perm[i] = uint32_t(key[i]);
  }
  return FOT_OK;
}
extern "C" fot_status fot_reorder_to_partition(uint32_t *partition,
                                               uint32_t num_cells,
                                               uint32_t *inv_perm) {
  return fot::guarded("fot_reorder_to_partition", [&] {
    return fot_reorder_to_partition_impl(partition, num_cells, inv_perm);
  });
}

static fot_status fot_point_renumbering_by_partition_impl(
    const uint32_t *c2n, uint32_t num_cells, uint32_t mesh_dim,
    uint32_t num_points, const uint32_t *partition, uint32_t *perm) {
  if ((!c2n || !partition) && num_cells) return fail(FOT_ERR_INVALID_ARGUMENT, "null");
  if (!perm && num_points) return fail(FOT_ERR_INVALID_ARGUMENT, "null perm");
  // getPointToPartition (mesh.hpp:346-363): sorted unique parts per point
  std::vector<uint64_t> pairs;
  pairs.reserve(size_t(num_cells) * mesh_dim);
  for (uint32_t i = 0; i < num_cells; ++i)
    for (uint32_t j = 0; j < mesh_dim; ++j) {
      const uint32_t p = c2n[size_t(mesh_dim) * i + j];
      if (p >= num_points) return fail(FOT_ERR_INVALID_INPUT, "point out of range");
      pairs.push_back((uint64_t(p) << 32) | partition[i]);
    }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  std::vector<uint32_t> off(size_t(num_points) + 1, 0), ids(pairs.size());
  for (size_t r = 0; r < pairs.size(); ++r) {
    ids[r] = uint32_t(pairs[r]);
    ++off[(pairs[r] >> 32) + 1];
  }
  for (uint32_t p = 0; p < num_points; ++p) off[p + 1] += off[p];
  // getPointRenumberingPermutation2 (mesh.hpp:365-388): stable sort by number
  // of parts (descending), then part list (lexicographically descending).
  // Untouched points compare equal among themselves and keep their order.
  std::vector<uint32_t> inv(num_points);
  std::iota(inv.begin(), inv.end(), 0u);
  std::stable_sort(inv.begin(), inv.end(), [&](uint32_t a, uint32_t b) {
    const uint32_t na = off[a + 1] - off[a], nb = off[b + 1] - off[b];
    if (na != nb) return na > nb;
    return std::lexicographical_compare(ids.begin() + off[b], ids.begin() + off[b + 1],
                                        ids.begin() + off[a], ids.begin() + off[a + 1]);
  });
  for (uint32_t i = 0; i < num_points; ++i) perm[inv[i]] = i;
  return FOT_OK;
}
extern "C" fot_status fot_point_renumbering_by_partition(
    const uint32_t *c2n, uint32_t num_cells, uint32_t mesh_dim,
    uint32_t num_points, const uint32_t *partition, uint32_t *perm) {
  return fot::guarded("fot_point_renumbering_by_partition", [&] {
    return fot_point_renumbering_by_partition_impl(c2n, num_cells, mesh_dim, num_points, partition,
                                                   perm);
  });
}

static fot_status fot_point_renumbering_first_touch_impl(const uint32_t *c2n,
                                                        uint32_t num_cells,
                                                        uint32_t mesh_dim,
                                                        uint32_t num_points,
                                                        uint32_t *perm) {
  if ((!c2n && num_cells) || (!perm && num_points))
    return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  // mesh.hpp:247-270
  std::fill(perm, perm + num_points, num_points);
  uint32_t next = 0;
  for (size_t k = 0; k < size_t(num_cells) * mesh_dim; ++k) {
    const uint32_t p = c2n[k];
    if (p >= num_points) return fail(FOT_ERR_INVALID_INPUT, "point out of range");
    if (perm[p] == num_points) perm[p] = next++;
  }
  for (uint32_t p = 0; p < num_points; ++p)
    if (perm[p] == num_points) perm[p] = next++;
  return FOT_OK;
}
extern "C" fot_status fot_point_renumbering_first_touch(const uint32_t *c2n,
                                                        uint32_t num_cells,
                                                        uint32_t mesh_dim,
                                                        uint32_t num_points,
                                                        uint32_t *perm) {
  return fot::guarded("fot_point_renumbering_first_touch", [&] {
    return fot_point_renumbering_first_touch_impl(c2n, num_cells, mesh_dim, num_points, perm);
  });
}

static fot_status fot_mesh_reorder_cells_impl(uint32_t *c2n0, uint32_t num_cells,
                                             uint32_t mesh_dim,
                                             const uint32_t *point_perm,
                                             uint32_t *inv_perm) {
  if ((!c2n0 || !point_perm || !inv_perm) && num_cells)
    return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  // mesh.hpp:199-225: renumber, then sort cells by (sorted tuple, index)
  const size_t len = size_t(num_cells) * mesh_dim;
  for (size_t k = 0; k < len; ++k) c2n0[k] = point_perm[c2n0[k]];
  std::vector<uint32_t> tuples(c2n0, c2n0 + len);
  for (uint32_t i = 0; i < num_cells; ++i)
    std::sort(tuples.begin() + size_t(i) * mesh_dim, tuples.begin() + size_t(i + 1) * mesh_dim);
  std::iota(inv_perm, inv_perm + num_cells, 0u);
  std::sort(inv_perm, inv_perm + num_cells, [&](uint32_t a, uint32_t b) {
    const uint32_t *ta = &tuples[size_t(a) * mesh_dim], *tb = &tuples[size_t(b) * mesh_dim];
    for (uint32_t k = 0; k < mesh_dim; ++k)
      if (ta[k] != tb[k]) return ta[k] < tb[k];
    return a < b;
  });
  std::vector<uint32_t> old(c2n0, c2n0 + len);
  for (uint32_t i = 0; i < num_cells; ++i)
    std::copy_n(&old[size_t(inv_perm[i]) * mesh_dim], mesh_dim, c2n0 + size_t(i) * mesh_dim);
  return FOT_OK;
}
extern "C" fot_status fot_mesh_reorder_cells(uint32_t *c2n0, uint32_t num_cells,
                                             uint32_t mesh_dim,
                                             const uint32_t *point_perm,
                                             uint32_t *inv_perm) {
  return fot::guarded("fot_mesh_reorder_cells", [&] {
    return fot_mesh_reorder_cells_impl(c2n0, num_cells, mesh_dim, point_perm, inv_perm);
  });
}

static fot_status fot_graph_csr_impl(const uint32_t *c2n, uint32_t num_cells,
                                    uint32_t mesh_dim, uint32_t num_points,
                                    uint32_t *indices, uint32_t *endpoints,
                                    uint32_t *num_arcs) {
  if (!indices || !endpoints || !num_arcs || (!c2n && num_cells))
    return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  // GraphCSR (reorder.hpp:39-83): per point, incident cells ascending; other
  // points of the cell appended unless already in this point's row.
  std::vector<uint32_t> off, cells;
  for (size_t k = 0; k < size_t(num_cells) * mesh_dim; ++k)
    if (c2n[k] >= num_points) return fail(FOT_ERR_INVALID_INPUT, "point out of range");
  point_cells(c2n, num_cells, mesh_dim, num_points, off, cells);
  uint32_t arcs = 0;
  indices[0] = 0;
  for (uint32_t p = 0; p < num_points; ++p) {
    const uint32_t row = arcs;
    for (uint32_t r = off[p]; r < off[p + 1]; ++r) {
      const uint32_t cell = cells[r];
      for (uint32_t j = 0; j < mesh_dim; ++j) {
        const uint32_t q = c2n[size_t(mesh_dim) * cell + j];
        if (q == p) continue;
        if (std::find(endpoints + row, endpoints + arcs, q) == endpoints + arcs)
          endpoints[arcs++] = q;
      }
    }
    indices[p + 1] = arcs;
  }
  *num_arcs = arcs;
  return FOT_OK;
}
extern "C" fot_status fot_graph_csr(const uint32_t *c2n, uint32_t num_cells,
                                    uint32_t mesh_dim, uint32_t num_points,
                                    uint32_t *indices, uint32_t *endpoints,
                                    uint32_t *num_arcs) {
  return fot::guarded("fot_graph_csr", [&] {
    return fot_graph_csr_impl(c2n, num_cells, mesh_dim, num_points, indices, endpoints, num_arcs);
  });
}

static fot_status fot_cell_to_cell_csr_impl(const uint32_t *const *c2n,
                                           const uint32_t *mesh_dim,
                                           uint32_t num_mappings,
                                           uint32_t num_cells,
                                           const uint32_t *num_points,
                                           uint32_t *indices, uint32_t *endpoints,
                                           uint64_t *num_arcs) {
  if (!c2n || !mesh_dim || !num_points || !indices || !num_arcs)
    return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  // Mesh::getCellToCellGraph (mesh.hpp:317-344) then GraphCSR of the pair
  // mesh: each row holds the cell's distinct neighbours in ascending order.
  std::vector<std::vector<uint32_t>> poff(num_mappings), pcell(num_mappings);
  for (uint32_t m = 0; m < num_mappings; ++m) {
    for (size_t k = 0; k < size_t(num_cells) * mesh_dim[m]; ++k)
      if (c2n[m][k] >= num_points[m]) return fail(FOT_ERR_INVALID_INPUT, "point out of range");
    point_cells(c2n[m], num_cells, mesh_dim[m], num_points[m], poff[m], pcell[m]);
  }
  std::vector<uint32_t> row;
  uint64_t arcs = 0;
  indices[0] = 0;
  for (uint32_t i = 0; i < num_cells; ++i) {
    row.clear();
    for (uint32_t m = 0; m < num_mappings; ++m)
      for (uint32_t o = 0; o < mesh_dim[m]; ++o) {
        const uint32_t p = c2n[m][size_t(mesh_dim[m]) * i + o];
        for (uint32_t r = poff[m][p]; r < poff[m][p + 1]; ++r)
          if (pcell[m][r] != i) row.push_back(pcell[m][r]);
      }
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    if (endpoints) std::copy(row.begin(), row.end(), endpoints + arcs);
    arcs += row.size();
    if (arcs > 0xffffffffull) return fail(FOT_ERR_UNSUPPORTED, "graph too large");
    indices[i + 1] = uint32_t(arcs);
  }
  *num_arcs = arcs;
  return FOT_OK;
}
extern "C" fot_status fot_cell_to_cell_csr(const uint32_t *const *c2n,
                                           const uint32_t *mesh_dim,
                                           uint32_t num_mappings,
                                           uint32_t num_cells,
                                           const uint32_t *num_points,
                                           uint32_t *indices, uint32_t *endpoints,
                                           uint64_t *num_arcs) {
  return fot::guarded("fot_cell_to_cell_csr", [&] {
    return fot_cell_to_cell_csr_impl(c2n, mesh_dim, num_mappings, num_cells, num_points, indices,
                                     endpoints, num_arcs);
  });
}

// countCacheLinesForBlock (problem.hpp:654-672) with the line size as a
// parameter: distinct lines of `line_bytes` holding the points' data (AOS:
// the dim values of each point; SOA: one line set per component, so dim x
// the lines of the ids).  The reference counts 32-byte sectors; CDNA's L2
// line is 128 bytes.
static fot_status fot_count_cache_lines_impl(const uint32_t *points, uint32_t n, uint32_t dim,
                                            uint32_t type_size, uint32_t line_bytes, int soa,
                                            uint64_t *lines) {
  if ((!points && n) || !lines || dim == 0 || type_size == 0 || line_bytes < type_size)
    return fail(FOT_ERR_INVALID_ARGUMENT, "bad cache-line query");
  const uint64_t per_line = line_bytes / type_size;
  std::vector<uint64_t> ids;
  ids.reserve(size_t(n) * (soa ? 1 : dim));
  for (uint32_t i = 0; i < n; ++i) {
    if (soa) ids.push_back(points[i] / per_line);
    else
      for (uint32_t d = 0; d < dim; ++d) ids.push_back((uint64_t(points[i]) * dim + d) / per_line);
  }
  std::sort(ids.begin(), ids.end());
  const uint64_t distinct = uint64_t(std::unique(ids.begin(), ids.end()) - ids.begin());
  *lines = (soa ? dim : 1) * distinct;
  return FOT_OK;
}
extern "C" fot_status fot_count_cache_lines(const uint32_t *points, uint32_t n, uint32_t dim,
                                            uint32_t type_size, uint32_t line_bytes, int soa,
                                            uint64_t *lines) {
  return fot::guarded("fot_count_cache_lines", [&] {
    return fot_count_cache_lines_impl(points, n, dim, type_size, line_bytes, soa, lines);
  });
}
