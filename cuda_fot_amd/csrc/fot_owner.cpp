// fot_owner.cpp -- owner-compute hierarchical plan.
//
// The reference orders the writeback of blocks that share points by block
// colouring: one launch per block colour (problem.hpp:273-304).  On MI355X a
// block colour is a small launch whose dependent loads leave the CUs idle,
// and ordering blocks inside one launch needs L2 writeback / invalidate on
// every block.  This plan removes the need for any order between blocks:
// each mapping-0 point gets one owner block, every cell touching an owned
// point runs in the owner (cells on a block boundary run in each of their
// owners), and a block writes back only its owned points.  Inside a block the
// reference's thread colouring still orders the LDS increments; it only has
// to separate cells sharing an owned point.  No reference counterpart.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "fot_internal.hpp"

using fot::fail;

namespace fot {

fot_status build_owner_plan(const uint32_t *c2n0, uint32_t num_cells, uint32_t md,
                            uint32_t num_points, const uint32_t *partition,
                            uint32_t block_size, const uint32_t *point_partition,
                            OwnerPlan &out) {
  if (num_cells == 0 || block_size == 0 || md == 0 || md > 16)
    return fail(FOT_ERR_INVALID_ARGUMENT, "owner plan: empty mesh or bad shape");
  out = OwnerPlan{};
  out.mesh_dim = md;
  for (size_t k = 0; k < size_t(num_cells) * md; ++k)
    if (c2n0[k] >= num_points) return fail(FOT_ERR_INVALID_INPUT, "point out of range");
  std::vector<uint32_t> order, blocks, own_part;
  const uint32_t kNone = 0xffffffffu;
  const uint32_t kMaxCells = 1024;
  std::vector<uint32_t> owner(num_points, kNone);
  uint32_t nb0 = 0;
  if (!partition && !point_partition) {
    // no partition given: parts of at most block_size points of the point
    // graph of mapping 0
    std::vector<uint32_t> ind(size_t(num_points) + 1),
        end(std::max<size_t>(1, size_t(num_cells) * md * (md > 1 ? md - 1 : 1)));
    uint32_t arcs = 0, nparts = 0;
    fot_status st = fot_graph_csr(c2n0, num_cells, md, num_points, ind.data(), end.data(), &arcs);
    if (st != FOT_OK) return st;
    own_part.resize(num_points);
    st = fot_kway_partition(ind.data(), end.data(), num_points, block_size, own_part.data(),
                            &nparts);
    if (st != FOT_OK) return st;
    point_partition = own_part.data();
  }
  if (point_partition) {
    // owners given per point: blocks in ascending part id, cells in plan
    // order sorted by the owner of their first point
    std::vector<uint32_t> ids(point_partition, point_partition + num_points);
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    nb0 = uint32_t(ids.size());
    for (uint32_t p = 0; p < num_points; ++p)
      owner[p] = uint32_t(std::lower_bound(ids.begin(), ids.end(), point_partition[p]) -
                          ids.begin());
    std::vector<uint64_t> key(num_cells);
    for (uint32_t i = 0; i < num_cells; ++i) key[i] = (uint64_t(owner[c2n0[size_t(md) * i]]) << 32) | i;
    std::sort(key.begin(), key.end());
    order.resize(num_cells);
    for (uint32_t i = 0; i < num_cells; ++i) order[i] = uint32_t(key[i]);
  } else {
    // owner = first block (in plan order) of the cell split (partition)
    // whose cells touch the point
    split_blocks(num_cells, partition, block_size, order, blocks);
    nb0 = uint32_t(blocks.size() - 1);
    for (uint32_t b = 0; b < nb0; ++b)
      for (uint32_t i = blocks[b]; i < blocks[b + 1]; ++i)
        for (uint32_t j = 0; j < md; ++j) {
          const uint32_t p = c2n0[size_t(md) * order[i] + j];
          if (owner[p] == kNone) owner[p] = b;
        }
  }
  // A block runs every cell touching its owned points, which can exceed one
  // workgroup (1024 threads) when the partition is poor: such blocks split
  // their owned points, in ascending order, into sub-blocks of at most
  // kMaxCells cells.
  std::vector<uint32_t> poff, pcell;
  {
    poff.assign(size_t(num_points) + 1, 0);
    for (size_t k = 0; k < size_t(num_cells) * md; ++k) ++poff[c2n0[k] + 1];
    for (uint32_t p = 0; p < num_points; ++p) poff[p + 1] += poff[p];
    pcell.resize(poff[num_points]);
    std::vector<uint32_t> fill(poff.begin(), poff.end() - 1);
    for (size_t k = 0; k < size_t(num_cells) * md; ++k) pcell[fill[c2n0[k]]++] = uint32_t(k / md);
  }
  uint32_t nb = nb0;
  {
    std::vector<std::vector<uint32_t>> pts(nb0);
    for (uint32_t p = 0; p < num_points; ++p)
      if (owner[p] != kNone) pts[owner[p]].push_back(p);
    std::vector<uint32_t> mark(num_cells, kNone);
    uint32_t next = 0;
    for (uint32_t b = 0; b < nb0; ++b) {
      uint32_t count = 0;
      for (uint32_t p : pts[b]) {
        if (poff[p + 1] - poff[p] > kMaxCells)
          return fail(FOT_ERR_UNSUPPORTED, "point %u touches more than %u cells", p, kMaxCells);
        uint32_t fresh = 0;
        for (uint32_t r = poff[p]; r < poff[p + 1]; ++r) fresh += mark[pcell[r]] != next;
        if (count > 0 && count + fresh > kMaxCells) {
          ++next;
          count = 0;
        }
        for (uint32_t r = poff[p]; r < poff[p + 1]; ++r)
          if (mark[pcell[r]] != next) {
            mark[pcell[r]] = next;
            ++count;
          }
        owner[p] = next;
      }
      if (!pts[b].empty()) ++next;
    }
    nb = next;
  }
  // owned points per block, ascending (counting sort by owner)
  std::vector<uint32_t> own_off(nb + 1, 0), own_pts;
  for (uint32_t p = 0; p < num_points; ++p)
    if (owner[p] != kNone) ++own_off[owner[p] + 1];
  for (uint32_t b = 0; b < nb; ++b) own_off[b + 1] += own_off[b];
  own_pts.resize(own_off[nb]);
  {
    std::vector<uint32_t> fill(own_off.begin(), own_off.end() - 1);
    for (uint32_t p = 0; p < num_points; ++p)
      if (owner[p] != kNone) own_pts[fill[owner[p]]++] = p;
  }
  // cells of each block: every cell touching an owned point, in plan order
  std::vector<uint32_t> cell_off(nb + 1, 0), cell_list;
  {
    std::vector<uint32_t> owners;
    owners.reserve(md);
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<uint32_t> fill;
      if (pass == 1) {
        for (uint32_t b = 0; b < nb; ++b) cell_off[b + 1] += cell_off[b];
        cell_list.resize(cell_off[nb]);
        fill.assign(cell_off.begin(), cell_off.end() - 1);
      }
      for (uint32_t i = 0; i < num_cells; ++i) {
        const uint32_t c = order[i];
        owners.clear();
        for (uint32_t j = 0; j < md; ++j) owners.push_back(owner[c2n0[size_t(md) * c + j]]);
        std::sort(owners.begin(), owners.end());
        owners.erase(std::unique(owners.begin(), owners.end()), owners.end());
        for (uint32_t b : owners) {
          if (pass == 0) ++cell_off[b + 1];
          else cell_list[fill[b]++] = c;
        }
      }
    }
  }
  std::vector<uint32_t> halo, local, col_ids, perm;
  std::vector<uint8_t> tcol;
  std::vector<uint32_t> local_of(num_points, kNone);
  for (uint32_t b = 0; b < nb; ++b) {
    const uint32_t c0 = cell_off[b], n = cell_off[b + 1] - c0;
    const uint32_t no = own_off[b + 1] - own_off[b];
    // local ids: owned points first, then halo points, each ascending
    for (uint32_t k = 0; k < no; ++k) local_of[own_pts[own_off[b] + k]] = k;
    halo.clear();
    for (uint32_t i = 0; i < n; ++i)
      for (uint32_t j = 0; j < md; ++j) {
        const uint32_t p = c2n0[size_t(md) * cell_list[c0 + i] + j];
        if (owner[p] != b) halo.push_back(p);
      }
    std::sort(halo.begin(), halo.end());
    halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
    for (uint32_t k = 0; k < halo.size(); ++k) local_of[halo[k]] = no + k;
    if (no + halo.size() > 0xffffu)
      return fail(FOT_ERR_UNSUPPORTED, "owner block with more than 65535 points");
    local.resize(size_t(n) * md);
    col_ids.resize(size_t(n) * md);
    // colouring ids: owned slots keep their id, halo slots get fresh ids so
    // they never conflict (their increments are not accumulated)
    uint32_t fresh = no;
    for (uint32_t i = 0; i < n; ++i)
      for (uint32_t j = 0; j < md; ++j) {
        const uint32_t l = local_of[c2n0[size_t(md) * cell_list[c0 + i] + j]];
        local[size_t(i) * md + j] = l;
        col_ids[size_t(i) * md + j] = l < no ? l : fresh++;
      }
    tcol.resize(n);
    const int nt = colour_block(col_ids.data(), n, md, tcol.data());
    if (nt < 0) return fail(FOT_ERR_COLOUR_OVERFLOW, "more than 128 thread colours");
    perm.resize(n);
    {
      uint32_t cnt[257] = {0};
      for (uint32_t i = 0; i < n; ++i) ++cnt[tcol[i] + 1];
      for (int k = 0; k < 256; ++k) cnt[k + 1] += cnt[k];
      for (uint32_t i = 0; i < n; ++i) perm[cnt[tcol[i]]++] = i;
    }
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t s = perm[i];
      out.cells.push_back(cell_list[c0 + s]);
      out.colour.push_back(tcol[s]);
      for (uint32_t j = 0; j < md; ++j) out.loc.push_back(local[size_t(s) * md + j]);
    }
    out.cached.insert(out.cached.end(), own_pts.begin() + own_off[b],
                      own_pts.begin() + own_off[b + 1]);
    out.cached.insert(out.cached.end(), halo.begin(), halo.end());
    out.blk_owned.push_back(no);
    out.blk_ncol.push_back(uint8_t(nt));
    out.blk_cell_off.push_back(uint32_t(out.cells.size()));
    out.blk_cache_off.push_back(uint32_t(out.cached.size()));
    for (uint32_t k = 0; k < no; ++k) local_of[own_pts[own_off[b] + k]] = kNone;
    for (uint32_t p : halo) local_of[p] = kNone;
  }
  return FOT_OK;
}

} // namespace fot

struct fot_oplan {
  fot::OwnerPlan plan;
};

static fot_status fot_oplan_build_impl(const uint32_t *c2n0, uint32_t num_cells,
                                      uint32_t mesh_dim, uint32_t num_points,
                                      const uint32_t *partition, uint32_t block_size,
                                      const uint32_t *point_partition, fot_oplan **out) {
  if (!c2n0 || !out) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  fot_oplan *p = new (std::nothrow) fot_oplan;
  if (!p) return fail(FOT_ERR_OUT_OF_MEMORY, "fot_oplan_build");
  const fot_status s = fot::build_owner_plan(c2n0, num_cells, mesh_dim, num_points, partition,
                                             block_size, point_partition, p->plan);
  if (s != FOT_OK) {
    delete p;
    return s;
  }
  *out = p;
  return FOT_OK;
}
extern "C" fot_status fot_oplan_build(const uint32_t *c2n0, uint32_t num_cells,
                                      uint32_t mesh_dim, uint32_t num_points,
                                      const uint32_t *partition, uint32_t block_size,
                                      const uint32_t *point_partition, fot_oplan **out) {
  return fot::guarded("fot_oplan_build", [&] {
    return fot_oplan_build_impl(c2n0, num_cells, mesh_dim, num_points, partition, block_size,
                                point_partition, out);
  });
}

extern "C" void fot_oplan_free(fot_oplan *plan) { delete plan; }

static fot_status fot_oplan_get_impl(const fot_oplan *plan, fot_oplan_view *v) {
  if (!plan || !v) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  const fot::OwnerPlan &p = plan->plan;
  v->mesh_dim = p.mesh_dim;
  v->num_blocks = uint32_t(p.blk_owned.size());
  v->num_positions = uint32_t(p.cells.size());
  v->num_cached = uint32_t(p.cached.size());
  v->blk_cell_off = p.blk_cell_off.data();
  v->blk_cache_off = p.blk_cache_off.data();
  v->blk_owned = p.blk_owned.data();
  v->cached = p.cached.data();
  v->cells = p.cells.data();
  v->loc = p.loc.data();
  v->colour = p.colour.data();
  v->blk_ncol = p.blk_ncol.data();
  return FOT_OK;
}
extern "C" fot_status fot_oplan_get(const fot_oplan *plan, fot_oplan_view *v) {
  return fot::guarded("fot_oplan_get", [&] { return fot_oplan_get_impl(plan, v); });
}
