// fot_internal.hpp -- shared internals of libfot (not installed).
#ifndef FOT_INTERNAL_HPP
#define FOT_INTERNAL_HPP

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "fot_capi.h"

namespace fot {

// Records the message of the last failing call on this thread and returns
// the status, so call sites read `return fail(FOT_ERR_..., "...")`.
fot_status fail(fot_status status, const char *fmt, ...);

// Every exported entry point runs its body through this, so no C++ exception
// crosses the C ABI (the host planner allocates through the standard
// library): out of host memory becomes FOT_ERR_OUT_OF_MEMORY, anything else
// FOT_ERR_INVALID_ARGUMENT, with the message kept as for any failure.
template <class Fn> fot_status guarded(const char *name, Fn &&fn) noexcept {
  try {
    return fn();
  } catch (const std::bad_alloc &) {
    return fail(FOT_ERR_OUT_OF_MEMORY, "%s: out of host memory", name);
  } catch (const std::exception &e) {
    return fail(FOT_ERR_INVALID_ARGUMENT, "%s: %s", name, e.what());
  } catch (...) {
    return fail(FOT_ERR_INVALID_ARGUMENT, "%s: unexpected exception", name);
  }
}

// 128-bit colour set, the width of the reference's Mesh::colourset_t
// (mesh.hpp:26).
struct Colours {
  uint64_t w[2] = {0, 0};
  bool test(unsigned i) const { return (w[i >> 6] >> (i & 63)) & 1u; }
  void set(unsigned i) { w[i >> 6] |= uint64_t(1) << (i & 63); }
  Colours operator|(const Colours &o) const {
    Colours r;
    r.w[0] = w[0] | o.w[0];
    r.w[1] = w[1] | o.w[1];
    return r;
  }
  Colours &operator|=(const Colours &o) {
    w[0] |= o.w[0];
    w[1] |= o.w[1];
    return *this;
  }
  // ~occupied & used
  static Colours available(const Colours &used, const Colours &occupied) {
    Colours r;
    r.w[0] = used.w[0] & ~occupied.w[0];
    r.w[1] = used.w[1] & ~occupied.w[1];
    return r;
  }
  bool none() const { return (w[0] | w[1]) == 0; }
  // lowest set bit, 128 if none
  unsigned first() const {
    if (w[0]) return unsigned(__builtin_ctzll(w[0]));
    if (w[1]) return 64u + unsigned(__builtin_ctzll(w[1]));
    return 128u;
  }
};

// Host side of the hierarchical plan, shaped like the reference's
// HierarchicalColourMemory::MemoryOfOneColour (colouring.hpp:24-41).
struct HostColour {
  std::vector<uint32_t> points_to_be_cached;
  std::vector<uint32_t> points_to_be_cached_offsets{0};
  std::vector<uint32_t> cell_list[FOT_MAX_MAPPINGS]; // SOA once built
  std::vector<uint8_t> cell_colours;
  std::vector<uint8_t> num_cell_colours;
  std::vector<uint32_t> block_offsets{0};
  std::vector<uint32_t> cell_inv_permutation;
  uint32_t shared_size = 0;
};

} // namespace fot

struct fot_hplan {
  uint32_t num_mappings = 0;
  uint32_t mesh_dim[FOT_MAX_MAPPINGS] = {};
  uint32_t num_cells = 0;
  std::vector<fot::HostColour> colours;
};

namespace fot {
// colourCellsWithOrdering on block-local ids; returns colours or -1.
int colour_block(const uint32_t *local, uint32_t nc, uint32_t md, uint8_t *colours);
void split_blocks(uint32_t num_cells, const uint32_t *partition, uint32_t block_size,
                  std::vector<uint32_t> &order, std::vector<uint32_t> &blocks);

// Owner-compute plan: every mapping-0 point is owned by one block; a block
// runs every cell touching a point it owns (cells on a block boundary run in
// each owner block) and writes back only its owned points.
struct OwnerPlan {
  uint32_t mesh_dim = 0;
  std::vector<uint32_t> blk_cell_off{0};  // [nb+1] ranges of plan positions
  std::vector<uint32_t> blk_cache_off{0}; // [nb+1] ranges of `cached`
  std::vector<uint32_t> blk_owned;        // [nb] owned points (a prefix of the range)
  std::vector<uint32_t> cached;           // owned points ascending, then halo ascending
  std::vector<uint32_t> cells;            // original cell of each plan position
  std::vector<uint32_t> loc;              // [positions * mesh_dim] block-local ids
  std::vector<uint8_t> colour;            // thread colour per position
  std::vector<uint8_t> blk_ncol;          // thread colours per block
};
fot_status build_owner_plan(const uint32_t *c2n0, uint32_t num_cells, uint32_t mesh_dim,
                            uint32_t num_points, const uint32_t *partition,
                            uint32_t block_size, const uint32_t *point_partition,
                            OwnerPlan &out);

fot_status build_hplan(const uint32_t *const *c2n, const uint32_t *mesh_dim,
                       uint32_t num_mappings, uint32_t num_cells,
                       uint32_t num_points0, const uint32_t *partition,
                       uint32_t block_size, fot_hplan &out);
} // namespace fot

#endif
