// fot_dispatch.hpp -- maps a fot_functor description to the compiled
// functor type.
#ifndef FOT_DISPATCH_HPP
#define FOT_DISPATCH_HPP

#include <type_traits>

#include "fot_functors.hpp"
#include "fot_internal.hpp"

namespace fot {

// Calls body.template operator()<F, SOA>() for the functor `f` names.
template <class Body>
fot_status dispatch(const fot_functor *f, int soa, Body &&body) {
  if (!f) return fail(FOT_ERR_INVALID_ARGUMENT, "null functor");
#define FOT_SOA(...)                                                                   \
  return soa ? body.template operator()<__VA_ARGS__, true>()                          \
             : body.template operator()<__VA_ARGS__, false>()
  switch (f->kernel) {
  case FOT_KERNEL_RES_CALC:
    FOT_SOA(ResCalcF);
  case FOT_KERNEL_HEX_SCATTER:
    FOT_SOA(HexScatterF);
  case FOT_KERNEL_VOLNA:
    FOT_SOA(VolnaF);
  case FOT_KERNEL_MINI_AERO:
    FOT_SOA(MiniAeroF);
  case FOT_KERNEL_GETACC:
    FOT_SOA(GetaccF);
  case FOT_KERNEL_LULESH:
    FOT_SOA(LuleshF);
  case FOT_KERNEL_MINE: {
#define FOT_MINE_PC(MD_, P_)                                                           \
  if (f->mesh_dim == MD_ && f->point_dim == P_) {                                      \
    if (f->cell_dim == 1) { FOT_SOA(MineF<MD_, P_, 1>); }                              \
    if (f->cell_dim == P_) { FOT_SOA(MineF<MD_, P_, P_>); }                            \
  }
    FOT_MINE_PC(2, 1)
    FOT_MINE_PC(2, 2)
    FOT_MINE_PC(2, 4)
    FOT_MINE_PC(2, 8)
    FOT_MINE_PC(4, 1)
    FOT_MINE_PC(4, 2)
    FOT_MINE_PC(4, 4)
    FOT_MINE_PC(4, 8)
    FOT_MINE_PC(8, 1)
    FOT_MINE_PC(8, 2)
    FOT_MINE_PC(8, 4)
    FOT_MINE_PC(8, 8)
#undef FOT_MINE_PC
    return fail(FOT_ERR_UNSUPPORTED, "mine: MeshDim %u PointDim %u CellDim %u not built",
                f->mesh_dim, f->point_dim, f->cell_dim);
  }
  default:
    return fail(FOT_ERR_UNSUPPORTED, "kernel %u is not built", f->kernel);
  }
#undef FOT_SOA
}

// Launch bound of the instantiations that run a functor's small blocks.
// Under a 1024-thread bound a kernel may use 128 VGPRs, where mini_aero's
// face flux spills up to 100 B per lane (none in the AOS two-phase and
// per-colour forms); under 512 it may use 256, and the face flux needs
// about 145-165.  Hierarchical and owner blocks of at most
// this many threads run a kernel bounded at it, and direct launches (whose
// results do not depend on the block size) always do.  (lulesh's element
// forces, at the cap too, measured the same at 512.)
template <class F>
constexpr int kSmallBound = std::is_same_v<F, MiniAeroF> ? 512 : 1024;
// Waves per SIMD the small-bound hierarchical instantiation is compiled for
// (at most 512 / waves VGPRs each).  mini_aero at 4 (128 VGPRs, spill-free
// in AOS) measured 5-10 % slower than at 3 on blocks of 128 and 256 faces,
// faster only on 512.
template <class F>
constexpr int kSmallBoundWaves = std::is_same_v<F, MiniAeroF> ? 3 : 1;

// Functors whose two-phase launches have an instantiation that loads their
// read-once cell data with nontemporal hints (StreamCellArgs): the hex
// corner forces (192 B per element), mini_aero's face values (96 B per
// face) and lulesh's stresses.  (getacc's corner data and mine's edge
// weights measured 21 % and 13 % slower that way at 1024^2, and volna's edge
// values in the 64-mesh batch 1.5 % slower, so they have none.)
template <class F>
constexpr bool kStreamsCells = std::is_same_v<F, HexScatterF> || std::is_same_v<F, MiniAeroF> ||
                               std::is_same_v<F, LuleshF>;

} // namespace fot

#endif
