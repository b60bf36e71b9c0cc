// Kernel structs of the C++ drop-in: the reference's functor API
// (<app>::StepSeq / StepOMP / StepGPUGlobal / StepGPUHierarchical with their
// `call<SOA>` signatures) over libfot.  The CPU structs run the user function
// on the host; the GPU structs launch libfot's gfx950 kernels through the
// C-ABI with exactly the reference's argument lists, and name the functor
// that Problem's engine loops (problem.hpp) plan and launch.
#ifndef FOT_KERNELS_HPP
#define FOT_KERNELS_HPP

#include <type_traits>

#include "data.hpp"

namespace fot {

// StepGPUGlobal::call (kernels/res_calc.hpp:104-117): one launch over the
// cells of one colour; device arrays of device pointers.
template <class Desc> struct GPUGlobal {
  static fot_functor functor() { return Desc::functor(); }
  template <bool SOA>
  static void call(const void **point_data, void *point_data_out, void **cell_data,
                   const MY_SIZE **cell_to_node, MY_SIZE num_cells, MY_SIZE *point_stride,
                   MY_SIZE cell_stride, MY_SIZE num_blocks, MY_SIZE block_size) {
    const fot_functor f = Desc::functor();
    check(fot_step_gpu_global(&f, SOA, point_data, point_data_out, cell_data, cell_to_node,
                              num_cells, point_stride, cell_stride, num_blocks, block_size,
                              nullptr),
          "fot_step_gpu_global");
  }
};

// The same arguments over all cells, FP64 global atomics (no reference
// counterpart: the launch behind loopGPUEdgeCentred).
template <class Desc> struct GPUAtomic {
  static fot_functor functor() { return Desc::functor(); }
  template <bool SOA>
  static void call(const void **point_data, void *point_data_out, void **cell_data,
                   const MY_SIZE **cell_to_node, MY_SIZE num_cells, MY_SIZE *point_stride,
                   MY_SIZE cell_stride, MY_SIZE num_blocks, MY_SIZE block_size) {
    const fot_functor f = Desc::functor();
    check(fot_step_gpu_atomic(&f, SOA, point_data, point_data_out, cell_data, cell_to_node,
                              num_cells, point_stride, cell_stride, num_blocks, block_size,
                              nullptr),
          "fot_step_gpu_atomic");
  }
};

// StepGPUHierarchical::call (kernels/res_calc.hpp:197-216): one launch over
// the blocks of one block colour of a HierarchicalColourMemory plan.
template <class Desc> struct GPUHierarchical {
  static fot_functor functor() { return Desc::functor(); }
  template <bool SOA>
  static void call(const void **point_data, void *point_data_out,
                   const MY_SIZE *points_to_be_cached, const MY_SIZE *points_to_be_cached_offsets,
                   void **cell_data, const MY_SIZE **cell_to_node,
                   const std::uint8_t *num_cell_colours, const std::uint8_t *cell_colours,
                   const MY_SIZE *block_offsets, MY_SIZE num_cells, const MY_SIZE *point_stride,
                   MY_SIZE cell_stride, MY_SIZE num_blocks, unsigned block_size,
                   unsigned cache_size) {
    const fot_functor f = Desc::functor();
    check(fot_step_gpu_hierarchical(&f, SOA, point_data, point_data_out, points_to_be_cached,
                                    points_to_be_cached_offsets, cell_data, cell_to_node,
                                    num_cell_colours, cell_colours, block_offsets, num_cells,
                                    point_stride, cell_stride, num_blocks, block_size, cache_size,
                                    nullptr),
          "fot_step_gpu_hierarchical");
  }
};

inline fot_functor make_functor(fot_kernel k, std::uint32_t md, std::uint32_t pd,
                                std::uint32_t cd) {
  fot_functor f{};
  check(fot_functor_init(&f, k, md, pd, cd), "fot_functor_init");
  return f;
}

}  // namespace fot

// ---- res_calc (kernels/res_calc.hpp, res_calc_func.hpp) -----------------
namespace res_calc {
constexpr unsigned MAPPING_DIM = 2;
constexpr unsigned Q_DIM = 4, X_DIM = 2, RES_DIM = 4, ADT_DIM = 1;

struct Desc {
  static fot_functor functor() { return fot::make_functor(FOT_KERNEL_RES_CALC, 2, 4, 0); }
};

// The airfoil edge flux (res_calc_func.hpp:11-57), evaluated in the same
// operation order as the device kernel so both give the same bits.
inline void user_func_host(const double *x1, const double *x2, const double *q1,
                           const double *q2, double adt1, double adt2, double *res1,
                           unsigned res1_stride, double *res2, unsigned res2_stride, unsigned xs,
                           unsigned qs) {
  const double gm1 = 0.4, eps = 0.05;
  const double dx = x1[0] - x2[0];
  const double dy = x1[xs] - x2[xs];
  double ri = 1.0 / q1[0];
  const double p1 = gm1 * (q1[3 * qs] - 0.5 * ri * (q1[qs] * q1[qs] + q1[2 * qs] * q1[2 * qs]));
  const double vol1 = ri * (q1[qs] * dy - q1[2 * qs] * dx);
  ri = 1.0 / q2[0];
  const double p2 = gm1 * (q2[3 * qs] - 0.5 * ri * (q2[qs] * q2[qs] + q2[2 * qs] * q2[2 * qs]));
  const double vol2 = ri * (q2[qs] * dy - q2[2 * qs] * dx);
  const double mu = 0.5 * (adt1 + adt2) * eps;
  double f[4];
  f[0] = 0.5 * (vol1 * q1[0] + vol2 * q2[0]) + mu * (q1[0] - q2[0]);
  f[1] = 0.5 * (vol1 * q1[qs] + p1 * dy + vol2 * q2[qs] + p2 * dy) + mu * (q1[qs] - q2[qs]);
  f[2] = 0.5 * (vol1 * q1[2 * qs] - p1 * dx + vol2 * q2[2 * qs] - p2 * dx) +
         mu * (q1[2 * qs] - q2[2 * qs]);
  f[3] = 0.5 * (vol1 * (q1[3 * qs] + p1) + vol2 * (q2[3 * qs] + p2)) +
         mu * (q1[3 * qs] - q2[3 * qs]);
  for (unsigned d = 0; d < 4; ++d) {
    res1[d * res1_stride] += f[d];
    res2[d * res2_stride] -= f[d];
  }
}

// StepSeq::call (kernels/res_calc.hpp:29-89): point data 1..3 = x, q, adt;
// x through mapping 1, q / adt / res through mapping 0.
struct StepSeq : Desc {
  template <bool SOA>
  static void call(const void **point_data, void *point_data_out, void **,
                   const MY_SIZE **cell_to_node, MY_SIZE ind, const unsigned *point_stride,
                   unsigned) {
    const double *x = static_cast<const double *>(point_data[1]);
    const double *q = static_cast<const double *>(point_data[2]);
    const double *adt = static_cast<const double *>(point_data[3]);
    double *res = static_cast<double *>(point_data_out);
    const MY_SIZE c0 = cell_to_node[0][2 * ind], c1 = cell_to_node[0][2 * ind + 1];
    const MY_SIZE n0 = cell_to_node[1][2 * ind], n1 = cell_to_node[1][2 * ind + 1];
    const unsigned xs = SOA ? point_stride[1] : 1, qs = SOA ? point_stride[2] : 1;
    const unsigned rs = SOA ? point_stride[0] : 1;
    user_func_host(x + (SOA ? n0 : X_DIM * n0), x + (SOA ? n1 : X_DIM * n1),
                   q + (SOA ? c0 : Q_DIM * c0), q + (SOA ? c1 : Q_DIM * c1), adt[c0], adt[c1],
                   res + (SOA ? c0 : RES_DIM * c0), rs, res + (SOA ? c1 : RES_DIM * c1), rs, xs,
                   qs);
  }
};
using StepOMP = StepSeq;
using StepGPUGlobal = fot::GPUGlobal<Desc>;
using StepGPUAtomic = fot::GPUAtomic<Desc>;
using StepGPUHierarchical = fot::GPUHierarchical<Desc>;
}  // namespace res_calc

// ---- mine (kernels/mine.hpp, mine_func.hpp) -----------------------------
namespace mine {
template <unsigned MeshDim, unsigned PointDim, unsigned CellDim, class DataType = double>
struct Desc {
  static_assert(std::is_same<DataType, double>::value, "libfot runs the loops in FP64");
  static_assert(MeshDim == 2 || MeshDim == 4 || MeshDim == 8, "MeshDim 2, 4 or 8");
  static_assert(CellDim == 1 || CellDim == PointDim, "CellDim 1 or PointDim");
  static fot_functor functor() {
    return fot::make_functor(FOT_KERNEL_MINE, MeshDim, PointDim, CellDim);
  }
};

// MeshDim 2: out_left += 2 * right * w, out_right += 2 * left * w;
// MeshDim 4 / 8: out_k += (p_{k+1} + p_{k-1}) * w (mine_func.hpp:11-99).
template <unsigned MeshDim, unsigned PointDim, unsigned CellDim, class DataType = double>
struct StepSeq : Desc<MeshDim, PointDim, CellDim, DataType> {
  static constexpr unsigned MESH_DIM = MeshDim;
  template <bool SOA>
  static void call(const void **point_data, void *point_data_out, void **cell_data,
                   const MY_SIZE **cell_to_node, MY_SIZE ind, const unsigned *point_stride,
                   unsigned cell_stride) {
    const double *in = static_cast<const double *>(point_data[0]);
    double *out = static_cast<double *>(point_data_out);
    const double *w = static_cast<const double *>(cell_data[0]) + ind;
    const unsigned s = SOA ? point_stride[0] : 1;
    MY_SIZE pt[MeshDim];
    for (unsigned k = 0; k < MeshDim; ++k) pt[k] = cell_to_node[0][MeshDim * ind + k];
    auto at = [&](MY_SIZE p, unsigned d) { return std::size_t(SOA ? p : PointDim * p) + d * s; };
    for (unsigned d = 0; d < PointDim; ++d) {
      const double wd = w[std::size_t(CellDim == PointDim ? d : 0) * cell_stride];
      double inc[MeshDim];
      if (MeshDim == 2) {
        inc[0] = 2 * in[at(pt[1], d)] * wd;
        inc[1] = 2 * in[at(pt[0], d)] * wd;
      } else {
        for (unsigned k = 0; k < MeshDim; ++k)
          inc[k] = (in[at(pt[(k + 1) % MeshDim], d)] + in[at(pt[(k + MeshDim - 1) % MeshDim], d)]) *
                   wd;
      }
      for (unsigned k = 0; k < MeshDim; ++k) out[at(pt[k], d)] += inc[k];
    }
  }
};
template <unsigned M, unsigned P, unsigned C, class T = double> using StepOMP = StepSeq<M, P, C, T>;
template <unsigned M, unsigned P, unsigned C, class T = double>
using StepGPUGlobal = fot::GPUGlobal<Desc<M, P, C, T>>;
template <unsigned M, unsigned P, unsigned C, class T = double>
using StepGPUAtomic = fot::GPUAtomic<Desc<M, P, C, T>>;
template <unsigned M, unsigned P, unsigned C, class T = double>
using StepGPUHierarchical = fot::GPUHierarchical<Desc<M, P, C, T>>;
}  // namespace mine

// ---- hex corner-force scatter (the scatter stage of kernels/lulesh.hpp:
// 23-57: node j += inc[i * MESH_DIM + j]) with the 24 corner forces given as
// cell data ----------------------------------------------------------------
namespace hex_scatter {
constexpr unsigned MESH_DIM = 8, POINT_DIM = 3, CELL_DIM = 24;
struct Desc {
  static fot_functor functor() { return fot::make_functor(FOT_KERNEL_HEX_SCATTER, 8, 3, 24); }
};
struct StepSeq : Desc {
  template <bool SOA>
  static void call(const void **, void *point_data_out, void **cell_data,
                   const MY_SIZE **cell_to_node, MY_SIZE ind, const unsigned *point_stride,
                   unsigned cell_stride) {
    double *f = static_cast<double *>(point_data_out);
    const double *F = static_cast<const double *>(cell_data[0]) + ind;
    const unsigned s = SOA ? point_stride[0] : 1;
    for (unsigned j = 0; j < MESH_DIM; ++j) {
      const MY_SIZE p = cell_to_node[0][MESH_DIM * ind + j];
      for (unsigned i = 0; i < POINT_DIM; ++i)
        f[std::size_t(SOA ? p : POINT_DIM * p) + i * s] +=
            F[std::size_t(i * MESH_DIM + j) * cell_stride];
    }
  }
};
using StepOMP = StepSeq;
using StepGPUGlobal = fot::GPUGlobal<Desc>;
using StepGPUAtomic = fot::GPUAtomic<Desc>;
using StepGPUHierarchical = fot::GPUHierarchical<Desc>;
}  // namespace hex_scatter

// ---- lulesh (kernels/lulesh.hpp, lulesh_func.hpp) -------------------------------
// point data 0 the node forces (dim 3, target), point data 1 the node
// coordinates (dim 3, read through mapping 0 as the reference does); cell
// data 0 the diagonal stresses (3), cell data 1 the element volume, which
// the user function writes.  The CPU struct runs libfot's host copy of the
// element function, the same code as the kernels.
namespace lulesh {
constexpr unsigned MESH_DIM = 8, POINT_DIM = 3, CELL_DIM0 = 3, CELL_DIM1 = 1;
struct Desc {
  static fot_functor functor() { return fot::make_functor(FOT_KERNEL_LULESH, 8, 3, 3); }
};
struct StepSeq : Desc {
  template <bool SOA>
  static void call(const void **point_data, void *point_data_out, void **cell_data,
                   const MY_SIZE **cell_to_node, MY_SIZE ind, const unsigned *point_stride,
                   unsigned cell_stride) {
    const double *xyz = static_cast<const double *>(point_data[1]);
    double *f = static_cast<double *>(point_data_out);
    const double *sig = static_cast<const double *>(cell_data[0]) + ind;
    double *determ = static_cast<double *>(cell_data[1]) + ind;
    const unsigned xs = SOA ? point_stride[1] : 1, fs = SOA ? point_stride[0] : 1;
    double x[MESH_DIM], y[MESH_DIM], z[MESH_DIM], inc[POINT_DIM * MESH_DIM];
    for (unsigned k = 0; k < MESH_DIM; ++k) {
      const MY_SIZE p = cell_to_node[0][MESH_DIM * ind + k];
      const double *c = xyz + (SOA ? p : POINT_DIM * p);
      x[k] = c[0], y[k] = c[xs], z[k] = c[2 * std::size_t(xs)];
    }
    fot::check(fot_lulesh_element(x, y, z, sig, cell_stride, inc, determ),
               "fot_lulesh_element");
    for (unsigned j = 0; j < MESH_DIM; ++j) {
      const MY_SIZE p = cell_to_node[0][MESH_DIM * ind + j];
      for (unsigned i = 0; i < POINT_DIM; ++i)
        f[std::size_t(SOA ? p : POINT_DIM * p) + i * fs] += inc[i * MESH_DIM + j];
    }
  }
};
using StepOMP = StepSeq;
using StepGPUGlobal = fot::GPUGlobal<Desc>;
using StepGPUAtomic = fot::GPUAtomic<Desc>;
using StepGPUHierarchical = fot::GPUHierarchical<Desc>;
}  // namespace lulesh

// ---- volna (kernels/volna.hpp, volna_func.hpp) in FP64 ---------------------
// point data 0 the state (dim 4, target), point data 1 the cell volume;
// cell data fluxes (3), bathymetry source (2), normals (2) and the
// right-boundary flag stored as 0.0 / 1.0.
namespace volna {
constexpr unsigned MESH_DIM = 2, INC_DIM = 4, FLUX_DIM = 3, CELL_DIM = 2;
struct Desc {
  static fot_functor functor() { return fot::make_functor(FOT_KERNEL_VOLNA, 2, 4, 3); }
};
struct StepSeq : Desc {
  template <bool SOA>
  static void call(const void **point_data, void *point_data_out, void **cell_data,
                   const MY_SIZE **cell_to_node, MY_SIZE ind, const unsigned *point_stride,
                   unsigned cell_stride) {
    const double *vol = static_cast<const double *>(point_data[1]);
    double *out = static_cast<double *>(point_data_out);
    const double *fl = static_cast<const double *>(cell_data[0]) + ind;
    const double *bs = static_cast<const double *>(cell_data[1]) + ind;
    const double *nm = static_cast<const double *>(cell_data[2]) + ind;
    const double boundary = static_cast<const double *>(cell_data[3])[ind];
    const MY_SIZE l = cell_to_node[0][2 * ind], r = cell_to_node[0][2 * ind + 1];
    const unsigned cs = cell_stride, s = SOA ? point_stride[0] : 1;
    double inc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    inc[0] -= fl[0] / vol[l];
    inc[1] -= (fl[cs] + bs[0] * nm[0]) / vol[l];
    inc[2] -= (fl[2 * cs] + bs[0] * nm[cs]) / vol[l];
    if (boundary == 0.0) {
      inc[4] += fl[0] / vol[r];
      inc[5] += (fl[cs] + bs[cs] * nm[0]) / vol[r];
      inc[6] += (fl[2 * cs] + bs[cs] * nm[cs]) / vol[r];
    }
    double *ol = out + (SOA ? l : INC_DIM * l), *orr = out + (SOA ? r : INC_DIM * r);
    for (unsigned i = 0; i < INC_DIM; ++i) {
      ol[i * s] += inc[i];
      orr[i * s] += inc[INC_DIM + i];
    }
  }
};
using StepOMP = StepSeq;
using StepGPUGlobal = fot::GPUGlobal<Desc>;
using StepGPUAtomic = fot::GPUAtomic<Desc>;
using StepGPUHierarchical = fot::GPUHierarchical<Desc>;
}  // namespace volna

// ---- mini_aero (kernels/mini_aero.hpp) -----------------------------------------
// point data 0 the cell flux (dim 5, target), point data 1 the 28 cell values
// (5 conservatives, 3 coordinates, 15 gradients, 5 limiters); cell data 0 the
// 12 face values.  The CPU struct runs libfot's host copy of the user
// function, the same code as the kernels.
namespace mini_aero {
constexpr unsigned MESH_DIM = 2, POINT_DIM0 = 5, POINT_DIM1 = 28, CELL_DIM = 12;
struct Desc {
  static fot_functor functor() { return fot::make_functor(FOT_KERNEL_MINI_AERO, 2, 5, 12); }
};
struct StepSeq : Desc {
  template <bool SOA>
  static void call(const void **point_data, void *point_data_out, void **cell_data,
                   const MY_SIZE **cell_to_node, MY_SIZE ind, const unsigned *point_stride,
                   unsigned cell_stride) {
    const double *cv = static_cast<const double *>(point_data[1]);
    double *out = static_cast<double *>(point_data_out);
    const double *face = static_cast<const double *>(cell_data[0]) + ind;
    const MY_SIZE l = cell_to_node[0][2 * ind], r = cell_to_node[0][2 * ind + 1];
    const unsigned s1 = SOA ? point_stride[1] : 1, s0 = SOA ? point_stride[0] : 1;
    double flux[5];
    fot::check(fot_mini_aero_face_flux(cv + (SOA ? l : POINT_DIM1 * l),
                                       cv + (SOA ? r : POINT_DIM1 * r), s1, face, cell_stride,
                                       flux),
               "fot_mini_aero_face_flux");
    double *ol = out + (SOA ? l : POINT_DIM0 * l), *orr = out + (SOA ? r : POINT_DIM0 * r);
    for (unsigned i = 0; i < POINT_DIM0; ++i) {
      ol[i * s0] -= flux[i];
      orr[i * s0] += flux[i];
    }
  }
};
using StepOMP = StepSeq;
using StepGPUGlobal = fot::GPUGlobal<Desc>;
using StepGPUAtomic = fot::GPUAtomic<Desc>;
using StepGPUHierarchical = fot::GPUHierarchical<Desc>;
}  // namespace mini_aero

#endif  // FOT_KERNELS_HPP
