// fot_functors.hpp -- device user functions of the loop engine and the
// per-kernel gather / increment code shared by every launch strategy.
//
// The whole library is compiled with -ffp-contract=off: each increment is the
// IEEE result of the reference's expression order, bit-identical to the
// host's, so only the summation order (atomics, colours, blocks) separates
// the strategies.
#ifndef FOT_FUNCTORS_HPP
#define FOT_FUNCTORS_HPP

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fot_capi.h"

#define FOT_DEV __device__ __forceinline__

namespace fot {

// index<SOA> (data_t.hpp:104-115)
template <bool SOA>
FOT_DEV size_t pidx(uint32_t n, uint32_t i, uint32_t dim, uint32_t d) {
  return SOA ? size_t(d) * n + i : size_t(i) * dim + d;
}

// kernels/res_calc_func.hpp:11-57 (same expression order; the reference's
// 0.5f / 1.0f literals are exact in double).
FOT_DEV void res_calc_user(const double *__restrict__ x1, const double *__restrict__ x2,
                           const double *__restrict__ q1, const double *__restrict__ q2,
                           const double *__restrict__ adt1,
                           const double *__restrict__ adt2, double *__restrict__ res1,
                           double *__restrict__ res2, uint32_t xs, uint32_t qs) {
  const double gm1 = 0.4, eps = 0.05;
  const double dx = x1[0] - x2[0];
  const double dy = x1[xs] - x2[xs];
  double ri = 1.0 / q1[0];
  const double p1 = gm1 * (q1[3 * qs] - 0.5 * ri * (q1[qs] * q1[qs] + q1[2 * qs] * q1[2 * qs]));
  const double vol1 = ri * (q1[qs] * dy - q1[2 * qs] * dx);
  ri = 1.0 / q2[0];
  const double p2 = gm1 * (q2[3 * qs] - 0.5 * ri * (q2[qs] * q2[qs] + q2[2 * qs] * q2[2 * qs]));
  const double vol2 = ri * (q2[qs] * dy - q2[2 * qs] * dx);
  const double mu = 0.5 * ((*adt1) + (*adt2)) * eps;
  double f = 0.5 * (vol1 * q1[0] + vol2 * q2[0]) + mu * (q1[0] - q2[0]);
  res1[0] = f;
  res2[0] = -f;
  f = 0.5 * (vol1 * q1[qs] + p1 * dy + vol2 * q2[qs] + p2 * dy) + mu * (q1[qs] - q2[qs]);
  res1[1] = f;
  res2[1] = -f;
  f = 0.5 * (vol1 * q1[2 * qs] - p1 * dx + vol2 * q2[2 * qs] - p2 * dx) +
      mu * (q1[2 * qs] - q2[2 * qs]);
  res1[2] = f;
  res2[2] = -f;
  f = 0.5 * (vol1 * (q1[3 * qs] + p1) + vol2 * (q2[3 * qs] + p2)) + mu * (q1[3 * qs] - q2[3 * qs]);
  res1[3] = f;
  res2[3] = -f;
}

// Arguments of the reference-shaped launches: device arrays of device
// pointers / strides exactly as Problem builds them (problem.hpp:158-176,
// 258-265).
struct StepArgs {
  const void *const *point_data;
  double *out;
  void *const *cell_data;
  const uint32_t *const *c2n;
  uint32_t num_cells;
  const uint32_t *point_stride;
  uint32_t cell_stride;
  FOT_DEV const double *pd(uint32_t k) const { return static_cast<const double *>(point_data[k]); }
  FOT_DEV const double *cd(uint32_t k) const { return static_cast<const double *>(cell_data[k]); }
  FOT_DEV const uint32_t *map(uint32_t k) const { return c2n[k]; }
  FOT_DEV uint32_t ps(uint32_t k) const { return point_stride[k]; }
  FOT_DEV uint32_t cs() const { return cell_stride; }
};

// Arguments of the engine's launches: everything by value in the kernel
// argument segment (no pointer chasing).
struct EngineArgs {
  const double *point_data[FOT_MAX_MAPPINGS];
  const double *cell_data[FOT_MAX_CELL_DATA];
  const uint32_t *maps[FOT_MAX_MAPPINGS];
  uint32_t num_points[FOT_MAX_MAPPINGS];
  double *out;
  uint32_t num_cells;
  uint32_t cell_stride;
  FOT_DEV const double *pd(uint32_t k) const { return point_data[k]; }
  FOT_DEV const double *cd(uint32_t k) const { return cell_data[k]; }
  FOT_DEV const uint32_t *map(uint32_t k) const { return maps[k]; }
  FOT_DEV uint32_t ps(uint32_t k) const { return num_points[k]; }
  FOT_DEV uint32_t cs() const { return cell_stride; }
};

// ---- functor traits --------------------------------------------------
// MD target slots of TD components; SLOT_OUTER selects the reference's add
// order (component-major for res_calc / mine, slot-major for the hex
// scatter).  compute() gets the cell's mapping-0 point ids `tgt` and reads
// everything else through the argument block: mappings >= 1 as AOS rows of
// cell `ind`, cell data as SOA with the block's cell stride.

struct ResCalcF {
  static constexpr uint32_t MD = 2, TD = 4;
  static constexpr bool SLOT_OUTER = false;
  // res_calc::StepSeq / StepGPUGlobal (kernels/res_calc.hpp:29-89,
  // 179-242): x through mapping 1; q, adt and res through mapping 0.
  template <bool SOA, class A>
  FOT_DEV static void compute(const A &a, uint32_t ind, const uint32_t *tgt, double *inc) {
    const uint32_t *m1 = a.map(1);
    const double *x = a.pd(1), *q = a.pd(2), *adt = a.pd(3);
    const uint32_t n0 = m1[2 * ind], n1 = m1[2 * ind + 1];
    const uint32_t xs = SOA ? a.ps(1) : 1, qs = SOA ? a.ps(2) : 1;
    const uint32_t ux = SOA ? 1 : 2, uq = SOA ? 1 : 4;
    res_calc_user(x + size_t(ux) * n0, x + size_t(ux) * n1, q + size_t(uq) * tgt[0],
                  q + size_t(uq) * tgt[1], adt + tgt[0], adt + tgt[1], inc, inc + 4, xs, qs);
  }
};

// mine (kernels/mine.hpp:22-178, mine_func.hpp:11-99): MeshDim 2 gives
// out_left = 2*right*w, out_right = 2*left*w; MeshDim 4/8 give
// out_k = (p_{k+1} + p_{k-1}) * w.
template <uint32_t MDv, uint32_t P, uint32_t C>
struct MineF {
  static constexpr uint32_t MD = MDv, TD = P;
  static constexpr bool SLOT_OUTER = false;
  template <bool SOA>
  FOT_DEV static void user(const double *in, uint32_t ps, const double *w, uint32_t cs,
                           const uint32_t *pt, double *inc) {
    const double *p[MD];
#pragma unroll
    for (uint32_t k = 0; k < MD; ++k) p[k] = in + size_t(SOA ? 1 : P) * pt[k];
    const uint32_t s = SOA ? ps : 1;
#pragma unroll
    for (uint32_t d = 0; d < P; ++d) {
      const double wd = w[size_t(C == P ? d : 0) * cs];
      if (MD == 2) {
        inc[0 * P + d] = 2 * p[1][d * s] * wd;
        inc[1 * P + d] = 2 * p[0][d * s] * wd;
      } else {
#pragma unroll
        for (uint32_t k = 0; k < MD; ++k)
          inc[k * P + d] = (p[(k + 1) % MD][d * s] + p[(k + MD - 1) % MD][d * s]) * wd;
      }
    }
  }
  template <bool SOA, class A>
  FOT_DEV static void compute(const A &a, uint32_t ind, const uint32_t *tgt, double *inc) {
    user<SOA>(a.pd(0), a.ps(0), a.cd(0) + ind, a.cs(), tgt, inc);
  }
};

// Scatter-only hex functor behind the lulesh struct API: 24 corner forces
// per element as cell data [i * 8 + j] (component i, corner j), added slot by
// slot (kernels/lulesh.hpp:52-56).
struct HexScatterF {
  static constexpr uint32_t MD = 8, TD = 3;
  static constexpr bool SLOT_OUTER = true;
  FOT_DEV static void user(const double *F, uint32_t cs, double *inc) {
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j)
#pragma unroll
      for (uint32_t i = 0; i < 3; ++i) inc[j * 3 + i] = F[size_t(i * 8 + j) * cs];
  }
  template <bool SOA, class A>
  FOT_DEV static void compute(const A &a, uint32_t ind, const uint32_t *, double *inc) {
    user(a.cd(0) + ind, a.cs(), inc);
  }
};

// volna (kernels/volna.hpp:27-78, volna_func.hpp:16-36) in FP64: the cell
// volumes are point data 1, read through mapping-0 ids as the reference
// does; cell data 0..3 = edge fluxes (3), bathymetry source (2), edge
// normals (2), right-boundary flag (0.0 / 1.0).  The increments start at 0
// and are updated with the reference's `-=` / `+=`, so even signed zeros
// match.  The fourth state component gets no increment.
struct VolnaF {
  static constexpr uint32_t MD = 2, TD = 4;
  static constexpr bool SLOT_OUTER = false;
  template <bool SOA, class A>
  FOT_DEV static void compute(const A &a, uint32_t ind, const uint32_t *tgt, double *inc) {
    const double *vol = a.pd(1);
    const uint32_t cs = a.cs();
    const double *fl = a.cd(0) + ind, *bs = a.cd(1) + ind, *nm = a.cd(2) + ind;
    const double boundary = a.cd(3)[ind];
    const double v0 = vol[tgt[0]], v1 = vol[tgt[1]];
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) inc[k] = 0.0;
    inc[0] -= fl[0] / v0;
    inc[1] -= (fl[cs] + bs[0] * nm[0]) / v0;
    inc[2] -= (fl[2 * cs] + bs[0] * nm[cs]) / v0;
    if (boundary == 0.0) {
      inc[4] += fl[0] / v1;
      inc[5] += (fl[cs] + bs[cs] * nm[0]) / v1;
      inc[6] += (fl[2 * cs] + bs[cs] * nm[cs]) / v1;
    }
  }
};

// Mapping-0 ids of cell `ind` (AOS row).
template <class F, class A>
FOT_DEV void targets(const A &a, uint32_t ind, uint32_t *tgt) {
#pragma unroll
  for (uint32_t k = 0; k < F::MD; ++k) tgt[k] = a.map(0)[F::MD * ind + k];
}

// FP64 global atomic add: hardware global atomic add (no compare-and-swap
// loop).
FOT_DEV void atomic_add_f64(double *p, double v) { unsafeAtomicAdd(p, v); }

// Adds the increments of one cell to `out` in the functor's order.
template <class F, bool SOA, bool ATOMIC>
FOT_DEV void add_increments(double *out, uint32_t n, const uint32_t *tgt, const double *inc) {
  if (F::SLOT_OUTER) {
#pragma unroll
    for (uint32_t k = 0; k < F::MD; ++k)
#pragma unroll
      for (uint32_t d = 0; d < F::TD; ++d) {
        double *p = out + pidx<SOA>(n, tgt[k], F::TD, d);
        if (ATOMIC) atomic_add_f64(p, inc[k * F::TD + d]);
        else *p += inc[k * F::TD + d];
      }
  } else {
#pragma unroll
    for (uint32_t d = 0; d < F::TD; ++d)
#pragma unroll
      for (uint32_t k = 0; k < F::MD; ++k) {
        double *p = out + pidx<SOA>(n, tgt[k], F::TD, d);
        if (ATOMIC) atomic_add_f64(p, inc[k * F::TD + d]);
        else *p += inc[k * F::TD + d];
      }
  }
}

// Adds the increments of one cell to an LDS cache (SOA, stride m) at
// block-local ids `loc`, in the functor's order.
template <class F>
FOT_DEV void add_to_cache(double *cache, uint32_t m, const uint32_t *loc, const double *inc) {
  if (F::SLOT_OUTER) {
#pragma unroll
    for (uint32_t k = 0; k < F::MD; ++k)
#pragma unroll
      for (uint32_t d = 0; d < F::TD; ++d) cache[d * m + loc[k]] += inc[k * F::TD + d];
  } else {
#pragma unroll
    for (uint32_t d = 0; d < F::TD; ++d)
#pragma unroll
      for (uint32_t k = 0; k < F::MD; ++k) cache[d * m + loc[k]] += inc[k * F::TD + d];
  }
}

// As add_to_cache, for owner-compute blocks: only slots of owned points
// (local id below `owned`) accumulate; halo increments belong to another
// block.
template <class F>
FOT_DEV void add_to_cache_owned(double *cache, uint32_t m, const uint32_t *loc, const double *inc,
                                uint32_t owned) {
  if (F::SLOT_OUTER) {
#pragma unroll
    for (uint32_t k = 0; k < F::MD; ++k)
#pragma unroll
      for (uint32_t d = 0; d < F::TD; ++d)
        if (loc[k] < owned) cache[d * m + loc[k]] += inc[k * F::TD + d];
  } else {
#pragma unroll
    for (uint32_t d = 0; d < F::TD; ++d)
#pragma unroll
      for (uint32_t k = 0; k < F::MD; ++k)
        if (loc[k] < owned) cache[d * m + loc[k]] += inc[k * F::TD + d];
  }
}

} // namespace fot

#endif
