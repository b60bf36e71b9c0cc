// fot_step.hip -- the reference-shaped single launches behind the kernel
// functors' call<SOA> (StepGPUGlobal, StepGPUHierarchical) plus the atomics
// launch behind loopGPUEdgeCentred.
#include <hip/hip_runtime.h>

#include "fot_dispatch.hpp"

using namespace fot;

namespace {

// Global-colouring / atomics step: one thread per cell of the launch.
template <class F, bool SOA, bool ATOMIC, int LB = 1024>
__global__ __launch_bounds__(LB) void step_direct_kernel(StepArgs a) {
  const uint32_t ind = blockIdx.x * blockDim.x + threadIdx.x;
  if (ind >= a.num_cells) return;
  uint32_t tgt[F::MD];
  double inc[F::MD * F::TD];
  targets<F>(a, ind, tgt);
  F::template compute<SOA>(a, ind, tgt, inc);
  add_increments<F, SOA, ATOMIC>(a.out, SOA ? a.point_stride[0] : 0u, tgt, inc);
}

// Gather of the reference's hierarchical kernels: mapping k of plan cell
// `ind` is the SOA list c2n[k] with stride n; mapping 0 holds block-local
// ids into `cached` (this block's points_to_be_cached).
template <class F, bool SOA> struct HierGather;

template <bool SOA> struct HierGather<ResCalcF, SOA> {
  // kernels/res_calc.hpp:343-375: x through mapping 1, q and adt through
  // mapping 2.
  FOT_DEV static void run(const StepArgs &a, uint32_t ind, uint32_t n, const uint32_t *,
                          uint32_t *loc, double *inc) {
    const double *x = a.pd(1), *q = a.pd(2), *adt = a.pd(3);
    loc[0] = a.c2n[0][ind];
    loc[1] = a.c2n[0][ind + n];
    const uint32_t n0 = a.c2n[1][ind], n1 = a.c2n[1][ind + n];
    const uint32_t c0 = a.c2n[2][ind], c1 = a.c2n[2][ind + n];
    const uint32_t xs = SOA ? a.ps(1) : 1, qs = SOA ? a.ps(2) : 1;
    const uint32_t ux = SOA ? 1 : 2, uq = SOA ? 1 : 4;
    res_calc_user(x + size_t(ux) * n0, x + size_t(ux) * n1, q + size_t(uq) * c0,
                  q + size_t(uq) * c1, adt + c0, adt + c1, inc, inc + 4, xs, qs);
  }
};

template <uint32_t MDv, uint32_t P, uint32_t C, bool SOA>
struct HierGather<MineF<MDv, P, C>, SOA> {
  using F = MineF<MDv, P, C>;
  // mine::stepGPUHierarchical (kernels/mine.hpp:484-735): input point data
  // of the block's points, cell data at the plan position.
  FOT_DEV static void run(const StepArgs &a, uint32_t ind, uint32_t n, const uint32_t *cached,
                          uint32_t *loc, double *inc) {
    uint32_t g[F::MD];
#pragma unroll
    for (uint32_t k = 0; k < F::MD; ++k) {
      loc[k] = a.c2n[0][ind + k * n];
      g[k] = cached[loc[k]];
    }
    F::template user<SOA>(a.pd(0), a.ps(0), a.cd(0) + ind, a.cs(), g, inc);
  }
};

template <bool SOA> struct HierGather<HexScatterF, SOA> {
  FOT_DEV static void run(const StepArgs &a, uint32_t ind, uint32_t n, const uint32_t *,
                          uint32_t *loc, double *inc) {
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) loc[k] = a.c2n[0][ind + k * n];
    HexScatterF::user(a.cd(0) + ind, a.cs(), inc);
  }
};

// functors that read their other point data through the global ids of
// mapping 0 (volna's volumes, mini_aero's cell values: mapping 1 is the same
// map): the block's cached ids give them back
template <class F, bool SOA> struct HierGatherGlobalIds {
  FOT_DEV static void run(const StepArgs &a, uint32_t ind, uint32_t n, const uint32_t *cached,
                          uint32_t *loc, double *inc) {
    uint32_t g[F::MD];
#pragma unroll
    for (uint32_t k = 0; k < F::MD; ++k) {
      loc[k] = a.c2n[0][ind + k * n];
      g[k] = cached[loc[k]];
    }
    F::template compute<SOA>(a, ind, g, inc);
  }
};
template <bool SOA> struct HierGather<VolnaF, SOA> : HierGatherGlobalIds<VolnaF, SOA> {};
template <bool SOA> struct HierGather<MiniAeroF, SOA> : HierGatherGlobalIds<MiniAeroF, SOA> {};
template <bool SOA> struct HierGather<GetaccF, SOA> : HierGatherGlobalIds<GetaccF, SOA> {};
template <bool SOA> struct HierGather<LuleshF, SOA> : HierGatherGlobalIds<LuleshF, SOA> {};

__device__ uint32_t g_lds_overflow; // set when a block's cache exceeds cache_size

// stepGPUHierarchical (kernels/res_calc.hpp:278-443 and its siblings): the
// block's increments accumulate in an LDS cache of its points, one thread
// colour at a time, then the cache is added to the output.  Blocks of one
// launch share no target point (block colouring), so the writeback is a
// plain read-modify-write.
template <class F, bool SOA, int LB = 1024>
__global__ __launch_bounds__(LB) void step_hier_kernel(
    StepArgs a, const uint32_t *__restrict__ ptbc, const uint32_t *__restrict__ ptbc_off,
    const uint8_t *__restrict__ num_cell_colours, const uint8_t *__restrict__ cell_colours,
    const uint32_t *__restrict__ block_offsets, uint32_t cache_size) {
  extern __shared__ double cache[];
  const uint32_t bid = blockIdx.x, tid = threadIdx.x;
  const uint32_t b0 = block_offsets[bid];
  const uint32_t bs = block_offsets[bid + 1] - b0;
  const uint32_t co = ptbc_off[bid];
  const uint32_t m = ptbc_off[bid + 1] - co;
  if (size_t(m) * F::TD * sizeof(double) > cache_size) {
    if (tid == 0) atomicOr(&g_lds_overflow, 1u);
    return;
  }
  const uint32_t ind = b0 + tid;
  uint32_t loc[F::MD];
  double inc[F::MD * F::TD];
  uint32_t colour = 0xffu;
  if (tid < bs) {
    colour = cell_colours[ind];
    HierGather<F, SOA>::run(a, ind, a.num_cells, ptbc + co, loc, inc);
  }
  for (uint32_t i = tid; i < m * F::TD; i += blockDim.x) cache[i] = 0.0;
  __syncthreads();
  const uint32_t ncol = num_cell_colours[bid];
  for (uint32_t c = 0; c < ncol; ++c) {
    if (colour == c) add_to_cache<F>(cache, m, loc, inc);
    __syncthreads();
  }
  const uint32_t n = SOA ? a.point_stride[0] : 0u;
  for (uint32_t i = tid; i < m; i += blockDim.x) {
    const uint32_t g = ptbc[co + i];
#pragma unroll
    for (uint32_t d = 0; d < F::TD; ++d) {
      double *p = a.out + pidx<SOA>(n, g, F::TD, d);
      *p = *p + cache[d * m + i];
    }
  }
}

fot_status check_step(const void *const *pd, void *out, const uint32_t *const *c2n,
                      const uint32_t *ps, uint32_t bs) {
  if (!pd || !out || !c2n || !ps) return fail(FOT_ERR_INVALID_ARGUMENT, "null device array");
  if (bs == 0 || bs > 1024) return fail(FOT_ERR_INVALID_ARGUMENT, "block size %u", bs);
  return FOT_OK;
}

fot_status launched() {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(FOT_ERR_HIP, "launch: %s", hipGetErrorString(e));
  return FOT_OK;
}

} // namespace

static fot_status fot_functor_init_impl(fot_functor *f, fot_kernel kernel, uint32_t mesh_dim,
                                       uint32_t point_dim, uint32_t cell_dim) {
  if (!f) return fail(FOT_ERR_INVALID_ARGUMENT, "null functor");
  f->kernel = kernel;
  switch (kernel) {
  case FOT_KERNEL_RES_CALC: // kernels/res_calc.hpp:85-86
    f->mesh_dim = 2, f->point_dim = 4, f->cell_dim = 0;
    return FOT_OK;
  case FOT_KERNEL_HEX_SCATTER: // kernels/lulesh.hpp:7-10
    f->mesh_dim = 8, f->point_dim = 3, f->cell_dim = 24;
    return FOT_OK;
  case FOT_KERNEL_MINI_AERO: // kernels/mini_aero.hpp:16-19
    f->mesh_dim = 2, f->point_dim = 5, f->cell_dim = 12;
    return FOT_OK;
  case FOT_KERNEL_GETACC: // kernels/getacc_scatter.hpp:22-23 (cell_dim: cnmass)
    f->mesh_dim = 4, f->point_dim = 4, f->cell_dim = 4;
    return FOT_OK;
  case FOT_KERNEL_LULESH: // kernels/lulesh.hpp:7-10 (cell_dim: sigma)
    f->mesh_dim = 8, f->point_dim = 3, f->cell_dim = 3;
    return FOT_OK;
  case FOT_KERNEL_VOLNA: // kernels/volna.hpp:16-19 (cell_dim: the fluxes)
    f->mesh_dim = 2, f->point_dim = 4, f->cell_dim = 3;
    return FOT_OK;
  case FOT_KERNEL_MINE:
    if ((mesh_dim != 2 && mesh_dim != 4 && mesh_dim != 8) || point_dim == 0 ||
        (cell_dim != 1 && cell_dim != point_dim))
      return fail(FOT_ERR_INVALID_ARGUMENT, "mine: CellDim must be 1 or PointDim");
    f->mesh_dim = mesh_dim, f->point_dim = point_dim, f->cell_dim = cell_dim;
    return FOT_OK;
  }
  return fail(FOT_ERR_INVALID_ARGUMENT, "unknown kernel %d", int(kernel));
}
extern "C" fot_status fot_functor_init(fot_functor *f, fot_kernel kernel, uint32_t mesh_dim,
                                       uint32_t point_dim, uint32_t cell_dim) {
  return fot::guarded("fot_functor_init", [&] {
    return fot_functor_init_impl(f, kernel, mesh_dim, point_dim, cell_dim);
  });
}

static fot_status step_direct(bool atomic, const fot_functor *f, int soa,
                              const void *const *pd, void *out, void *const *cd,
                              const uint32_t *const *c2n, uint32_t num_cells,
                              const uint32_t *ps, uint32_t cs, uint32_t nb, uint32_t bs,
                              void *stream) {
  fot_status st = check_step(pd, out, c2n, ps, bs);
  if (st != FOT_OK) return st;
  if (num_cells == 0) return FOT_OK;
  if (uint64_t(nb) * bs < num_cells) return fail(FOT_ERR_INVALID_ARGUMENT, "grid too small");
  const StepArgs a{pd, static_cast<double *>(out), cd, c2n, num_cells, ps, cs};
  const hipStream_t s = static_cast<hipStream_t>(stream);
  return dispatch(f, soa, [&]<class F, bool SOA>() {
    // one thread per cell: the cells a launch covers do not depend on its
    // shape, so a functor with a smaller bound runs blocks of that size
    constexpr int LB = kSmallBound<F>;
    const uint32_t dbs = bs <= uint32_t(LB) ? bs : uint32_t(LB);
    const uint32_t dnb = dbs == bs ? nb : (num_cells + dbs - 1) / dbs;
    if (atomic)
      hipLaunchKernelGGL((step_direct_kernel<F, SOA, true, LB>), dim3(dnb), dim3(dbs), 0, s, a);
    else
      hipLaunchKernelGGL((step_direct_kernel<F, SOA, false, LB>), dim3(dnb), dim3(dbs), 0, s, a);
    return launched();
  });
}

static fot_status fot_step_gpu_global_impl(const fot_functor *f, int soa,
                                          const void *const *pd, void *out,
                                          void *const *cd, const uint32_t *const *c2n,
                                          uint32_t num_cells, const uint32_t *ps,
                                          uint32_t cs, uint32_t nb, uint32_t bs,
                                          void *stream) {
  return step_direct(false, f, soa, pd, out, cd, c2n, num_cells, ps, cs, nb, bs, stream);
}
extern "C" fot_status fot_step_gpu_global(const fot_functor *f, int soa,
                                          const void *const *pd, void *out,
                                          void *const *cd, const uint32_t *const *c2n,
                                          uint32_t num_cells, const uint32_t *ps,
                                          uint32_t cs, uint32_t nb, uint32_t bs,
                                          void *stream) {
  return fot::guarded("fot_step_gpu_global", [&] {
    return fot_step_gpu_global_impl(f, soa, pd, out, cd, c2n, num_cells, ps, cs, nb, bs, stream);
  });
}

static fot_status fot_step_gpu_atomic_impl(const fot_functor *f, int soa,
                                          const void *const *pd, void *out,
                                          void *const *cd, const uint32_t *const *c2n,
                                          uint32_t num_cells, const uint32_t *ps,
                                          uint32_t cs, uint32_t nb, uint32_t bs,
                                          void *stream) {
  return step_direct(true, f, soa, pd, out, cd, c2n, num_cells, ps, cs, nb, bs, stream);
}
extern "C" fot_status fot_step_gpu_atomic(const fot_functor *f, int soa,
                                          const void *const *pd, void *out,
                                          void *const *cd, const uint32_t *const *c2n,
                                          uint32_t num_cells, const uint32_t *ps,
                                          uint32_t cs, uint32_t nb, uint32_t bs,
                                          void *stream) {
  return fot::guarded("fot_step_gpu_atomic", [&] {
    return fot_step_gpu_atomic_impl(f, soa, pd, out, cd, c2n, num_cells, ps, cs, nb, bs, stream);
  });
}

static fot_status fot_step_gpu_hierarchical_impl(
    const fot_functor *f, int soa, const void *const *pd, void *out,
    const uint32_t *ptbc, const uint32_t *ptbc_off, void *const *cd,
    const uint32_t *const *c2n, const uint8_t *ncc, const uint8_t *cc,
    const uint32_t *bo, uint32_t num_cells, const uint32_t *ps, uint32_t cs,
    uint32_t nb, uint32_t bs, uint32_t cache_size, void *stream) {
  fot_status st = check_step(pd, out, c2n, ps, bs);
  if (st != FOT_OK) return st;
  if (!ptbc || !ptbc_off || !ncc || !cc || !bo)
    return fail(FOT_ERR_INVALID_ARGUMENT, "null plan array");
  if (nb == 0) return FOT_OK;
  if (cache_size > 160u * 1024u) return fail(FOT_ERR_INVALID_ARGUMENT, "cache_size > LDS");
  const StepArgs a{pd, static_cast<double *>(out), cd, c2n, num_cells, ps, cs};
  const hipStream_t s = static_cast<hipStream_t>(stream);
  return dispatch(f, soa, [&]<class F, bool SOA>() {
    if constexpr (kSmallBound<F> < 1024) {
      if (bs <= uint32_t(kSmallBound<F>)) {
        hipLaunchKernelGGL((step_hier_kernel<F, SOA, kSmallBound<F>>), dim3(nb), dim3(bs),
                           cache_size, s, a, ptbc, ptbc_off, ncc, cc, bo, cache_size);
        return launched();
      }
    }
    hipLaunchKernelGGL((step_hier_kernel<F, SOA>), dim3(nb), dim3(bs), cache_size, s, a, ptbc,
                       ptbc_off, ncc, cc, bo, cache_size);
    return launched();
  });
}
extern "C" fot_status fot_step_gpu_hierarchical(
    const fot_functor *f, int soa, const void *const *pd, void *out,
    const uint32_t *ptbc, const uint32_t *ptbc_off, void *const *cd,
    const uint32_t *const *c2n, const uint8_t *ncc, const uint8_t *cc,
    const uint32_t *bo, uint32_t num_cells, const uint32_t *ps, uint32_t cs,
    uint32_t nb, uint32_t bs, uint32_t cache_size, void *stream) {
  return fot::guarded("fot_step_gpu_hierarchical", [&] {
    return fot_step_gpu_hierarchical_impl(f, soa, pd, out, ptbc, ptbc_off, cd, c2n, ncc, cc, bo,
                                          num_cells, ps, cs, nb, bs, cache_size, stream);
  });
}

static fot_status fot_lds_overflow_flag_impl(uint32_t *flag, int clear) {
  if (!flag) return fail(FOT_ERR_INVALID_ARGUMENT, "null flag");
  hipError_t e = hipMemcpyFromSymbol(flag, HIP_SYMBOL(g_lds_overflow), sizeof(uint32_t));
  if (e == hipSuccess && clear) {
    const uint32_t zero = 0;
    e = hipMemcpyToSymbol(HIP_SYMBOL(g_lds_overflow), &zero, sizeof(uint32_t));
  }
  if (e != hipSuccess) return fail(FOT_ERR_HIP, "%s", hipGetErrorString(e));
  return FOT_OK;
}
extern "C" fot_status fot_lds_overflow_flag(uint32_t *flag, int clear) {
  return fot::guarded("fot_lds_overflow_flag", [&] {
    return fot_lds_overflow_flag_impl(flag, clear);
  });
}

// The lulesh element function on the host (the code the kernels run): the
// C++ drop-in's lulesh::StepSeq calls it per element.
static fot_status fot_lulesh_element_impl(const double *x, const double *y, const double *z,
                                         const double *sig, uint32_t sig_stride, double *f,
                                         double *determ) {
  if (!x || !y || !z || !sig || !f || !determ)
    return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  fot::lu::element(x, y, z, sig[0], sig[sig_stride], sig[2 * size_t(sig_stride)], f, determ);
  return FOT_OK;
}
extern "C" fot_status fot_lulesh_element(const double *x, const double *y, const double *z,
                                         const double *sig, uint32_t sig_stride, double *f,
                                         double *determ) {
  return fot::guarded("fot_lulesh_element", [&] {
    return fot_lulesh_element_impl(x, y, z, sig, sig_stride, f, determ);
  });
}

// The mini_aero user function on the host (the code the kernels run): the
// C++ drop-in's mini_aero::StepSeq calls it per face.
static fot_status fot_mini_aero_face_flux_impl(const double *left, const double *right,
                                              uint32_t cell_stride, const double *face,
                                              uint32_t face_stride, double *flux) {
  if (!left || !right || !face || !flux) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  fot::ma::face_flux(left, right, cell_stride, face, face_stride, flux);
  return FOT_OK;
}
extern "C" fot_status fot_mini_aero_face_flux(const double *left, const double *right,
                                              uint32_t cell_stride, const double *face,
                                              uint32_t face_stride, double *flux) {
  return fot::guarded("fot_mini_aero_face_flux", [&] {
    return fot_mini_aero_face_flux_impl(left, right, cell_stride, face, face_stride, flux);
  });
}
