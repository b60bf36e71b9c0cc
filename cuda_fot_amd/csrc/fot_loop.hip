// fot_loop.hip -- the MI355X-native loop engine: plan once, then run
// iterations as FP64-atomics launches, global-colour launches or
// hierarchical LDS-block launches, optionally replayed from a hipGraph.
//
// Replaces Problem<SOA>::loopGPUCellCentred / loopGPUHierarchical
// (problem.hpp:89-339) and adds the atomics loop (loopGPUEdgeCentred).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "fot_dispatch.hpp"

using namespace fot;

#define FOT_HIP(call)                                                          \
  do {                                                                         \
    const hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess)                                                      \
      return fail(e_ == hipErrorOutOfMemory ? FOT_ERR_OUT_OF_MEMORY            \
                                            : FOT_ERR_HIP,                     \
                  "%s: %s", #call, hipGetErrorString(e_));                     \
  } while (0)

namespace {

constexpr uint32_t kNumXcd = 8; // MI355X: 8 XCDs of 32 CUs

// Workgroups are dispatched round-robin over the XCDs; this bijection gives
// each XCD a contiguous run of plan blocks, so neighbouring blocks (which
// share points) meet in the same L2.
FOT_DEV uint32_t xcd_block(uint32_t i, uint32_t n) {
  const uint32_t q = n / kNumXcd, r = n % kNumXcd;
  const uint32_t x = i % kNumXcd, slot = i / kNumXcd;
  return x * q + min(x, r) + slot;
}

struct HierDev {
  const uint32_t *blk_cell_off;  // [nb+1] plan positions
  const uint32_t *blk_cache_off; // [nb+1] into cached
  const uint32_t *cached;        // mapping-0 point ids per block, ascending
  const uint16_t *loc;           // [num_cells * MD] block-local ids (AOS)
  const uint8_t *colour;         // [num_cells] thread colour
  const uint8_t *blk_ncol;       // [nb] thread colours per block
};

template <uint32_t MD>
FOT_DEV void load_loc(const uint16_t *__restrict__ p, uint32_t *loc) {
  if constexpr (MD == 2) {
    const uint32_t v = *reinterpret_cast<const uint32_t *>(p);
    loc[0] = v & 0xffffu;
    loc[1] = v >> 16;
  } else if constexpr (MD == 8) {
    const uint4 v = *reinterpret_cast<const uint4 *>(p);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      loc[2 * k] = w[k] & 0xffffu;
      loc[2 * k + 1] = w[k] >> 16;
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < MD; ++k) loc[k] = p[k];
  }
}

// ---- direct launches (atomics, global colour) -------------------------
template <class F, bool SOA, bool ATOMIC>
__global__ __launch_bounds__(1024) void eng_direct_kernel(EngineArgs a) {
  const uint32_t ind = blockIdx.x * blockDim.x + threadIdx.x;
  if (ind >= a.num_cells) return;
  uint32_t tgt[F::MD];
  double inc[F::MD * F::TD];
  targets<F>(a, ind, tgt);
  F::template compute<SOA>(a, ind, tgt, inc);
  add_increments<F, SOA, ATOMIC>(a.out, a.num_points[0], tgt, inc);
}

// res_calc over all edges with FP64 global atomics: the edge's two map rows
// are 8-byte loads, consecutive lanes read consecutive rows (coalesced).
template <bool SOA>
__global__ __launch_bounds__(1024) void res_calc_atomic_kernel(EngineArgs a) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.num_cells) return;
  const uint2 c = reinterpret_cast<const uint2 *>(a.maps[0])[e];
  const uint2 n = reinterpret_cast<const uint2 *>(a.maps[1])[e];
  const double *x = a.point_data[1], *q = a.point_data[2], *adt = a.point_data[3];
  double x1[2], x2[2], q1[4], q2[4];
  if (SOA) {
    const uint32_t nx = a.num_points[1], nq = a.num_points[2];
    x1[0] = x[n.x], x1[1] = x[nx + n.x], x2[0] = x[n.y], x2[1] = x[nx + n.y];
#pragma unroll
    for (uint32_t d = 0; d < 4; ++d) q1[d] = q[size_t(d) * nq + c.x], q2[d] = q[size_t(d) * nq + c.y];
  } else {
    const double2 xa = reinterpret_cast<const double2 *>(x)[n.x];
    const double2 xb = reinterpret_cast<const double2 *>(x)[n.y];
    x1[0] = xa.x, x1[1] = xa.y, x2[0] = xb.x, x2[1] = xb.y;
    const double2 *q2p = reinterpret_cast<const double2 *>(q);
    const double2 a0 = q2p[2 * size_t(c.x)], a1 = q2p[2 * size_t(c.x) + 1];
    const double2 b0 = q2p[2 * size_t(c.y)], b1 = q2p[2 * size_t(c.y) + 1];
    q1[0] = a0.x, q1[1] = a0.y, q1[2] = a1.x, q1[3] = a1.y;
    q2[0] = b0.x, q2[1] = b0.y, q2[2] = b1.x, q2[3] = b1.y;
  }
  const double adt1 = adt[c.x], adt2 = adt[c.y];
  double inc[8];
  res_calc_user(x1, x2, q1, q2, &adt1, &adt2, inc, inc + 4, 1, 1);
  const uint32_t np = a.num_points[0];
#pragma unroll
  for (uint32_t d = 0; d < 4; ++d) {
    atomic_add_f64(a.out + pidx<SOA>(np, c.x, 4, d), inc[d]);
    atomic_add_f64(a.out + pidx<SOA>(np, c.y, 4, d), inc[4 + d]);
  }
}

// ---- hierarchical launches --------------------------------------------
// One workgroup per plan block: increments accumulate in an LDS cache of the
// block's points thread colour by thread colour (one barrier each), then the
// cache is added to the output.  One launch covers one block colour, whose
// blocks share no point, so the writeback is a plain read-modify-write and
// the summation order is fixed by the plan.
template <class F, bool SOA>
__global__ __launch_bounds__(1024) void eng_hier_kernel(EngineArgs a, HierDev h,
                                                        uint32_t blk_begin, uint32_t nbc) {
  extern __shared__ double cache[];
  const uint32_t b = blk_begin + xcd_block(blockIdx.x, nbc);
  const uint32_t tid = threadIdx.x;
  const uint32_t e0 = h.blk_cell_off[b], ne = h.blk_cell_off[b + 1] - e0;
  const uint32_t c0 = h.blk_cache_off[b], m = h.blk_cache_off[b + 1] - c0;
  for (uint32_t i = tid; i < m * F::TD; i += blockDim.x) cache[i] = 0.0;
  uint32_t loc[F::MD], tgt[F::MD];
  double inc[F::MD * F::TD];
  uint32_t colour = 0xffu;
  if (tid < ne) {
    const uint32_t pos = e0 + tid;
    colour = h.colour[pos];
    load_loc<F::MD>(h.loc + size_t(pos) * F::MD, loc);
#pragma unroll
    for (uint32_t k = 0; k < F::MD; ++k) tgt[k] = h.cached[c0 + loc[k]];
    F::template compute<SOA>(a, pos, tgt, inc);
  }
  __syncthreads();
  const uint32_t ncol = h.blk_ncol[b];
  for (uint32_t c = 0; c < ncol; ++c) {
    if (colour == c) add_to_cache<F>(cache, m, loc, inc);
    __syncthreads();
  }
  const uint32_t np = a.num_points[0];
  for (uint32_t i = tid; i < m; i += blockDim.x) {
    const uint32_t g = h.cached[c0 + i];
#pragma unroll
    for (uint32_t d = 0; d < F::TD; ++d) {
      double *p = a.out + pidx<SOA>(np, g, F::TD, d);
      *p = *p + cache[d * m + i];
    }
  }
}

// res_calc hierarchical with the block's cells' q and adt staged in LDS next
// to the res cache (they are read through the same mapping-0 points), so
// the edge loop gathers only x from HBM.
template <bool SOA>
__global__ __launch_bounds__(1024) void res_calc_hier_staged_kernel(EngineArgs a, HierDev h,
                                                                    uint32_t blk_begin,
                                                                    uint32_t nbc) {
  extern __shared__ double lds[];
  const uint32_t b = blk_begin + xcd_block(blockIdx.x, nbc);
  const uint32_t tid = threadIdx.x;
  const uint32_t e0 = h.blk_cell_off[b], ne = h.blk_cell_off[b + 1] - e0;
  const uint32_t c0 = h.blk_cache_off[b], m = h.blk_cache_off[b + 1] - c0;
  double *res_s = lds, *q_s = lds + 4 * m, *adt_s = lds + 8 * m;
  const double *x = a.point_data[1], *q = a.point_data[2], *adt = a.point_data[3];
  for (uint32_t i = tid; i < m; i += blockDim.x) {
    const uint32_t g = h.cached[c0 + i];
    if (SOA) {
      const uint32_t nq = a.num_points[2];
#pragma unroll
      for (uint32_t d = 0; d < 4; ++d) q_s[d * m + i] = q[size_t(d) * nq + g];
    } else {
      const double2 *q2p = reinterpret_cast<const double2 *>(q);
      const double2 v0 = q2p[2 * size_t(g)], v1 = q2p[2 * size_t(g) + 1];
      q_s[0 * m + i] = v0.x, q_s[1 * m + i] = v0.y, q_s[2 * m + i] = v1.x, q_s[3 * m + i] = v1.y;
    }
    adt_s[i] = adt[g];
#pragma unroll
    for (uint32_t d = 0; d < 4; ++d) res_s[d * m + i] = 0.0;
  }
  uint32_t colour = 0xffu, l0 = 0, l1 = 0;
  uint2 nn = make_uint2(0, 0);
  if (tid < ne) {
    const uint32_t pos = e0 + tid;
    colour = h.colour[pos];
    const uint32_t lv = reinterpret_cast<const uint32_t *>(h.loc)[pos];
    l0 = lv & 0xffffu, l1 = lv >> 16;
    nn = reinterpret_cast<const uint2 *>(a.maps[1])[pos];
  }
  double x1[2], x2[2];
  if (tid < ne) {
    if (SOA) {
      const uint32_t nx = a.num_points[1];
      x1[0] = x[nn.x], x1[1] = x[nx + nn.x], x2[0] = x[nn.y], x2[1] = x[nx + nn.y];
    } else {
      const double2 xa = reinterpret_cast<const double2 *>(x)[nn.x];
      const double2 xb = reinterpret_cast<const double2 *>(x)[nn.y];
      x1[0] = xa.x, x1[1] = xa.y, x2[0] = xb.x, x2[1] = xb.y;
    }
  }
  __syncthreads();
  double inc[8];
  if (tid < ne) {
    double q1[4], q2[4];
#pragma unroll
    for (uint32_t d = 0; d < 4; ++d) q1[d] = q_s[d * m + l0], q2[d] = q_s[d * m + l1];
    const double a1 = adt_s[l0], a2 = adt_s[l1];
    res_calc_user(x1, x2, q1, q2, &a1, &a2, inc, inc + 4, 1, 1);
  }
  const uint32_t ncol = h.blk_ncol[b];
  for (uint32_t c = 0; c < ncol; ++c) {
    if (colour == c) {
#pragma unroll
      for (uint32_t d = 0; d < 4; ++d) {
        res_s[d * m + l0] += inc[d];
        res_s[d * m + l1] += inc[4 + d];
      }
    }
    __syncthreads();
  }
  const uint32_t np = a.num_points[0];
  for (uint32_t i = tid; i < m; i += blockDim.x) {
    const uint32_t g = h.cached[c0 + i];
    if (SOA) {
#pragma unroll
      for (uint32_t d = 0; d < 4; ++d) {
        double *p = a.out + size_t(d) * np + g;
        *p = *p + res_s[d * m + i];
      }
    } else {
      double2 *o = reinterpret_cast<double2 *>(a.out) + 2 * size_t(g);
      double2 v0 = o[0], v1 = o[1];
      v0.x = v0.x + res_s[0 * m + i];
      v0.y = v0.y + res_s[1 * m + i];
      v1.x = v1.x + res_s[2 * m + i];
      v1.y = v1.y + res_s[3 * m + i];
      o[0] = v0;
      o[1] = v1;
    }
  }
}

// dst[d * n_dst + i] = src[d * n_src + idx[i]] (SOA gather of cell data)
__global__ __launch_bounds__(256) void gather_soa_kernel(double *__restrict__ dst,
                                                         const double *__restrict__ src,
                                                         const uint32_t *__restrict__ idx,
                                                         uint32_t n_dst, uint32_t n_src,
                                                         uint32_t dim) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_dst) return;
  const uint32_t s = idx[i];
  for (uint32_t d = 0; d < dim; ++d) dst[size_t(d) * n_dst + i] = src[size_t(d) * n_src + s];
}

// mappings >= 1 each functor reads through compute()
uint32_t mapping_mask(const fot_functor &f) {
  return f.kernel == FOT_KERNEL_RES_CALC ? 0x2u : 0u;
}
// functors whose increments read the iterate point_data[0] (Jacobi copy)
bool reads_iterate(const fot_functor &f) { return f.kernel == FOT_KERNEL_MINE; }

} // namespace

struct fot_loop {
  fot_functor f{};
  fot_problem_desc desc{};
  fot_loop_config cfg{};
  std::vector<void *> owned;
  EngineArgs base{};
  // global colour
  std::vector<EngineArgs> colour_args;
  // hierarchical
  HierDev hier{};
  EngineArgs plan_args{};
  std::vector<uint32_t> colour_blk_off;
  uint32_t lds_bytes = 0, threads = 0;
  bool staged = false;
  // graph of one iteration
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  fot_loop_stats stats{};

  ~fot_loop() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void *p : owned) (void)hipFree(p);
  }

  template <class T>
  fot_status upload(const T *host, size_t n, T **dev) {
    *dev = nullptr;
    if (n == 0) return FOT_OK;
    void *p = nullptr;
    FOT_HIP(hipMalloc(&p, n * sizeof(T)));
    owned.push_back(p);
    stats.plan_device_bytes += n * sizeof(T);
    FOT_HIP(hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice));
    *dev = static_cast<T *>(p);
    return FOT_OK;
  }
  fot_status alloc(size_t bytes, void **dev) {
    *dev = nullptr;
    if (bytes == 0) return FOT_OK;
    FOT_HIP(hipMalloc(dev, bytes));
    owned.push_back(*dev);
    stats.plan_device_bytes += bytes;
    return FOT_OK;
  }
  // SOA gather of every cell data array into a new buffer of n_dst cells
  fot_status gather_cell_data(const uint32_t *host_idx, uint32_t n_dst, EngineArgs &args) {
    if (desc.num_cell_data == 0 || n_dst == 0) return FOT_OK;
    uint32_t *d_idx = nullptr;
    void *tmp = nullptr;
    FOT_HIP(hipMalloc(&tmp, sizeof(uint32_t) * n_dst));
    d_idx = static_cast<uint32_t *>(tmp);
    hipError_t e = hipMemcpy(d_idx, host_idx, sizeof(uint32_t) * n_dst, hipMemcpyHostToDevice);
    for (uint32_t k = 0; k < desc.num_cell_data && e == hipSuccess; ++k) {
      void *dst = nullptr;
      const size_t bytes = sizeof(double) * size_t(n_dst) * desc.cell_dim[k];
      e = hipMalloc(&dst, bytes);
      if (e != hipSuccess) break;
      owned.push_back(dst);
      stats.plan_device_bytes += bytes;
      hipLaunchKernelGGL(gather_soa_kernel, dim3((n_dst + 255) / 256), dim3(256), 0, nullptr,
                         static_cast<double *>(dst), desc.cell_data[k], d_idx, n_dst,
                         desc.num_cells, desc.cell_dim[k]);
      e = hipGetLastError();
      args.cell_data[k] = static_cast<const double *>(dst);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(d_idx);
    if (e != hipSuccess) return fail(FOT_ERR_HIP, "cell data gather: %s", hipGetErrorString(e));
    args.cell_stride = n_dst;
    return FOT_OK;
  }

  fot_status build_atomic();
  fot_status build_global_colour();
  fot_status build_hierarchical();
  fot_status enqueue(hipStream_t s);
};

fot_status fot_loop::build_atomic() {
  for (uint32_t m = 0; m < desc.num_mappings; ++m) {
    uint32_t *d = nullptr;
    fot_status st = upload(desc.cell_to_node[m], size_t(desc.num_cells) * desc.mesh_dim[m], &d);
    if (st != FOT_OK) return st;
    base.maps[m] = d;
  }
  stats.num_launches_per_iter = 1;
  return FOT_OK;
}

fot_status fot_loop::build_global_colour() {
  // Mesh::colourCells then per-colour compacted maps and cell data
  // (problem.hpp:91-157)
  std::vector<uint8_t> col(desc.num_cells);
  uint32_t ncol = 0;
  fot_status st = fot_colour_cells(desc.cell_to_node[0], desc.num_cells, desc.mesh_dim[0],
                                   desc.num_points[0], col.data(), &ncol);
  if (st != FOT_OK) return st;
  std::vector<std::vector<uint32_t>> cells(ncol);
  for (uint32_t i = 0; i < desc.num_cells; ++i) cells[col[i]].push_back(i);
  std::vector<uint32_t> compact;
  for (uint32_t c = 0; c < ncol; ++c) {
    EngineArgs a = base;
    const uint32_t n = uint32_t(cells[c].size());
    for (uint32_t m = 0; m < desc.num_mappings; ++m) {
      const uint32_t md = desc.mesh_dim[m];
      compact.resize(size_t(n) * md);
      for (uint32_t i = 0; i < n; ++i)
        std::memcpy(&compact[size_t(i) * md], desc.cell_to_node[m] + size_t(cells[c][i]) * md,
                    md * sizeof(uint32_t));
      uint32_t *d = nullptr;
      st = upload(compact.data(), compact.size(), &d);
      if (st != FOT_OK) return st;
      a.maps[m] = d;
    }
    st = gather_cell_data(cells[c].data(), n, a);
    if (st != FOT_OK) return st;
    a.num_cells = n;
    colour_args.push_back(a);
  }
  stats.num_colours = ncol;
  stats.num_launches_per_iter = ncol;
  return FOT_OK;
}

fot_status fot_loop::build_hierarchical() {
  fot_hplan plan;
  fot_status st = build_hplan(desc.cell_to_node, desc.mesh_dim, desc.num_mappings,
                              desc.num_cells, desc.num_points[0], cfg.partition,
                              cfg.block_size, plan);
  if (st != FOT_OK) return st;
  const uint32_t md = desc.mesh_dim[0];
  std::vector<uint32_t> blk_cell_off{0}, blk_cache_off{0}, cached, plan_cell;
  std::vector<uint16_t> loc;
  std::vector<uint8_t> colour, blk_ncol;
  colour_blk_off.assign(1, 0);
  uint32_t max_cached = 0, max_tc = 0;
  double tc_sum = 0;
  const uint32_t mask = mapping_mask(f);
  std::vector<std::vector<uint32_t>> plan_maps(desc.num_mappings);
  for (const HostColour &hc : plan.colours) {
    const uint32_t nb = uint32_t(hc.num_cell_colours.size());
    const uint32_t n = uint32_t(hc.cell_colours.size());
    const uint32_t cell_base = uint32_t(plan_cell.size());
    for (uint32_t b = 0; b < nb; ++b) {
      blk_cell_off.push_back(cell_base + hc.block_offsets[b + 1]);
      const uint32_t cs = hc.points_to_be_cached_offsets[b + 1] - hc.points_to_be_cached_offsets[b];
      max_cached = std::max(max_cached, cs);
      blk_cache_off.push_back(blk_cache_off.back() + cs);
      blk_ncol.push_back(hc.num_cell_colours[b]);
      max_tc = std::max<uint32_t>(max_tc, hc.num_cell_colours[b]);
      tc_sum += hc.num_cell_colours[b];
    }
    cached.insert(cached.end(), hc.points_to_be_cached.begin(), hc.points_to_be_cached.end());
    for (uint32_t i = 0; i < n; ++i) {
      for (uint32_t k = 0; k < md; ++k) {
        const uint32_t l = hc.cell_list[0][size_t(k) * n + i];
        if (l > 0xffffu) return fail(FOT_ERR_UNSUPPORTED, "block cache above 65536 points");
        loc.push_back(uint16_t(l));
      }
      colour.push_back(hc.cell_colours[i]);
      plan_cell.push_back(hc.cell_inv_permutation[i]);
      for (uint32_t m = 1; m < desc.num_mappings; ++m) {
        if (!((mask >> m) & 1u)) continue;
        for (uint32_t j = 0; j < desc.mesh_dim[m]; ++j)
          plan_maps[m].push_back(hc.cell_list[m][size_t(j) * n + i]);
      }
    }
    colour_blk_off.push_back(colour_blk_off.back() + nb);
  }
  const uint32_t nb_total = colour_blk_off.back();
  uint32_t *d32 = nullptr;
  uint16_t *d16 = nullptr;
  uint8_t *d8 = nullptr;
  if ((st = upload(blk_cell_off.data(), blk_cell_off.size(), &d32)) != FOT_OK) return st;
  hier.blk_cell_off = d32;
  if ((st = upload(blk_cache_off.data(), blk_cache_off.size(), &d32)) != FOT_OK) return st;
  hier.blk_cache_off = d32;
  if ((st = upload(cached.data(), cached.size(), &d32)) != FOT_OK) return st;
  hier.cached = d32;
  if ((st = upload(loc.data(), loc.size(), &d16)) != FOT_OK) return st;
  hier.loc = d16;
  if ((st = upload(colour.data(), colour.size(), &d8)) != FOT_OK) return st;
  hier.colour = d8;
  if ((st = upload(blk_ncol.data(), blk_ncol.size(), &d8)) != FOT_OK) return st;
  hier.blk_ncol = d8;
  plan_args = base;
  for (uint32_t m = 1; m < desc.num_mappings; ++m) {
    if (!((mask >> m) & 1u)) continue;
    if ((st = upload(plan_maps[m].data(), plan_maps[m].size(), &d32)) != FOT_OK) return st;
    plan_args.maps[m] = d32;
  }
  if ((st = gather_cell_data(plan_cell.data(), desc.num_cells, plan_args)) != FOT_OK) return st;
  const uint32_t td = desc.point_dim[0];
  staged = (f.kernel == FOT_KERNEL_RES_CALC) && (cfg.flags & FOT_LOOP_STAGE_INPUTS);
  lds_bytes = uint32_t(sizeof(double)) * max_cached * (staged ? 9u : td);
  if (lds_bytes > 160u * 1024u)
    return fail(FOT_ERR_UNSUPPORTED, "block cache needs %u B of LDS (max 160 KiB)", lds_bytes);
  uint32_t max_block = 0;
  for (size_t b = 0; b + 1 < blk_cell_off.size(); ++b)
    max_block = std::max(max_block, blk_cell_off[b + 1] - blk_cell_off[b]);
  threads = std::max<uint32_t>(64, (max_block + 63) / 64 * 64);
  if (threads > 1024) return fail(FOT_ERR_UNSUPPORTED, "blocks above 1024 cells");
  stats.num_colours = uint32_t(plan.colours.size());
  stats.num_blocks = nb_total;
  stats.num_launches_per_iter = stats.num_colours;
  stats.max_cached = max_cached;
  stats.lds_bytes = lds_bytes;
  stats.max_thread_colours = max_tc;
  stats.avg_thread_colours = nb_total ? tc_sum / nb_total : 0.0;
  stats.reuse_factor = double(cached.size()) / (double(md) * desc.num_cells);
  return FOT_OK;
}

fot_status fot_loop::enqueue(hipStream_t s) {
  const fot_status st = dispatch(&f, desc.soa, [&]<class F, bool SOA>() -> fot_status {
    const uint32_t bs = cfg.block_size;
    if (cfg.strategy == FOT_STRATEGY_ATOMIC) {
      const uint32_t nb = (desc.num_cells + bs - 1) / bs;
      if constexpr (std::is_same<F, ResCalcF>::value)
        hipLaunchKernelGGL((res_calc_atomic_kernel<SOA>), dim3(nb), dim3(bs), 0, s, base);
      else
        hipLaunchKernelGGL((eng_direct_kernel<F, SOA, true>), dim3(nb), dim3(bs), 0, s, base);
    } else if (cfg.strategy == FOT_STRATEGY_GLOBAL_COLOUR) {
      for (const EngineArgs &a : colour_args) {
        if (a.num_cells == 0) continue;
        hipLaunchKernelGGL((eng_direct_kernel<F, SOA, false>), dim3((a.num_cells + bs - 1) / bs),
                           dim3(bs), 0, s, a);
      }
    } else {
      for (size_t c = 0; c + 1 < colour_blk_off.size(); ++c) {
        const uint32_t b0 = colour_blk_off[c], nbc = colour_blk_off[c + 1] - b0;
        if (nbc == 0) continue;
        if constexpr (std::is_same<F, ResCalcF>::value) {
          if (staged) {
            hipLaunchKernelGGL((res_calc_hier_staged_kernel<SOA>), dim3(nbc), dim3(threads),
                               lds_bytes, s, plan_args, hier, b0, nbc);
            continue;
          }
        }
        hipLaunchKernelGGL((eng_hier_kernel<F, SOA>), dim3(nbc), dim3(threads), lds_bytes, s,
                           plan_args, hier, b0, nbc);
      }
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FOT_ERR_HIP, "launch: %s", hipGetErrorString(e));
    return FOT_OK;
  });
  return st;
}

extern "C" fot_status fot_loop_create(const fot_functor *f, const fot_problem_desc *desc,
                                      const fot_loop_config *cfg, fot_loop **out) {
  if (!f || !desc || !cfg || !out) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (desc->num_mappings == 0 || desc->num_mappings > FOT_MAX_MAPPINGS ||
      desc->num_cell_data > FOT_MAX_CELL_DATA || !desc->point_data_out)
    return fail(FOT_ERR_INVALID_ARGUMENT, "bad problem description");
  if (desc->mesh_dim[0] != f->mesh_dim || desc->point_dim[0] != f->point_dim)
    return fail(FOT_ERR_INVALID_ARGUMENT, "mapping 0 shape does not match the functor");
  for (uint32_t m = 0; m < desc->num_mappings; ++m)
    if (!desc->cell_to_node[m] && desc->num_cells)
      return fail(FOT_ERR_INVALID_ARGUMENT, "mapping %u has no host array", m);
  // every mapping entry must name a point of its set: the kernels do not
  // bounds-check their gathers
  for (uint32_t m = 0; m < desc->num_mappings; ++m)
    for (size_t k = 0; k < size_t(desc->num_cells) * desc->mesh_dim[m]; ++k)
      if (desc->cell_to_node[m][k] >= desc->num_points[m])
        return fail(FOT_ERR_INVALID_INPUT, "mapping %u entry %zu = %u >= %u points", m, k,
                    desc->cell_to_node[m][k], desc->num_points[m]);
  if (f->kernel == FOT_KERNEL_RES_CALC &&
      (desc->num_mappings < 4 || desc->point_dim[1] != 2 || desc->point_dim[2] != 4 ||
       desc->point_dim[3] != 1 || desc->mesh_dim[1] != 2 ||
       desc->num_points[2] < desc->num_points[0] || desc->num_points[3] < desc->num_points[0]))
    return fail(FOT_ERR_INVALID_ARGUMENT, "res_calc needs res/x/q/adt mappings");
  if ((f->kernel == FOT_KERNEL_MINE || f->kernel == FOT_KERNEL_HEX_SCATTER) &&
      (desc->num_cell_data < 1 || desc->cell_dim[0] != f->cell_dim))
    return fail(FOT_ERR_INVALID_ARGUMENT, "functor needs cell data of dim %u", f->cell_dim);
  if (cfg->block_size == 0 || cfg->block_size > 1024)
    return fail(FOT_ERR_INVALID_ARGUMENT, "block size %u", cfg->block_size);
  std::unique_ptr<fot_loop> L(new fot_loop);
  L->f = *f;
  L->desc = *desc;
  L->cfg = *cfg;
  L->cfg.partition = nullptr; // host array only used while planning
  for (uint32_t m = 0; m < desc->num_mappings; ++m) {
    L->base.point_data[m] = desc->point_data[m];
    L->base.num_points[m] = desc->num_points[m];
  }
  for (uint32_t k = 0; k < desc->num_cell_data; ++k) L->base.cell_data[k] = desc->cell_data[k];
  L->base.out = desc->point_data_out;
  L->base.num_cells = desc->num_cells;
  L->base.cell_stride = desc->num_cells;
  L->stats.strategy = cfg->strategy;
  L->stats.block_size = cfg->block_size;
  fot_status st;
  L->cfg.partition = cfg->partition;
  switch (cfg->strategy) {
  case FOT_STRATEGY_ATOMIC: st = L->build_atomic(); break;
  case FOT_STRATEGY_GLOBAL_COLOUR: st = L->build_global_colour(); break;
  case FOT_STRATEGY_HIERARCHICAL: st = L->build_hierarchical(); break;
  default: st = fail(FOT_ERR_INVALID_ARGUMENT, "unknown strategy %u", cfg->strategy);
  }
  L->cfg.partition = nullptr;
  if (st != FOT_OK) return st;
  if (cfg->use_graph) {
    hipStream_t cs;
    FOT_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      st = L->enqueue(cs);
      e = hipStreamEndCapture(cs, &L->graph);
    }
    if (e == hipSuccess && st == FOT_OK) e = hipGraphInstantiate(&L->exec, L->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (st != FOT_OK) return st;
    if (e != hipSuccess) return fail(FOT_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
  }
  *out = L.release();
  return FOT_OK;
}

extern "C" fot_status fot_loop_destroy(fot_loop *loop) {
  delete loop;
  return FOT_OK;
}

extern "C" fot_status fot_loop_step(fot_loop *loop, void *stream) {
  if (!loop) return fail(FOT_ERR_INVALID_ARGUMENT, "null loop");
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (loop->exec) {
    FOT_HIP(hipGraphLaunch(loop->exec, s));
    return FOT_OK;
  }
  return loop->enqueue(s);
}

extern "C" fot_status fot_loop_run(fot_loop *loop, uint32_t iters, void *stream) {
  if (!loop) return fail(FOT_ERR_INVALID_ARGUMENT, "null loop");
  const fot_problem_desc &d = loop->desc;
  double *in = const_cast<double *>(d.point_data[0]);
  const size_t bytes = sizeof(double) * size_t(d.num_points[0]) * d.point_dim[0];
  fot_status st = fot_copy(d.point_data_out, in, bytes, stream);
  if (st != FOT_OK) return st;
  const bool jacobi = reads_iterate(loop->f);
  for (uint32_t it = 0; it < iters; ++it) {
    if ((st = fot_loop_step(loop, stream)) != FOT_OK) return st;
    if (jacobi || it + 1 == iters)
      if ((st = fot_copy(in, d.point_data_out, bytes, stream)) != FOT_OK) return st;
  }
  FOT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return FOT_OK;
}

extern "C" fot_status fot_loop_stats_get(const fot_loop *loop, fot_loop_stats *stats) {
  if (!loop || !stats) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  *stats = loop->stats;
  return FOT_OK;
}
