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
#include <numeric>
#include <type_traits>
#include <vector>

#include "fot_dispatch.hpp"
#include "fot_hip.hpp"

using namespace fot;


namespace {

constexpr uint32_t kNumXcd = 8; // MI355X: 8 XCDs of 32 CUs
// MI355X's Infinity Cache (MALL): the memory-side cache in front of HBM
constexpr uint64_t kInfinityCacheBytes = uint64_t(256) << 20;
typedef double dvec2 __attribute__((ext_vector_type(2)));

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
  const uint32_t *blk_owned;     // [nb] owned points (owner-compute plans only)
  const uint32_t *packed;        // res_calc owner compute: l0 | l1 << 12 | colour << 24
  // two-phase writeback: per cached entry its partial slot, or kExclusive
  // when no other block touches the point; partial sums [slot][TD]
  const uint32_t *wslot;
  double *partial;
};

constexpr uint32_t kExclusive = 0xffffffffu;

// Two-phase writeback of a block's cache entry i (point g, values v): added
// in place when this block alone touches g, else stored in its partial slot
// for the merge launch.
template <uint32_t TD, bool SOA>
FOT_DEV void writeback_two_phase(const HierDev &h, double *out, uint32_t np, uint32_t entry,
                                 uint32_t g, const double *v) {
  const uint32_t s = h.wslot[entry];
  if (s == kExclusive) {
#pragma unroll
    for (uint32_t d = 0; d < TD; ++d) {
      double *p = out + pidx<SOA>(np, g, TD, d);
      *p = *p + v[d];
    }
  } else {
#pragma unroll
    for (uint32_t d = 0; d < TD; ++d) h.partial[size_t(s) * TD + d] = v[d];
  }
}

// Merge launch of the two-phase writeback: point pt[j] gets its partials
// slots[off[j]] .. slots[off[j+1]-1], which are in block-colour order, added
// one by one, as the per-colour launches add the blocks' caches.  (Slots are
// handed out block by block, so a block's partial stores are contiguous;
// handing them out point by point instead, contiguous for this launch's
// reads, measured 5 % slower on hex and mini_aero meshes.)
template <uint32_t TD, bool SOA>
__global__ __launch_bounds__(256) void merge_partials_kernel(
    double *__restrict__ out, uint32_t np, const uint32_t *__restrict__ pt,
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ slots,
    const double *__restrict__ partial, uint32_t n) {
  const uint32_t j = blockIdx.x * 256u + threadIdx.x;
  if (j >= n) return;
  const uint32_t g = pt[j], s0 = off[j], s1 = off[j + 1];
  double v[TD];
#pragma unroll
  for (uint32_t d = 0; d < TD; ++d) v[d] = out[pidx<SOA>(np, g, TD, d)];
  for (uint32_t s = s0; s < s1; ++s) {
    const double *q = partial + size_t(slots[s]) * TD;
#pragma unroll
    for (uint32_t d = 0; d < TD; ++d) v[d] = v[d] + q[d];
  }
#pragma unroll
  for (uint32_t d = 0; d < TD; ++d) out[pidx<SOA>(np, g, TD, d)] = v[d];
}

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
template <class F, bool SOA, bool ATOMIC, int LB = 1024>
__global__ __launch_bounds__(LB) void eng_direct_kernel(EngineArgs a) {
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
// cache is added to the output.  Blocks of one block colour share no point.
// The writeback order between colours is enforced either by launch
// boundaries (one launch per block colour) or, in a single launch, by
// per-block completion flags: a block writes back only after every
// lower-colour block it shares a point with has published its writeback.
// Both keep the summation order of the plan, so they give identical bits.

struct DepSync {
  uint32_t ticket;     // next plan block to hand out
  uint32_t done;       // blocks retired in this launch
  uint32_t generation; // launches completed
  uint32_t error;      // a bounded wait ran out
};

struct DepDev {
  const uint32_t *dep_off; // [nb+1]
  const uint32_t *deps;    // immediate lower-colour neighbours of each block
  uint32_t *flags;         // [nb] = generation + 1 once written back
  DepSync *sync;
  uint32_t nb;
};

// ~0.1 s of s_sleep polling; a wait that long means a broken plan.
constexpr uint32_t kSpinLimit = 1u << 21;

// The plan block of this workgroup.  Per-colour launches remap blockIdx so
// each XCD gets a contiguous run of blocks; the single launch hands out
// tickets in plan (colour-major) order, so every block a workgroup may wait
// on has already been taken by a running workgroup.
template <bool DEP>
FOT_DEV uint32_t pick_block(const DepDev &dd, uint32_t blk_begin, uint32_t nbc, uint32_t &gen) {
  if constexpr (!DEP) {
    gen = 0;
    return blk_begin + xcd_block(blockIdx.x, nbc);
  } else {
    __shared__ uint32_t s_block, s_gen;
    if (threadIdx.x == 0) {
      s_block = atomicAdd(&dd.sync->ticket, 1u);
      s_gen = __hip_atomic_load(&dd.sync->generation, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    gen = s_gen;
    return s_block;
  }
}

// Waits (first wave) for the lower-colour neighbours of block b.
template <bool DEP>
FOT_DEV void wait_predecessors(const DepDev &dd, uint32_t b, uint32_t gen) {
  if constexpr (DEP) {
    if (threadIdx.x < 64) {
      const uint32_t d0 = dd.dep_off[b], nd = dd.dep_off[b + 1] - d0;
      for (uint32_t k = threadIdx.x; k < nd; k += 64) {
        const uint32_t p = dd.deps[d0 + k];
        uint32_t spins = 0;
        while (__hip_atomic_load(&dd.flags[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) !=
               gen + 1u) {
          __builtin_amdgcn_s_sleep(1);
          if (++spins == kSpinLimit) {
            atomicOr(&dd.sync->error, 1u);
            break;
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
  }
}

// Publishes block b's writeback; the last block of the launch re-arms the
// ticket counter for the next launch.
template <bool DEP>
FOT_DEV void publish(const DepDev &dd, uint32_t b, uint32_t gen) {
  if constexpr (DEP) {
    __builtin_amdgcn_s_waitcnt(0); // this wave's stores are in L2
    __syncthreads();
    if (threadIdx.x == 0) {
      __hip_atomic_store(&dd.flags[b], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      if (atomicAdd(&dd.sync->done, 1u) == dd.nb - 1u) {
        __hip_atomic_store(&dd.sync->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&dd.sync->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&dd.sync->generation, gen + 1u, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

// MODE: 0 one launch per block colour, 1 dependency order, 2 two-phase
// writeback (every block in one launch; merge_partials_kernel after).  LB:
// the most threads a launch may use (a lower bound leaves the compiler more
// registers per thread).  STG: the functor's staged input (StagedInput) is
// loaded once per block point into LDS after the cache, and the functor reads
// it there by block-local id instead of gathering it per cell.  NTC: the
// functor's read-once cell data loaded with nontemporal hints
// (StreamCellArgs; the same values).
template <class F, bool SOA, int MODE, int LB = 1024, bool STG = false, bool NTC = false>
__global__ __launch_bounds__(LB, LB < 1024 ? kSmallBoundWaves<F> : 1) void eng_hier_kernel(
    std::conditional_t<NTC, StreamCellArgs, EngineArgs> a, HierDev h, DepDev dd,
    uint32_t blk_begin, uint32_t nbc) {
  constexpr bool DEP = MODE == 1;
  extern __shared__ double cache[];
  uint32_t gen;
  const uint32_t b = pick_block<DEP>(dd, blk_begin, nbc, gen);
  const uint32_t tid = threadIdx.x;
  const uint32_t e0 = h.blk_cell_off[b], ne = h.blk_cell_off[b + 1] - e0;
  const uint32_t c0 = h.blk_cache_off[b], m = h.blk_cache_off[b + 1] - c0;
  for (uint32_t i = tid; i < m * F::TD; i += blockDim.x) cache[i] = 0.0;
  uint32_t loc[F::MD], tgt[F::MD];
  double inc[F::MD * F::TD];
  uint32_t colour = 0xffu;
  if constexpr (STG) {
    using SI = StagedInput<F>;
    double *st = cache + size_t(m) * F::TD;
    const double *src = a.pd(SI::slot);
    const uint32_t n = a.ps(SI::slot);
    // a thread stages whole points: the id, then every value of the point
    // (two round trips, all of a point's loads in flight together)
    for (uint32_t i = tid; i < m; i += blockDim.x) {
      const uint32_t g = h.cached[c0 + i];
      double v[SI::dim];
#pragma unroll
      for (uint32_t k = 0; k < SI::dim; ++k)
        v[k] = src[SOA ? size_t(k) * n + g : size_t(g) * SI::dim + k];
#pragma unroll
      for (uint32_t k = 0; k < SI::dim; ++k) st[SOA ? k * m + i : i * SI::dim + k] = v[k];
    }
    __syncthreads();
    if (tid < ne) {
      const uint32_t pos = e0 + tid;
      colour = h.colour[pos];
      load_loc<F::MD>(h.loc + size_t(pos) * F::MD, loc);
      F::template compute<SOA>(LdsArgs<SI::slot>{a, st, m}, pos, loc, inc);
    }
  } else {
    if (tid < ne) {
      const uint32_t pos = e0 + tid;
      colour = h.colour[pos];
      load_loc<F::MD>(h.loc + size_t(pos) * F::MD, loc);
#pragma unroll
      for (uint32_t k = 0; k < F::MD; ++k) tgt[k] = h.cached[c0 + loc[k]];
      F::template compute<SOA>(a, pos, tgt, inc);
    }
  }
  __syncthreads();
  const uint32_t ncol = h.blk_ncol[b];
  for (uint32_t c = 0; c < ncol; ++c) {
    if (colour == c) add_to_cache<F>(cache, m, loc, inc);
    __syncthreads();
  }
  wait_predecessors<DEP>(dd, b, gen);
  // (loading the first two write-back entries per thread, their ids, slots
  // and in-place output values, before the accumulation measured slower on
  // hex 128^3: 150 -> 179 us at B = 512, two-phase)
  const uint32_t np = a.num_points[0];
  for (uint32_t i = tid; i < m; i += blockDim.x) {
    const uint32_t g = h.cached[c0 + i];
    if constexpr (MODE == 2) {
      double v[F::TD];
#pragma unroll
      for (uint32_t d = 0; d < F::TD; ++d) v[d] = cache[d * m + i];
      writeback_two_phase<F::TD, SOA>(h, a.out, np, c0 + i, g, v);
    } else {
#pragma unroll
      for (uint32_t d = 0; d < F::TD; ++d) {
        double *p = a.out + pidx<SOA>(np, g, F::TD, d);
        *p = *p + cache[d * m + i];
      }
    }
  }
  publish<DEP>(dd, b, gen);
}

// res_calc hierarchical with the block's cells' q and adt staged in LDS next
// to the res cache (they are read through the same mapping-0 points), so
// the edge loop gathers only x from HBM.
template <bool SOA, int MODE>
__global__ __launch_bounds__(1024) void res_calc_hier_staged_kernel(EngineArgs a, HierDev h,
                                                                    DepDev dd,
                                                                    uint32_t blk_begin,
                                                                    uint32_t nbc) {
  constexpr bool DEP = MODE == 1;
  extern __shared__ double lds[];
  uint32_t gen;
  const uint32_t b = pick_block<DEP>(dd, blk_begin, nbc, gen);
  const uint32_t tid = threadIdx.x, bd = blockDim.x;
  const uint32_t e0 = h.blk_cell_off[b], ne = h.blk_cell_off[b + 1] - e0;
  const uint32_t c0 = h.blk_cache_off[b], m = h.blk_cache_off[b + 1] - c0;
  double *res_s = lds, *q_s = lds + 4 * m, *adt_s = lds + 8 * m;
  const double *x = a.point_data[1], *q = a.point_data[2], *adt = a.point_data[3];
  const uint32_t np = a.num_points[0], nq = a.num_points[2], nx = a.num_points[1];
  // Two memory round trips before the barrier, not four: first the plan
  // entries (this thread's edge and its first staged cell), then every value
  // they name (x of the edge, q and adt of the cell).  Indices are clamped
  // into the block so the loads are unconditional; a block with more cached
  // cells than threads stages the rest in further rounds.  (Reading the res
  // to write back this early as well measured slower: 42.6 vs 39.3 us per
  // step on the GPS-ordered 1200x600 airfoil with 1024-edge blocks.)
  const uint32_t pos = e0 + min(tid, ne - 1);
  const uint32_t i0 = min(tid, m - 1);
  const uint32_t colour_e = h.colour[pos];
  const uint32_t lv = reinterpret_cast<const uint32_t *>(h.loc)[pos];
  const uint2 nn = reinterpret_cast<const uint2 *>(a.maps[1])[pos];
  const uint32_t g0 = h.cached[c0 + i0];
  double x1[2], x2[2], q0[4], adt0;
  if (SOA) {
    x1[0] = x[nn.x], x1[1] = x[nx + nn.x], x2[0] = x[nn.y], x2[1] = x[nx + nn.y];
#pragma unroll
    for (uint32_t d = 0; d < 4; ++d) q0[d] = q[size_t(d) * nq + g0];
  } else {
    const double2 xa = reinterpret_cast<const double2 *>(x)[nn.x];
    const double2 xb = reinterpret_cast<const double2 *>(x)[nn.y];
    x1[0] = xa.x, x1[1] = xa.y, x2[0] = xb.x, x2[1] = xb.y;
    const double2 *q2p = reinterpret_cast<const double2 *>(q);
    const double2 v0 = q2p[2 * size_t(g0)], v1 = q2p[2 * size_t(g0) + 1];
    q0[0] = v0.x, q0[1] = v0.y, q0[2] = v1.x, q0[3] = v1.y;
  }
  adt0 = adt[g0];
  if (tid < m) {
#pragma unroll
    for (uint32_t d = 0; d < 4; ++d) q_s[d * m + tid] = q0[d], res_s[d * m + tid] = 0.0;
    adt_s[tid] = adt0;
  }
  for (uint32_t i = tid + bd; i < m; i += bd) {
    const uint32_t g = h.cached[c0 + i];
#pragma unroll
    for (uint32_t d = 0; d < 4; ++d) q_s[d * m + i] = q[pidx<SOA>(nq, g, 4, d)];
    adt_s[i] = adt[g];
#pragma unroll
    for (uint32_t d = 0; d < 4; ++d) res_s[d * m + i] = 0.0;
  }
  const uint32_t colour = tid < ne ? colour_e : 0xffu;
  const uint32_t l0 = lv & 0xffffu, l1 = lv >> 16;
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
  wait_predecessors<DEP>(dd, b, gen);
  if constexpr (MODE == 2) {
    for (uint32_t i = tid; i < m; i += bd) {
      const uint32_t g = i == tid ? g0 : h.cached[c0 + i];
      double v[4];
#pragma unroll
      for (uint32_t d = 0; d < 4; ++d) v[d] = res_s[d * m + i];
      writeback_two_phase<4, SOA>(h, a.out, np, c0 + i, g, v);
    }
    return;
  }
  for (uint32_t i = tid; i < m; i += bd) {
    const uint32_t g = i == tid ? g0 : h.cached[c0 + i];
    double r[4];
    if (SOA) {
#pragma unroll
      for (uint32_t d = 0; d < 4; ++d) r[d] = a.out[size_t(d) * np + g];
    } else {
      const double2 *o = reinterpret_cast<const double2 *>(a.out) + 2 * size_t(g);
      const double2 u0 = o[0], u1 = o[1];
      r[0] = u0.x, r[1] = u0.y, r[2] = u1.x, r[3] = u1.y;
    }
#pragma unroll
    for (uint32_t d = 0; d < 4; ++d) r[d] = r[d] + res_s[d * m + i];
    if (SOA) {
#pragma unroll
      for (uint32_t d = 0; d < 4; ++d) a.out[size_t(d) * np + g] = r[d];
    } else {
      double2 *o = reinterpret_cast<double2 *>(a.out) + 2 * size_t(g);
      o[0] = make_double2(r[0], r[1]);
      o[1] = make_double2(r[2], r[3]);
    }
  }
  publish<DEP>(dd, b, gen);
}

// ---- owner-compute launches -----------------------------------------------
// One launch over all blocks.  A block's LDS cache holds the current values
// of its owned points, loaded while its inputs are staged; the increments of
// its cells are added thread colour by thread colour (halo slots belong to
// another block and are skipped) and the cache is stored back.  No block
// shares a written point with another, so there is no order between blocks.

// Whether an owner block loads its plan entries and first staged values
// before the functor's arithmetic, overlapping their round trips.  Off where
// that measured slower: mini_aero's face flux already fills the 128 registers
// a 1024-thread workgroup allows and spills, and getacc lost 1 %.
template <class F>
constexpr bool kEarlyStaging = !std::is_same_v<F, fot::MiniAeroF> && !std::is_same_v<F, fot::GetaccF>;

template <class F, bool SOA>
__device__ __forceinline__ void owner_block(const EngineArgs &a, const HierDev &h, uint32_t b,
                                            double *lds) {
  const uint32_t tid = threadIdx.x, bd = blockDim.x;
  const uint32_t e0 = h.blk_cell_off[b], ne = h.blk_cell_off[b + 1] - e0;
  const uint32_t c0 = h.blk_cache_off[b];
  const uint32_t nw = h.blk_owned[b];
  double *cache = lds;
  uint32_t *g_s = reinterpret_cast<uint32_t *>(lds + size_t(F::TD) * nw);
  const uint32_t np = a.num_points[0];
  uint32_t loc[F::MD], tgt[F::MD];
  double inc[F::MD * F::TD];
  uint32_t colour = 0xffu;
  if constexpr (kEarlyStaging<F>) {
    // Plan entries first (the first staging round's owned id, this thread's
    // cell and its local ids), then the values they name (the owned point's
    // current value, the cell's target ids), then the functor's own loads:
    // three dependent round trips instead of five.  Indices are clamped into
    // the block, so the loads are unconditional; the staged value is in LDS
    // before the functor computes, so it holds no register across it.
    // (a block owns at least one point; it may run no cell, when its points
    // are touched by none: then it reads position 0 and local id 0, both in
    // bounds, and computes nothing)
    const uint32_t g0 = h.cached[c0 + min(tid, nw - 1)];
    const bool cells = ne > 0;
    const uint32_t pos = cells ? e0 + min(tid, ne - 1) : 0u;
    const uint32_t colour_c = h.colour[pos];
    load_loc<F::MD>(h.loc + size_t(pos) * F::MD, loc);
#pragma unroll
    for (uint32_t k = 0; k < F::MD; ++k) loc[k] = cells ? loc[k] : 0u;
    double v0[F::TD];
#pragma unroll
    for (uint32_t d = 0; d < F::TD; ++d) v0[d] = a.out[pidx<SOA>(np, g0, F::TD, d)];
#pragma unroll
    for (uint32_t k = 0; k < F::MD; ++k) tgt[k] = h.cached[c0 + loc[k]];
    if (tid < nw) {
      g_s[tid] = g0;
#pragma unroll
      for (uint32_t d = 0; d < F::TD; ++d) cache[d * nw + tid] = v0[d];
    }
    for (uint32_t i = tid + bd; i < nw; i += bd) {
      const uint32_t g = h.cached[c0 + i];
      g_s[i] = g;
#pragma unroll
      for (uint32_t d = 0; d < F::TD; ++d) cache[d * nw + i] = a.out[pidx<SOA>(np, g, F::TD, d)];
    }
    if (tid < ne) {
      colour = colour_c;
      F::template compute<SOA>(a, pos, tgt, inc);
    }
  } else {
    // stage the owned points, then load the cell's plan entries and compute
    for (uint32_t i = tid; i < nw; i += bd) {
      const uint32_t g = h.cached[c0 + i];
      g_s[i] = g;
#pragma unroll
      for (uint32_t d = 0; d < F::TD; ++d) cache[d * nw + i] = a.out[pidx<SOA>(np, g, F::TD, d)];
    }
    if (tid < ne) {
      const uint32_t pos = e0 + tid;
      colour = h.colour[pos];
      load_loc<F::MD>(h.loc + size_t(pos) * F::MD, loc);
#pragma unroll
      for (uint32_t k = 0; k < F::MD; ++k) tgt[k] = h.cached[c0 + loc[k]];
      F::template compute<SOA>(a, pos, tgt, inc);
    }
  }
  __syncthreads();
  const uint32_t ncol = h.blk_ncol[b];
  for (uint32_t c = 0; c < ncol; ++c) {
    if (colour == c) add_to_cache_owned<F>(cache, nw, loc, inc, nw);
    __syncthreads();
  }
  for (uint32_t i = tid; i < nw; i += blockDim.x) {
    const uint32_t g = g_s[i];
#pragma unroll
    for (uint32_t d = 0; d < F::TD; ++d) a.out[pidx<SOA>(np, g, F::TD, d)] = cache[d * nw + i];
  }
}

template <class F, bool SOA, int LB = 1024>
__global__ __launch_bounds__(LB) void eng_owner_kernel(EngineArgs a, HierDev h, uint32_t nb) {
  extern __shared__ double lds[];
  owner_block<F, SOA>(a, h, xcd_block(blockIdx.x, nb), lds);
}

// One problem of a batch: its engine arguments and owner-compute plan.
struct BatchProblem {
  EngineArgs args;
  HierDev hier;
};

// Batched owner-compute step: a persistent grid whose workgroups take
// (problem, block) items from a device queue until it runs dry.  queue[0] is
// the next item, queue[1] counts finished workgroups; the last workgroup
// resets both, so the next launch (or graph replay) starts from zero with no
// memset.
template <class F, bool SOA>
__global__ __launch_bounds__(1024) void batch_owner_kernel(const BatchProblem *__restrict__ probs,
                                                           const uint2 *__restrict__ work,
                                                           uint32_t nwork,
                                                           uint32_t *__restrict__ queue) {
  extern __shared__ double lds[];
  // the next item is taken while the current one runs, so the queue
  // atomic's round trip overlaps the work instead of preceding each item
  __shared__ uint32_t item[2];
  if (threadIdx.x == 0) item[0] = atomicAdd(&queue[0], 1u);
  __syncthreads();
  uint32_t w = item[0], buf = 0;
  while (w < nwork) {
    if (threadIdx.x == 0) item[buf ^ 1u] = atomicAdd(&queue[0], 1u);
    const uint2 pb = work[w];
    owner_block<F, SOA>(probs[pb.x].args, probs[pb.x].hier, pb.y, lds);
    __syncthreads(); // this item's LDS no longer read; the next index visible
    buf ^= 1u;
    w = item[buf];
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
      atomicExch(&queue[0], 0u);
      atomicExch(&queue[1], 0u);
    }
  }
}

// Batched copies: item i copies count doubles (<= kCopyChunk) src -> dst.
struct CopyItem {
  double *dst;
  const double *src;
  uint32_t count, pad;
};
constexpr uint32_t kCopyChunk = 4096;

__global__ __launch_bounds__(256) void batch_copy_kernel(const CopyItem *__restrict__ items) {
  const CopyItem it = items[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < it.count; i += blockDim.x) it.dst[i] = it.src[i];
}

// Entries appended to the owner res_calc plan arrays: covers reads up to
// threads * max(cells, staged points) per thread past a block's start.
constexpr uint32_t kOwnerPad = 4096;

// Per-block record of the owner-compute res_calc kernel: one scalar load.
struct OwnerMeta {
  uint32_t cell_off, ncells, cache_off, ncached;
  uint32_t nowned, own_start, ncol, contiguous; // owned ids own_start + i when contiguous
  uint32_t node_off, nnodes, pad0, pad1;        // the block's mapping-1 points (x);
                                                // pad0: first owned-cell list entry
  uint32_t pad2, pad3, pad4, pad5;
};

// res_calc owner compute, latency-oriented: block record in one scalar load,
// owned points addressed directly when the mesh numbers them contiguously,
// EPT edges per thread so each wave keeps more gathers in flight.
template <bool SOA, int EPT, bool SX, int KS, bool PRE>
__global__ __launch_bounds__(1024) void res_calc_owner2_kernel(EngineArgs a,
                                                               const OwnerMeta *__restrict__ meta,
                                                               const uint32_t *__restrict__ cached,
                                                               const uint32_t *__restrict__ packed,
                                                               const uint32_t *__restrict__ nodes,
                                                               const uint32_t *__restrict__ nloc,
                                                               uint32_t nb) {
  extern __shared__ double lds[];
  const uint32_t b = xcd_block(blockIdx.x, nb);
  const OwnerMeta mt = meta[b];
  const uint32_t tid = threadIdx.x, bd = blockDim.x;
  const uint32_t m = mt.ncached, nw = mt.nowned, ne = mt.ncells;
  const bool contig = mt.contiguous != 0;
  const uint32_t nx_b = SX ? mt.nnodes : 0;
  // staged nodes first: dynamic LDS starts 16-byte aligned
  double2 *x_s = reinterpret_cast<double2 *>(lds);
  double *res_s = lds + 2 * nx_b, *q_s = res_s + 4 * nw, *adt_s = q_s + 4 * m;
  // with PRE, 1/rho and the pressure of every staged cell follow adt
  double *ri_s = adt_s + m, *p_s = ri_s + (PRE ? m : 0);
  uint32_t *g_s = reinterpret_cast<uint32_t *>(p_s + (PRE ? m : 0));
  const double *x = a.point_data[1], *q = a.point_data[2], *adt = a.point_data[3];
  const uint32_t np = a.num_points[0];
  const uint32_t *pcached = cached + mt.cache_off;
  uint32_t pk[EPT];
  uint2 nn[EPT];
  // plan arrays are padded: these loads need no bounds branch, and branch-free
  // loads keep the compiler from waiting on each one at a join
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const uint32_t t = tid + e * bd;
    const uint32_t pv = packed[mt.cell_off + t];
    pk[e] = t < ne ? pv : 0xff000000u;
    if (SX) {
      const uint32_t v = nloc[mt.cell_off + t];
      nn[e] = make_uint2(v & 0xffffu, v >> 16);
    } else {
      nn[e] = reinterpret_cast<const uint2 *>(a.maps[1])[mt.cell_off + t];
    }
  }
  // Staging in two rounds so that every wave has all its loads in flight at
  // once: first the point ids (KS per thread), then the data they name.
  uint32_t gq[KS], gx[KS];
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const uint32_t i = tid + k * bd;
    const uint32_t pc = pcached[i];
    gq[k] = (contig && i < nw) ? mt.own_start + i : (i < m ? pc : 0u);
    gx[k] = 0;
    if (SX) {
      const uint32_t pn = nodes[mt.node_off + i];
      gx[k] = i < nx_b ? pn : 0u;
    }
  }
  // every id is a valid point (point 0 past the block's end): unconditional.
  // res is read for owned slots only; the others re-read the first owned
  // cell (a line the block reads anyway) instead of a halo cell's res
  double2 vq0[KS], vq1[KS], vr0[KS], vr1[KS], vx[KS];
  double vadt[KS];
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const uint32_t g = gq[k];
    const uint32_t gr = (tid + k * bd < nw) ? g : mt.own_start;
    if (SOA) {
      const uint32_t nq = a.num_points[2];
      vq0[k] = make_double2(q[g], q[size_t(nq) + g]);
      vq1[k] = make_double2(q[2 * size_t(nq) + g], q[3 * size_t(nq) + g]);
      vr0[k] = make_double2(a.out[gr], a.out[size_t(np) + gr]);
      vr1[k] = make_double2(a.out[2 * size_t(np) + gr], a.out[3 * size_t(np) + gr]);
    } else {
      const double2 *q2p = reinterpret_cast<const double2 *>(q);
      vq0[k] = q2p[2 * size_t(g)], vq1[k] = q2p[2 * size_t(g) + 1];
      const double2 *o = reinterpret_cast<const double2 *>(a.out) + 2 * size_t(gr);
      vr0[k] = o[0], vr1[k] = o[1];
    }
    vadt[k] = adt[g];
    if (SX) {
      const uint32_t nxp = a.num_points[1];
      vx[k] = SOA ? make_double2(x[gx[k]], x[size_t(nxp) + gx[k]])
                  : reinterpret_cast<const double2 *>(x)[gx[k]];
    }
  }
  // one wait for all of them: without this use the compiler sinks each load
  // into the guarded LDS store below and waits at every guard
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    asm volatile("" ::"v"(vq0[k].x), "v"(vq0[k].y), "v"(vq1[k].x), "v"(vq1[k].y));
    asm volatile("" ::"v"(vr0[k].x), "v"(vr0[k].y), "v"(vr1[k].x), "v"(vr1[k].y), "v"(vadt[k]));
    if (SX) asm volatile("" ::"v"(vx[k].x), "v"(vx[k].y));
  }
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const uint32_t i = tid + k * bd;
    if (i < m) {
      q_s[0 * m + i] = vq0[k].x, q_s[1 * m + i] = vq0[k].y;
      q_s[2 * m + i] = vq1[k].x, q_s[3 * m + i] = vq1[k].y;
      adt_s[i] = vadt[k];
      if (PRE) {
        const double qq[4] = {vq0[k].x, vq0[k].y, vq1[k].x, vq1[k].y};
        res_calc_point_terms(qq, 1, &ri_s[i], &p_s[i]);
      }
      if (i < nw) {
        res_s[0 * nw + i] = vr0[k].x, res_s[1 * nw + i] = vr0[k].y;
        res_s[2 * nw + i] = vr1[k].x, res_s[3 * nw + i] = vr1[k].y;
        if (!contig) g_s[i] = gq[k];
      }
    }
    if (SX && i < nx_b) x_s[i] = vx[k];
  }
  double x1[EPT][2], x2[EPT][2];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    x1[e][0] = x1[e][1] = x2[e][0] = x2[e][1] = 0.0;
    if (!SX) { // nn is a valid node pair for every thread (padded map)
      if (SOA) {
        const uint32_t nx = a.num_points[1];
        x1[e][0] = x[nn[e].x], x1[e][1] = x[nx + nn[e].x];
        x2[e][0] = x[nn[e].y], x2[e][1] = x[nx + nn[e].y];
      } else {
        const double2 xa = reinterpret_cast<const double2 *>(x)[nn[e].x];
        const double2 xb = reinterpret_cast<const double2 *>(x)[nn[e].y];
        x1[e][0] = xa.x, x1[e][1] = xa.y, x2[e][0] = xb.x, x2[e][1] = xb.y;
      }
    }
  }
  __syncthreads();
  double inc[EPT][8];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    if (tid + e * bd < ne) {
      if (SX) {
        const double2 xa = x_s[nn[e].x], xb = x_s[nn[e].y];
        x1[e][0] = xa.x, x1[e][1] = xa.y, x2[e][0] = xb.x, x2[e][1] = xb.y;
      }
      const uint32_t l0 = pk[e] & 0xfffu, l1 = (pk[e] >> 12) & 0xfffu;
      double q1[4], q2[4];
#pragma unroll
      for (uint32_t d = 0; d < 4; ++d) q1[d] = q_s[d * m + l0], q2[d] = q_s[d * m + l1];
      const double a1 = adt_s[l0], a2 = adt_s[l1];
      if (PRE)
        res_calc_user_terms(x1[e], x2[e], q1, q2, ri_s[l0], ri_s[l1], p_s[l0], p_s[l1], a1, a2,
                            inc[e], inc[e] + 4);
      else
        res_calc_user(x1[e], x2[e], q1, q2, &a1, &a2, inc[e], inc[e] + 4, 1, 1);
    }
  }
  for (uint32_t c = 0; c < mt.ncol; ++c) {
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      if ((pk[e] >> 24) == c) {
        const uint32_t l0 = pk[e] & 0xfffu, l1 = (pk[e] >> 12) & 0xfffu;
#pragma unroll
        for (uint32_t d = 0; d < 4; ++d) {
          if (l0 < nw) res_s[d * nw + l0] += inc[e][d];
          if (l1 < nw) res_s[d * nw + l1] += inc[e][4 + d];
        }
      }
    }
    __syncthreads();
  }
  for (uint32_t i = tid; i < nw; i += bd) {
    const uint32_t g = contig ? mt.own_start + i : g_s[i];
    if (SOA) {
#pragma unroll
      for (uint32_t d = 0; d < 4; ++d) a.out[size_t(d) * np + g] = res_s[d * nw + i];
    } else {
      double2 *o = reinterpret_cast<double2 *>(a.out) + 2 * size_t(g);
      o[0] = make_double2(res_s[0 * nw + i], res_s[1 * nw + i]);
      o[1] = make_double2(res_s[2 * nw + i], res_s[3 * nw + i]);
    }
  }
}

// Owner-compute res_calc with a gather accumulation: each edge's flux f is
// computed, written to LDS (over the dead q / adt region), and each owned
// cell adds the +-f of its incident edges in thread-colour order, starting
// from its res held in registers.  Every owned cell receives at most one
// increment per colour, so this is the colour-ordered sum of the other
// owner kernels, bit for bit, with three barriers instead of one per
// colour, no res staging in LDS and a smaller LDS footprint.  plist holds
// per owned cell up to 2 * LW entries (edge position | slot << 15; LW 32-bit
// words, 2 or 4), sorted by colour, 0xffff-terminated.  Owned cells occupy
// the first KO staging rounds (KO = 1 when every block owns at most
// blockDim.x cells): later rounds stage halo q / adt only, with no res or
// list registers.
//
// A block runs in three steps: its ids (packed edge words, node pairs,
// staged cell ids, owned-cell lists), the data they name, then the body.
// res_calc_owner3_kernel runs one block per workgroup; the persistent
// res_calc_owner4_kernel loads the next block's ids while it runs the body
// of the current one.

// a block's ids, in registers
template <int EPT, int KS, int KO> struct Owner3Ids {
  uint32_t pk[EPT];
  uint2 nn[EPT];
  uint32_t gq[KS];
  uint32_t lst[KO][4];
};

// the data a block's ids name, in registers until the body stages them
template <int EPT, int KS, int KO> struct Owner3Data {
  double2 vq0[KS], vq1[KS], vr0[KO], vr1[KO];
  double vadt[KS];
  double x1[EPT][2], x2[EPT][2];
};

// plan arrays are padded: branch-free loads
template <int EPT, int KS, int LW, int KO>
FOT_DEV void owner3_load_ids(Owner3Ids<EPT, KS, KO> &id, const OwnerMeta &mt, const EngineArgs &a,
                             const uint32_t *__restrict__ cached,
                             const uint32_t *__restrict__ packed,
                             const uint32_t *__restrict__ plist, uint32_t tid, uint32_t bd) {
  typedef typename std::conditional<LW == 4, uint4, uint2>::type list_t;
  const list_t *lists = reinterpret_cast<const list_t *>(plist);
  const uint32_t *pcached = cached + mt.cache_off;
  const uint32_t m = mt.ncached, nw = mt.nowned, ne = mt.ncells;
  const bool contig = mt.contiguous != 0;
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const uint32_t t = tid + e * bd;
    const uint32_t pv = packed[mt.cell_off + t];
    id.pk[e] = t < ne ? pv : 0xff000000u;
    id.nn[e] = reinterpret_cast<const uint2 *>(a.maps[1])[mt.cell_off + t];
  }
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const uint32_t i = tid + k * bd;
    // a contiguous block's owned ids are own_start + i: their lanes, and the
    // lanes past the block's end, read one shared word (the first halo id,
    // the last id) instead of loading ids they do not use
    const uint32_t j = contig ? (i < nw ? nw : (i < m ? i : m - 1)) : i;
    const uint32_t pc = pcached[j];
    id.gq[k] = (contig && i < nw) ? mt.own_start + i : (i < m ? pc : 0u);
    if (LW > 0 && k < KO) {
      const list_t l = lists[mt.pad0 + i];
      id.lst[k][0] = l.x, id.lst[k][1] = l.y;
      if constexpr (LW == 4) id.lst[k][2] = l.z, id.lst[k][3] = l.w;
    }
  }
}

// every id is a valid point (point 0 past the block's end): unconditional.
// res is read for owned slots only; the others re-read the first owned cell
// (a line the block reads anyway) instead of a halo cell's res
// NT: the owned res values are loaded (and stored by owner3_body) with
// nontemporal hints, so that on a working set beyond the Infinity Cache the
// read-once, write-once res does not evict the x and q values other blocks
// read again
template <bool SOA, int EPT, int KS, int KO, bool NT = false>
FOT_DEV void owner3_load_data(Owner3Data<EPT, KS, KO> &d, const Owner3Ids<EPT, KS, KO> &id,
                              const OwnerMeta &mt, const EngineArgs &a, uint32_t tid,
                              uint32_t bd) {
  const double *x = a.point_data[1], *q = a.point_data[2], *adt = a.point_data[3];
  const uint32_t np = a.num_points[0], nw = mt.nowned;
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const uint32_t g = id.gq[k];
    const uint32_t gr = (tid + k * bd < nw) ? g : mt.own_start;
    if (SOA) {
      const uint32_t nq = a.num_points[2];
      d.vq0[k] = make_double2(q[g], q[size_t(nq) + g]);
      d.vq1[k] = make_double2(q[2 * size_t(nq) + g], q[3 * size_t(nq) + g]);
      if (k < KO) {
        double r[4];
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c)
          r[c] = NT ? __builtin_nontemporal_load(&a.out[c * size_t(np) + gr])
                    : a.out[c * size_t(np) + gr];
        d.vr0[k] = make_double2(r[0], r[1]), d.vr1[k] = make_double2(r[2], r[3]);
      }
    } else {
      const double2 *q2p = reinterpret_cast<const double2 *>(q);
      d.vq0[k] = q2p[2 * size_t(g)], d.vq1[k] = q2p[2 * size_t(g) + 1];
      if (k < KO) {
        if (NT) {
          const dvec2 *o = reinterpret_cast<const dvec2 *>(a.out) + 2 * size_t(gr);
          const dvec2 r0 = __builtin_nontemporal_load(o), r1 = __builtin_nontemporal_load(o + 1);
          d.vr0[k] = make_double2(r0.x, r0.y), d.vr1[k] = make_double2(r1.x, r1.y);
        } else {
          const double2 *o = reinterpret_cast<const double2 *>(a.out) + 2 * size_t(gr);
          d.vr0[k] = o[0], d.vr1[k] = o[1];
        }
      }
    }
    d.vadt[k] = adt[g];
  }
#pragma unroll
  for (int e = 0; e < EPT; ++e) { // nn is a valid node pair for every thread (padded map)
    if (SOA) {
      const uint32_t nx = a.num_points[1];
      d.x1[e][0] = x[id.nn[e].x], d.x1[e][1] = x[nx + id.nn[e].x];
      d.x2[e][0] = x[id.nn[e].y], d.x2[e][1] = x[nx + id.nn[e].y];
    } else {
      const double2 xa = reinterpret_cast<const double2 *>(x)[id.nn[e].x];
      const double2 xb = reinterpret_cast<const double2 *>(x)[id.nn[e].y];
      d.x1[e][0] = xa.x, d.x1[e][1] = xa.y, d.x2[e][0] = xb.x, d.x2[e][1] = xb.y;
    }
  }
}

// stage q / adt, fluxes, gather, store.  Starts with the LDS free; ends with
// the LDS still being read by other waves (callers reusing it sync first).
// With PRE, 1/rho and the pressure of each staged cell are computed once at
// staging (LDS) instead of once per edge side: the same operations on the
// same values, the same bits.
template <bool SOA, int EPT, int KS, int LW, int KO, bool PRE, bool NT = false>
FOT_DEV void owner3_body(double *lds, Owner3Data<EPT, KS, KO> &d, const Owner3Ids<EPT, KS, KO> &id,
                         const OwnerMeta &mt, const EngineArgs &a, uint32_t tid, uint32_t bd) {
  const uint32_t m = mt.ncached, nw = mt.nowned, ne = mt.ncells;
  const uint32_t np = a.num_points[0];
  // LDS holds double2 planes, so lanes on consecutive slots touch
  // consecutive 16-byte words: q as (q0, q1)[m] then (q2, q3)[m], and after
  // the second barrier the fluxes as (f0, f1)[ne] then (f2, f3)[ne]
  double *q_s = lds, *adt_s = q_s + 4 * m, *ri_s = adt_s + m, *p_s = ri_s + m;
  double2 *q2_s = reinterpret_cast<double2 *>(q_s);
  double2 *f2_s = reinterpret_cast<double2 *>(lds);
  // LW == 0: the owned cells' edge lists are built here instead of loaded:
  // slot_s[c * nw + i] is the position (| 0x8000 for the edge's second cell)
  // of owned cell i's edge of thread colour c, 0xffff where it has none.  A
  // cell has at most one edge per colour (the plan checks it), so reading its
  // slots in colour order is the order of the loaded lists.
  const uint32_t ncol = mt.ncol;
  uint16_t *slot_s = reinterpret_cast<uint16_t *>(
      lds + (PRE ? 7u * m > 4u * ne ? 7u * m : 4u * ne : 5u * m > 4u * ne ? 5u * m : 4u * ne));
  if (LW == 0)
    for (uint32_t i = tid; i < ncol * nw; i += bd) slot_s[i] = 0xffffu;
  // one wait for all the loads (see res_calc_owner2_kernel)
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    asm volatile("" ::"v"(d.vq0[k].x), "v"(d.vq0[k].y), "v"(d.vq1[k].x), "v"(d.vq1[k].y),
                 "v"(d.vadt[k]));
  }
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const uint32_t i = tid + k * bd;
    if (i < m) {
      q2_s[i] = d.vq0[k];
      q2_s[m + i] = d.vq1[k];
      adt_s[i] = d.vadt[k];
      if (PRE) {
        const double qq[4] = {d.vq0[k].x, d.vq0[k].y, d.vq1[k].x, d.vq1[k].y};
        res_calc_point_terms(qq, 1, &ri_s[i], &p_s[i]);
      }
    }
  }
  __syncthreads();
  double fl[EPT][4];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    if (tid + e * bd < ne) {
      const uint32_t l0 = id.pk[e] & 0xfffu, l1 = (id.pk[e] >> 12) & 0xfffu;
      const double2 a0 = q2_s[l0], a1 = q2_s[m + l0], b0 = q2_s[l1], b1 = q2_s[m + l1];
      const double q1[4] = {a0.x, a0.y, a1.x, a1.y}, q2[4] = {b0.x, b0.y, b1.x, b1.y};
      double r2[4];
      const double adt1 = adt_s[l0], adt2 = adt_s[l1];
      if (PRE)
        res_calc_user_terms(d.x1[e], d.x2[e], q1, q2, ri_s[l0], ri_s[l1], p_s[l0], p_s[l1], adt1,
                            adt2, fl[e], r2);
      else
        res_calc_user(d.x1[e], d.x2[e], q1, q2, &adt1, &adt2, fl[e], r2, 1, 1);
    }
  }
  __syncthreads(); // q and adt are dead: the fluxes take their place
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const uint32_t t = tid + e * bd;
    if (t < ne) {
      f2_s[t] = make_double2(fl[e][0], fl[e][1]);
      f2_s[ne + t] = make_double2(fl[e][2], fl[e][3]);
      if (LW == 0) {
        const uint32_t l0 = id.pk[e] & 0xfffu, l1 = (id.pk[e] >> 12) & 0xfffu;
        const uint32_t col = id.pk[e] >> 24;
        if (l0 < nw) slot_s[col * nw + l0] = uint16_t(t);
        if (l1 < nw) slot_s[col * nw + l1] = uint16_t(t | 0x8000u);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KO; ++k) {
    const uint32_t i = tid + k * bd;
    if (i < nw) {
      double r[4] = {d.vr0[k].x, d.vr0[k].y, d.vr1[k].x, d.vr1[k].y};
      auto add = [&](uint32_t ent) {
        const uint32_t pos = ent & 0x7fffu;
        const double2 f01 = f2_s[pos], f23 = f2_s[ne + pos];
        if (ent & 0x8000u) { // the edge's second cell: res2 = -f, and r + (-f) == r - f
          r[0] -= f01.x, r[1] -= f01.y, r[2] -= f23.x, r[3] -= f23.y;
        } else {
          r[0] += f01.x, r[1] += f01.y, r[2] += f23.x, r[3] += f23.y;
        }
      };
      if (LW == 0) {
        for (uint32_t c = 0; c < ncol; ++c) {
          const uint32_t ent = slot_s[c * nw + i];
          if (ent != 0xffffu) add(ent);
        }
      } else {
#pragma unroll
        for (int j = 0; j < 2 * LW; ++j) {
          const uint32_t ent = (id.lst[k][j >> 1] >> (16 * (j & 1))) & 0xffffu;
          if (ent == 0xffffu) break;
          add(ent);
        }
      }
      const uint32_t g = id.gq[k];
      if (SOA) {
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) {
          if (NT)
            __builtin_nontemporal_store(r[c], &a.out[size_t(c) * np + g]);
          else
            a.out[size_t(c) * np + g] = r[c];
        }
      } else if (NT) {
        dvec2 *o = reinterpret_cast<dvec2 *>(a.out) + 2 * size_t(g);
        __builtin_nontemporal_store(dvec2{r[0], r[1]}, o);
        __builtin_nontemporal_store(dvec2{r[2], r[3]}, o + 1);
      } else {
        double2 *o = reinterpret_cast<double2 *>(a.out) + 2 * size_t(g);
        o[0] = make_double2(r[0], r[1]);
        o[1] = make_double2(r[2], r[3]);
      }
    }
  }
}

template <bool SOA, int EPT, int KS, int LW, int KO, bool PRE>
__global__ __launch_bounds__(1024) void res_calc_owner3_kernel(EngineArgs a,
                                                               const OwnerMeta *__restrict__ meta,
                                                               const uint32_t *__restrict__ cached,
                                                               const uint32_t *__restrict__ packed,
                                                               const uint32_t *__restrict__ plist,
                                                               uint32_t nb) {
  extern __shared__ double lds[];
  const uint32_t b = xcd_block(blockIdx.x, nb);
  const OwnerMeta mt = meta[b];
  const uint32_t tid = threadIdx.x, bd = blockDim.x;
  Owner3Ids<EPT, KS, KO> id;
  Owner3Data<EPT, KS, KO> d;
  owner3_load_ids<EPT, KS, LW, KO>(id, mt, a, cached, packed, plist, tid, bd);
  owner3_load_data<SOA, EPT, KS, KO>(d, id, mt, a, tid, bd);
  owner3_body<SOA, EPT, KS, LW, KO, PRE>(lds, d, id, mt, a, tid, bd);
}

// The same with the owned res loaded and stored nontemporally (edge lists
// built in LDS, no per-cell terms): the same bits.
template <bool SOA, int EPT, int KS, int KO>
__global__ __launch_bounds__(1024) void res_calc_owner3_stream_kernel(
    EngineArgs a, const OwnerMeta *__restrict__ meta, const uint32_t *__restrict__ cached,
    const uint32_t *__restrict__ packed, const uint32_t *__restrict__ plist, uint32_t nb) {
  extern __shared__ double lds[];
  const uint32_t b = xcd_block(blockIdx.x, nb);
  const OwnerMeta mt = meta[b];
  const uint32_t tid = threadIdx.x, bd = blockDim.x;
  Owner3Ids<EPT, KS, KO> id;
  Owner3Data<EPT, KS, KO> d;
  owner3_load_ids<EPT, KS, 0, KO>(id, mt, a, cached, packed, plist, tid, bd);
  owner3_load_data<SOA, EPT, KS, KO, true>(d, id, mt, a, tid, bd);
  owner3_body<SOA, EPT, KS, 0, KO, false, true>(lds, d, id, mt, a, tid, bd);
}

// Persistent form: gridDim.x workgroups (a multiple of 8, at most the
// resident count) each run blocks j = blockIdx.x, + gridDim.x, ...; with
// blocks mapped by xcd_block every workgroup's blocks sit in one XCD's
// chunk.  The next block's record and ids load while the current block's
// body runs, so after its first block a workgroup waits for one round trip
// per block (the data), not two.  Same per-block arithmetic: the same bits.
template <bool SOA, int EPT, int KS, int LW, int KO, bool PRE>
__global__ __launch_bounds__(1024) void res_calc_owner4_kernel(EngineArgs a,
                                                               const OwnerMeta *__restrict__ meta,
                                                               const uint32_t *__restrict__ cached,
                                                               const uint32_t *__restrict__ packed,
                                                               const uint32_t *__restrict__ plist,
                                                               uint32_t nb) {
  extern __shared__ double lds[];
  const uint32_t tid = threadIdx.x, bd = blockDim.x, G = gridDim.x;
  uint32_t j = blockIdx.x;
  if (j >= nb) return;
  OwnerMeta mt = meta[xcd_block(j, nb)];
  Owner3Ids<EPT, KS, KO> id;
  owner3_load_ids<EPT, KS, LW, KO>(id, mt, a, cached, packed, plist, tid, bd);
  for (;;) {
    Owner3Data<EPT, KS, KO> d;
    owner3_load_data<SOA, EPT, KS, KO>(d, id, mt, a, tid, bd);
    const uint32_t jn = j + G;
    const bool more = jn < nb; // uniform
    // the last block loads its own entries again: branch-free, in bounds
    const OwnerMeta mtn = meta[xcd_block(more ? jn : j, nb)];
    Owner3Ids<EPT, KS, KO> idn;
    owner3_load_ids<EPT, KS, LW, KO>(idn, mtn, a, cached, packed, plist, tid, bd);
    owner3_body<SOA, EPT, KS, LW, KO, PRE>(lds, d, id, mt, a, tid, bd);
    if (!more) break;
    __syncthreads(); // the next block's staging overwrites f_s
    j = jn;
    mt = mtn;
    id = idn;
  }
}

// the gather kernel instantiation for a plan: one block per workgroup or
// persistent, edges per thread, staging rounds, list words, owned rounds
template <bool SOA, int E, int K, int W, int O>
const void *owner3_pick(bool persistent, bool pre) {
  if (pre)
    return persistent
               ? reinterpret_cast<const void *>(&res_calc_owner4_kernel<SOA, E, K, W, O, true>)
               : reinterpret_cast<const void *>(&res_calc_owner3_kernel<SOA, E, K, W, O, true>);
  return persistent
             ? reinterpret_cast<const void *>(&res_calc_owner4_kernel<SOA, E, K, W, O, false>)
             : reinterpret_cast<const void *>(&res_calc_owner3_kernel<SOA, E, K, W, O, false>);
}
template <bool SOA, int E, int K>
const void *owner3_pick_w(bool persistent, bool pre, uint32_t lw, bool ko1, bool nt) {
  if (lw == 0 && nt && !persistent && !pre)
    return ko1 ? reinterpret_cast<const void *>(&res_calc_owner3_stream_kernel<SOA, E, K, 1>)
               : reinterpret_cast<const void *>(&res_calc_owner3_stream_kernel<SOA, E, K, K>);
  if (lw == 0)
    return ko1 ? owner3_pick<SOA, E, K, 0, 1>(persistent, pre)
               : owner3_pick<SOA, E, K, 0, K>(persistent, pre);
  if (lw == 2)
    return ko1 ? owner3_pick<SOA, E, K, 2, 1>(persistent, pre)
               : owner3_pick<SOA, E, K, 2, K>(persistent, pre);
  return ko1 ? owner3_pick<SOA, E, K, 4, 1>(persistent, pre)
             : owner3_pick<SOA, E, K, 4, K>(persistent, pre);
}
template <bool SOA, int E>
const void *owner3_pick_k(bool persistent, bool pre, uint32_t ks, uint32_t lw, bool ko1,
                          bool nt) {
  switch (ks) {
  case 1: return owner3_pick_w<SOA, E, 1>(persistent, pre, lw, ko1, nt);
  case 2: return owner3_pick_w<SOA, E, 2>(persistent, pre, lw, ko1, nt);
  case 3: return owner3_pick_w<SOA, E, 3>(persistent, pre, lw, ko1, nt);
  default: return owner3_pick_w<SOA, E, 4>(persistent, pre, lw, ko1, nt);
  }
}
// nt: the streaming-output instantiation, which exists for the default form
// only (one block per workgroup, lists in LDS, no per-cell terms); the other
// forms ignore it
template <bool SOA>
const void *owner3_pick_kernel(bool persistent, bool pre, uint32_t ept, uint32_t ks, uint32_t lw,
                               bool ko1, bool nt) {
  return ept == 2 ? owner3_pick_k<SOA, 2>(persistent, pre, ks, lw, ko1, nt)
                  : owner3_pick_k<SOA, 1>(persistent, pre, ks, lw, ko1, nt);
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

// dst[d * n_dst + ids[i]] = src[d * n_src + pos[i]] (SOA scatter of a cell
// output back to the caller's cell order)
__global__ __launch_bounds__(256) void scatter_soa_kernel(double *__restrict__ dst,
                                                          const double *__restrict__ src,
                                                          const uint32_t *__restrict__ pos,
                                                          const uint32_t *__restrict__ ids,
                                                          uint32_t n, uint32_t n_dst,
                                                          uint32_t n_src, uint32_t dim) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = pos[i], c = ids[i];
  for (uint32_t d = 0; d < dim; ++d) dst[size_t(d) * n_dst + c] = src[size_t(d) * n_src + p];
}

// the cell data slot a functor writes (lulesh's element volume), or -1
int cell_output_slot(const fot_functor &f) { return f.kernel == FOT_KERNEL_LULESH ? 1 : -1; }

// mappings >= 1 each functor reads through compute()
uint32_t mapping_mask(const fot_functor &f) {
  return f.kernel == FOT_KERNEL_RES_CALC ? 0x2u : 0u;
}
// functors whose increments read the iterate point_data[0] (Jacobi copy)
bool reads_iterate(const fot_functor &f) { return f.kernel == FOT_KERNEL_MINE; }

} // namespace

// Launches eng_hier_kernel<F, SOA, MODE>: the small-bound instantiation when
// the blocks fit it, the staged-input one when asked and the functor has a
// staged input.
template <class F, bool SOA, int MODE>
void launch_eng_hier(bool stage, uint32_t grid, uint32_t threads, uint32_t lds, hipStream_t s,
                     const EngineArgs &a, const HierDev &h, const DepDev &dd, uint32_t b0,
                     uint32_t nbc, bool stream_cells = false) {
  constexpr bool CAN = StagedInput<F>::slot >= 0;
  constexpr int SB = kSmallBound<F>;
  if constexpr (MODE == 2 && kStreamsCells<F>) {
    if (stream_cells && !stage) {
      const StreamCellArgs sa{a};
      if (threads <= uint32_t(SB))
        hipLaunchKernelGGL((eng_hier_kernel<F, SOA, 2, SB, false, true>), dim3(grid),
                           dim3(threads), lds, s, sa, h, dd, b0, nbc);
      else
        hipLaunchKernelGGL((eng_hier_kernel<F, SOA, 2, 1024, false, true>), dim3(grid),
                           dim3(threads), lds, s, sa, h, dd, b0, nbc);
      return;
    }
  }
  if constexpr (SB < 1024) {
    if (threads <= uint32_t(SB)) {
      if constexpr (CAN) {
        if (stage) {
          hipLaunchKernelGGL((eng_hier_kernel<F, SOA, MODE, SB, true>), dim3(grid),
                             dim3(threads), lds, s, a, h, dd, b0, nbc);
          return;
        }
      }
      hipLaunchKernelGGL((eng_hier_kernel<F, SOA, MODE, SB, false>), dim3(grid), dim3(threads),
                         lds, s, a, h, dd, b0, nbc);
      return;
    }
  }
  if constexpr (CAN) {
    if (stage) {
      hipLaunchKernelGGL((eng_hier_kernel<F, SOA, MODE, 1024, true>), dim3(grid), dim3(threads),
                         lds, s, a, h, dd, b0, nbc);
      return;
    }
  }
  hipLaunchKernelGGL((eng_hier_kernel<F, SOA, MODE, 1024, false>), dim3(grid), dim3(threads),
                     lds, s, a, h, dd, b0, nbc);
}

// Values per point of a functor's staged input (0: none).
uint32_t staged_input_dim(const fot_functor &f) {
  uint32_t d = 0;
  (void)dispatch(&f, 0, [&]<class F, bool SOA>() -> fot_status {
    d = StagedInput<F>::dim;
    return FOT_OK;
  });
  return d;
}

// Whether the functor's two-phase launch has a streaming-cell-data form.
bool streams_cells(const fot_functor &f) {
  bool r = false;
  (void)dispatch(&f, 0, [&]<class F, bool SOA>() -> fot_status {
    r = kStreamsCells<F>;
    return FOT_OK;
  });
  return r;
}

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
  bool stage_pd = false; // generic hierarchical: the functor's staged input in LDS
  bool dep_order = false;
  bool owner = false;
  // two-phase writeback: the merge launch's shared points, their partial
  // slot lists (CSR) and the slots
  bool two_phase = false;
  uint32_t merge_n = 0;
  const uint32_t *merge_pt = nullptr, *merge_off = nullptr, *merge_slots = nullptr;
  const OwnerMeta *owner_meta = nullptr;
  const uint32_t *owner_nodes = nullptr, *owner_nloc = nullptr;
  bool stage_x = false;
  bool point_terms = false;      // owner res_calc: per-cell 1/rho and pressure
  bool owner_gather = false;     // owner res_calc: gather accumulation kernel
  const uint32_t *owner_plist = nullptr;
  uint32_t list_words = 4;       // owner gather: 32-bit words per owned-cell list
  bool lists_in_lds = false;     // owner gather: the kernel builds the lists in LDS
  uint32_t stage_per_thread = 1; // owner res_calc: staged points per thread
  bool owned_first_round = false; // owner res_calc: every block owns <= threads cells
  const void *owner3_fn = nullptr; // owner gather: the kernel instantiation to launch
  // nontemporal hints on read-once data: owner res_calc gather, the owned
  // res; two-phase, the streaming functors' cell data
  bool nontemporal = false;
  uint32_t launch_grid = 0;        // workgroups of the single-launch strategies
  uint32_t ept = 1;
  DepDev dep{};
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
  // a written cell data slot in a plan-order copy: (copy position, caller's
  // cell) pairs, one per distinct cell, for the scatter back
  struct CellOut {
    const double *src;
    uint32_t src_stride;
    const uint32_t *pos, *ids;
    uint32_t count;
  };
  std::vector<CellOut> cell_out;

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
    const int slot = cell_output_slot(f);
    if (slot >= 0 && uint32_t(slot) < desc.num_cell_data) {
      // owner plans repeat halo cells: scatter each cell from its first copy
      std::vector<uint8_t> seen(desc.num_cells, 0);
      std::vector<uint32_t> pos, ids;
      for (uint32_t i = 0; i < n_dst; ++i)
        if (!seen[host_idx[i]]) {
          seen[host_idx[i]] = 1;
          pos.push_back(i);
          ids.push_back(host_idx[i]);
        }
      uint32_t *dp = nullptr, *di = nullptr;
      fot_status st;
      if ((st = upload(pos.data(), pos.size(), &dp)) != FOT_OK) return st;
      if ((st = upload(ids.data(), ids.size(), &di)) != FOT_OK) return st;
      cell_out.push_back({args.cell_data[slot], n_dst, dp, di, uint32_t(pos.size())});
    }
    return FOT_OK;
  }

  // writes the functor's cell output from the plan-order copies back to the
  // caller's cell data (no-op when the loop works on the caller's array)
  fot_status scatter_cell_outputs(hipStream_t s) {
    const int slot = cell_output_slot(f);
    for (const CellOut &c : cell_out) {
      if (c.count == 0) continue;
      hipLaunchKernelGGL(scatter_soa_kernel, dim3((c.count + 255) / 256), dim3(256), 0, s,
                         const_cast<double *>(desc.cell_data[slot]), c.src, c.pos, c.ids, c.count,
                         desc.num_cells, c.src_stride, desc.cell_dim[slot]);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FOT_ERR_HIP, "cell output scatter: %s", hipGetErrorString(e));
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

// The gather kernel adds each owned cell's fluxes in thread-colour order
// through per-cell position lists, so the positions of a block can come in
// any order without changing a bit.  Ascending (first, second) block-local
// ids put neighbouring lanes of the flux step on neighbouring LDS words.
static void sort_positions_by_local_ids(const std::vector<uint32_t> &blk_cell_off,
                                        std::vector<uint32_t> &cells,
                                        std::vector<uint8_t> &colour,
                                        std::vector<uint16_t> &loc) {
  std::vector<uint32_t> idx, c2;
  std::vector<uint8_t> k2;
  std::vector<uint16_t> l2;
  for (size_t b = 0; b + 1 < blk_cell_off.size(); ++b) {
    const uint32_t p0 = blk_cell_off[b], n = blk_cell_off[b + 1] - p0;
    idx.resize(n);
    std::iota(idx.begin(), idx.end(), p0);
    auto key = [&](uint32_t i) {
      return (uint32_t(loc[2 * size_t(i)]) << 16) | loc[2 * size_t(i) + 1];
    };
    std::stable_sort(idx.begin(), idx.end(),
                     [&](uint32_t a, uint32_t c) { return key(a) < key(c); });
    c2.resize(n);
    k2.resize(n);
    l2.resize(2 * size_t(n));
    for (uint32_t i = 0; i < n; ++i) {
      c2[i] = cells[idx[i]];
      k2[i] = colour[idx[i]];
      l2[2 * size_t(i)] = loc[2 * size_t(idx[i])];
      l2[2 * size_t(i) + 1] = loc[2 * size_t(idx[i]) + 1];
    }
    std::copy(c2.begin(), c2.end(), cells.begin() + p0);
    std::copy(k2.begin(), k2.end(), colour.begin() + p0);
    std::copy(l2.begin(), l2.end(), loc.begin() + 2 * size_t(p0));
  }
}

fot_status fot_loop::build_hierarchical() {
  fot_status st;
  const uint32_t md = desc.mesh_dim[0];
  uint32_t max_nodes = 0;
  std::vector<uint32_t> blk_cell_off{0}, blk_cache_off{0}, cached, plan_cell, blk_owned;
  std::vector<uint16_t> loc;
  std::vector<uint8_t> colour, blk_ncol;
  colour_blk_off.assign(1, 0);
  uint32_t max_cached = 0, max_tc = 0, num_plan_colours = 0;
  double tc_sum = 0;
  const uint32_t mask = mapping_mask(f);
  std::vector<std::vector<uint32_t>> plan_maps(desc.num_mappings);
  owner = (cfg.flags & FOT_LOOP_OWNER_COMPUTE) != 0;
  if (owner && (cfg.flags & FOT_LOOP_DEPENDENCY_ORDER))
    return fail(FOT_ERR_INVALID_ARGUMENT, "owner compute needs no dependency order");
  two_phase = (cfg.flags & FOT_LOOP_TWO_PHASE) != 0;
  if (two_phase && (owner || (cfg.flags & FOT_LOOP_DEPENDENCY_ORDER)))
    return fail(FOT_ERR_INVALID_ARGUMENT,
                "the two-phase writeback is a block-colour plan's (no owner compute, no "
                "dependency order)");
  if (f.kernel == FOT_KERNEL_RES_CALC &&
      (owner || (cfg.flags & FOT_LOOP_STAGE_INPUTS))) {
    // these kernels stage q and adt through mapping-0 ids; res_calc reads its
    // res, q and adt mappings from three streams (res_calc.cu:13-20) that the
    // airfoil mesh fills with the one edge->cell map
    const size_t n = size_t(desc.num_cells) * 2 * sizeof(uint32_t);
    if (n && (std::memcmp(desc.cell_to_node[2], desc.cell_to_node[0], n) ||
              std::memcmp(desc.cell_to_node[3], desc.cell_to_node[0], n)))
      return fail(FOT_ERR_UNSUPPORTED,
                  "staged res_calc needs the q and adt mappings equal to the res mapping");
  }
  if (owner) {
    OwnerPlan op;
    const uint32_t *cell_part = cfg.partition, *point_part = nullptr;
    if (cfg.flags & FOT_LOOP_POINT_PARTITION) {
      if (!cfg.partition) return fail(FOT_ERR_INVALID_ARGUMENT, "point partition missing");
      point_part = cfg.partition;
      cell_part = nullptr;
    }
    st = build_owner_plan(desc.cell_to_node[0], desc.num_cells, md, desc.num_points[0],
                          cell_part, cfg.block_size, point_part, op);
    if (st != FOT_OK) return st;
    const uint32_t nb = uint32_t(op.blk_owned.size());
    blk_cell_off = op.blk_cell_off;
    blk_cache_off = op.blk_cache_off;
    blk_owned = op.blk_owned;
    cached = op.cached;
    plan_cell = op.cells;
    colour = op.colour;
    blk_ncol = op.blk_ncol;
    loc.assign(op.loc.begin(), op.loc.end());
    if (f.kernel == FOT_KERNEL_RES_CALC && (cfg.flags & FOT_LOOP_OWNER_GATHER) && md == 2)
      sort_positions_by_local_ids(blk_cell_off, plan_cell, colour, loc);
    for (uint32_t b = 0; b < nb; ++b) {
      max_cached = std::max(max_cached, blk_cache_off[b + 1] - blk_cache_off[b]);
      max_tc = std::max<uint32_t>(max_tc, blk_ncol[b]);
      tc_sum += blk_ncol[b];
    }
    for (uint32_t m = 1; m < desc.num_mappings; ++m) {
      if (!((mask >> m) & 1u)) continue;
      const uint32_t dm = desc.mesh_dim[m];
      plan_maps[m].resize(size_t(plan_cell.size()) * dm);
      for (size_t i = 0; i < plan_cell.size(); ++i)
        std::memcpy(&plan_maps[m][i * dm], desc.cell_to_node[m] + size_t(plan_cell[i]) * dm,
                    dm * sizeof(uint32_t));
    }
    colour_blk_off.push_back(nb);
    num_plan_colours = 1;
  } else {
  fot_hplan plan;
  st = build_hplan(desc.cell_to_node, desc.mesh_dim, desc.num_mappings,
                              desc.num_cells, desc.num_points[0], cfg.partition,
                              cfg.block_size, plan);
  if (st != FOT_OK) return st;
  num_plan_colours = uint32_t(plan.colours.size());
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
  }
  const uint32_t nb_total = colour_blk_off.back();
  uint32_t *d32 = nullptr;
  uint16_t *d16 = nullptr;
  uint8_t *d8 = nullptr;
  if ((st = upload(blk_cell_off.data(), blk_cell_off.size(), &d32)) != FOT_OK) return st;
  hier.blk_cell_off = d32;
  if ((st = upload(blk_cache_off.data(), blk_cache_off.size(), &d32)) != FOT_OK) return st;
  hier.blk_cache_off = d32;
  // owner res_calc loads plan arrays without bounds branches (clamped reads
  // past a block's end land in the next block or this padding)
  const bool owner_rc = owner && f.kernel == FOT_KERNEL_RES_CALC;
  point_terms = owner_rc && (cfg.flags & FOT_LOOP_POINT_TERMS);
  if (owner_rc) cached.resize(cached.size() + kOwnerPad, 0u);
  if ((st = upload(cached.data(), cached.size(), &d32)) != FOT_OK) return st;
  hier.cached = d32;
  if ((st = upload(loc.data(), loc.size(), &d16)) != FOT_OK) return st;
  hier.loc = d16;
  if ((st = upload(colour.data(), colour.size(), &d8)) != FOT_OK) return st;
  hier.colour = d8;
  if ((st = upload(blk_ncol.data(), blk_ncol.size(), &d8)) != FOT_OK) return st;
  hier.blk_ncol = d8;
  if (owner) {
    if ((st = upload(blk_owned.data(), blk_owned.size(), &d32)) != FOT_OK) return st;
    hier.blk_owned = d32;
    if (f.kernel == FOT_KERNEL_RES_CALC) {
      std::vector<uint32_t> packed(plan_cell.size());
      for (size_t i = 0; i < plan_cell.size(); ++i) {
        const uint32_t l0 = loc[2 * i], l1 = loc[2 * i + 1];
        if (l0 > 0xfffu || l1 > 0xfffu)
          return fail(FOT_ERR_UNSUPPORTED, "owner block with more than 4096 points");
        packed[i] = l0 | (l1 << 12) | (uint32_t(colour[i]) << 24);
      }
      packed.resize(packed.size() + kOwnerPad, 0xff000000u);
      if ((st = upload(packed.data(), packed.size(), &d32)) != FOT_OK) return st;
      hier.packed = d32;
      std::vector<OwnerMeta> meta(blk_owned.size());
      for (size_t b = 0; b < meta.size(); ++b) {
        OwnerMeta &mt = meta[b];
        mt.cell_off = blk_cell_off[b];
        mt.ncells = blk_cell_off[b + 1] - blk_cell_off[b];
        mt.cache_off = blk_cache_off[b];
        mt.ncached = blk_cache_off[b + 1] - blk_cache_off[b];
        mt.nowned = blk_owned[b];
        mt.own_start = mt.nowned ? cached[mt.cache_off] : 0;
        mt.ncol = blk_ncol[b];
        mt.contiguous = 1;
        for (uint32_t i = 1; i < mt.nowned; ++i)
          if (cached[mt.cache_off + i] != mt.own_start + i) {
            mt.contiguous = 0;
            break;
          }
      }
      // owned-cell edge lists for the gather kernel: (position | slot << 15)
      // sorted by the edge's thread colour, at most 8, 0xffff-terminated
      owner_gather = (cfg.flags & FOT_LOOP_OWNER_GATHER) != 0;
      if (owner_gather) {
        uint32_t total_owned = 0;
        for (const OwnerMeta &mt : meta) total_owned += mt.nowned;
        std::vector<uint16_t> plist((size_t(total_owned) + kOwnerPad) * 8, 0xffffu);
        uint32_t longest = 0;
        bool one_per_colour = true; // every owned cell has <= 1 edge per colour
        std::vector<std::vector<std::pair<uint32_t, uint16_t>>> lists;
        uint32_t own_off = 0;
        for (size_t b = 0; b < meta.size() && owner_gather; ++b) {
          OwnerMeta &mt = meta[b];
          mt.pad0 = own_off;
          lists.assign(mt.nowned, {});
          for (uint32_t t = 0; t < mt.ncells; ++t) {
            const uint32_t pv = packed[mt.cell_off + t];
            const uint32_t l0 = pv & 0xfffu, l1 = (pv >> 12) & 0xfffu, col = pv >> 24;
            if (l0 < mt.nowned) lists[l0].push_back({col, uint16_t(t)});
            if (l1 < mt.nowned) lists[l1].push_back({col, uint16_t(t | 0x8000u)});
          }
          for (uint32_t i = 0; i < mt.nowned && owner_gather; ++i) {
            if (lists[i].size() > 8 || mt.ncells > 0x8000u) {
              owner_gather = false;
              break;
            }
            std::sort(lists[i].begin(), lists[i].end());
            for (size_t j = 1; j < lists[i].size(); ++j)
              if (lists[i][j].first == lists[i][j - 1].first) one_per_colour = false;
            longest = std::max(longest, uint32_t(lists[i].size()));
            for (size_t j = 0; j < lists[i].size(); ++j)
              plist[(size_t(own_off) + i) * 8 + j] = lists[i][j].second;
          }
          own_off += mt.nowned;
        }
        if (owner_gather) {
          // 4 entries (8 bytes) per owned cell when every list fits, else 8;
          // none (0) when the kernel builds them in LDS
          list_words = longest <= 4 ? 2 : 4;
          lists_in_lds = one_per_colour && !(cfg.flags & FOT_LOOP_LISTS_IN_HBM);
          if (list_words == 2) {
            const size_t n = plist.size() / 8;
            for (size_t i = 0; i < n; ++i)
              for (int j = 0; j < 4; ++j) plist[i * 4 + j] = plist[i * 8 + j];
            plist.resize(n * 4);
          }
          void *pp = nullptr;
          if ((st = alloc(plist.size() * sizeof(uint16_t), &pp)) != FOT_OK) return st;
          FOT_HIP(hipMemcpy(pp, plist.data(), plist.size() * sizeof(uint16_t),
                            hipMemcpyHostToDevice));
          owner_plist = static_cast<const uint32_t *>(pp);
        }
      }
      // each block's mapping-1 points (x) for staging: ascending unique node
      // ids and per-cell local pairs
      std::vector<uint32_t> nodes_all, nloc(plan_cell.size()), tmp;
      stage_x = (cfg.flags & FOT_LOOP_GATHER_NODES) == 0 && !owner_gather;
      for (size_t b = 0; b < meta.size() && stage_x; ++b) {
        OwnerMeta &mt = meta[b];
        tmp.clear();
        for (uint32_t i = mt.cell_off; i < mt.cell_off + mt.ncells; ++i) {
          tmp.push_back(plan_maps[1][2 * size_t(i)]);
          tmp.push_back(plan_maps[1][2 * size_t(i) + 1]);
        }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        if (tmp.size() > 0xffffu) {
          stage_x = false;
          break;
        }
        mt.node_off = uint32_t(nodes_all.size());
        mt.nnodes = uint32_t(tmp.size());
        max_nodes = std::max(max_nodes, mt.nnodes);
        for (uint32_t i = mt.cell_off; i < mt.cell_off + mt.ncells; ++i) {
          const uint32_t a0 = uint32_t(std::lower_bound(tmp.begin(), tmp.end(),
                                                        plan_maps[1][2 * size_t(i)]) - tmp.begin());
          const uint32_t a1 = uint32_t(std::lower_bound(tmp.begin(), tmp.end(),
                                                        plan_maps[1][2 * size_t(i) + 1]) -
                                       tmp.begin());
          nloc[i] = a0 | (a1 << 16);
        }
        nodes_all.insert(nodes_all.end(), tmp.begin(), tmp.end());
      }
      if (stage_x) {
        nodes_all.resize(nodes_all.size() + kOwnerPad, 0u);
        nloc.resize(nloc.size() + kOwnerPad, 0u);
        if ((st = upload(nodes_all.data(), nodes_all.size(), &d32)) != FOT_OK) return st;
        owner_nodes = d32;
        if ((st = upload(nloc.data(), nloc.size(), &d32)) != FOT_OK) return st;
        owner_nloc = d32;
      }
      OwnerMeta *dm = nullptr;
      if ((st = upload(meta.data(), meta.size(), &dm)) != FOT_OK) return st;
      owner_meta = dm;
    }
  }
  plan_args = base;
  for (uint32_t m = 1; m < desc.num_mappings; ++m) {
    if (!((mask >> m) & 1u)) continue;
    if (owner_rc) plan_maps[m].resize(plan_maps[m].size() + 2 * size_t(kOwnerPad), 0u);
    if ((st = upload(plan_maps[m].data(), plan_maps[m].size(), &d32)) != FOT_OK) return st;
    plan_args.maps[m] = d32;
  }
  if ((st = gather_cell_data(plan_cell.data(), uint32_t(plan_cell.size()), plan_args)) != FOT_OK) return st;
  if (two_phase) {
    // Blocks in plan (colour-major) order; a point's partial slots are
    // handed out in that order, so ascending slots are ascending block
    // colours.  Slots of one block are consecutive (its shared points in
    // cache order): its partial stores are contiguous.
    std::vector<uint32_t> touch(desc.num_points[0], 0);
    for (uint32_t i = 0; i < blk_cache_off[nb_total]; ++i) ++touch[cached[i]];
    std::vector<uint32_t> rank(desc.num_points[0], 0xffffffffu), pt, off{0};
    for (uint32_t g = 0; g < desc.num_points[0]; ++g)
      if (touch[g] > 1) {
        rank[g] = uint32_t(pt.size());
        pt.push_back(g);
        off.push_back(off.back() + touch[g]);
      }
    std::vector<uint32_t> wslot(blk_cache_off[nb_total]), slots(off.back()), fill(pt.size(), 0);
    uint32_t next = 0;
    for (uint32_t i = 0; i < blk_cache_off[nb_total]; ++i) {
      const uint32_t r = rank[cached[i]];
      if (r == 0xffffffffu) {
        wslot[i] = kExclusive;
      } else {
        wslot[i] = next;
        slots[off[r] + fill[r]++] = next++;
      }
    }
    merge_n = uint32_t(pt.size());
    if ((st = upload(wslot.data(), wslot.size(), &d32)) != FOT_OK) return st;
    hier.wslot = d32;
    if (merge_n) {
      if ((st = upload(pt.data(), pt.size(), &d32)) != FOT_OK) return st;
      merge_pt = d32;
      if ((st = upload(off.data(), off.size(), &d32)) != FOT_OK) return st;
      merge_off = d32;
      if ((st = upload(slots.data(), slots.size(), &d32)) != FOT_OK) return st;
      merge_slots = d32;
      void *pp = nullptr;
      if ((st = alloc(sizeof(double) * size_t(next) * desc.point_dim[0], &pp)) != FOT_OK)
        return st;
      hier.partial = static_cast<double *>(pp);
    }
  }
  if (cfg.flags & FOT_LOOP_DEPENDENCY_ORDER) {
    // immediate lower-colour neighbour of each block through each point: in
    // colour-major plan order the last block seen on a point is it
    std::vector<uint32_t> last(desc.num_points[0], 0xffffffffu), dep_off{0}, deps, mine;
    for (uint32_t b = 0; b < nb_total; ++b) {
      mine.clear();
      for (uint32_t i = blk_cache_off[b]; i < blk_cache_off[b + 1]; ++i) {
        const uint32_t q = last[cached[i]];
        if (q != 0xffffffffu) mine.push_back(q);
      }
      std::sort(mine.begin(), mine.end());
      mine.erase(std::unique(mine.begin(), mine.end()), mine.end());
      deps.insert(deps.end(), mine.begin(), mine.end());
      dep_off.push_back(uint32_t(deps.size()));
      for (uint32_t i = blk_cache_off[b]; i < blk_cache_off[b + 1]; ++i) last[cached[i]] = b;
    }
    if ((st = upload(dep_off.data(), dep_off.size(), &d32)) != FOT_OK) return st;
    dep.dep_off = d32;
    if (deps.empty()) deps.push_back(0);
    if ((st = upload(deps.data(), deps.size(), &d32)) != FOT_OK) return st;
    dep.deps = d32;
    void *pf = nullptr;
    if ((st = alloc(sizeof(uint32_t) * nb_total + sizeof(DepSync), &pf)) != FOT_OK) return st;
    FOT_HIP(hipMemset(pf, 0, sizeof(uint32_t) * nb_total + sizeof(DepSync)));
    FOT_HIP(hipDeviceSynchronize());
    dep.sync = static_cast<DepSync *>(pf);
    dep.flags = reinterpret_cast<uint32_t *>(static_cast<char *>(pf) + sizeof(DepSync));
    dep.nb = nb_total;
    dep_order = true;
  }
  const uint32_t td = desc.point_dim[0];
  staged = (f.kernel == FOT_KERNEL_RES_CALC) && (cfg.flags & FOT_LOOP_STAGE_INPUTS);
  const uint32_t sdim = staged_input_dim(f);
  stage_pd = !owner && f.kernel != FOT_KERNEL_RES_CALC && (cfg.flags & FOT_LOOP_STAGE_INPUTS) &&
             sdim > 0;
  lds_bytes = uint32_t(sizeof(double)) * max_cached * (staged ? 9u : td + (stage_pd ? sdim : 0u));
  if (stage_pd && lds_bytes > 160u * 1024u) {
    // the staged input does not fit beside the cache: gather it per cell
    // instead (the same values, the same bits)
    stage_pd = false;
    lds_bytes = uint32_t(sizeof(double)) * max_cached * td;
  }
  if (owner) {
    // owned values (+ ids), and for res_calc the staged q / adt of all points
    lds_bytes = 0;
    for (size_t b = 0; b + 1 < blk_cache_off.size(); ++b) {
      const uint32_t m = blk_cache_off[b + 1] - blk_cache_off[b], nw = blk_owned[b];
      const uint32_t bytes = f.kernel == FOT_KERNEL_RES_CALC
                                 ? 8u * (4u * nw + 5u * m) + 4u * nw
                                 : 8u * td * nw + 4u * nw;
      lds_bytes = std::max(lds_bytes, bytes);
    }
    if (f.kernel == FOT_KERNEL_RES_CALC && stage_x) lds_bytes += 16u * max_nodes; // staged x
    if (f.kernel == FOT_KERNEL_RES_CALC && owner_gather) {
      // q / adt (+ 1/rho and pressure) of the staged cells, later the edges'
      // fluxes in the same place
      const uint32_t per_cell = point_terms ? 7u : 5u;
      lds_bytes = 0;
      for (size_t b = 0; b + 1 < blk_cache_off.size(); ++b) {
        const uint32_t m = blk_cache_off[b + 1] - blk_cache_off[b];
        const uint32_t ne = blk_cell_off[b + 1] - blk_cell_off[b];
        // + the edge-list slots (2 bytes per owned cell and thread colour)
        const uint32_t slots = lists_in_lds ? 2u * uint32_t(blk_ncol[b]) * blk_owned[b] : 0u;
        lds_bytes = std::max(lds_bytes, 8u * std::max(per_cell * m, 4u * ne) + slots);
      }
    }
    if (f.kernel == FOT_KERNEL_RES_CALC && point_terms && !owner_gather) {
      uint32_t mm = 0; // + 1/rho and pressure of the largest block's cells
      for (size_t b = 0; b + 1 < blk_cache_off.size(); ++b)
        mm = std::max(mm, blk_cache_off[b + 1] - blk_cache_off[b]);
      lds_bytes += 16u * mm;
    }
  }
  if (lds_bytes > 160u * 1024u)
    return fail(FOT_ERR_UNSUPPORTED, "block cache needs %u B of LDS (max 160 KiB)", lds_bytes);
  uint32_t max_block = 0;
  for (size_t b = 0; b + 1 < blk_cell_off.size(); ++b)
    max_block = std::max(max_block, blk_cell_off[b + 1] - blk_cell_off[b]);
  ept = (owner && f.kernel == FOT_KERNEL_RES_CALC && (cfg.flags & FOT_LOOP_TWO_PER_THREAD)) ? 2 : 1;
  threads = std::max<uint32_t>(64, ((max_block + ept - 1) / ept + 63) / 64 * 64);
  if (threads > 1024) return fail(FOT_ERR_UNSUPPORTED, "blocks above 1024 cells");
  if (owner && f.kernel == FOT_KERNEL_RES_CALC) {
    // the staging rounds cover max(cached, nodes) of every block
    uint32_t most = max_nodes;
    for (size_t b = 0; b + 1 < blk_cache_off.size(); ++b)
      most = std::max(most, blk_cache_off[b + 1] - blk_cache_off[b]);
    // (widening the block so that one round covers every staged point frees
    // registers, 8 waves per SIMD instead of 6, but measured slower: 22.0 vs
    // 21.3 us on the 1200x600 airfoil, 104.7 vs 96.9 us on 2400x1200)
    stage_per_thread = (most + threads - 1) / threads;
    if (stage_per_thread > 4) {
      // widen the block rather than stage more than 4 points per thread
      threads = std::min<uint32_t>(1024, ((most + 3) / 4 + 63) / 64 * 64);
      stage_per_thread = (most + threads - 1) / threads;
      if (stage_per_thread > 4)
        return fail(FOT_ERR_UNSUPPORTED, "owner block stages %u points (max %u)", most,
                    4u * 1024u);
    }
    stage_per_thread = std::max<uint32_t>(1, stage_per_thread);
    uint32_t most_owned = 0;
    for (uint32_t nw : blk_owned) most_owned = std::max(most_owned, nw);
    owned_first_round = most_owned <= threads;
  }
  launch_grid = (dep_order || owner || two_phase) ? nb_total : 0;
  // nontemporal hints when asked, or by default when the step's data (point
  // arrays, output, cell data and the plan) exceed the Infinity Cache:
  // inside it they slow a step down, since every array is read again the
  // next step
  uint64_t step_bytes = stats.plan_device_bytes;
  for (uint32_t k = 0; k < desc.num_mappings; ++k)
    if (desc.point_data[k])
      step_bytes += uint64_t(desc.num_points[k]) * desc.point_dim[k] * sizeof(double);
  step_bytes += uint64_t(desc.num_points[0]) * desc.point_dim[0] * sizeof(double); // out
  for (uint32_t k = 0; k < desc.num_cell_data; ++k)
    step_bytes += uint64_t(desc.num_cells) * desc.cell_dim[k] * sizeof(double);
  const bool want_nt = (cfg.flags & FOT_LOOP_NONTEMPORAL) != 0 ||
                       (!(cfg.flags & FOT_LOOP_CACHED) && step_bytes > kInfinityCacheBytes);
  nontemporal = want_nt && two_phase && !stage_pd && streams_cells(f);
  if (owner && f.kernel == FOT_KERNEL_RES_CALC && owner_gather) {
    const bool persistent = (cfg.flags & FOT_LOOP_PERSISTENT) != 0;
    const uint32_t lw = lists_in_lds ? 0u : list_words;
    nontemporal = want_nt && lw == 0 && !persistent && !point_terms;
    owner3_fn = desc.soa ? owner3_pick_kernel<true>(persistent, point_terms, ept,
                                                    stage_per_thread, lw, owned_first_round,
                                                    nontemporal)
                         : owner3_pick_kernel<false>(persistent, point_terms, ept,
                                                     stage_per_thread, lw, owned_first_round,
                                                     nontemporal);
    if (persistent && nb_total) {
      // as many workgroups as are resident at once, a multiple of the 8 XCDs
      int dev = 0, cus = 0, per_cu = 0;
      FOT_HIP(hipGetDevice(&dev));
      FOT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      FOT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, owner3_fn, int(threads),
                                                           size_t(lds_bytes)));
      if (per_cu < 1) return fail(FOT_ERR_UNSUPPORTED, "persistent owner kernel does not fit a CU");
      uint64_t g = std::min<uint64_t>(nb_total, uint64_t(per_cu) * uint64_t(cus));
      if (g >= kNumXcd) g -= g % kNumXcd;
      launch_grid = uint32_t(g);
    }
  }
  stats.num_colours = num_plan_colours;
  stats.num_blocks = nb_total;
  stats.num_launches_per_iter =
      (dep_order || owner) ? 1u : two_phase ? (merge_n ? 2u : 1u) : stats.num_colours;
  stats.max_cached = max_cached;
  stats.lds_bytes = lds_bytes;
  stats.threads_per_block = threads;
  stats.stage_per_thread = stage_per_thread;
  stats.launch_grid = launch_grid;
  stats.nontemporal = nontemporal ? 1u : 0u;
  stats.max_thread_colours = max_tc;
  stats.avg_thread_colours = nb_total ? tc_sum / nb_total : 0.0;
  // (cached may carry padding at its end: count the plan's entries only)
  stats.reuse_factor = double(blk_cache_off.back()) / (double(md) * desc.num_cells);
  if (owner) stats.redundancy = double(plan_cell.size()) / desc.num_cells;
  double lines = 0;
  for (size_t b = 0; b + 1 < blk_cache_off.size(); ++b) {
    uint64_t l = 0;
    if ((st = fot_count_cache_lines(cached.data() + blk_cache_off[b],
                                    blk_cache_off[b + 1] - blk_cache_off[b], td, 8, 128,
                                    desc.soa, &l)) != FOT_OK)
      return st;
    lines += double(l);
  }
  stats.avg_cache_lines = nb_total ? lines / nb_total : 0.0;
  return FOT_OK;
}

fot_status fot_loop::enqueue(hipStream_t s) {
  if (desc.num_cells == 0) return FOT_OK; // nothing to launch
  const fot_status st = dispatch(&f, desc.soa, [&]<class F, bool SOA>() -> fot_status {
    const uint32_t bs = cfg.block_size;
    // direct launches: one thread per cell, whatever the block size
    constexpr int LB = kSmallBound<F>;
    const uint32_t dbs = std::min<uint32_t>(bs, LB);
    if (cfg.strategy == FOT_STRATEGY_ATOMIC) {
      if constexpr (std::is_same<F, ResCalcF>::value)
        hipLaunchKernelGGL((res_calc_atomic_kernel<SOA>), dim3((desc.num_cells + bs - 1) / bs),
                           dim3(bs), 0, s, base);
      else
        hipLaunchKernelGGL((eng_direct_kernel<F, SOA, true, LB>),
                           dim3((desc.num_cells + dbs - 1) / dbs), dim3(dbs), 0, s, base);
    } else if (cfg.strategy == FOT_STRATEGY_GLOBAL_COLOUR) {
      for (const EngineArgs &a : colour_args) {
        if (a.num_cells == 0) continue;
        hipLaunchKernelGGL((eng_direct_kernel<F, SOA, false, LB>),
                           dim3((a.num_cells + dbs - 1) / dbs), dim3(dbs), 0, s, a);
      }
    } else if (owner) {
      const uint32_t nb = colour_blk_off.back();
      // no cells: nothing to add (and the staging reads assume point 0 exists)
      if (nb == 0 || desc.num_cells == 0) return FOT_OK;
      if constexpr (std::is_same<F, ResCalcF>::value) {
#define FOT_OWNER2P(E_, X_, K_, P_)                                                     \
  hipLaunchKernelGGL((res_calc_owner2_kernel<SOA, E_, X_, K_, P_>), dim3(nb), dim3(threads), \
                     lds_bytes, s, plan_args, owner_meta, hier.cached, hier.packed,          \
                     owner_nodes, owner_nloc, nb)
#define FOT_OWNER2(E_, X_, K_)                                                          \
  do {                                                                                  \
    if (point_terms) FOT_OWNER2P(E_, X_, K_, true);                                     \
    else FOT_OWNER2P(E_, X_, K_, false);                                                \
  } while (0)
#define FOT_OWNER2_K(E_, X_)                                                            \
  switch (stage_per_thread) {                                                           \
  case 1: FOT_OWNER2(E_, X_, 1); break;                                                 \
  case 2: FOT_OWNER2(E_, X_, 2); break;                                                 \
  case 3: FOT_OWNER2(E_, X_, 3); break;                                                 \
  default: FOT_OWNER2(E_, X_, 4); break;                                                \
  }
        if (owner_gather) {
          EngineArgs pa = plan_args;
          const OwnerMeta *om = owner_meta;
          const uint32_t *pc = hier.cached, *pp = hier.packed, *pl = owner_plist;
          uint32_t nbv = nb;
          void *kargs[] = {&pa, &om, &pc, &pp, &pl, &nbv};
          FOT_HIP(hipLaunchKernel(owner3_fn, dim3(launch_grid), dim3(threads), kargs,
                                  size_t(lds_bytes), s));
        } else if (ept == 2) {
          if (stage_x) FOT_OWNER2_K(2, true)
          else FOT_OWNER2_K(2, false)
        } else {
          if (stage_x) FOT_OWNER2_K(1, true)
          else FOT_OWNER2_K(1, false)
        }
#undef FOT_OWNER2_K
#undef FOT_OWNER2
#undef FOT_OWNER2P
      } else {
        bool done = false;
        if constexpr (kSmallBound<F> < 1024) {
          if (threads <= uint32_t(kSmallBound<F>)) {
            hipLaunchKernelGGL((eng_owner_kernel<F, SOA, kSmallBound<F>>), dim3(nb), dim3(threads),
                               lds_bytes, s, plan_args, hier, nb);
            done = true;
          }
        }
        if (!done)
          hipLaunchKernelGGL((eng_owner_kernel<F, SOA>), dim3(nb), dim3(threads), lds_bytes, s,
                             plan_args, hier, nb);
      }
    } else if (dep_order) {
      // one launch over every plan block
      const uint32_t nb = colour_blk_off.back();
      bool done = false;
      if constexpr (std::is_same<F, ResCalcF>::value) {
        if (staged) {
          hipLaunchKernelGGL((res_calc_hier_staged_kernel<SOA, 1>), dim3(nb), dim3(threads),
                             lds_bytes, s, plan_args, hier, dep, 0u, nb);
          done = true;
        }
      }
      if (!done)
        launch_eng_hier<F, SOA, 1>(stage_pd, nb, threads, lds_bytes, s, plan_args, hier, dep,
                                   0u, nb);
    } else if (two_phase) {
      // every block in one launch, then the merge of the shared points
      const uint32_t nb = colour_blk_off.back();
      bool done = false;
      if constexpr (std::is_same<F, ResCalcF>::value) {
        if (staged) {
          hipLaunchKernelGGL((res_calc_hier_staged_kernel<SOA, 2>), dim3(nb), dim3(threads),
                             lds_bytes, s, plan_args, hier, dep, 0u, nb);
          done = true;
        }
      }
      if (!done)
        launch_eng_hier<F, SOA, 2>(stage_pd, nb, threads, lds_bytes, s, plan_args, hier, dep,
                                   0u, nb, nontemporal);
      if (merge_n)
        hipLaunchKernelGGL((merge_partials_kernel<F::TD, SOA>), dim3((merge_n + 255) / 256),
                           dim3(256), 0, s, plan_args.out, plan_args.num_points[0], merge_pt,
                           merge_off, merge_slots, hier.partial, merge_n);
    } else {
      for (size_t c = 0; c + 1 < colour_blk_off.size(); ++c) {
        const uint32_t b0 = colour_blk_off[c], nbc = colour_blk_off[c + 1] - b0;
        if (nbc == 0) continue;
        if constexpr (std::is_same<F, ResCalcF>::value) {
          if (staged) {
            hipLaunchKernelGGL((res_calc_hier_staged_kernel<SOA, 0>), dim3(nbc),
                               dim3(threads), lds_bytes, s, plan_args, hier, dep, b0, nbc);
            continue;
          }
        }
        launch_eng_hier<F, SOA, 0>(stage_pd, nbc, threads, lds_bytes, s, plan_args, hier, dep,
                                   b0, nbc);
      }
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FOT_ERR_HIP, "launch: %s", hipGetErrorString(e));
    return FOT_OK;
  });
  return st;
}

static fot_status fot_loop_create_impl(const fot_functor *f, const fot_problem_desc *desc,
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
  {
    // the point data every functor reads (the kernels do not check pointers)
    const uint32_t reads = f->kernel == FOT_KERNEL_RES_CALC ? 4u
                           : (f->kernel == FOT_KERNEL_LULESH || f->kernel == FOT_KERNEL_MINI_AERO ||
                              f->kernel == FOT_KERNEL_VOLNA)
                               ? 2u
                               : 1u;
    for (uint32_t m = 0; m < reads && m < desc->num_mappings; ++m)
      if (!desc->point_data[m] && desc->num_cells)
        return fail(FOT_ERR_INVALID_ARGUMENT, "point data %u is missing", m);
    for (uint32_t k = 0; k < desc->num_cell_data; ++k)
      if (!desc->cell_data[k] && desc->num_cells)
        return fail(FOT_ERR_INVALID_ARGUMENT, "cell data %u is missing", k);
  }
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
  if (f->kernel == FOT_KERNEL_LULESH &&
      (desc->num_mappings < 2 || desc->point_dim[1] != 3 ||
       desc->num_points[1] != desc->num_points[0] || !desc->point_data[1] ||
       desc->num_cell_data < 2 || desc->cell_dim[0] != 3 || desc->cell_dim[1] != 1 ||
       !desc->cell_data[1]))
    return fail(FOT_ERR_INVALID_ARGUMENT,
                "lulesh needs node coordinates (point data 1, dim 3, the mapping-0 points), "
                "sigma(3) and volume(1) cell data");
  if (f->kernel == FOT_KERNEL_GETACC &&
      (desc->num_cell_data < 5 || desc->cell_dim[0] != 4 || desc->cell_dim[1] != 1 ||
       desc->cell_dim[2] != 4 || desc->cell_dim[3] != 4 || desc->cell_dim[4] != 4))
    return fail(FOT_ERR_INVALID_ARGUMENT,
                "getacc needs cnmass(4), rho(1), cnwt(4), cnfx(4), cnfy(4) cell data");
  if (f->kernel == FOT_KERNEL_MINI_AERO &&
      (desc->num_mappings < 2 || desc->point_dim[1] != 28 ||
       desc->num_points[1] < desc->num_points[0] || desc->num_cell_data < 1 ||
       desc->cell_dim[0] != 12))
    return fail(FOT_ERR_INVALID_ARGUMENT,
                "mini_aero needs flux/cell-value mappings (dims 5, 28) and 12 face values");
  if (f->kernel == FOT_KERNEL_VOLNA &&
      (desc->num_mappings < 2 || desc->point_dim[1] != 1 ||
       desc->num_points[1] < desc->num_points[0] || desc->num_cell_data < 4 ||
       desc->cell_dim[0] != 3 || desc->cell_dim[1] != 2 || desc->cell_dim[2] != 2 ||
       desc->cell_dim[3] != 1))
    return fail(FOT_ERR_INVALID_ARGUMENT,
                "volna needs state/volume mappings and fluxes(3), bathy(2), normals(2), "
                "boundary(1) cell data");
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
  if (cfg->strategy > FOT_STRATEGY_HIERARCHICAL)
    return fail(FOT_ERR_INVALID_ARGUMENT, "unknown strategy %u", cfg->strategy);
  if (desc->num_cells == 0) {
    // a loop over no cells is a no-op: no plan, no launches, no graph
    // (fot_loop_run still copies point data 0 through, as the reference does)
    L->colour_blk_off.assign(1, 0);
    L->threads = 64;
    *out = L.release();
    return FOT_OK;
  }
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
extern "C" fot_status fot_loop_create(const fot_functor *f, const fot_problem_desc *desc,
                                      const fot_loop_config *cfg, fot_loop **out) {
  return fot::guarded("fot_loop_create", [&] { return fot_loop_create_impl(f, desc, cfg, out); });
}

static fot_status fot_loop_destroy_impl(fot_loop *loop) {
  delete loop;
  return FOT_OK;
}
extern "C" fot_status fot_loop_destroy(fot_loop *loop) {
  return fot::guarded("fot_loop_destroy", [&] { return fot_loop_destroy_impl(loop); });
}

static fot_status fot_loop_step_impl(fot_loop *loop, void *stream) {
  if (!loop) return fail(FOT_ERR_INVALID_ARGUMENT, "null loop");
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (loop->exec) {
    FOT_HIP(hipGraphLaunch(loop->exec, s));
    return FOT_OK;
  }
  return loop->enqueue(s);
}
extern "C" fot_status fot_loop_step(fot_loop *loop, void *stream) {
  return fot::guarded("fot_loop_step", [&] { return fot_loop_step_impl(loop, stream); });
}

static fot_status fot_loop_run_impl(fot_loop *loop, uint32_t iters, void *stream) {
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
  if (iters && (st = loop->scatter_cell_outputs(static_cast<hipStream_t>(stream))) != FOT_OK)
    return st;
  FOT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return fot_loop_check(loop);
}
extern "C" fot_status fot_loop_run(fot_loop *loop, uint32_t iters, void *stream) {
  return fot::guarded("fot_loop_run", [&] { return fot_loop_run_impl(loop, iters, stream); });
}

static fot_status fot_loop_check_impl(const fot_loop *loop) {
  if (!loop) return fail(FOT_ERR_INVALID_ARGUMENT, "null loop");
  if (!loop->dep_order) return FOT_OK;
  DepSync host{};
  FOT_HIP(hipMemcpy(&host, loop->dep.sync, sizeof(DepSync), hipMemcpyDeviceToHost));
  if (host.error)
    return fail(FOT_ERR_TIMEOUT, "a block waited too long for its predecessors");
  return FOT_OK;
}
extern "C" fot_status fot_loop_check(const fot_loop *loop) {
  return fot::guarded("fot_loop_check", [&] { return fot_loop_check_impl(loop); });
}

static fot_status fot_loop_stats_get_impl(const fot_loop *loop, fot_loop_stats *stats) {
  if (!loop || !stats) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  *stats = loop->stats;
  return FOT_OK;
}
extern "C" fot_status fot_loop_stats_get(const fot_loop *loop, fot_loop_stats *stats) {
  return fot::guarded("fot_loop_stats_get", [&] { return fot_loop_stats_get_impl(loop, stats); });
}

// ---------------------------------------------------------------------
// Batched loops: many problems of one functor, one persistent launch per
// step (no reference counterpart).
// ---------------------------------------------------------------------
struct fot_batch {
  fot_functor f{};
  int soa = 0;
  std::vector<fot_loop *> loops;
  std::vector<void *> owned;
  BatchProblem *probs = nullptr;
  uint2 *work = nullptr;
  uint32_t nwork = 0;
  uint32_t *queue = nullptr;
  CopyItem *copy_in = nullptr, *copy_back = nullptr; // out <- in, in <- out
  uint32_t ncopy = 0;
  uint32_t grid = 0, threads = 0, lds_bytes = 0;
  bool jacobi = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  ~fot_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (fot_loop *l : loops) delete l;
    for (void *p : owned) (void)hipFree(p);
  }
  template <class T> fot_status upload(const std::vector<T> &v, T **dev) {
    void *p = nullptr;
    FOT_HIP(hipMalloc(&p, std::max<size_t>(1, v.size()) * sizeof(T)));
    owned.push_back(p);
    if (!v.empty()) FOT_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    *dev = static_cast<T *>(p);
    return FOT_OK;
  }
  fot_status launch_step(hipStream_t s) {
    if (nwork == 0) return FOT_OK;
    return dispatch(&f, soa, [&]<class F, bool SOA>() -> fot_status {
      hipLaunchKernelGGL((batch_owner_kernel<F, SOA>), dim3(grid), dim3(threads), lds_bytes, s,
                         probs, work, nwork, queue);
      FOT_HIP(hipGetLastError());
      return FOT_OK;
    });
  }
  fot_status launch_copy(const CopyItem *items, hipStream_t s) {
    if (ncopy == 0) return FOT_OK;
    hipLaunchKernelGGL(batch_copy_kernel, dim3(ncopy), dim3(256), 0, s, items);
    FOT_HIP(hipGetLastError());
    return FOT_OK;
  }
};

static fot_status fot_batch_create_impl(const fot_functor *f, uint32_t num_problems,
                                       const fot_problem_desc *descs,
                                       const fot_loop_config *cfgs, uint32_t use_graph,
                                       fot_batch **out) {
  if (!f || !descs || !cfgs || !out || num_problems == 0)
    return fail(FOT_ERR_INVALID_ARGUMENT, "null argument or empty batch");
  *out = nullptr;
  std::unique_ptr<fot_batch> B(new fot_batch);
  B->f = *f;
  B->soa = descs[0].soa;
  B->jacobi = reads_iterate(*f);
  fot_status st;
  std::vector<BatchProblem> host(num_problems);
  std::vector<CopyItem> cin, cback;
  for (uint32_t i = 0; i < num_problems; ++i) {
    if (descs[i].soa != B->soa)
      return fail(FOT_ERR_INVALID_ARGUMENT, "problem %u: every problem needs the same layout", i);
    fot_loop_config c = cfgs[i];
    c.strategy = FOT_STRATEGY_HIERARCHICAL;
    c.flags = (c.flags & FOT_LOOP_POINT_PARTITION) | FOT_LOOP_OWNER_COMPUTE;
    c.use_graph = 0;
    fot_loop *l = nullptr;
    if ((st = fot_loop_create(f, &descs[i], &c, &l)) != FOT_OK) return st;
    B->loops.push_back(l);
    host[i].args = l->plan_args;
    host[i].hier = l->hier;
    B->threads = std::max(B->threads, l->threads);
    B->lds_bytes = std::max(B->lds_bytes, l->lds_bytes);
    const size_t n = size_t(descs[i].num_points[0]) * descs[i].point_dim[0];
    double *in = const_cast<double *>(descs[i].point_data[0]);
    for (size_t o = 0; o < n; o += kCopyChunk) {
      const uint32_t cnt = uint32_t(std::min<size_t>(kCopyChunk, n - o));
      cin.push_back({descs[i].point_data_out + o, in + o, cnt, 0});
      cback.push_back({in + o, descs[i].point_data_out + o, cnt, 0});
    }
  }
  // work items, largest problems first (longest jobs first balances the tail)
  std::vector<uint32_t> order(num_problems);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return descs[a].num_cells > descs[b].num_cells;
  });
  std::vector<uint2> work;
  for (uint32_t i : order) {
    const uint32_t nb = B->loops[i]->colour_blk_off.back();
    for (uint32_t b = 0; b < nb; ++b) work.push_back(make_uint2(i, b));
  }
  B->nwork = uint32_t(work.size());
  B->ncopy = uint32_t(cin.size());
  if ((st = B->upload(host, &B->probs)) != FOT_OK) return st;
  if ((st = B->upload(work, &B->work)) != FOT_OK) return st;
  if ((st = B->upload(cin, &B->copy_in)) != FOT_OK) return st;
  if ((st = B->upload(cback, &B->copy_back)) != FOT_OK) return st;
  if ((st = B->upload(std::vector<uint32_t>{0u, 0u}, &B->queue)) != FOT_OK) return st;
  if (B->lds_bytes > 160u * 1024u)
    return fail(FOT_ERR_UNSUPPORTED, "batch needs %u B of LDS", B->lds_bytes);
  // persistent grid: as many workgroups as fit on the device at once
  int dev = 0, cus = 0, per_cu = 0;
  FOT_HIP(hipGetDevice(&dev));
  FOT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  st = dispatch(f, B->soa, [&]<class F, bool SOA>() -> fot_status {
    FOT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void *>(&batch_owner_kernel<F, SOA>), int(B->threads),
        size_t(B->lds_bytes)));
    return FOT_OK;
  });
  if (st != FOT_OK) return st;
  B->grid = std::max<uint32_t>(1, std::min<uint32_t>(B->nwork, uint32_t(std::max(1, per_cu) * cus)));
  FOT_HIP(hipDeviceSynchronize());
  if (use_graph && (B->nwork > 0 || B->jacobi)) {
    // one timestep: the batched increments, then in <- out for every problem
    // when the loop reads its own iterate (else the copy runs once, after the
    // last step, as in the eager run).  A step with nothing to launch is not
    // captured: the eager run does nothing for it either.
    hipStream_t cs = nullptr;
    FOT_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      st = B->launch_step(cs);
      if (st == FOT_OK && B->jacobi) st = B->launch_copy(B->copy_back, cs);
      e = hipStreamEndCapture(cs, &B->graph);
    }
    if (e == hipSuccess && st == FOT_OK) e = hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (st != FOT_OK) return st;
    if (e != hipSuccess) return fail(FOT_ERR_HIP, "batch graph: %s", hipGetErrorString(e));
  }
  *out = B.release();
  return FOT_OK;
}
extern "C" fot_status fot_batch_create(const fot_functor *f, uint32_t num_problems,
                                       const fot_problem_desc *descs,
                                       const fot_loop_config *cfgs, uint32_t use_graph,
                                       fot_batch **out) {
  return fot::guarded("fot_batch_create", [&] {
    return fot_batch_create_impl(f, num_problems, descs, cfgs, use_graph, out);
  });
}

static fot_status fot_batch_destroy_impl(fot_batch *b) {
  delete b;
  return FOT_OK;
}
extern "C" fot_status fot_batch_destroy(fot_batch *b) {
  return fot::guarded("fot_batch_destroy", [&] { return fot_batch_destroy_impl(b); });
}

static fot_status fot_batch_step_impl(fot_batch *b, void *stream) {
  if (!b) return fail(FOT_ERR_INVALID_ARGUMENT, "null batch");
  return b->launch_step(static_cast<hipStream_t>(stream));
}
extern "C" fot_status fot_batch_step(fot_batch *b, void *stream) {
  return fot::guarded("fot_batch_step", [&] { return fot_batch_step_impl(b, stream); });
}

static fot_status fot_batch_run_impl(fot_batch *b, uint32_t iters, void *stream) {
  if (!b) return fail(FOT_ERR_INVALID_ARGUMENT, "null batch");
  hipStream_t s = static_cast<hipStream_t>(stream);
  fot_status st = b->launch_copy(b->copy_in, s);
  for (uint32_t it = 0; it < iters && st == FOT_OK; ++it) {
    if (b->exec) {
      FOT_HIP(hipGraphLaunch(b->exec, s)); // copies back itself when jacobi
    } else {
      st = b->launch_step(s);
      if (st == FOT_OK && b->jacobi) st = b->launch_copy(b->copy_back, s);
    }
  }
  if (st == FOT_OK && iters && !b->jacobi) st = b->launch_copy(b->copy_back, s);
  for (fot_loop *l : b->loops)
    if (st == FOT_OK && iters) st = l->scatter_cell_outputs(s);
  if (st != FOT_OK) return st;
  FOT_HIP(hipStreamSynchronize(s));
  return FOT_OK;
}
extern "C" fot_status fot_batch_run(fot_batch *b, uint32_t iters, void *stream) {
  return fot::guarded("fot_batch_run", [&] { return fot_batch_run_impl(b, iters, stream); });
}

static fot_status fot_batch_stats_get_impl(const fot_batch *b, fot_batch_stats *s) {
  if (!b || !s) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  s->num_problems = uint32_t(b->loops.size());
  s->num_work_items = b->nwork;
  s->grid = b->grid;
  s->threads = b->threads;
  s->lds_bytes = b->lds_bytes;
  s->uses_graph = b->exec ? 1u : 0u;
  return FOT_OK;
}
extern "C" fot_status fot_batch_stats_get(const fot_batch *b, fot_batch_stats *s) {
  return fot::guarded("fot_batch_stats_get", [&] { return fot_batch_stats_get_impl(b, s); });
}

// ---------------------------------------------------------------------
// Context layer: plans, one-shot loops, graphs and batches on a context's
// stream (the context and its buffers are in fot_runtime.hip).
// ---------------------------------------------------------------------
namespace {
fot_status ctx_stream(fot_ctx *ctx, hipStream_t *s) {
  void *p = nullptr;
  const fot_status st = fot_ctx_stream(ctx, &p);
  *s = static_cast<hipStream_t>(p);
  return st;
}
} // namespace

static fot_status fot_plan_upload_impl(fot_ctx *ctx, const fot_functor *f,
                                      const fot_problem_desc *desc, const fot_loop_config *cfg,
                                      fot_plan **plan) {
  if (!ctx || !plan) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  *plan = nullptr;
  return fot_loop_create(f, desc, cfg, plan);
}
extern "C" fot_status fot_plan_upload(fot_ctx *ctx, const fot_functor *f,
                                      const fot_problem_desc *desc, const fot_loop_config *cfg,
                                      fot_plan **plan) {
  return fot::guarded("fot_plan_upload", [&] {
    return fot_plan_upload_impl(ctx, f, desc, cfg, plan);
  });
}

static fot_status fot_plan_free_impl(fot_plan *plan) { return fot_loop_destroy(plan); }
extern "C" fot_status fot_plan_free(fot_plan *plan) {
  return fot::guarded("fot_plan_free", [&] { return fot_plan_free_impl(plan); });
}

static fot_status fot_loop_launch_impl(fot_ctx *ctx, const fot_functor *f,
                                      const fot_problem_desc *desc, const fot_loop_config *cfg,
                                      uint32_t iters) {
  hipStream_t s = nullptr;
  fot_status st = ctx ? ctx_stream(ctx, &s) : fail(FOT_ERR_INVALID_ARGUMENT, "null context");
  if (st != FOT_OK) return st;
  fot_loop *loop = nullptr;
  if ((st = fot_loop_create(f, desc, cfg, &loop)) != FOT_OK) return st;
  st = fot_loop_run(loop, iters, s); // waits for the stream
  fot_loop_destroy(loop);
  return st;
}
extern "C" fot_status fot_loop_launch(fot_ctx *ctx, const fot_functor *f,
                                      const fot_problem_desc *desc, const fot_loop_config *cfg,
                                      uint32_t iters) {
  return fot::guarded("fot_loop_launch", [&] {
    return fot_loop_launch_impl(ctx, f, desc, cfg, iters);
  });
}

struct fot_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~fot_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

static fot_status fot_graph_capture_impl(fot_ctx *ctx, fot_plan *plan, fot_graph **out) {
  if (!ctx || !plan || !out) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  std::unique_ptr<fot_graph> G(new fot_graph);
  if (plan->desc.num_cells > 0) {
    // the plan's own enqueue (not a replay of a graph it may hold), captured
    // on a private stream as fot_loop_create does for use_graph
    hipStream_t cs = nullptr;
    FOT_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    fot_status st = FOT_OK;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      st = plan->enqueue(cs);
      e = hipStreamEndCapture(cs, &G->graph);
    }
    if (e == hipSuccess && st == FOT_OK) e = hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (st != FOT_OK) return st;
    if (e != hipSuccess) return fail(FOT_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
  }
  *out = G.release();
  return FOT_OK;
}
extern "C" fot_status fot_graph_capture(fot_ctx *ctx, fot_plan *plan, fot_graph **out) {
  return fot::guarded("fot_graph_capture", [&] { return fot_graph_capture_impl(ctx, plan, out); });
}

static fot_status fot_graph_launch_impl(fot_ctx *ctx, fot_graph *graph) {
  if (!ctx || !graph) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  if (!graph->exec) return FOT_OK; // a plan over no cells
  hipStream_t s = nullptr;
  const fot_status st = ctx_stream(ctx, &s);
  if (st != FOT_OK) return st;
  FOT_HIP(hipGraphLaunch(graph->exec, s));
  return FOT_OK;
}
extern "C" fot_status fot_graph_launch(fot_ctx *ctx, fot_graph *graph) {
  return fot::guarded("fot_graph_launch", [&] { return fot_graph_launch_impl(ctx, graph); });
}

static fot_status fot_graph_free_impl(fot_graph *graph) {
  delete graph;
  return FOT_OK;
}
extern "C" fot_status fot_graph_free(fot_graph *graph) {
  return fot::guarded("fot_graph_free", [&] { return fot_graph_free_impl(graph); });
}

static fot_status fot_batch_launch_impl(fot_ctx *ctx, fot_batch *batch, uint32_t iters) {
  if (!ctx || !batch) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  hipStream_t s = nullptr;
  const fot_status st = ctx_stream(ctx, &s);
  if (st != FOT_OK) return st;
  return fot_batch_run(batch, iters, s);
}
extern "C" fot_status fot_batch_launch(fot_ctx *ctx, fot_batch *batch, uint32_t iters) {
  return fot::guarded("fot_batch_launch", [&] { return fot_batch_launch_impl(ctx, batch, iters); });
}

// ---------------------------------------------------------------------
// Autotuner measurement: each candidate config planned and timed on a
// private stream with HIP events (no reference counterpart).
// ---------------------------------------------------------------------
static fot_status fot_autotune_measure_impl(const fot_functor *f, const fot_problem_desc *desc,
                                           const fot_loop_config *cands, uint32_t n,
                                           uint32_t warmup, uint32_t steps, float *times_ms) {
  if (!f || !desc || !cands || !times_ms || n == 0 || steps == 0)
    return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  FOT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  std::unique_ptr<std::remove_pointer<hipStream_t>::type, hipError_t (*)(hipStream_t)> sg(
      s, hipStreamDestroy);
  using EventGuard = std::unique_ptr<std::remove_pointer<hipEvent_t>::type,
                                     hipError_t (*)(hipEvent_t)>;
  FOT_HIP(hipEventCreate(&e0));
  EventGuard g0(e0, hipEventDestroy);
  FOT_HIP(hipEventCreate(&e1));
  EventGuard g1(e1, hipEventDestroy);
  fot_status result = FOT_OK;
  for (uint32_t i = 0; i < n && result == FOT_OK; ++i) {
    times_ms[i] = -1.f;
    fot_loop_config c = cands[i];
    c.use_graph = 0;
    fot_loop *l = nullptr;
    const fot_status st = fot_loop_create(f, desc, &c, &l);
    if (st == FOT_ERR_COLOUR_OVERFLOW || st == FOT_ERR_UNSUPPORTED) continue; // not plannable
    if (st != FOT_OK) {
      result = st;
      break;
    }
    std::unique_ptr<fot_loop> lg(l);
    for (uint32_t w = 0; w < warmup && result == FOT_OK; ++w) result = l->enqueue(s);
    hipError_t he = hipEventRecord(e0, s);
    for (uint32_t k = 0; k < steps && result == FOT_OK && he == hipSuccess; ++k)
      result = l->enqueue(s);
    if (he == hipSuccess) he = hipEventRecord(e1, s);
    if (he == hipSuccess) he = hipEventSynchronize(e1);
    float ms = 0.f;
    if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
    if (he != hipSuccess) result = fail(FOT_ERR_HIP, "autotune timing: %s", hipGetErrorString(he));
    if (result == FOT_OK) result = fot_loop_check(l);
    times_ms[i] = ms / float(steps);
  }
  return result;
}
extern "C" fot_status fot_autotune_measure(const fot_functor *f, const fot_problem_desc *desc,
                                           const fot_loop_config *cands, uint32_t n,
                                           uint32_t warmup, uint32_t steps, float *times_ms) {
  return fot::guarded("fot_autotune_measure", [&] {
    return fot_autotune_measure_impl(f, desc, cands, n, warmup, steps, times_ms);
  });
}
