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
#include <type_traits>

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

// The per-point terms of res_calc_user (1/rho and the pressure), with the
// same operations, so a kernel can compute them once per cell of its block.
FOT_DEV void res_calc_point_terms(const double *q, uint32_t qs, double *ri, double *p) {
  const double gm1 = 0.4;
  const double r = 1.0 / q[0];
  *ri = r;
  *p = gm1 * (q[3 * qs] - 0.5 * r * (q[qs] * q[qs] + q[2 * qs] * q[2 * qs]));
}

// res_calc_user with the per-point terms given: the same operations on the
// same values, so the same bits.
FOT_DEV void res_calc_user_terms(const double *__restrict__ x1, const double *__restrict__ x2,
                                 const double *__restrict__ q1, const double *__restrict__ q2,
                                 double ri1, double ri2, double p1, double p2, double adt1,
                                 double adt2, double *__restrict__ res1,
                                 double *__restrict__ res2) {
  const double eps = 0.05;
  const double dx = x1[0] - x2[0];
  const double dy = x1[1] - x2[1];
  const double vol1 = ri1 * (q1[1] * dy - q1[2] * dx);
  const double vol2 = ri2 * (q2[1] * dy - q2[2] * dx);
  const double mu = 0.5 * (adt1 + adt2) * eps;
  double f = 0.5 * (vol1 * q1[0] + vol2 * q2[0]) + mu * (q1[0] - q2[0]);
  res1[0] = f;
  res2[0] = -f;
  f = 0.5 * (vol1 * q1[1] + p1 * dy + vol2 * q2[1] + p2 * dy) + mu * (q1[1] - q2[1]);
  res1[1] = f;
  res2[1] = -f;
  f = 0.5 * (vol1 * q1[2] - p1 * dx + vol2 * q2[2] - p2 * dx) + mu * (q1[2] - q2[2]);
  res1[2] = f;
  res2[2] = -f;
  f = 0.5 * (vol1 * (q1[3] + p1) + vol2 * (q2[3] + p2)) + mu * (q1[3] - q2[3]);
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
  static constexpr bool kStreamCells = false;
};

// Arguments of the engine's launches: everything by value in the kernel
// argument segment (no pointer chasing).  kStreamCells: the functor loads
// its read-once cell data with nontemporal hints (StreamCellArgs).
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
  static constexpr bool kStreamCells = false;
};

// The same arguments, with the functors that stream their cell data (the
// hex corner forces, mini_aero's face values, lulesh's stresses: read once
// per step) loading it
// with nontemporal hints, so that on data beyond the Infinity Cache it does
// not evict the point data and outputs other blocks touch again.  The same
// values: the same bits.
struct StreamCellArgs : EngineArgs {
  static constexpr bool kStreamCells = true;
};

// One value of read-once data: a nontemporal load when NT.
template <bool NT> __host__ __device__ __forceinline__ double load_once(const double *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

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
  // sequential cache adds: batching them costs this light kernel occupancy
  // (69 -> 80..110 VGPRs, 7 -> 6..4 waves per SIMD), measured slower on
  // GPS-ordered blocks and within 2 % (chunks of 4 and 8) on compact
  // bisection blocks with 9 thread colours
  static constexpr uint32_t ADD_CHUNK = 1;
  template <bool NT = false>
  FOT_DEV static void user(const double *F, uint32_t cs, double *inc) {
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j)
#pragma unroll
      for (uint32_t i = 0; i < 3; ++i) inc[j * 3 + i] = load_once<NT>(&F[size_t(i * 8 + j) * cs]);
  }
  template <bool SOA, class A>
  FOT_DEV static void compute(const A &a, uint32_t ind, const uint32_t *, double *inc) {
    user<A::kStreamCells>(a.cd(0) + ind, a.cs(), inc);
  }
};

// ---- mini_aero face flux (kernels/mini_aero_func.hpp:11-85 with
// kernels/mini_aero_utils.hpp) ------------------------------------------------
// Cell state reconstructed to the face with limited gradients, a Roe flux
// with the entropy fix, minus the viscous flux of the face-averaged state.
// Every operation is in the reference's order and association (the constants
// (gamma - 1), 0.5 * a * b, the zero-initialised accumulators included), so
// host and device give the same bits.  Host + device: libfot's CPU entry
// point (fot_cpu_cell) runs the same code.
#define FOT_HD __host__ __device__ __forceinline__
namespace ma {

FOT_HD void to_primitive(const double *U, uint32_t s, double *V) {
  const double gamma = 1.4, rgas = 287.05;
  const double r = U[0];
  const double ri = 1.0 / r;
  const double u = U[s] * ri, v = U[2 * s] * ri, w = U[3 * s] * ri;
  const double k = 0.5 * (u * u + v * v + w * w);
  const double e = U[4 * s] * ri - k;
  V[0] = r, V[1] = u, V[2] = v, V[3] = w;
  V[4] = e * (gamma - 1.0) / rgas;
}

// y = A x for a 5x5 row-major A, accumulated from zero
FOT_HD void mat5(const double *A, const double *x, double *y) {
  for (int i = 0; i < 5; ++i) {
    double acc = 0.0 * 0.0;
    for (int j = 0; j < 5; ++j) acc += 1.0 * A[5 * i + j] * x[j];
    y[i] = acc;
  }
}

FOT_HD void roe_flux(const double *pl, const double *pr, const double *n, const double *t,
                     const double *b, double *flux) {
  const double efix_u = 0.1, efix_c = 0.1, gm1 = 0.4;
  const double rl = pl[0], ul = pl[1], vl = pl[2], wl = pl[3];
  const double rr = pr[0], ur = pr[1], vr = pr[2], wr = pr[3];
  const double pres_l = rl * 287.05 * pl[4], pres_r = rr * 287.05 * pr[4];
  const double h_l = 1004.0 * pl[4], h_r = 1004.0 * pr[4];
  const double H_l = h_l + 0.5 * (ul * ul + vl * vl + wl * wl);
  const double H_r = h_r + 0.5 * (ur * ur + vr * vr + wr * wr);
  const double mf_l = rl * (n[0] * ul + n[1] * vl + n[2] * wl);
  const double mf_r = rr * (n[0] * ur + n[1] * vr + n[2] * wr);
  const double psum = pres_l + pres_r;
  flux[0] = 0.5 * (mf_l + mf_r);
  flux[1] = 0.5 * (mf_l * ul + mf_r * ur + n[0] * psum);
  flux[2] = 0.5 * (mf_l * vl + mf_r * vr + n[1] * psum);
  flux[3] = 0.5 * (mf_l * wl + mf_r * wr + n[2] * psum);
  flux[4] = 0.5 * (mf_l * H_l + mf_r * H_r);
  const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  const double tn = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  const double bn = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  const double nu[3] = {n[0] / nn, n[1] / nn, n[2] / nn};
  const double tu[3] = {t[0] / tn, t[1] / tn, t[2] / tn};
  const double bu[3] = {b[0] / bn, b[1] / bn, b[2] / bn};
  const double denom = 1.0 / (sqrt(rl) + sqrt(rr));
  const double alpha = sqrt(rl) * denom;
  const double beta = 1.0 - alpha;
  const double u = alpha * ul + beta * ur, v = alpha * vl + beta * vr, w = alpha * wl + beta * wr;
  const double h = alpha * h_l + beta * h_r +
                   0.5 * alpha * beta *
                       ((ur - ul) * (ur - ul) + (vr - vl) * (vr - vl) + (wr - wl) * (wr - wl));
  const double c = sqrt(gm1 * h);
  const double vn = u * nu[0] + v * nu[1] + w * nu[2];
  const double vt = u * tu[0] + v * tu[1] + w * tu[2];
  const double vb = u * bu[0] + v * bu[1] + w * bu[2];
  const double ke = 0.5 * (u * u + v * v + w * w);
  const double c2i = 1.0 / (c * c);
  const double hc2i = 0.5 * c2i;
  double dq[5];
  dq[0] = rr - rl;
  dq[1] = rr * ur - rl * ul;
  dq[2] = rr * vr - rl * vl;
  dq[3] = rr * wr - rl * wl;
  dq[4] = (rr * H_r - pres_r) - (rl * H_l - pres_l);
  const double cbar = c * nn;
  const double ubar = u * n[0] + v * n[1] + w * n[2];
  const double cfl = fabs(ubar) + cbar;
  const double eig1 = ubar + cbar, eig2 = ubar - cbar, eig3 = ubar;
  double a1 = fabs(eig1), a2 = fabs(eig2), a3 = fabs(eig3);
  const double epuc = efix_u * cfl, epcc = efix_c * cfl;
  if (a1 < epcc) a1 = 0.5 * (eig1 * eig1 + epcc * epcc) / epcc;
  if (a2 < epcc) a2 = 0.5 * (eig2 * eig2 + epcc * epcc) / epcc;
  if (a3 < epuc) a3 = 0.5 * (eig3 * eig3 + epuc * epuc) / epuc;
  const double ep3 = 0.5 * (eig3 + a3), em3 = 0.5 * (eig3 - a3);
  const double eigp[5] = {0.5 * (eig1 + a1), 0.5 * (eig2 + a2), ep3, ep3, ep3};
  const double eigm[5] = {0.5 * (eig1 - a1), 0.5 * (eig2 - a2), em3, em3, em3};
  const double kh = ke - h;
  const double L[25] = {gm1 * kh + c * (c - vn), c * nu[0] - gm1 * u, c * nu[1] - gm1 * v,
                        c * nu[2] - gm1 * w, gm1,
                        gm1 * kh + c * (c + vn), -c * nu[0] - gm1 * u, -c * nu[1] - gm1 * v,
                        -c * nu[2] - gm1 * w, gm1,
                        kh, -u, -v, -w, 1.0,
                        -vt, tu[0], tu[1], tu[2], 0.0,
                        -vb, bu[0], bu[1], bu[2], 0.0};
  double ldq[5];
  mat5(L, dq, ldq);
  for (int j = 0; j < 5; ++j) ldq[j] = (eigp[j] - eigm[j]) * ldq[j];
  const double he = h + ke;
  const double R[25] = {hc2i, hc2i, -gm1 * c2i, 0.0, 0.0,
                        (u + nu[0] * c) * hc2i, (u - nu[0] * c) * hc2i, -gm1 * u * c2i, tu[0],
                        bu[0],
                        (v + nu[1] * c) * hc2i, (v - nu[1] * c) * hc2i, -gm1 * v * c2i, tu[1],
                        bu[1],
                        (w + nu[2] * c) * hc2i, (w - nu[2] * c) * hc2i, -gm1 * w * c2i, tu[2],
                        bu[2],
                        (he + c * vn) * hc2i, (he - c * vn) * hc2i, (c * c - gm1 * he) * c2i, vt,
                        vb};
  double rl_dq[5];
  mat5(R, ldq, rl_dq);
  for (int i = 0; i < 5; ++i) flux[i] -= 0.5 * rl_dq[i];
}

FOT_HD void viscous_flux(const double g[5][3], const double *prim, const double *a,
                         double *vf) {
  const double T = prim[4];
  const double mu = 1.458e-6 * T * sqrt(T) / (T + 110.4);
  const double kc = mu * 1006.0 / 0.71;
  for (int i = 0; i < 5; ++i) vf[i] = 0.0;
  double div = 0.0;
  for (int i = 0; i < 3; ++i) div += g[i + 1][i];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      const double delta = (i == j) ? 1.0 : 0.0;
      const double S = 0.5 * (g[i + 1][j] + g[j + 1][i]);
      const double tij = S - div * delta / 3.;
      vf[1 + i] += (2 * mu * tij) * a[j];
      vf[4] += (2 * mu * tij) * prim[i + 1] * a[j];
    }
    vf[4] += kc * g[4][i] * a[i];
  }
}

// flux of face `f`: cell rows L / R of 28 values (5 conservatives, 3
// coordinates, 15 gradients, 5 limiters) with element stride s; face row
// `fd` of 12 values (coordinates, normal, tangent, binormal), stride fs.
// NT: the face values loaded with nontemporal hints (the same values).
template <bool NT = false>
FOT_HD void face_flux(const double *L, const double *R, uint32_t s, const double *fd,
                      uint32_t fs, double *flux) {
  double cl[5], cr[5], pl[5], pr[5];
  for (int k = 0; k < 5; ++k) cl[k] = L[k * s], cr[k] = R[k * s];
  to_primitive(cl, 1, pl);
  to_primitive(cr, 1, pr);
  for (int k = 0; k < 5; ++k) {
    double gl = 0, gr = 0;
    for (int d = 0; d < 3; ++d) {
      gl += (load_once<NT>(&fd[d * fs]) - L[(5 + d) * s]) * L[(8 + k * 3 + d) * s];
      gr += (load_once<NT>(&fd[d * fs]) - R[(5 + d) * s]) * R[(8 + k * 3 + d) * s];
    }
    pl[k] += gl * L[(23 + k) * s];
    pr[k] += gr * R[(23 + k) * s];
  }
  double n[3], t[3], b[3];
  for (int d = 0; d < 3; ++d)
    n[d] = load_once<NT>(&fd[(3 + d) * fs]), t[d] = load_once<NT>(&fd[(6 + d) * fs]),
    b[d] = load_once<NT>(&fd[(9 + d) * fs]);
  // the viscous flux (independent of the Roe flux) first, so that the
  // gradients are dead before the Roe flux: the same operations, fewer
  // values live at once
  double pf[5], gf[5][3];
  for (int k = 0; k < 5; ++k) {
    pf[k] = 0.5 * (pl[k] + pr[k]);
    for (int d = 0; d < 3; ++d) gf[k][d] = 0.5 * (L[(8 + k * 3 + d) * s] + R[(8 + k * 3 + d) * s]);
  }
  double vf[5];
  viscous_flux(gf, pf, n, vf);
  roe_flux(pl, pr, n, t, b, flux);
  for (int k = 0; k < 5; ++k) flux[k] -= vf[k];
}

} // namespace ma

// mini_aero (kernels/mini_aero.hpp:43-117): point data 0 the cell flux (dim
// 5, target), point data 1 the 28 cell values (through mapping-0 ids as the
// reference reads them), cell data 0 the 12 face values.  left -= flux,
// right += flux: the increments are -flux and +flux (x - f == x + (-f)).
struct MiniAeroF {
  static constexpr uint32_t MD = 2, TD = 5;
  static constexpr bool SLOT_OUTER = false;
  template <bool SOA, class A>
  FOT_DEV static void compute(const A &a, uint32_t ind, const uint32_t *tgt, double *inc) {
    const double *cv = a.pd(1);
    const uint32_t s = SOA ? a.ps(1) : 1;
    const double *L = cv + (SOA ? size_t(tgt[0]) : size_t(28) * tgt[0]);
    const double *R = cv + (SOA ? size_t(tgt[1]) : size_t(28) * tgt[1]);
    double flux[5];
    ma::face_flux<A::kStreamCells>(L, R, s, a.cd(0) + ind, a.cs(), flux);
#pragma unroll
    for (int k = 0; k < 5; ++k) inc[k] = -flux[k], inc[5 + k] = flux[k];
  }
};

// volna (kernels/volna.hpp:27-78, volna_func.hpp:16-36) in FP64: the cell
// volumes are point data 1, read through mapping-0 ids as the reference
// does; cell data 0..3 = edge fluxes (3), bathymetry source (2), edge
// normals (2), right-boundary flag (0.0 / 1.0).  The increments start at 0
// and are updated with the reference's `-=` / `+=`, so even signed zeros
// match.  The fourth state component gets no increment.
// ---- lulesh element forces (kernels/lulesh_func.hpp:13-38 with
// kernels/lulesh_utils.hpp) ----------------------------------------------------
// The element volume from the shape-function derivatives at the centre (of
// which the reference keeps only the volume: its derivative array is then
// overwritten with the node normals), the node normals (each face adds a
// quarter of the cross product of its bisectors to its four nodes, faces and
// nodes in the reference's order) and f = -(sigma * normal).  The
// reference's association throughout, so host, device and the oracle agree
// bit for bit.  f[i * 8 + j]: component i of corner j.
namespace lu {

template <int A, int B, int C, int D>
FOT_HD void face_normal(const double *x, const double *y, const double *z, double *nx,
                        double *ny, double *nz) {
  const double b0x = 0.5 * (x[D] + x[C] - x[B] - x[A]);
  const double b0y = 0.5 * (y[D] + y[C] - y[B] - y[A]);
  const double b0z = 0.5 * (z[D] + z[C] - z[B] - z[A]);
  const double b1x = 0.5 * (x[C] + x[B] - x[D] - x[A]);
  const double b1y = 0.5 * (y[C] + y[B] - y[D] - y[A]);
  const double b1z = 0.5 * (z[C] + z[B] - z[D] - z[A]);
  const double ax = 0.25 * (b0y * b1z - b0z * b1y);
  const double ay = 0.25 * (b0z * b1x - b0x * b1z);
  const double az = 0.25 * (b0x * b1y - b0y * b1x);
  nx[A] += ax, nx[B] += ax, nx[C] += ax, nx[D] += ax;
  ny[A] += ay, ny[B] += ay, ny[C] += ay, ny[D] += ay;
  nz[A] += az, nz[B] += az, nz[C] += az, nz[D] += az;
}

FOT_HD void element(const double *x, const double *y, const double *z, double sxx, double syy,
                    double szz, double *f, double *determ) {
  const double fjxxi = .125 * ((x[6] - x[0]) + (x[5] - x[3]) - (x[7] - x[1]) - (x[4] - x[2]));
  const double fjxet = .125 * ((x[6] - x[0]) - (x[5] - x[3]) + (x[7] - x[1]) - (x[4] - x[2]));
  const double fjxze = .125 * ((x[6] - x[0]) + (x[5] - x[3]) + (x[7] - x[1]) + (x[4] - x[2]));
  const double fjyxi = .125 * ((y[6] - y[0]) + (y[5] - y[3]) - (y[7] - y[1]) - (y[4] - y[2]));
  const double fjyet = .125 * ((y[6] - y[0]) - (y[5] - y[3]) + (y[7] - y[1]) - (y[4] - y[2]));
  const double fjyze = .125 * ((y[6] - y[0]) + (y[5] - y[3]) + (y[7] - y[1]) + (y[4] - y[2]));
  const double fjzxi = .125 * ((z[6] - z[0]) + (z[5] - z[3]) - (z[7] - z[1]) - (z[4] - z[2]));
  const double fjzet = .125 * ((z[6] - z[0]) - (z[5] - z[3]) + (z[7] - z[1]) - (z[4] - z[2]));
  const double fjzze = .125 * ((z[6] - z[0]) + (z[5] - z[3]) + (z[7] - z[1]) + (z[4] - z[2]));
  const double cjxet = -(fjyxi * fjzze) + (fjzxi * fjyze);
  const double cjyet = (fjxxi * fjzze) - (fjzxi * fjxze);
  const double cjzet = -(fjxxi * fjyze) + (fjyxi * fjxze);
  *determ = 8. * (fjxet * cjxet + fjyet * cjyet + fjzet * cjzet);
  double nx[8], ny[8], nz[8];
  for (int i = 0; i < 8; ++i) nx[i] = ny[i] = nz[i] = 0.0;
  face_normal<0, 1, 2, 3>(x, y, z, nx, ny, nz);
  face_normal<0, 4, 5, 1>(x, y, z, nx, ny, nz);
  face_normal<1, 5, 6, 2>(x, y, z, nx, ny, nz);
  face_normal<2, 6, 7, 3>(x, y, z, nx, ny, nz);
  face_normal<3, 7, 4, 0>(x, y, z, nx, ny, nz);
  face_normal<4, 7, 6, 5>(x, y, z, nx, ny, nz);
  for (int i = 0; i < 8; ++i) {
    f[0 * 8 + i] = -(sxx * nx[i]);
    f[1 * 8 + i] = -(syy * ny[i]);
    f[2 * 8 + i] = -(szz * nz[i]);
  }
}

} // namespace lu

// lulesh::StepSeq / StepGPU* (kernels/lulesh.hpp:23-313): xyz (point data 1)
// and the force target (point data 0) through mapping 0; cell data 0 the
// diagonal stresses, cell data 1 the element volume, which the kernel
// writes (into the loop's copy of the cell data: see cell_output_slot).
struct LuleshF {
  static constexpr uint32_t MD = 8, TD = 3;
  static constexpr bool SLOT_OUTER = true;
  template <bool SOA, class A>
  FOT_DEV static void compute(const A &a, uint32_t ind, const uint32_t *tgt, double *inc) {
    const double *xyz = a.pd(1);
    const uint32_t xs = SOA ? a.ps(1) : 1, ux = SOA ? 1 : 3;
    double x[8], y[8], z[8];
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
      const double *c = xyz + size_t(ux) * tgt[k];
      x[k] = c[0], y[k] = c[xs], z[k] = c[2 * size_t(xs)];
    }
    const double *sig = a.cd(0) + ind;
    const uint32_t cs = a.cs();
    double f[24], det;
    constexpr bool NT = A::kStreamCells;
    lu::element(x, y, z, load_once<NT>(sig), load_once<NT>(&sig[cs]),
                load_once<NT>(&sig[2 * size_t(cs)]), f, &det);
    const_cast<double *>(a.cd(1))[ind] = det;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j)
#pragma unroll
      for (uint32_t i = 0; i < 3; ++i) inc[j * 3 + i] = f[i * 8 + j];
  }
};

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

// getacc_scatter (kernels/getacc_scatter.hpp:27-78, getacc_scatter_func.hpp:
// 17-94): corner jj of a quad adds its mass (cnmass[jj], else the previous
// corner's, else rho * cnwt[jj], with the 1e-40 cut), cnwt, cnfx and cnfy.
// Increments start at zero and are added with +=, as the reference does.
struct GetaccF {
  static constexpr uint32_t MD = 4, TD = 4;
  static constexpr bool SLOT_OUTER = false;
  FOT_DEV static void user(const double *cnmass, const double *rho, const double *cnwt,
                           const double *cnfx, const double *cnfy, uint32_t cs, double *inc) {
    const double zerocut = 1.0e-40;
#pragma unroll
    for (uint32_t jj = 0; jj < 4; ++jj) {
      const uint32_t ii = (jj + 3) & 3u;
      double m = 0.0;
      if (cnmass[jj * cs] > zerocut) m += cnmass[jj * cs];
      else if (cnmass[ii * cs] > zerocut) m += cnmass[ii * cs];
      else m += *rho * cnwt[jj * cs];
      inc[4 * jj + 0] = m;
      inc[4 * jj + 1] = 0.0 + cnwt[jj * cs];
      inc[4 * jj + 2] = 0.0 + cnfx[jj * cs];
      inc[4 * jj + 3] = 0.0 + cnfy[jj * cs];
    }
  }
  template <bool SOA, class A>
  FOT_DEV static void compute(const A &a, uint32_t ind, const uint32_t *, double *inc) {
    user(a.cd(0) + ind, a.cd(1) + ind, a.cd(2) + ind, a.cd(3) + ind, a.cd(4) + ind, a.cs(), inc);
  }
};

// The point data a functor reads through its cells' mapping-0 ids (`tgt` in
// compute), which a hierarchical block can stage in LDS with its cache:
// slot (-1: none) and values per point.
template <class F> struct StagedInput {
  static constexpr int slot = -1;
  static constexpr uint32_t dim = 0;
};
template <> struct StagedInput<MiniAeroF> { // the 28 cell values
  static constexpr int slot = 1;
  static constexpr uint32_t dim = 28;
};
template <> struct StagedInput<LuleshF> { // node coordinates
  static constexpr int slot = 1;
  static constexpr uint32_t dim = 3;
};
template <> struct StagedInput<VolnaF> { // cell volumes
  static constexpr int slot = 1;
  static constexpr uint32_t dim = 1;
};
template <uint32_t MDv, uint32_t P, uint32_t C> struct StagedInput<MineF<MDv, P, C>> {
  static constexpr int slot = 0; // the iterate
  static constexpr uint32_t dim = P;
};

// The engine arguments seen by a functor whose staged input is in LDS: point
// data `SLOT` is the block's staged copy (AOS rows of the staged dim, or SOA
// with stride m), indexed by block-local ids; everything else as in `a`.
template <int SLOT> struct LdsArgs {
  const EngineArgs &a;
  const double *staged;
  uint32_t m;
  static constexpr bool kStreamCells = false;
  FOT_DEV const double *pd(uint32_t k) const { return k == uint32_t(SLOT) ? staged : a.pd(k); }
  FOT_DEV uint32_t ps(uint32_t k) const { return k == uint32_t(SLOT) ? m : a.ps(k); }
  FOT_DEV const double *cd(uint32_t k) const { return a.cd(k); }
  FOT_DEV uint32_t cs() const { return a.cs(); }
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

// Whether a cell's block-local ids are pairwise distinct (every cell of
// the meshes here; a degenerate one takes the sequential adds).
template <class F> FOT_DEV bool distinct_slots(const uint32_t *loc) {
  bool ok = true;
#pragma unroll
  for (uint32_t k = 0; k < F::MD; ++k)
#pragma unroll
    for (uint32_t l = k + 1; l < F::MD; ++l) ok = ok && loc[k] != loc[l];
  return ok;
}

// With pairwise distinct slots a cell's adds can be batched: read up to
// kAddChunk cache values, then write their sums, one LDS round trip per
// chunk instead of one per value (the compiler cannot prove the slots
// distinct, so it serialises plain +=).  The same additions, the same bits.
// Slots at or above `owned` read slot 0 and write nothing.
constexpr uint32_t kAddChunk = 8;
// A functor's chunk: kAddChunk unless it sets ADD_CHUNK (1: sequential adds).
template <class F, class = void> struct AddChunk {
  static constexpr uint32_t value = kAddChunk;
};
template <class F> struct AddChunk<F, std::void_t<decltype(F::ADD_CHUNK)>> {
  static constexpr uint32_t value = F::ADD_CHUNK;
};

template <class F>
FOT_DEV void add_cell_batched(double *cache, uint32_t m, const uint32_t *loc, const double *inc,
                              uint32_t owned) {
  constexpr uint32_t S = F::MD * F::TD, CH = AddChunk<F>::value;
#pragma unroll
  for (uint32_t c0 = 0; c0 < S; c0 += CH) {
    double v[CH];
#pragma unroll
    for (uint32_t j = 0; j < CH; ++j)
      if (c0 + j < S) {
        const uint32_t k = (c0 + j) / F::TD, d = (c0 + j) % F::TD;
        v[j] = cache[d * m + (loc[k] < owned ? loc[k] : 0u)];
      }
#pragma unroll
    for (uint32_t j = 0; j < CH; ++j)
      if (c0 + j < S) {
        const uint32_t k = (c0 + j) / F::TD, d = (c0 + j) % F::TD;
        if (loc[k] < owned) cache[d * m + loc[k]] = v[j] + inc[c0 + j];
      }
  }
}

// Adds the increments of one cell to an LDS cache (SOA, stride m) at
// block-local ids `loc`, in the functor's order (batched when the ids are
// distinct).
template <class F>
FOT_DEV void add_to_cache(double *cache, uint32_t m, const uint32_t *loc, const double *inc) {
  if (AddChunk<F>::value > 1 && distinct_slots<F>(loc)) {
    add_cell_batched<F>(cache, m, loc, inc, 0xffffffffu);
    return;
  }
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
  if (owned == 0) return;
  if (AddChunk<F>::value > 1 && distinct_slots<F>(loc)) {
    add_cell_batched<F>(cache, m, loc, inc, owned);
    return;
  }
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
