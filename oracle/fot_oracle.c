/*
 * fot_oracle.c -- CPU oracle for the edge-centred indirect-increment loop.
 *
 * TEST INFRASTRUCTURE ONLY (see fot_oracle.h).  Built with -ffp-contract=off
 * so every increment is the IEEE result of the reference's expression order.
 * The reference cannot be built in this image (it needs cuda.h, METIS and
 * Scotch), so this restatement is pinned by the reference's own known-answer
 * tests, by invariants and by closed forms (tests/test_oracle_*.py).
 */
#include "fot_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Planted mistakes.  ORACLE_PLANT is 0 in every normal build.  A nonzero
 * value flips exactly one rule of the restatement, the kind of slip a
 * restatement could plausibly make; tests/test_oracle_planted.py builds each
 * variant into oracle/lib/planted/ and shows that a known-answer test fails
 * on it, so every such rule is pinned by a test that notices it flipped. */
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif
enum {
  PLANT_NONE = 0,
  PLANT_BLOCK_COLOUR_FIRST_FIT = 1, /* first available block colour, not least used */
  PLANT_PART_CHANGE_KEEPS_TID = 2,  /* no tid reset where the partition changes */
  PLANT_THREAD_ORDER_REVERSED = 3,  /* positions descending inside a thread colour */
  PLANT_CACHE_DESCENDING = 4,       /* points_to_be_cached in descending order */
  PLANT_MINE_NO_FACTOR_2 = 5,       /* mine_func<2> without its factor 2 */
  PLANT_MINE_SAME_SIDE = 6,         /* mine_func<2>: out_L from L, not from R */
  PLANT_MINE_RING_NEIGHBOUR = 7,    /* mine_func<4/8>: neighbours k+1 and k+2 */
  PLANT_GLOBAL_COLOUR_FIRST_FIT = 8 /* Mesh::colourCells first fit */
};

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* 128-bit colour sets: Mesh::colourset_t = std::bitset<128> (mesh.hpp:26)  */
/* ------------------------------------------------------------------------ */
typedef struct {
  uint64_t w[2];
} cset;

static inline int cset_test(cset s, unsigned i) {
  return (int)((s.w[i >> 6] >> (i & 63)) & 1u);
}
static inline void cset_set(cset *s, unsigned i) {
  s->w[i >> 6] |= (uint64_t)1 << (i & 63);
}
static inline cset cset_or(cset a, cset b) {
  cset r = {{a.w[0] | b.w[0], a.w[1] | b.w[1]}};
  return r;
}
static inline cset cset_andnot(cset used, cset occ) { /* ~occ & used */
  cset r = {{~occ.w[0] & used.w[0], ~occ.w[1] & used.w[1]}};
  return r;
}
static inline int cset_none(cset a) { return (a.w[0] | a.w[1]) == 0; }
static inline cset cset_shl1_set0(cset a) { /* (a << 1) | 1 */
  cset r;
  r.w[1] = (a.w[1] << 1) | (a.w[0] >> 63);
  r.w[0] = (a.w[0] << 1) | 1u;
  return r;
}

/* Mesh::getAvailableColour<true> (mesh.hpp:287-315): the available colour
 * with the smallest set size; the first such index wins (strict '>'). */
static unsigned avail_min_size(cset avail, const uint32_t *set_sizes,
                               unsigned n) {
  unsigned colour = n;
  for (unsigned i = 0; i < n; ++i) {
    if (cset_test(avail, i)) {
      if (colour >= n || set_sizes[colour] > set_sizes[i]) colour = i;
    }
  }
  return colour;
}
/* Mesh::getAvailableColour<false>: the first available colour. */
static unsigned avail_first(cset avail) {
  for (unsigned i = 0; i < 128; ++i)
    if (cset_test(avail, i)) return i;
  return 128;
}

/* ------------------------------------------------------------------------ */
/* User functions                                                            */
/* ------------------------------------------------------------------------ */

/* kernels/res_calc_func.hpp:11-57, same expression order.  The reference's
 * 0.5f and 1.0f literals are exact in double. */
void oracle_res_calc_func(const double *x1, const double *x2, const double *q1,
                          const double *q2, const double *adt1,
                          const double *adt2, double *res1, double *res2,
                          unsigned x_stride, unsigned q_stride) {
  const double gm1 = 0.4, eps = 0.05;
  double dx, dy, mu, ri, p1, vol1, p2, vol2, f;
  dx = x1[0 * x_stride] - x2[0 * x_stride];
  dy = x1[1 * x_stride] - x2[1 * x_stride];
  ri = 1.0 / q1[0];
  p1 = gm1 * (q1[3 * q_stride] -
              0.5 * ri *
                  (q1[1 * q_stride] * q1[1 * q_stride] +
                   q1[2 * q_stride] * q1[2 * q_stride]));
  vol1 = ri * (q1[1 * q_stride] * dy - q1[2 * q_stride] * dx);
  ri = 1.0 / q2[0];
  p2 = gm1 * (q2[3 * q_stride] -
              0.5 * ri *
                  (q2[1 * q_stride] * q2[1 * q_stride] +
                   q2[2 * q_stride] * q2[2 * q_stride]));
  vol2 = ri * (q2[1 * q_stride] * dy - q2[2 * q_stride] * dx);
  mu = 0.5 * ((*adt1) + (*adt2)) * eps;
  f = 0.5 * (vol1 * q1[0] + vol2 * q2[0]) + mu * (q1[0] - q2[0]);
  res1[0] = f;
  res2[0] = -f;
  f = 0.5 * (vol1 * q1[1 * q_stride] + p1 * dy + vol2 * q2[1 * q_stride] +
             p2 * dy) +
      mu * (q1[1 * q_stride] - q2[1 * q_stride]);
  res1[1] = f;
  res2[1] = -f;
  f = 0.5 * (vol1 * q1[2 * q_stride] - p1 * dx + vol2 * q2[2 * q_stride] -
             p2 * dx) +
      mu * (q1[2 * q_stride] - q2[2 * q_stride]);
  res1[2] = f;
  res2[2] = -f;
  f = 0.5 * (vol1 * (q1[3 * q_stride] + p1) + vol2 * (q2[3 * q_stride] + p2)) +
      mu * (q1[3 * q_stride] - q2[3 * q_stride]);
  res1[3] = f;
  res2[3] = -f;
}

/* ---- mini_aero (kernels/mini_aero_func.hpp:11-85 and
 * kernels/mini_aero_utils.hpp:1-375), restated step by step ---------------- */
static double ma_pressure(const double *V) { /* mini_aero_utils.hpp:6-12 */
  const double Rgas = 287.05;
  return V[0] * Rgas * V[4];
}

static double ma_enthalpy(const double *V) { /* :14-18 */
  const double Cp = 1004.0;
  return Cp * V[4];
}

static void ma_matvec5(double alpha, const double *A, const double *x, double beta,
                       double *y) { /* :20-29 */
  for (int i = 0; i < 5; ++i) {
    y[i] *= beta;
    for (int j = 0; j < 5; ++j) y[i] += alpha * A[5 * i + j] * x[j];
  }
}

static void ma_inviscid(const double *Pl, const double *Pr, double *flux, const double *n,
                        const double *t, const double *b) { /* :31-242, unit strides */
  const double efix_u = 0.1, efix_c = 0.1, gm1 = 0.4;
  const double rho_l = Pl[0], u_l = Pl[1], v_l = Pl[2], w_l = Pl[3];
  const double p_l = ma_pressure(Pl), h_l = ma_enthalpy(Pl);
  const double ht_l = h_l + 0.5 * (u_l * u_l + v_l * v_l + w_l * w_l);
  const double mf_l = rho_l * (n[0] * u_l + n[1] * v_l + n[2] * w_l);
  const double rho_r = Pr[0], u_r = Pr[1], v_r = Pr[2], w_r = Pr[3];
  const double p_r = ma_pressure(Pr), h_r = ma_enthalpy(Pr);
  const double ht_r = h_r + 0.5 * (u_r * u_r + v_r * v_r + w_r * w_r);
  const double mf_r = rho_r * (n[0] * u_r + n[1] * v_r + n[2] * w_r);
  const double psum = p_l + p_r;
  flux[0] = 0.5 * (mf_l + mf_r);
  flux[1] = 0.5 * (mf_l * u_l + mf_r * u_r + n[0] * psum);
  flux[2] = 0.5 * (mf_l * v_l + mf_r * v_r + n[1] * psum);
  flux[3] = 0.5 * (mf_l * w_l + mf_r * w_r + n[2] * psum);
  flux[4] = 0.5 * (mf_l * ht_l + mf_r * ht_r);
  const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  const double tn = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  const double bn = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  const double nu[3] = {n[0] / nn, n[1] / nn, n[2] / nn};
  const double tu[3] = {t[0] / tn, t[1] / tn, t[2] / tn};
  const double bu[3] = {b[0] / bn, b[1] / bn, b[2] / bn};
  const double denom = 1.0 / (sqrt(rho_l) + sqrt(rho_r));
  const double alpha = sqrt(rho_l) * denom;
  const double beta = 1.0 - alpha;
  const double ur = alpha * u_l + beta * u_r;
  const double vr = alpha * v_l + beta * v_r;
  const double wr = alpha * w_l + beta * w_r;
  const double hr = alpha * h_l + beta * h_r +
                    0.5 * alpha * beta *
                        ((u_r - u_l) * (u_r - u_l) + (v_r - v_l) * (v_r - v_l) +
                         (w_r - w_l) * (w_r - w_l));
  const double cr = sqrt(gm1 * hr);
  double M[25];
  const double vn = ur * nu[0] + vr * nu[1] + wr * nu[2];
  const double vt = ur * tu[0] + vr * tu[1] + wr * tu[2];
  const double vb = ur * bu[0] + vr * bu[1] + wr * bu[2];
  const double ke = 0.5 * (ur * ur + vr * vr + wr * wr);
  const double c2i = 1.0 / (cr * cr);
  const double hc2i = 0.5 * c2i;
  double dq[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  dq[0] = rho_r - rho_l;
  dq[1] = rho_r * u_r - rho_l * u_l;
  dq[2] = rho_r * v_r - rho_l * v_l;
  dq[3] = rho_r * w_r - rho_l * w_l;
  dq[4] = (rho_r * ht_r - p_r) - (rho_l * ht_l - p_l);
  const double cbar = cr * nn;
  const double ubar = ur * n[0] + vr * n[1] + wr * n[2];
  const double cfl = fabs(ubar) + cbar;
  const double e1 = ubar + cbar, e2 = ubar - cbar, e3 = ubar;
  double a1 = fabs(e1), a2 = fabs(e2), a3 = fabs(e3);
  const double epuc = efix_u * cfl, epcc = efix_c * cfl;
  if (a1 < epcc) a1 = 0.5 * (e1 * e1 + epcc * epcc) / epcc;
  if (a2 < epcc) a2 = 0.5 * (e2 * e2 + epcc * epcc) / epcc;
  if (a3 < epuc) a3 = 0.5 * (e3 * e3 + epuc * epuc) / epuc;
  double eigp[5] = {0.5 * (e1 + a1), 0.5 * (e2 + a2), 0.5 * (e3 + a3), 0.0, 0.0};
  eigp[3] = eigp[4] = eigp[2];
  double eigm[5] = {0.5 * (e1 - a1), 0.5 * (e2 - a2), 0.5 * (e3 - a3), 0.0, 0.0};
  eigm[3] = eigm[4] = eigm[2];
  double ldq[5] = {0, 0, 0, 0, 0}, rlldq[5] = {0, 0, 0, 0, 0};
  M[0] = gm1 * (ke - hr) + cr * (cr - vn);
  M[1] = cr * nu[0] - gm1 * ur;
  M[2] = cr * nu[1] - gm1 * vr;
  M[3] = cr * nu[2] - gm1 * wr;
  M[4] = gm1;
  M[5] = gm1 * (ke - hr) + cr * (cr + vn);
  M[6] = -cr * nu[0] - gm1 * ur;
  M[7] = -cr * nu[1] - gm1 * vr;
  M[8] = -cr * nu[2] - gm1 * wr;
  M[9] = gm1;
  M[10] = ke - hr;
  M[11] = -ur;
  M[12] = -vr;
  M[13] = -wr;
  M[14] = 1.0;
  M[15] = -vt;
  M[16] = tu[0];
  M[17] = tu[1];
  M[18] = tu[2];
  M[19] = 0.0;
  M[20] = -vb;
  M[21] = bu[0];
  M[22] = bu[1];
  M[23] = bu[2];
  M[24] = 0.0;
  ma_matvec5(1.0, M, dq, 0.0, ldq);
  for (int j = 0; j < 5; ++j) ldq[j] = (eigp[j] - eigm[j]) * ldq[j];
  M[0] = hc2i;
  M[1] = hc2i;
  M[2] = -gm1 * c2i;
  M[3] = 0.0;
  M[4] = 0.0;
  M[5] = (ur + nu[0] * cr) * hc2i;
  M[6] = (ur - nu[0] * cr) * hc2i;
  M[7] = -gm1 * ur * c2i;
  M[8] = tu[0];
  M[9] = bu[0];
  M[10] = (vr + nu[1] * cr) * hc2i;
  M[11] = (vr - nu[1] * cr) * hc2i;
  M[12] = -gm1 * vr * c2i;
  M[13] = tu[1];
  M[14] = bu[1];
  M[15] = (wr + nu[2] * cr) * hc2i;
  M[16] = (wr - nu[2] * cr) * hc2i;
  M[17] = -gm1 * wr * c2i;
  M[18] = tu[2];
  M[19] = bu[2];
  M[20] = (hr + ke + cr * vn) * hc2i;
  M[21] = (hr + ke - cr * vn) * hc2i;
  M[22] = (cr * cr - gm1 * (hr + ke)) * c2i;
  M[23] = vt;
  M[24] = vb;
  ma_matvec5(1.0, M, ldq, 0.0, rlldq);
  for (int i = 0; i < 5; ++i) flux[i] -= 0.5 * rlldq[i];
}

static void ma_viscous(double g[5][3], const double *P, const double *a,
                       double *vf) { /* :244-300 */
  const double s0 = 1.458e-6, s1 = 110.4;
  const double T = P[4];
  const double mu = s0 * T * sqrt(T) / (T + s1);
  const double Pr = 0.71, Cp = 1006.0;
  const double kc = mu * Cp / Pr;
  double div = 0;
  for (int c = 0; c < 5; ++c) vf[c] = 0.0;
  for (int i = 0; i < 3; i++) div += g[i + 1][i];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      const double dij = (i == j) ? 1 : 0;
      const double Sij = 0.5 * (g[i + 1][j] + g[j + 1][i]);
      const double tij = Sij - div * dij / 3.;
      vf[1 + i] += (2 * mu * tij) * a[j];
      vf[4] += (2 * mu * tij) * P[i + 1] * a[j];
    }
    vf[4] += kc * g[4][i] * a[i];
  }
}

static void ma_primitives(const double *U, double *V) { /* :302-322 */
  double gamma = 1.4, Rgas = 287.05;
  double r = U[0], ri = 1.0 / r;
  double u = U[1] * ri, v = U[2] * ri, w = U[3] * ri;
  double k = 0.5 * (u * u + v * v + w * w);
  double e = U[4] * ri - k;
  double T = e * (gamma - 1.0) / Rgas;
  V[0] = r, V[1] = u, V[2] = v, V[3] = w, V[4] = T;
}

/* mini_aero_func.hpp:11-85: L, R cell rows of 28 (stride s), face row of 12
 * (stride fs) */
static void ma_face(const double *L, const double *R, uint32_t s, const double *F, uint32_t fs,
                    double *flux) {
  double cl[5], cr[5], pl[5], pr[5];
  for (int c = 0; c < 5; ++c) cl[c] = L[c * s], cr[c] = R[c * s];
  ma_primitives(cl, pl);
  ma_primitives(cr, pr);
  for (int c = 0; c < 5; ++c) {
    double gl = 0, gr = 0;
    for (int d = 0; d < 3; ++d) {
      gl += (F[d * fs] - L[(5 + d) * s]) * L[(8 + c * 3 + d) * s];
      gr += (F[d * fs] - R[(5 + d) * s]) * R[(8 + c * 3 + d) * s];
    }
    pl[c] += gl * L[(23 + c) * s];
    pr[c] += gr * R[(23 + c) * s];
  }
  double n[3], t[3], b[3];
  for (int d = 0; d < 3; ++d) {
    n[d] = F[(3 + d) * fs];
    t[d] = F[(6 + d) * fs];
    b[d] = F[(9 + d) * fs];
  }
  ma_inviscid(pl, pr, flux, n, t, b);
  double pf[5], gf[5][3];
  for (int c = 0; c < 5; ++c) {
    pf[c] = 0.5 * (pl[c] + pr[c]);
    for (int d = 0; d < 3; ++d) gf[c][d] = 0.5 * (L[(8 + c * 3 + d) * s] + R[(8 + c * 3 + d) * s]);
  }
  double vf[5];
  ma_viscous(gf, pf, n, vf);
  for (int c = 0; c < 5; ++c) flux[c] -= vf[c];
}

/* lulesh (kernels/lulesh_func.hpp:13-38 with kernels/lulesh_utils.hpp).
 * The volume comes from the shape-function derivatives at the element
 * centre (CalcElemShapeFunctionDerivatives, lulesh_utils.hpp:4-93), whose
 * derivative values the reference then overwrites with the node normals:
 * only the three cofactors of the eta column enter the volume.  Node normals
 * (CalcElemNodeNormals, :95-183): each of the six faces, in the reference's
 * order and node order, adds a quarter of the cross product of its two
 * bisectors to each of its four nodes.  Forces (SumElemStressesToNodeForces,
 * :185-197): f = -(sigma * normal), component by component.  Expressions
 * keep the reference's association so the results are bit-identical. */
void oracle_lulesh_func(const double *x, const double *y, const double *z, double sxx,
                        double syy, double szz, double *f, double *determ) {
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

  static const int face[6][4] = {{0, 1, 2, 3}, {0, 4, 5, 1}, {1, 5, 6, 2},
                                 {2, 6, 7, 3}, {3, 7, 4, 0}, {4, 7, 6, 5}};
  double nx[8], ny[8], nz[8];
  for (int i = 0; i < 8; ++i) nx[i] = ny[i] = nz[i] = 0.0;
  for (int q = 0; q < 6; ++q) {
    const int a = face[q][0], b = face[q][1], c = face[q][2], d = face[q][3];
    const double b0x = 0.5 * (x[d] + x[c] - x[b] - x[a]);
    const double b0y = 0.5 * (y[d] + y[c] - y[b] - y[a]);
    const double b0z = 0.5 * (z[d] + z[c] - z[b] - z[a]);
    const double b1x = 0.5 * (x[c] + x[b] - x[d] - x[a]);
    const double b1y = 0.5 * (y[c] + y[b] - y[d] - y[a]);
    const double b1z = 0.5 * (z[c] + z[b] - z[d] - z[a]);
    const double ax = 0.25 * (b0y * b1z - b0z * b1y);
    const double ay = 0.25 * (b0z * b1x - b0x * b1z);
    const double az = 0.25 * (b0x * b1y - b0y * b1x);
    const int nodes[4] = {a, b, c, d};
    for (int k = 0; k < 4; ++k) {
      nx[nodes[k]] += ax;
      ny[nodes[k]] += ay;
      nz[nodes[k]] += az;
    }
  }
  for (int i = 0; i < 8; ++i) {
    f[0 * 8 + i] = -(sxx * nx[i]);
    f[1 * 8 + i] = -(syy * ny[i]);
    f[2 * 8 + i] = -(szz * nz[i]);
  }
}

/* Kernels whose increments the reference adds slot by slot (corner-major):
 * lulesh's StepSeq (kernels/lulesh.hpp:52-56) and the scatter-only hex
 * functor behind the same API. */
static inline int slot_outer(const oracle_problem *p) {
  return p->kernel == ORACLE_HEX_SCATTER || p->kernel == ORACLE_LULESH;
}

/* Index of element (point i, component d) of a point set (data_t.hpp:104-115). */
static inline size_t pidx(int soa, uint32_t n, uint32_t i, uint32_t dim,
                          uint32_t d) {
  return soa ? (size_t)d * n + i : (size_t)i * dim + d;
}

/* Per-kernel increment: inc[slot * tdim + d] for the mesh_dim[0] target
 * slots of cell `ind`, reading the iteration's input point data. */
static void compute_inc(const oracle_problem *p, uint32_t ind, double *inc) {
  const int soa = p->soa;
  switch (p->kernel) {
  case ORACLE_RES_CALC: {
    /* res_calc::StepSeq (kernels/res_calc.hpp:29-89): x through mapping 1,
     * q, adt and res through mapping 0. */
    const uint32_t *m0 = p->cell_to_node[0], *m1 = p->cell_to_node[1];
    const double *x = p->point_data[1], *q = p->point_data[2],
                 *adt = p->point_data[3];
    const uint32_t ux = soa ? 1 : 2, uq = soa ? 1 : 4;
    const unsigned xs = soa ? p->num_points[1] : 1;
    const unsigned qs = soa ? p->num_points[2] : 1;
    oracle_res_calc_func(x + (size_t)ux * m1[2 * ind + 0],
                         x + (size_t)ux * m1[2 * ind + 1],
                         q + (size_t)uq * m0[2 * ind + 0],
                         q + (size_t)uq * m0[2 * ind + 1],
                         adt + m0[2 * ind + 0], adt + m0[2 * ind + 1], inc,
                         inc + 4, xs, qs);
    break;
  }
  case ORACLE_MINE2:
  case ORACLE_MINE4:
  case ORACLE_MINE8: {
    /* mine::StepSeq<MeshDim> (kernels/mine.hpp:22-178) with mine_func
     * (kernels/mine_func.hpp:11-99). */
    const uint32_t md = p->mesh_dim[0], P = p->point_dim[0];
    const uint32_t C = p->cell_dim[0];
    const uint32_t *m0 = p->cell_to_node[0];
    const double *in = p->point_data[0];
    const double *w = p->cell_data[0] + ind;
    const uint32_t n = p->num_points[0], cs = p->num_cells;
    const double *pt[8];
    for (uint32_t k = 0; k < md; ++k)
      pt[k] = in + (size_t)(soa ? 1 : P) * m0[md * ind + k];
    const uint32_t ps = soa ? n : 1;
    for (uint32_t d = 0; d < P; ++d) {
      const uint32_t cd = (C == P) ? d : 0;
      const double wd = w[(size_t)cd * cs];
      if (md == 2) {
        const double f = ORACLE_PLANT == PLANT_MINE_NO_FACTOR_2 ? 1 : 2;
        const int same = ORACLE_PLANT == PLANT_MINE_SAME_SIDE;
        inc[0 * P + d] = f * pt[same ? 0 : 1][d * ps] * wd;
        inc[1 * P + d] = f * pt[same ? 1 : 0][d * ps] * wd;
      } else {
        for (uint32_t k = 0; k < md; ++k) {
          const uint32_t a = (k + 1) % md;
          const uint32_t b = ORACLE_PLANT == PLANT_MINE_RING_NEIGHBOUR ? (k + 2) % md
                                                                      : (k + md - 1) % md;
          inc[k * P + d] = (pt[a][d * ps] + pt[b][d * ps]) * wd;
        }
      }
    }
    break;
  }
  case ORACLE_HEX_SCATTER: {
    /* Scatter-only functor behind the lulesh struct API: the 24 corner
     * forces of an element are cell data laid out as the reference's
     * f_out[i * MESH_DIM + j] (kernels/lulesh.hpp:52-56). */
    const double *F = p->cell_data[0] + ind;
    const uint32_t cs = p->num_cells;
    for (uint32_t j = 0; j < 8; ++j)
      for (uint32_t i = 0; i < 3; ++i)
        inc[j * 3 + i] = F[(size_t)(i * 8 + j) * cs];
    break;
  }
  case ORACLE_LULESH: {
    /* lulesh::StepSeq (kernels/lulesh.hpp:23-58): xyz (point data 1) and the
     * force target through mapping 0; cell data 0 = sigma (xx, yy, zz), cell
     * data 1 = the element volume, written by the user function. */
    const uint32_t *m0 = p->cell_to_node[0];
    const double *xyz = p->point_data[1];
    const uint32_t xs = soa ? p->num_points[1] : 1, ux = soa ? 1 : 3;
    const uint32_t cs = p->num_cells;
    double xv[8], yv[8], zv[8], fc[24], determ;
    for (int k = 0; k < 8; ++k) {
      const double *c = xyz + (size_t)ux * m0[8 * ind + k];
      xv[k] = c[0];
      yv[k] = c[xs];
      zv[k] = c[2 * (size_t)xs];
    }
    const double *sig = p->cell_data[0] + ind;
    oracle_lulesh_func(xv, yv, zv, sig[0], sig[cs], sig[2 * (size_t)cs], fc, &determ);
    ((double *)p->cell_data[1])[ind] = determ; /* the kernel's cell output */
    for (int j = 0; j < 8; ++j)
      for (int i = 0; i < 3; ++i) inc[j * 3 + i] = fc[i * 8 + j];
    break;
  }
  case ORACLE_GETACC: {
    /* getacc::StepSeq (kernels/getacc_scatter.hpp:27-78) with the user
     * function (getacc_scatter_func.hpp:17-94): cell data cnmass, rho, cnwt,
     * cnfx, cnfy (SOA, stride num_cells); corner jj's mass falls back to the
     * previous corner's (3 before 0), then to rho * cnwt[jj]. */
    const uint32_t cs = p->num_cells;
    const double *cnmass = p->cell_data[0] + ind, *rho = p->cell_data[1] + ind;
    const double *cnwt = p->cell_data[2] + ind, *cnfx = p->cell_data[3] + ind;
    const double *cnfy = p->cell_data[4] + ind;
    const double zerocut = 1.0e-40;
    for (int k = 0; k < 16; ++k) inc[k] = 0;
    for (int jj = 0; jj < 4; ++jj) {
      double *pt = inc + 4 * jj;
      if (cnmass[jj * cs] > zerocut) {
        pt[0] += cnmass[jj * cs];
      } else {
        int ii = jj - 1;
        if (ii == -1) ii = 3;
        if (cnmass[ii * cs] > zerocut) pt[0] += cnmass[ii * cs];
        else pt[0] += *rho * cnwt[jj * cs];
      }
      pt[1] += cnwt[jj * cs];
      pt[2] += cnfx[jj * cs];
      pt[3] += cnfy[jj * cs];
    }
    break;
  }
  case ORACLE_MINI_AERO: {
    /* mini_aero::StepSeq (kernels/mini_aero.hpp:43-117): cell values (28)
     * through mapping-0 ids, the face's 12 values as SOA cell data; left
     * -= flux, right += flux, i.e. increments -flux and +flux. */
    const uint32_t *m0 = p->cell_to_node[0];
    const double *cv = p->point_data[1];
    const uint32_t s = soa ? p->num_points[1] : 1, u = soa ? 1 : 28;
    double flux[5];
    ma_face(cv + (size_t)u * m0[2 * ind], cv + (size_t)u * m0[2 * ind + 1], s,
            p->cell_data[0] + ind, p->num_cells, flux);
    for (int c = 0; c < 5; ++c) inc[c] = -flux[c], inc[5 + c] = flux[c];
    break;
  }
  case ORACLE_VOLNA: {
    /* volna::StepSeq (kernels/volna.hpp:27-78) with volna_func
     * (kernels/volna_func.hpp:16-36), in FP64: volumes (point data 1)
     * through mapping-0 ids; cell data fluxes (3), bathymetry source (2),
     * edge normals (2) and the right-boundary flag (0.0 / 1.0).  The
     * increments start at zero and take the reference's -= / +=. */
    const uint32_t *m0 = p->cell_to_node[0];
    const double *vol = p->point_data[1];
    const uint32_t cs = p->num_cells;
    const double *fl = p->cell_data[0] + ind, *bs = p->cell_data[1] + ind;
    const double *nm = p->cell_data[2] + ind;
    const double v0 = vol[m0[2 * ind]], v1 = vol[m0[2 * ind + 1]];
    for (int k = 0; k < 8; ++k) inc[k] = 0.0;
    inc[0] -= fl[0] / v0;
    inc[1] -= (fl[cs] + bs[0] * nm[0]) / v0;
    inc[2] -= (fl[2 * cs] + bs[0] * nm[cs]) / v0;
    if (p->cell_data[3][ind] == 0.0) {
      inc[4] += fl[0] / v1;
      inc[5] += (fl[cs] + bs[cs] * nm[0]) / v1;
      inc[6] += (fl[2 * cs] + bs[cs] * nm[cs]) / v1;
    }
    break;
  }
  default:
    break;
  }
}

/* Adds the increments of cell `ind` into `out` in the reference's order;
 * abs (nullable) accumulates |inc|. */
static void add_inc(const oracle_problem *p, uint32_t ind, const double *inc,
                    double *out, double *abs_acc) {
  const uint32_t md = p->mesh_dim[0], P = p->point_dim[0];
  const uint32_t n = p->num_points[0];
  const uint32_t *m0 = p->cell_to_node[0];
  if (slot_outer(p)) { /* slot outer (lulesh.hpp:52-56) */
    for (uint32_t k = 0; k < md; ++k)
      for (uint32_t d = 0; d < P; ++d) {
        const size_t g = pidx(p->soa, n, m0[md * ind + k], P, d);
        out[g] += inc[k * P + d];
        if (abs_acc) abs_acc[g] += fabs(inc[k * P + d]);
      }
  } else { /* component outer, slot inner (res_calc.hpp:144-147) */
    for (uint32_t d = 0; d < P; ++d)
      for (uint32_t k = 0; k < md; ++k) {
        const size_t g = pidx(p->soa, n, m0[md * ind + k], P, d);
        out[g] += inc[k * P + d];
        if (abs_acc) abs_acc[g] += fabs(inc[k * P + d]);
      }
  }
}

static size_t target_len(const oracle_problem *p) {
  return (size_t)p->num_points[0] * p->point_dim[0];
}

/* problem.hpp:362-380 (serial loop, cells in index order). */
int oracle_loop_serial(const oracle_problem *p, uint32_t iters,
                       double *abs_sum, double *loop_seconds) {
  const size_t len = target_len(p);
  double *temp = (double *)malloc(len * sizeof(double));
  double inc[64];
  if (!temp) return -1;
  memcpy(temp, p->point_data[0], len * sizeof(double));
  if (abs_sum)
    for (size_t i = 0; i < len; ++i) abs_sum[i] = fabs(temp[i]);
  const double t0 = now_seconds();
  for (uint32_t it = 0; it < iters; ++it) {
    for (uint32_t c = 0; c < p->num_cells; ++c) {
      compute_inc(p, c, inc);
      add_inc(p, c, inc, temp, abs_sum);
    }
    memcpy(p->point_data[0], temp, len * sizeof(double));
  }
  if (loop_seconds) *loop_seconds = now_seconds() - t0;
  free(temp);
  return 0;
}

/* problem.hpp:382-437 (colour-major loop; OpenMP inside a colour). */
int oracle_loop_colour(const oracle_problem *p, uint32_t iters, int nthreads,
                       double *loop_seconds) {
  const size_t len = target_len(p);
  const uint32_t nc = p->num_cells;
  uint8_t *col = (uint8_t *)malloc(nc ? nc : 1);
  uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * (nc ? nc : 1));
  uint32_t off[66] = {0};
  double *temp = (double *)malloc(len * sizeof(double));
  int ncol;
  if (!col || !order || !temp) return -1;
  ncol = oracle_colour_cells(p->cell_to_node[0], nc, p->mesh_dim[0],
                             p->num_points[0], col);
  if (ncol < 0) {
    free(col), free(order), free(temp);
    return -1;
  }
  for (uint32_t c = 0; c < nc; ++c) off[col[c] + 1]++;
  for (int k = 0; k < ncol; ++k) off[k + 1] += off[k];
  {
    uint32_t fill[65];
    memcpy(fill, off, sizeof(fill));
    for (uint32_t c = 0; c < nc; ++c) order[fill[col[c]]++] = c;
  }
  memcpy(temp, p->point_data[0], len * sizeof(double));
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  /* timed like the reference: the iteration loop only (problem.hpp:410-436) */
  const double t0 = now_seconds();
  for (uint32_t it = 0; it < iters; ++it) {
    for (int k = 0; k < ncol; ++k) {
      const int64_t b = off[k], e = off[k + 1];
#pragma omp parallel for schedule(static) if (nthreads > 1)
      for (int64_t i = b; i < e; ++i) {
        double inc[64];
        compute_inc(p, order[i], inc);
        add_inc(p, order[i], inc, temp, NULL);
      }
    }
    memcpy(p->point_data[0], temp, len * sizeof(double));
  }
  if (loop_seconds) *loop_seconds = now_seconds() - t0;
  free(col), free(order), free(temp);
  return 0;
}

/* Hierarchical summation order (res_calc.hpp:377-442): zeroed cache,
 * thread colours in order, then out += cache, blocks in plan order. */
int oracle_loop_hier_replay(const oracle_problem *p, const oracle_hplan *h,
                            uint32_t iters) {
  const size_t len = target_len(p);
  const uint32_t md = p->mesh_dim[0], P = p->point_dim[0];
  const uint32_t n = p->num_points[0];
  double *temp = (double *)malloc(len * sizeof(double));
  uint32_t maxc = 0;
  double *cache;
  double inc[64];
  if (!temp) return -1;
  for (uint32_t c = 0; c < h->num_colours; ++c) {
    const uint32_t bb = h->colour_block_offset[c] + c;
    const uint32_t nb = h->colour_block_offset[c + 1] - h->colour_block_offset[c];
    for (uint32_t b = 0; b < nb; ++b) {
      const uint32_t m = h->points_to_be_cached_offsets[bb + b + 1] -
                         h->points_to_be_cached_offsets[bb + b];
      if (m > maxc) maxc = m;
    }
  }
  cache = (double *)malloc(sizeof(double) * (size_t)(maxc ? maxc : 1) * P);
  if (!cache) {
    free(temp);
    return -1;
  }
  memcpy(temp, p->point_data[0], len * sizeof(double));
  for (uint32_t it = 0; it < iters; ++it) {
    uint32_t blk = 0;
    for (uint32_t c = 0; c < h->num_colours; ++c) {
      const uint32_t c0 = h->colour_cell_offset[c];
      const uint32_t ncell = h->colour_cell_offset[c + 1] - c0;
      const uint32_t bb = h->colour_block_offset[c] + c; /* nb_c+1 entries */
      const uint32_t nb = h->colour_block_offset[c + 1] - h->colour_block_offset[c];
      const uint32_t cache0 = h->colour_cache_offset[c];
      for (uint32_t b = 0; b < nb; ++b, ++blk) {
        const uint32_t from = h->block_offsets[bb + b];
        const uint32_t to = h->block_offsets[bb + b + 1];
        const uint32_t po = h->points_to_be_cached_offsets[bb + b];
        const uint32_t m = h->points_to_be_cached_offsets[bb + b + 1] - po;
        const uint32_t ntc = h->num_cell_colours[blk];
        memset(cache, 0, sizeof(double) * (size_t)m * P);
        for (uint32_t tc = 0; tc < ntc; ++tc) {
          for (uint32_t t = from; t < to; ++t) {
            const uint32_t pos = c0 + t;
            if (h->cell_colours[pos] != tc) continue;
            const uint32_t cell = h->cell_inv_permutation[pos];
            compute_inc(p, cell, inc);
            if (slot_outer(p)) {
              for (uint32_t k = 0; k < md; ++k)
                for (uint32_t d = 0; d < P; ++d) {
                  const uint32_t l = h->cell_list[0][(size_t)md * c0 +
                                                     (size_t)k * ncell + t];
                  cache[(size_t)d * m + l] += inc[k * P + d];
                }
            } else {
              for (uint32_t d = 0; d < P; ++d)
                for (uint32_t k = 0; k < md; ++k) {
                  const uint32_t l = h->cell_list[0][(size_t)md * c0 +
                                                     (size_t)k * ncell + t];
                  cache[(size_t)d * m + l] += inc[k * P + d];
                }
            }
          }
        }
        for (uint32_t i = 0; i < m; ++i) {
          const uint32_t g = h->points_to_be_cached[cache0 + po + i];
          for (uint32_t d = 0; d < P; ++d) {
            const size_t gi = pidx(p->soa, n, g, P, d);
            temp[gi] = temp[gi] + cache[(size_t)d * m + i];
          }
        }
      }
    }
    memcpy(p->point_data[0], temp, len * sizeof(double));
  }
  free(cache);
  free(temp);
  return 0;
}

int oracle_loop_owner_replay(const oracle_problem *p, const oracle_oplan *o,
                             uint32_t iters) {
  const size_t len = target_len(p);
  const uint32_t md = p->mesh_dim[0], P = p->point_dim[0];
  const uint32_t n = p->num_points[0];
  double *temp = (double *)malloc(len * sizeof(double));
  uint32_t maxo = 1;
  double inc[64];
  if (!temp || md != o->mesh_dim) {
    free(temp);
    return -1;
  }
  for (uint32_t b = 0; b < o->num_blocks; ++b)
    if (o->blk_owned[b] > maxo) maxo = o->blk_owned[b];
  double *cache = (double *)malloc(sizeof(double) * (size_t)maxo * P);
  if (!cache) {
    free(temp);
    return -1;
  }
  memcpy(temp, p->point_data[0], len * sizeof(double));
  for (uint32_t it = 0; it < iters; ++it) {
    for (uint32_t b = 0; b < o->num_blocks; ++b) {
      const uint32_t no = o->blk_owned[b], c0 = o->blk_cache_off[b];
      for (uint32_t i = 0; i < no; ++i)
        for (uint32_t d = 0; d < P; ++d)
          cache[(size_t)d * no + i] = temp[pidx(p->soa, n, o->cached[c0 + i], P, d)];
      for (uint32_t tc = 0; tc < o->blk_ncol[b]; ++tc)
        for (uint32_t pos = o->blk_cell_off[b]; pos < o->blk_cell_off[b + 1]; ++pos) {
          if (o->colour[pos] != tc) continue;
          compute_inc(p, o->cells[pos], inc);
          const uint32_t *l = o->loc + (size_t)pos * md;
          if (slot_outer(p)) {
            for (uint32_t k = 0; k < md; ++k)
              for (uint32_t d = 0; d < P; ++d)
                if (l[k] < no) cache[(size_t)d * no + l[k]] += inc[k * P + d];
          } else {
            for (uint32_t d = 0; d < P; ++d)
              for (uint32_t k = 0; k < md; ++k)
                if (l[k] < no) cache[(size_t)d * no + l[k]] += inc[k * P + d];
          }
        }
      for (uint32_t i = 0; i < no; ++i) {
        const uint32_t g = o->cached[c0 + i];
        for (uint32_t d = 0; d < P; ++d)
          temp[pidx(p->soa, n, g, P, d)] = cache[(size_t)d * no + i];
      }
    }
    memcpy(p->point_data[0], temp, len * sizeof(double));
  }
  free(cache);
  free(temp);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Planning                                                                  */
/* ------------------------------------------------------------------------ */

/* mesh.hpp:117-162 */
int oracle_colour_cells(const uint32_t *c2n, uint32_t num_cells,
                        uint32_t mesh_dim, uint32_t num_points,
                        uint8_t *colour_out) {
  cset *pc = (cset *)calloc(num_points ? num_points : 1, sizeof(cset));
  uint32_t set_sizes[64] = {0};
  cset used = {{0, 0}};
  int ncol = 0;
  if (!pc) return -1;
  for (uint32_t i = 0; i < num_cells; ++i) {
    cset occ = {{0, 0}}, avail;
    unsigned colour;
    for (uint32_t j = 0; j < mesh_dim; ++j)
      occ = cset_or(occ, pc[c2n[(size_t)mesh_dim * i + j]]);
    avail = cset_andnot(used, occ);
    if (cset_none(avail)) {
      used = cset_shl1_set0(used);
      ++ncol;
      avail = cset_andnot(used, occ);
    }
    colour = ORACLE_PLANT == PLANT_GLOBAL_COLOUR_FIRST_FIT ? avail_first(avail)
                                                           : avail_min_size(avail, set_sizes, 64);
    if (colour >= 64 || ncol > 64) {
      free(pc);
      return -1; /* the reference warns "not enough colour" and overruns */
    }
    colour_out[i] = (uint8_t)colour;
    for (uint32_t j = 0; j < mesh_dim; ++j)
      cset_set(&pc[c2n[(size_t)mesh_dim * i + j]], colour);
    ++set_sizes[colour];
  }
  free(pc);
  return ncol;
}

/* getCellToCell (partition.cu:281-301) as CSR: for each cell, for each slot,
 * the other cells on that point in ascending order (std::multimap equal_range
 * keeps insertion order); duplicates kept. */
static void cell_to_cell_local(const uint32_t *mapping, uint32_t nc,
                               uint32_t md, uint32_t **adj_off_out,
                               uint32_t **adj_out) {
  const uint32_t len = nc * md;
  uint32_t maxp = 0;
  for (uint32_t k = 0; k < len; ++k)
    if (mapping[k] + 1 > maxp) maxp = mapping[k] + 1;
  uint32_t *poff = (uint32_t *)calloc(maxp + 1, sizeof(uint32_t));
  uint32_t *pcell = (uint32_t *)malloc(sizeof(uint32_t) * (len ? len : 1));
  uint32_t *fill = (uint32_t *)malloc(sizeof(uint32_t) * (maxp + 1));
  for (uint32_t k = 0; k < len; ++k) poff[mapping[k] + 1]++;
  for (uint32_t q = 0; q < maxp; ++q) poff[q + 1] += poff[q];
  memcpy(fill, poff, sizeof(uint32_t) * (maxp + 1));
  for (uint32_t k = 0; k < len; ++k) pcell[fill[mapping[k]]++] = k / md;
  uint32_t *aoff = (uint32_t *)calloc(nc + 1, sizeof(uint32_t));
  for (uint32_t i = 0; i < nc; ++i) {
    uint32_t cnt = 0;
    for (uint32_t o = 0; o < md; ++o) {
      const uint32_t pt = mapping[md * i + o];
      for (uint32_t r = poff[pt]; r < poff[pt + 1]; ++r)
        if (pcell[r] != i) ++cnt;
    }
    aoff[i + 1] = aoff[i] + cnt;
  }
  uint32_t *adj = (uint32_t *)malloc(sizeof(uint32_t) * (aoff[nc] ? aoff[nc] : 1));
  for (uint32_t i = 0; i < nc; ++i) {
    uint32_t w = aoff[i];
    for (uint32_t o = 0; o < md; ++o) {
      const uint32_t pt = mapping[md * i + o];
      for (uint32_t r = poff[pt]; r < poff[pt + 1]; ++r)
        if (pcell[r] != i) adj[w++] = pcell[r];
    }
  }
  free(poff), free(pcell), free(fill);
  *adj_off_out = aoff;
  *adj_out = adj;
}

/* colouring.cu:12-102 */
int oracle_colour_cells_with_ordering(const uint32_t *mapping,
                                      uint32_t num_cells, uint32_t mapping_dim,
                                      uint8_t *colours_out) {
  uint32_t *aoff, *adj;
  const uint32_t nc = num_cells;
  if (nc == 0) return 0;
  cell_to_cell_local(mapping, nc, mapping_dim, &aoff, &adj);
  /* getCellOrder: smallest-last, ties -> last index (colouring.cu:20-32) */
  uint32_t *deg = (uint32_t *)malloc(sizeof(uint32_t) * nc);
  uint8_t *done = (uint8_t *)calloc(nc, 1);
  uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * nc);
  for (uint32_t i = 0; i < nc; ++i) deg[i] = aoff[i + 1] - aoff[i];
  for (uint32_t it = 0; it < nc; ++it) {
    uint32_t mi = nc;
    for (uint32_t j = 0; j < nc; ++j) {
      if (done[j]) continue;
      if (mi == nc || deg[mi] >= deg[j]) mi = j;
    }
    done[mi] = 1;
    order[it] = mi;
    /* each occurrence of a neighbour decrements its degree once */
    for (uint32_t r = aoff[mi]; r < aoff[mi + 1]; ++r) --deg[adj[r]];
  }
  /* colourCellToCell: reverse order, first-fit (colouring.cu:63-92) */
  cset *cc = (cset *)calloc(nc, sizeof(cset));
  cset used = {{0, 0}};
  int ncol = 0;
  for (uint32_t i = nc; i-- > 0;) {
    const uint32_t v = order[i];
    cset occ = {{0, 0}}, avail;
    for (uint32_t r = aoff[v]; r < aoff[v + 1]; ++r) occ = cset_or(occ, cc[adj[r]]);
    avail = cset_andnot(used, occ);
    if (cset_none(avail)) {
      if (ncol >= 128) {
        ncol = -1;
        break;
      }
      cset_set(&used, (unsigned)ncol++);
      avail = cset_andnot(used, occ);
    }
    const unsigned colour = avail_first(avail);
    colours_out[v] = (uint8_t)colour;
    cset_set(&cc[v], colour);
  }
  free(deg), free(done), free(order), free(cc), free(aoff), free(adj);
  return ncol;
}

typedef struct {
  uint32_t point, local_cell, slot;
} pt_ref;

static int cmp_ptref(const void *a, const void *b) {
  const pt_ref *x = (const pt_ref *)a, *y = (const pt_ref *)b;
  const int asc = ORACLE_PLANT == PLANT_CACHE_DESCENDING ? 1 : -1;
  if (x->point != y->point) return x->point < y->point ? asc : -asc;
  if (x->local_cell != y->local_cell) return x->local_cell < y->local_cell ? -1 : 1;
  if (x->slot != y->slot) return x->slot < y->slot ? -1 : 1;
  return 0;
}

typedef struct {
  uint64_t key; /* partition << 32 | cell */
} pc_pair;
static int cmp_u64(const void *a, const void *b) {
  const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* HierarchicalColourMemory ctor (colouring.hpp:44-147), colourBlock
 * (:169-235), sortCellsByColours (:246-272), cellListSOA (:300-308). */
int oracle_hplan_build(const uint32_t *const *c2n, const uint32_t *mesh_dim,
                       uint32_t num_mappings, uint32_t num_cells,
                       uint32_t num_points0, const uint32_t *partition,
                       uint32_t block_size, oracle_hplan *out) {
  const uint32_t nc = num_cells, md0 = mesh_dim[0];
  uint32_t *pos_cell = (uint32_t *)malloc(sizeof(uint32_t) * nc); /* cwip */
  uint32_t *blocks = (uint32_t *)malloc(sizeof(uint32_t) * (nc + 2));
  uint32_t nblk = 0;
  memset(out, 0, sizeof(*out));
  if (nc == 0 || block_size == 0) return -1;
  if (partition) {
    uint64_t *pairs = (uint64_t *)malloc(sizeof(uint64_t) * nc);
    for (uint32_t i = 0; i < nc; ++i)
      pairs[i] = ((uint64_t)partition[i] << 32) | i;
    qsort(pairs, nc, sizeof(uint64_t), cmp_u64);
    uint32_t tid = 0;
    blocks[nblk++] = 0;
    pos_cell[0] = (uint32_t)pairs[0];
    for (uint32_t i = 1; i < nc; ++i) {
      pos_cell[i] = (uint32_t)pairs[i];
      if (++tid == block_size) {
        tid = 0;
        blocks[nblk++] = i;
        continue;
      }
      if ((pairs[i - 1] >> 32) != (pairs[i] >> 32)) {
        if (ORACLE_PLANT != PLANT_PART_CHANGE_KEEPS_TID) tid = 0;
        blocks[nblk++] = i;
      }
    }
    blocks[nblk++] = nc;
    free(pairs);
  } else {
    for (uint32_t i = 0; i < nc; i += block_size) blocks[nblk++] = i;
    blocks[nblk++] = nc;
    for (uint32_t i = 0; i < nc; ++i) pos_cell[i] = i;
  }
  const uint32_t nb = nblk - 1;
  /* pass 1: block colours, thread colouring, sorting (per block) */
  cset *pcol = (cset *)calloc(num_points0 ? num_points0 : 1, sizeof(cset));
  cset used = {{0, 0}};
  uint32_t set_sizes[128] = {0};
  uint32_t ncol = 0;
  uint8_t *bcol = (uint8_t *)malloc(nb ? nb : 1);
  uint8_t *tcol = (uint8_t *)malloc(nc);        /* by position */
  uint8_t *ntc = (uint8_t *)malloc(nb ? nb : 1);
  uint32_t *local = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nc * md0);
  uint32_t *nuniq = (uint32_t *)malloc(sizeof(uint32_t) * (nb ? nb : 1));
  uint32_t *uniq = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nc * md0);
  pt_ref *refs = (pt_ref *)malloc(sizeof(pt_ref) * (size_t)block_size * md0);
  uint32_t *tmp_local = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)block_size * md0);
  uint8_t *tmp_col = (uint8_t *)malloc(block_size);
  uint32_t *perm = (uint32_t *)malloc(sizeof(uint32_t) * block_size);
  uint32_t *tmp_pc = (uint32_t *)malloc(sizeof(uint32_t) * block_size);
  uint32_t uniq_total = 0;
  int err = 0;
  for (uint32_t b = 0; b < nb && !err; ++b) {
    const uint32_t from = blocks[b], to = blocks[b + 1], bs = to - from;
    cset occ = {{0, 0}}, avail;
    if (bs > block_size) {
      err = 1;
      break;
    }
    for (uint32_t i = from; i < to; ++i)
      for (uint32_t j = 0; j < md0; ++j)
        occ = cset_or(occ, pcol[c2n[0][(size_t)md0 * pos_cell[i] + j]]);
    avail = cset_andnot(used, occ);
    if (cset_none(avail)) {
      if (ncol >= 128) {
        err = 1;
        break;
      }
      used = cset_shl1_set0(used);
      set_sizes[ncol++] = 0;
      avail = cset_andnot(used, occ);
    }
    const unsigned colour = ORACLE_PLANT == PLANT_BLOCK_COLOUR_FIRST_FIT
                                ? avail_first(avail)
                                : avail_min_size(avail, set_sizes, ncol);
    ++set_sizes[colour];
    bcol[b] = (uint8_t)colour;
    /* colourBlock: colour written points, sorted unique points, local ids */
    uint32_t nr = 0;
    for (uint32_t i = from; i < to; ++i)
      for (uint32_t j = 0; j < md0; ++j) {
        const uint32_t pt = c2n[0][(size_t)md0 * pos_cell[i] + j];
        cset_set(&pcol[pt], colour);
        refs[nr].point = pt;
        refs[nr].local_cell = i - from;
        refs[nr].slot = j;
        ++nr;
      }
    qsort(refs, nr, sizeof(pt_ref), cmp_ptref);
    uint32_t nu = 0;
    for (uint32_t r = 0; r < nr; ++r) {
      if (r == 0 || refs[r].point != refs[r - 1].point) uniq[uniq_total + nu++] = refs[r].point;
      tmp_local[md0 * refs[r].local_cell + refs[r].slot] = nu - 1;
    }
    nuniq[b] = nu;
    uniq_total += nu;
    int nt = oracle_colour_cells_with_ordering(tmp_local, bs, md0, tmp_col);
    if (nt < 0) {
      err = 1;
      break;
    }
    ntc[b] = (uint8_t)nt;
    /* sortCellsByColours: stable by (colour, local index) */
    {
      uint32_t cnt[257] = {0};
      for (uint32_t i = 0; i < bs; ++i) cnt[tmp_col[i] + 1]++;
      for (int k = 0; k < 256; ++k) cnt[k + 1] += cnt[k];
      for (uint32_t i = 0; i < bs; ++i) perm[cnt[tmp_col[i]]++] = i;
      if (ORACLE_PLANT == PLANT_THREAD_ORDER_REVERSED)
        for (uint32_t k = 0, lo = 0; k < 256 && lo < bs; lo = cnt[k++])
          for (uint32_t a = lo, z = cnt[k]; a + 1 < z; ++a, --z) {
            const uint32_t t = perm[a];
            perm[a] = perm[z - 1];
            perm[z - 1] = t;
          }
    }
    for (uint32_t i = 0; i < bs; ++i) {
      tcol[from + i] = tmp_col[perm[i]];
      tmp_pc[i] = pos_cell[from + perm[i]];
      for (uint32_t j = 0; j < md0; ++j)
        local[(size_t)md0 * (from + i) + j] = tmp_local[md0 * perm[i] + j];
    }
    memcpy(pos_cell + from, tmp_pc, sizeof(uint32_t) * bs);
  }
  free(refs), free(tmp_local), free(tmp_col), free(perm), free(tmp_pc), free(pcol);
  if (err) {
    free(pos_cell), free(blocks), free(bcol), free(tcol), free(ntc), free(local);
    free(nuniq), free(uniq);
    return -1;
  }
  /* pass 2: group by block colour (blocks keep their order within a colour) */
  out->num_colours = ncol;
  out->num_mappings = num_mappings;
  for (uint32_t m = 0; m < num_mappings; ++m) out->mesh_dim[m] = mesh_dim[m];
  out->num_cells = nc;
  out->num_blocks = nb;
  out->num_cached = uniq_total;
  out->colour_cell_offset = (uint32_t *)calloc(ncol + 1, sizeof(uint32_t));
  out->colour_block_offset = (uint32_t *)calloc(ncol + 1, sizeof(uint32_t));
  out->colour_cache_offset = (uint32_t *)calloc(ncol + 1, sizeof(uint32_t));
  uint32_t *uniq_off = (uint32_t *)malloc(sizeof(uint32_t) * (nb + 1));
  uniq_off[0] = 0;
  for (uint32_t b = 0; b < nb; ++b) {
    uniq_off[b + 1] = uniq_off[b] + nuniq[b];
    out->colour_cell_offset[bcol[b] + 1] += blocks[b + 1] - blocks[b];
    out->colour_block_offset[bcol[b] + 1] += 1;
    out->colour_cache_offset[bcol[b] + 1] += nuniq[b];
  }
  for (uint32_t c = 0; c < ncol; ++c) {
    out->colour_cell_offset[c + 1] += out->colour_cell_offset[c];
    out->colour_block_offset[c + 1] += out->colour_block_offset[c];
    out->colour_cache_offset[c + 1] += out->colour_cache_offset[c];
  }
  for (uint32_t m = 0; m < num_mappings; ++m)
    out->cell_list[m] = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nc * mesh_dim[m]);
  out->cell_colours = (uint8_t *)malloc(nc);
  out->num_cell_colours = (uint8_t *)malloc(nb);
  out->block_offsets = (uint32_t *)malloc(sizeof(uint32_t) * (nb + ncol));
  out->points_to_be_cached = (uint32_t *)malloc(sizeof(uint32_t) * (uniq_total ? uniq_total : 1));
  out->points_to_be_cached_offsets = (uint32_t *)malloc(sizeof(uint32_t) * (nb + ncol));
  out->cell_inv_permutation = (uint32_t *)malloc(sizeof(uint32_t) * nc);
  {
    uint32_t blk = 0;
    for (uint32_t c = 0; c < ncol; ++c) {
      const uint32_t c0 = out->colour_cell_offset[c];
      const uint32_t ncell = out->colour_cell_offset[c + 1] - c0;
      const uint32_t bo = out->colour_block_offset[c] + c;
      uint32_t cell_w = 0, cache_w = 0, k = 0;
      out->block_offsets[bo] = 0;
      out->points_to_be_cached_offsets[bo] = 0;
      for (uint32_t b = 0; b < nb; ++b) {
        if (bcol[b] != c) continue;
        const uint32_t from = blocks[b], to = blocks[b + 1];
        for (uint32_t i = from; i < to; ++i, ++cell_w) {
          const uint32_t pos = c0 + cell_w;
          const uint32_t cell = pos_cell[i];
          out->cell_inv_permutation[pos] = cell;
          out->cell_colours[pos] = tcol[i];
          for (uint32_t j = 0; j < md0; ++j)
            out->cell_list[0][(size_t)md0 * c0 + (size_t)j * ncell + cell_w] =
                local[(size_t)md0 * i + j];
          for (uint32_t m = 1; m < num_mappings; ++m) {
            const uint32_t mdm = mesh_dim[m];
            for (uint32_t j = 0; j < mdm; ++j)
              out->cell_list[m][(size_t)mdm * c0 + (size_t)j * ncell + cell_w] =
                  c2n[m][(size_t)mdm * cell + j];
          }
        }
        memcpy(out->points_to_be_cached + out->colour_cache_offset[c] + cache_w,
               uniq + uniq_off[b], sizeof(uint32_t) * nuniq[b]);
        cache_w += nuniq[b];
        out->num_cell_colours[out->colour_block_offset[c] + k] = ntc[b];
        ++k;
        out->block_offsets[bo + k] = cell_w;
        out->points_to_be_cached_offsets[bo + k] = cache_w;
        ++blk;
      }
    }
    (void)blk;
  }
  free(uniq_off), free(pos_cell), free(blocks), free(bcol), free(tcol), free(ntc);
  free(local), free(nuniq), free(uniq);
  return 0;
}

void oracle_hplan_free(oracle_hplan *p) {
  if (!p) return;
  free(p->colour_cell_offset), free(p->colour_block_offset);
  free(p->colour_cache_offset);
  for (uint32_t m = 0; m < ORACLE_MAX_MAPPINGS; ++m) free(p->cell_list[m]);
  free(p->cell_colours), free(p->num_cell_colours), free(p->block_offsets);
  free(p->points_to_be_cached), free(p->points_to_be_cached_offsets);
  free(p->cell_inv_permutation);
  memset(p, 0, sizeof(*p));
}

/* mesh.hpp:227-245 */
void oracle_reorder_to_partition(uint32_t *partition, uint32_t num_cells,
                                 uint32_t *inv_perm_out) {
  uint64_t *pairs = (uint64_t *)malloc(sizeof(uint64_t) * (num_cells ? num_cells : 1));
  for (uint32_t i = 0; i < num_cells; ++i)
    pairs[i] = ((uint64_t)partition[i] << 32) | i;
  qsort(pairs, num_cells, sizeof(uint64_t), cmp_u64);
  for (uint32_t i = 0; i < num_cells; ++i) {
    partition[i] = (uint32_t)(pairs[i] >> 32);
    inv_perm_out[i] = (uint32_t)pairs[i];
  }
  free(pairs);
}

/* Sorted unique partition ids per point (mesh.hpp:346-363), CSR. */
typedef struct {
  const uint32_t *off, *ids;
} ptpart;
static const ptpart *g_ptpart; /* qsort has no context argument */

/* Comparator of getPointRenumberingPermutation2 (mesh.hpp:373-385) made a
 * strict weak order: more partitions first, then lexicographically larger
 * set first; two untouched points compare equal (the reference's lambda
 * returns true both ways for them, leaving their order to the library). */
static int cmp_ptpart(const void *a, const void *b) {
  const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  const uint32_t nx = g_ptpart->off[x + 1] - g_ptpart->off[x];
  const uint32_t ny = g_ptpart->off[y + 1] - g_ptpart->off[y];
  if (nx != ny) return nx > ny ? -1 : 1;
  for (uint32_t k = 0; k < nx; ++k) {
    const uint32_t u = g_ptpart->ids[g_ptpart->off[x] + k];
    const uint32_t v = g_ptpart->ids[g_ptpart->off[y] + k];
    if (u != v) return u > v ? -1 : 1;
  }
  return x < y ? -1 : (x > y ? 1 : 0); /* stability */
}

void oracle_point_renumbering_by_partition(const uint32_t *c2n,
                                           uint32_t num_cells,
                                           uint32_t mesh_dim,
                                           uint32_t num_points,
                                           const uint32_t *partition,
                                           uint32_t *perm_out) {
  const size_t len = (size_t)num_cells * mesh_dim;
  uint64_t *pairs = (uint64_t *)malloc(sizeof(uint64_t) * (len ? len : 1));
  uint32_t *off = (uint32_t *)calloc(num_points + 1, sizeof(uint32_t));
  uint32_t *ids = (uint32_t *)malloc(sizeof(uint32_t) * (len ? len : 1));
  uint32_t *inv = (uint32_t *)malloc(sizeof(uint32_t) * (num_points ? num_points : 1));
  size_t n = 0;
  for (uint32_t i = 0; i < num_cells; ++i)
    for (uint32_t j = 0; j < mesh_dim; ++j)
      pairs[n++] = ((uint64_t)c2n[(size_t)mesh_dim * i + j] << 32) | partition[i];
  qsort(pairs, n, sizeof(uint64_t), cmp_u64);
  size_t w = 0;
  for (size_t r = 0; r < n; ++r) {
    if (r > 0 && pairs[r] == pairs[r - 1]) continue;
    ids[w++] = (uint32_t)pairs[r];
    off[(pairs[r] >> 32) + 1]++;
  }
  for (uint32_t q = 0; q < num_points; ++q) off[q + 1] += off[q];
  ptpart pp = {off, ids};
  g_ptpart = &pp;
  for (uint32_t q = 0; q < num_points; ++q) inv[q] = q;
  qsort(inv, num_points, sizeof(uint32_t), cmp_ptpart);
  for (uint32_t i = 0; i < num_points; ++i) perm_out[inv[i]] = i;
  free(pairs), free(off), free(ids), free(inv);
}

/* mesh.hpp:247-270 */
void oracle_point_renumbering_first_touch(const uint32_t *c2n,
                                          uint32_t num_cells,
                                          uint32_t mesh_dim,
                                          uint32_t num_points,
                                          uint32_t *perm) {
  uint32_t next = 0;
  for (uint32_t q = 0; q < num_points; ++q) perm[q] = num_points;
  for (size_t k = 0; k < (size_t)num_cells * mesh_dim; ++k)
    if (perm[c2n[k]] == num_points) perm[c2n[k]] = next++;
  for (uint32_t q = 0; q < num_points; ++q)
    if (perm[q] == num_points) perm[q] = next++;
}

static uint32_t g_md;
static const uint32_t *g_tuples;
static int cmp_tuple(const void *a, const void *b) {
  const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  for (uint32_t k = 0; k < g_md; ++k) {
    const uint32_t u = g_tuples[(size_t)(g_md) * x + k];
    const uint32_t v = g_tuples[(size_t)(g_md) * y + k];
    if (u != v) return u < v ? -1 : 1;
  }
  return x < y ? -1 : (x > y ? 1 : 0);
}
static int cmp_u32(const void *a, const void *b) {
  const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* mesh.hpp:199-225 (mapping 0 part) */
void oracle_mesh_reorder_cells(uint32_t *c2n0, uint32_t num_cells,
                               uint32_t mesh_dim, const uint32_t *point_perm,
                               uint32_t *inv_perm_out) {
  const size_t len = (size_t)num_cells * mesh_dim;
  uint32_t *tuples = (uint32_t *)malloc(sizeof(uint32_t) * (len ? len : 1));
  uint32_t *old = (uint32_t *)malloc(sizeof(uint32_t) * (len ? len : 1));
  for (size_t k = 0; k < len; ++k) c2n0[k] = point_perm[c2n0[k]];
  memcpy(tuples, c2n0, sizeof(uint32_t) * len);
  for (uint32_t i = 0; i < num_cells; ++i)
    qsort(tuples + (size_t)mesh_dim * i, mesh_dim, sizeof(uint32_t), cmp_u32);
  g_md = mesh_dim;
  g_tuples = tuples;
  for (uint32_t i = 0; i < num_cells; ++i) inv_perm_out[i] = i;
  qsort(inv_perm_out, num_cells, sizeof(uint32_t), cmp_tuple);
  memcpy(old, c2n0, sizeof(uint32_t) * len);
  for (uint32_t i = 0; i < num_cells; ++i)
    memcpy(c2n0 + (size_t)mesh_dim * i, old + (size_t)mesh_dim * inv_perm_out[i],
           sizeof(uint32_t) * mesh_dim);
  free(tuples), free(old);
}

/* reorder.hpp:39-83: points in ascending order; per point its incident
 * cells in ascending order (multimap insertion order); neighbour appended
 * unless already present in this point's row. */
uint32_t oracle_graph_csr(const uint32_t *c2n, uint32_t num_cells,
                          uint32_t mesh_dim, uint32_t num_points,
                          uint32_t *indices, uint32_t *endpoints) {
  const size_t len = (size_t)num_cells * mesh_dim;
  uint32_t *poff = (uint32_t *)calloc(num_points + 1, sizeof(uint32_t));
  uint32_t *pcell = (uint32_t *)malloc(sizeof(uint32_t) * (len ? len : 1));
  uint32_t *fill = (uint32_t *)malloc(sizeof(uint32_t) * (num_points + 1));
  uint32_t arcs = 0;
  for (size_t k = 0; k < len; ++k) poff[c2n[k] + 1]++;
  for (uint32_t q = 0; q < num_points; ++q) poff[q + 1] += poff[q];
  memcpy(fill, poff, sizeof(uint32_t) * (num_points + 1));
  for (size_t k = 0; k < len; ++k) pcell[fill[c2n[k]]++] = (uint32_t)(k / mesh_dim);
  indices[0] = 0;
  for (uint32_t q = 0; q < num_points; ++q) {
    const uint32_t row = arcs;
    for (uint32_t r = poff[q]; r < poff[q + 1]; ++r) {
      const uint32_t cell = pcell[r];
      for (uint32_t j = 0; j < mesh_dim; ++j) {
        const uint32_t other = c2n[(size_t)mesh_dim * cell + j];
        if (other == q) continue;
        int found = 0;
        for (uint32_t s = row; s < arcs; ++s)
          if (endpoints[s] == other) {
            found = 1;
            break;
          }
        if (!found) endpoints[arcs++] = other;
      }
    }
    indices[q + 1] = arcs;
  }
  free(poff), free(pcell), free(fill);
  return arcs;
}

/* mesh.hpp:317-344 */
uint64_t oracle_cell_to_cell_graph(const uint32_t *const *c2n,
                                   const uint32_t *mesh_dim,
                                   uint32_t num_mappings, uint32_t num_cells,
                                   const uint32_t *num_points,
                                   uint32_t *pairs_out, uint64_t max_pairs) {
  uint64_t cap = 1024, n = 0;
  uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * cap);
  for (uint32_t m = 0; m < num_mappings; ++m) {
    const uint32_t md = mesh_dim[m], np = num_points[m];
    const size_t len = (size_t)num_cells * md;
    uint32_t *poff = (uint32_t *)calloc(np + 1, sizeof(uint32_t));
    uint32_t *pcell = (uint32_t *)malloc(sizeof(uint32_t) * (len ? len : 1));
    uint32_t *fill = (uint32_t *)malloc(sizeof(uint32_t) * (np + 1));
    for (size_t k = 0; k < len; ++k) poff[c2n[m][k] + 1]++;
    for (uint32_t q = 0; q < np; ++q) poff[q + 1] += poff[q];
    memcpy(fill, poff, sizeof(uint32_t) * (np + 1));
    for (size_t k = 0; k < len; ++k) pcell[fill[c2n[m][k]]++] = (uint32_t)(k / md);
    for (uint32_t i = 0; i < num_cells; ++i)
      for (uint32_t o = 0; o < md; ++o) {
        const uint32_t pt = c2n[m][(size_t)md * i + o];
        for (uint32_t r = poff[pt]; r < poff[pt + 1]; ++r)
          if (pcell[r] > i) {
            if (n == cap) {
              cap *= 2;
              keys = (uint64_t *)realloc(keys, sizeof(uint64_t) * cap);
            }
            keys[n++] = ((uint64_t)i << 32) | pcell[r];
          }
      }
    free(poff), free(pcell), free(fill);
  }
  qsort(keys, n, sizeof(uint64_t), cmp_u64);
  uint64_t w = 0;
  for (uint64_t r = 0; r < n; ++r) {
    if (r > 0 && keys[r] == keys[r - 1]) continue;
    if (w < max_pairs) {
      pairs_out[2 * w] = (uint32_t)(keys[r] >> 32);
      pairs_out[2 * w + 1] = (uint32_t)keys[r];
    }
    ++w;
  }
  free(keys);
  return w;
}
