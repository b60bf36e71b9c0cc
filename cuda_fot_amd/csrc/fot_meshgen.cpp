// fot_meshgen.cpp -- synthetic meshes and data for the loop engine.
// All randomness is SplitMix64 (Steele, Lea, Flood 2014); each stream is
// seeded from (seed, stream id) so outputs do not depend on call order.
#include <cmath>
#include <numeric>
#include <vector>

#include "fot_internal.hpp"

using fot::fail;

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return double(next() >> 11) * 0x1.0p-53; } // [0,1)
  // Box-Muller, one normal per pair of uniforms
  double normal() {
    const double u1 = 1.0 - uniform(); // (0,1]
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  uint32_t below(uint32_t n) { return uint32_t((uint64_t(uint32_t(next() >> 32)) * n) >> 32); }
};

uint64_t stream_seed(uint64_t seed, uint64_t stream) {
  SplitMix64 g(seed ^ (0x632be59bd9b4e019ull * (stream + 1)));
  return g.next();
}

void permutation(uint32_t n, uint64_t seed, uint32_t *perm) {
  std::iota(perm, perm + n, 0u);
  SplitMix64 g(seed);
  for (uint32_t i = n; i > 1; --i) {
    const uint32_t j = g.below(i);
    std::swap(perm[i - 1], perm[j]);
  }
}

} // namespace

static fot_status fot_gen_permutation_impl(uint32_t n, uint64_t seed, uint32_t *perm) {
  if (!perm && n) return fail(FOT_ERR_INVALID_ARGUMENT, "null perm");
  permutation(n, seed, perm);
  return FOT_OK;
}
extern "C" fot_status fot_gen_permutation(uint32_t n, uint64_t seed, uint32_t *perm) {
  return fot::guarded("fot_gen_permutation", [&] {
    return fot_gen_permutation_impl(n, seed, perm);
  });
}

static fot_status fot_gen_uniform_impl(uint64_t n, uint64_t seed, double lo, double hi,
                                      double *out) {
  if (!out && n) return fail(FOT_ERR_INVALID_ARGUMENT, "null out");
  SplitMix64 g(seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * g.uniform();
  return FOT_OK;
}
extern "C" fot_status fot_gen_uniform(uint64_t n, uint64_t seed, double lo, double hi,
                                      double *out) {
  return fot::guarded("fot_gen_uniform", [&] {
    return fot_gen_uniform_impl(n, seed, lo, hi, out);
  });
}

static fot_status fot_gen_normal_impl(uint64_t n, uint64_t seed, double mean, double sd,
                                     double *out) {
  if (!out && n) return fail(FOT_ERR_INVALID_ARGUMENT, "null out");
  SplitMix64 g(seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = mean + sd * g.normal();
  return FOT_OK;
}
extern "C" fot_status fot_gen_normal(uint64_t n, uint64_t seed, double mean, double sd,
                                     double *out) {
  return fot::guarded("fot_gen_normal", [&] {
    return fot_gen_normal_impl(n, seed, mean, sd, out);
  });
}

// Grid::fillEdgeList (grid.hpp:77-104): for each row but the last, a
// (down, right) edge pair per column but the last, then the last column's
// down edge (listed lower point first); finally the last row's right edges.
static fot_status fot_gen_grid_edges_impl(uint32_t N, uint32_t M, uint32_t *e) {
  if (!e || N < 2 || M < 2) return fail(FOT_ERR_INVALID_ARGUMENT, "grid needs N,M >= 2");
  size_t k = 0;
  uint32_t upper = 0, lower = M;
  for (uint32_t r = 0; r + 1 < N; ++r) {
    for (uint32_t c = 0; c + 1 < M; ++c) {
      e[k++] = upper;
      e[k++] = lower;
      e[k++] = upper;
      e[k++] = ++upper;
      ++lower;
    }
    e[k++] = lower++;
    e[k++] = upper++;
  }
  for (uint32_t c = 0; c + 1 < M; ++c) {
    e[k++] = upper;
    e[k++] = ++upper;
  }
  return FOT_OK;
}
extern "C" fot_status fot_gen_grid_edges(uint32_t N, uint32_t M, uint32_t *e) {
  return fot::guarded("fot_gen_grid_edges", [&] { return fot_gen_grid_edges_impl(N, M, e); });
}

// Grid::fillCellList3D (grid.hpp:415-460): corner order (0,0,0) (0,1,0)
// (1,1,0) (1,0,0) then the same on the layer above.
static fot_status fot_gen_hex_grid_impl(uint32_t N1, uint32_t N2, uint32_t N3,
                                       uint32_t *cells) {
  if (!cells || N1 < 2 || N2 < 2 || N3 < 2)
    return fail(FOT_ERR_INVALID_ARGUMENT, "hex grid needs N >= 2");
  size_t k = 0;
  for (uint32_t l = 0; l + 1 < N3; ++l)
    for (uint32_t r = 0; r + 1 < N1; ++r)
      for (uint32_t c = 0; c + 1 < N2; ++c) {
        const uint32_t bot = l * N1 * N2, top = (l + 1) * N1 * N2;
        const uint32_t in = r * N2 + c, out = (r + 1) * N2 + c;
        cells[k++] = bot + in;
        cells[k++] = bot + in + 1;
        cells[k++] = bot + out + 1;
        cells[k++] = bot + out;
        cells[k++] = top + in;
        cells[k++] = top + in + 1;
        cells[k++] = top + out + 1;
        cells[k++] = top + out;
      }
  return FOT_OK;
}
extern "C" fot_status fot_gen_hex_grid(uint32_t N1, uint32_t N2, uint32_t N3,
                                       uint32_t *cells) {
  return fot::guarded("fot_gen_hex_grid", [&] { return fot_gen_hex_grid_impl(N1, N2, N3, cells); });
}

// Grid::fillCellList (grid.hpp:386-411): quads of an N x M point grid
static fot_status fot_gen_quad_grid_impl(uint32_t N, uint32_t M, uint32_t *cells) {
  if (!cells || N < 2 || M < 2) return fail(FOT_ERR_INVALID_ARGUMENT, "quad grid needs N,M >= 2");
  size_t k = 0;
  for (uint32_t r = 0; r + 1 < N; ++r)
    for (uint32_t c = 0; c + 1 < M; ++c) {
      const uint32_t top_left = r * M + c, bottom_left = (r + 1) * M + c;
      cells[k++] = top_left;
      cells[k++] = top_left + 1;
      cells[k++] = bottom_left;
      cells[k++] = bottom_left + 1;
    }
  return FOT_OK;
}
extern "C" fot_status fot_gen_quad_grid(uint32_t N, uint32_t M, uint32_t *cells) {
  return fot::guarded("fot_gen_quad_grid", [&] { return fot_gen_quad_grid_impl(N, M, cells); });
}

// Airfoil-like structured quad mesh with interior edges only.  Edge order
// before shuffling follows the cells row by row: for cell (i, j) the edge
// below it (j > 0), then the edge left of it (i > 0).
static fot_status fot_gen_airfoil_impl(uint32_t nx, uint32_t ny, uint64_t seed,
                                      uint32_t shuffle, uint32_t *pedge,
                                      uint32_t *pecell, double *x, double *q,
                                      double *adt) {
  if (nx < 2 || ny < 2) return fail(FOT_ERR_INVALID_ARGUMENT, "airfoil needs nx,ny >= 2");
  const uint32_t nn = (nx + 1) * (ny + 1), nc = nx * ny;
  const uint32_t ne = nx * (ny - 1) + (nx - 1) * ny;
  std::vector<uint32_t> node_perm(nn), cell_perm(nc), edge_perm(ne);
  if (shuffle & 1u) permutation(nn, stream_seed(seed, 1), node_perm.data());
  else std::iota(node_perm.begin(), node_perm.end(), 0u);
  if (shuffle & 2u) permutation(nc, stream_seed(seed, 2), cell_perm.data());
  else std::iota(cell_perm.begin(), cell_perm.end(), 0u);
  if (shuffle & 4u) permutation(ne, stream_seed(seed, 3), edge_perm.data());
  else std::iota(edge_perm.begin(), edge_perm.end(), 0u);
  auto node = [&](uint32_t i, uint32_t j) { return node_perm[j * (nx + 1) + i]; };
  auto cell = [&](uint32_t i, uint32_t j) { return cell_perm[j * nx + i]; };
  uint32_t k = 0;
  for (uint32_t j = 0; j < ny; ++j)
    for (uint32_t i = 0; i < nx; ++i) {
      if (j > 0) { // horizontal edge, normal +y: left node first
        const uint32_t e = edge_perm[k++];
        if (pedge) {
          pedge[2 * e] = node(i, j);
          pedge[2 * e + 1] = node(i + 1, j);
        }
        if (pecell) {
          pecell[2 * e] = cell(i, j - 1);
          pecell[2 * e + 1] = cell(i, j);
        }
      }
      if (i > 0) { // vertical edge, normal +x: upper node first
        const uint32_t e = edge_perm[k++];
        if (pedge) {
          pedge[2 * e] = node(i, j + 1);
          pedge[2 * e + 1] = node(i, j);
        }
        if (pecell) {
          pecell[2 * e] = cell(i - 1, j);
          pecell[2 * e + 1] = cell(i, j);
        }
      }
    }
  if (x) {
    SplitMix64 g(stream_seed(seed, 4));
    for (uint32_t j = 0; j <= ny; ++j)
      for (uint32_t i = 0; i <= nx; ++i) {
        const uint32_t n = node(i, j);
        const double dx = -0.2 + 0.4 * g.uniform();
        const double dy = -0.2 + 0.4 * g.uniform();
        x[2 * n] = double(i) + dx;
        x[2 * n + 1] = double(j) + dy;
      }
  }
  if (q) {
    static const double freestream[4] = {1.0, 0.4, 0.0, 2.5};
    SplitMix64 g(stream_seed(seed, 5));
    for (uint32_t c = 0; c < nc; ++c) {
      const uint32_t cp = cell_perm[c];
      for (int d = 0; d < 4; ++d) q[4 * size_t(cp) + d] = freestream[d] * (1.0 + 1e-2 * g.normal());
    }
  }
  if (adt) {
    SplitMix64 g(stream_seed(seed, 6));
    for (uint32_t c = 0; c < nc; ++c) adt[cell_perm[c]] = 0.5 + g.uniform();
  }
  return FOT_OK;
}
extern "C" fot_status fot_gen_airfoil(uint32_t nx, uint32_t ny, uint64_t seed,
                                      uint32_t shuffle, uint32_t *pedge,
                                      uint32_t *pecell, double *x, double *q,
                                      double *adt) {
  return fot::guarded("fot_gen_airfoil", [&] {
    return fot_gen_airfoil_impl(nx, ny, seed, shuffle, pedge, pecell, x, q, adt);
  });
}

namespace {
struct Pt {
  double x, y;
};
} // namespace

// Triangle mesh of the shallow-water edge loop (volna, kernels/volna.hpp):
// an (nx+1) x (ny+1) node grid perturbed by U(-0.2, 0.2), every quad split
// along a random diagonal.  Per quad (i, j) with corners a=(i,j), b=(i+1,j),
// c=(i+1,j+1), d=(i,j+1): diagonal a-c gives T0=(a,b,c), T1=(a,c,d);
// diagonal b-d gives T0=(a,b,d), T1=(b,c,d).  Edges in order: horizontal,
// vertical, diagonal.  An interior edge maps to (left, right) with its unit
// normal pointing from left to right; a boundary edge maps to (cell, cell)
// with is_boundary 1 and the outward normal.
static fot_status fot_gen_tri_mesh_impl(uint32_t nx, uint32_t ny, uint64_t seed, uint32_t shuffle,
                                       uint32_t *edge_cells, double *is_boundary,
                                       double *normals, double *fluxes, double *bathy,
                                       double *volume, double *state) {
  if (nx < 1 || ny < 1) return fail(FOT_ERR_INVALID_ARGUMENT, "tri mesh needs nx,ny >= 1");
  const uint64_t nc64 = 2ull * nx * ny;
  const uint64_t ne64 = uint64_t(nx) * (ny + 1) + uint64_t(nx + 1) * ny + uint64_t(nx) * ny;
  if (ne64 > 0xffffffffull) return fail(FOT_ERR_INVALID_ARGUMENT, "tri mesh too large");
  const uint32_t nc = uint32_t(nc64), ne = uint32_t(ne64);
  std::vector<uint32_t> cell_perm(nc), edge_perm(ne);
  if (shuffle & 2u) permutation(nc, stream_seed(seed, 2), cell_perm.data());
  else std::iota(cell_perm.begin(), cell_perm.end(), 0u);
  if (shuffle & 4u) permutation(ne, stream_seed(seed, 3), edge_perm.data());
  else std::iota(edge_perm.begin(), edge_perm.end(), 0u);
  std::vector<Pt> X(size_t(nx + 1) * (ny + 1));
  {
    SplitMix64 g(stream_seed(seed, 4));
    for (uint32_t j = 0; j <= ny; ++j)
      for (uint32_t i = 0; i <= nx; ++i) {
        const double dx = -0.2 + 0.4 * g.uniform();
        const double dy = -0.2 + 0.4 * g.uniform();
        X[size_t(j) * (nx + 1) + i] = {double(i) + dx, double(j) + dy};
      }
  }
  std::vector<uint8_t> diag(size_t(nx) * ny);
  {
    SplitMix64 g(stream_seed(seed, 7));
    for (auto &d : diag) d = uint8_t(g.next() >> 63);
  }
  auto node = [&](uint32_t i, uint32_t j) { return X[size_t(j) * (nx + 1) + i]; };
  // triangle t of quad (i, j): vertices and centroid
  auto tri = [&](uint32_t i, uint32_t j, int t, Pt v[3]) {
    const Pt a = node(i, j), b = node(i + 1, j), c = node(i + 1, j + 1), d = node(i, j + 1);
    if (diag[size_t(j) * nx + i] == 0) {
      if (t == 0) v[0] = a, v[1] = b, v[2] = c;
      else v[0] = a, v[1] = c, v[2] = d;
    } else {
      if (t == 0) v[0] = a, v[1] = b, v[2] = d;
      else v[0] = b, v[1] = c, v[2] = d;
    }
  };
  auto centroid = [&](uint32_t i, uint32_t j, int t) {
    Pt v[3];
    tri(i, j, t, v);
    return Pt{(v[0].x + v[1].x + v[2].x) / 3.0, (v[0].y + v[1].y + v[2].y) / 3.0};
  };
  auto cid = [&](uint32_t i, uint32_t j, int t) { return cell_perm[2 * (j * nx + i) + t]; };
  // which triangle of quad (i, j) holds its right / left edge
  auto right_tri = [&](uint32_t i, uint32_t j) { return diag[size_t(j) * nx + i] == 0 ? 0 : 1; };
  auto left_tri = [&](uint32_t i, uint32_t j) { return diag[size_t(j) * nx + i] == 0 ? 1 : 0; };
  SplitMix64 gf(stream_seed(seed, 8)), gb(stream_seed(seed, 9));
  uint32_t k = 0;
  auto emit = [&](Pt p, Pt q, bool boundary, uint32_t lc, Pt lcen, uint32_t rc, Pt rcen) {
    const uint32_t e = edge_perm[k++];
    double nxv = q.y - p.y, nyv = -(q.x - p.x);
    const double len = std::sqrt(nxv * nxv + nyv * nyv);
    nxv /= len, nyv /= len;
    // interior: towards the right cell; boundary: away from the cell
    const Pt to = boundary ? Pt{0.5 * (p.x + q.x), 0.5 * (p.y + q.y)} : rcen;
    if (nxv * (to.x - lcen.x) + nyv * (to.y - lcen.y) < 0) nxv = -nxv, nyv = -nyv;
    if (edge_cells) {
      edge_cells[2 * size_t(e)] = lc;
      edge_cells[2 * size_t(e) + 1] = boundary ? lc : rc;
    }
    if (is_boundary) is_boundary[e] = boundary ? 1.0 : 0.0;
    if (normals) normals[e] = nxv, normals[size_t(ne) + e] = nyv;
    // fluxes and bathymetry source N(0, 0.1), drawn in generation order
    for (int d = 0; d < 3; ++d) {
      const double v = 0.1 * gf.normal();
      if (fluxes) fluxes[size_t(d) * ne + e] = v;
    }
    for (int d = 0; d < 2; ++d) {
      const double v = 0.1 * gb.normal();
      if (bathy) bathy[size_t(d) * ne + e] = v;
    }
  };
  const Pt none{0, 0};
  for (uint32_t j = 0; j <= ny; ++j) // horizontal edges: below T1 of (i,j-1), above T0 of (i,j)
    for (uint32_t i = 0; i < nx; ++i) {
      const Pt p = node(i, j), q = node(i + 1, j);
      if (j == 0) emit(p, q, true, cid(i, 0, 0), centroid(i, 0, 0), 0, none);
      else if (j == ny) emit(p, q, true, cid(i, ny - 1, 1), centroid(i, ny - 1, 1), 0, none);
      else
        emit(p, q, false, cid(i, j - 1, 1), centroid(i, j - 1, 1), cid(i, j, 0),
             centroid(i, j, 0));
    }
  for (uint32_t j = 0; j < ny; ++j) // vertical edges: left quad's right triangle first
    for (uint32_t i = 0; i <= nx; ++i) {
      const Pt p = node(i, j), q = node(i, j + 1);
      if (i == 0) {
        const int t = left_tri(0, j);
        emit(p, q, true, cid(0, j, t), centroid(0, j, t), 0, none);
      } else if (i == nx) {
        const int t = right_tri(nx - 1, j);
        emit(p, q, true, cid(nx - 1, j, t), centroid(nx - 1, j, t), 0, none);
      } else {
        const int tl = right_tri(i - 1, j), tr = left_tri(i, j);
        emit(p, q, false, cid(i - 1, j, tl), centroid(i - 1, j, tl), cid(i, j, tr),
             centroid(i, j, tr));
      }
    }
  for (uint32_t j = 0; j < ny; ++j) // diagonals: T0 left, T1 right
    for (uint32_t i = 0; i < nx; ++i) {
      const bool ac = diag[size_t(j) * nx + i] == 0;
      const Pt p = ac ? node(i, j) : node(i + 1, j), q = ac ? node(i + 1, j + 1) : node(i, j + 1);
      emit(p, q, false, cid(i, j, 0), centroid(i, j, 0), cid(i, j, 1), centroid(i, j, 1));
    }
  if (volume) {
    for (uint32_t j = 0; j < ny; ++j)
      for (uint32_t i = 0; i < nx; ++i)
        for (int t = 0; t < 2; ++t) {
          Pt v[3];
          tri(i, j, t, v);
          const double a2 = (v[1].x - v[0].x) * (v[2].y - v[0].y) -
                            (v[2].x - v[0].x) * (v[1].y - v[0].y);
          volume[cid(i, j, t)] = 0.5 * std::fabs(a2);
        }
  }
  if (state) {
    SplitMix64 g(stream_seed(seed, 10));
    for (uint32_t c = 0; c < nc; ++c) {
      const uint32_t cp = cell_perm[c];
      state[4 * size_t(cp) + 0] = 0.5 + 1.5 * g.uniform();
      state[4 * size_t(cp) + 1] = 0.1 * g.normal();
      state[4 * size_t(cp) + 2] = 0.1 * g.normal();
      state[4 * size_t(cp) + 3] = -g.uniform();
    }
  }
  return FOT_OK;
}
extern "C" fot_status fot_gen_tri_mesh(uint32_t nx, uint32_t ny, uint64_t seed, uint32_t shuffle,
                                       uint32_t *edge_cells, double *is_boundary,
                                       double *normals, double *fluxes, double *bathy,
                                       double *volume, double *state) {
  return fot::guarded("fot_gen_tri_mesh", [&] {
    return fot_gen_tri_mesh_impl(nx, ny, seed, shuffle, edge_cells, is_boundary, normals, fluxes,
                                 bathy, volume, state);
  });
}

// mini_aero-style hex face mesh (declared in fot_capi.h)
static fot_status fot_gen_hex_faces_impl(uint32_t n1, uint32_t n2, uint32_t n3, uint64_t seed,
                                        uint32_t shuffle, uint32_t *face_cells, double *face,
                                        double *cells) {
  if (n1 < 1 || n2 < 1 || n3 < 1) return fail(FOT_ERR_INVALID_ARGUMENT, "hex faces need n >= 1");
  const uint64_t nc64 = uint64_t(n1) * n2 * n3;
  const uint64_t nf64 = uint64_t(n1 - 1) * n2 * n3 + uint64_t(n1) * (n2 - 1) * n3 +
                        uint64_t(n1) * n2 * (n3 - 1);
  if (nf64 > 0xffffffffull || nc64 > 0xffffffffull)
    return fail(FOT_ERR_INVALID_ARGUMENT, "hex face mesh too large");
  const uint32_t nc = uint32_t(nc64), nf = uint32_t(nf64);
  std::vector<uint32_t> cell_perm(nc), face_perm(nf);
  if (shuffle & 2u) permutation(nc, stream_seed(seed, 2), cell_perm.data());
  else std::iota(cell_perm.begin(), cell_perm.end(), 0u);
  if (shuffle & 4u) permutation(nf, stream_seed(seed, 3), face_perm.data());
  else std::iota(face_perm.begin(), face_perm.end(), 0u);
  auto cid = [&](uint32_t i, uint32_t j, uint32_t k) {
    return cell_perm[(size_t(k) * n2 + j) * n1 + i];
  };
  uint32_t f = 0;
  // axis a: faces between cell (i,j,k) and its +a neighbour
  for (int a = 0; a < 3; ++a) {
    const uint32_t m1 = n1 - (a == 0), m2 = n2 - (a == 1), m3 = n3 - (a == 2);
    for (uint32_t k = 0; k < m3; ++k)
      for (uint32_t j = 0; j < m2; ++j)
        for (uint32_t i = 0; i < m1; ++i) {
          const uint32_t e = face_perm[f++];
          if (face_cells) {
            face_cells[2 * size_t(e)] = cid(i, j, k);
            face_cells[2 * size_t(e) + 1] = cid(i + (a == 0), j + (a == 1), k + (a == 2));
          }
          if (face) {
            const double c[3] = {i + 0.5 + 0.5 * (a == 0), j + 0.5 + 0.5 * (a == 1),
                                 k + 0.5 + 0.5 * (a == 2)};
            // normal e_a, tangent e_{a+1}, binormal e_{a+2}
            for (int d = 0; d < 3; ++d) {
              face[size_t(d) * nf + e] = c[d];
              face[size_t(3 + d) * nf + e] = d == a ? 1.0 : 0.0;
              face[size_t(6 + d) * nf + e] = d == (a + 1) % 3 ? 1.0 : 0.0;
              face[size_t(9 + d) * nf + e] = d == (a + 2) % 3 ? 1.0 : 0.0;
            }
          }
        }
  }
  if (cells) {
    const double rho = 1.0, u = 50.0, v = 10.0, w = 5.0, T = 300.0;
    const double e_int = 287.05 * T / 0.4;
    const double U0[5] = {rho, rho * u, rho * v, rho * w,
                          rho * (e_int + 0.5 * (u * u + v * v + w * w))};
    SplitMix64 gq(stream_seed(seed, 5)), gg(stream_seed(seed, 6)), gl(stream_seed(seed, 7));
    for (uint32_t k = 0; k < n3; ++k)
      for (uint32_t j = 0; j < n2; ++j)
        for (uint32_t i = 0; i < n1; ++i) {
          double *r = cells + size_t(cid(i, j, k)) * 28;
          for (int d = 0; d < 5; ++d) r[d] = U0[d] * (1.0 + 1e-2 * gq.normal());
          r[5] = i + 0.5, r[6] = j + 0.5, r[7] = k + 0.5;
          for (int d = 0; d < 15; ++d) r[8 + d] = 1e-3 * gg.normal();
          for (int d = 0; d < 5; ++d) r[23 + d] = gl.uniform();
        }
  }
  return FOT_OK;
}
extern "C" fot_status fot_gen_hex_faces(uint32_t n1, uint32_t n2, uint32_t n3, uint64_t seed,
                                        uint32_t shuffle, uint32_t *face_cells, double *face,
                                        double *cells) {
  return fot::guarded("fot_gen_hex_faces", [&] {
    return fot_gen_hex_faces_impl(n1, n2, n3, seed, shuffle, face_cells, face, cells);
  });
}
