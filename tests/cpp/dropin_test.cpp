// Drives the C++ drop-in (include/fot/problem.hpp) the way the reference's
// drivers drive Problem<SOA> (res_calc.cu, tests.hpp): builds problems,
// runs the CPU or GPU loops through the kernel structs and writes each
// result, in original point order, as raw little-endian doubles into
// <outdir>/<name>.bin for the Python side to check against the oracle.
//   dropin_test cpu <outdir>   CPU loops, pre-passes and IO round trips
//   dropin_test gpu <outdir>   GPU loops (needs a device)
// Exit status 0 on success; a failed self-check prints and exits 1.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>

#include "fot/problem.hpp"

namespace {

std::string g_out;

#define CHECK(cond)                                                        \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
      std::exit(1);                                                        \
    }                                                                      \
  } while (0)

template <bool SOA> void dump(const Problem<SOA> &p, const std::string &name) {
  const MY_SIZE n = p.mesh.numPoints(0), dim = p.point_weights[0].getDim();
  std::vector<double> v(std::size_t(n) * dim);
  for (MY_SIZE i = 0; i < n; ++i) {
    const MY_SIZE q = p.applied_permutation.empty() ? i : p.applied_permutation[i];
    for (MY_SIZE j = 0; j < dim; ++j)
      v[std::size_t(i) * dim + j] = p.point_weights[0].template operator[]<double>(
          index<SOA>(n, q, dim, j));
  }
  std::ofstream f(g_out + "/" + name + ".bin", std::ios::binary);
  f.write(reinterpret_cast<const char *>(v.data()), std::streamsize(v.size() * sizeof(double)));
  std::printf("wrote %s (%zu doubles)\n", name.c_str(), v.size());
}

// the airfoil res_calc problem (res_calc.cu:10-40 shapes), 40 x 24 cells,
// seed 3, everything shuffled; res0 = U[-1,1) with seed 11
template <bool SOA> Problem<SOA> airfoil(MY_SIZE block_size) {
  const std::uint32_t nx = 40, ny = 24;
  const std::uint32_t nn = (nx + 1) * (ny + 1), nc = nx * ny;
  const std::uint32_t ne = nx * (ny - 1) + (nx - 1) * ny;
  std::vector<std::uint32_t> pedge(2 * ne), pecell(2 * ne);
  std::vector<double> x(2 * nn), q(4 * nc), adt(nc), res(4 * nc);
  fot::check(fot_gen_airfoil(nx, ny, 3, 7, pedge.data(), pecell.data(), x.data(), q.data(),
                             adt.data()),
             "fot_gen_airfoil");
  fot::check(fot_gen_uniform(res.size(), 11, -1.0, 1.0, res.data()), "fot_gen_uniform");
  Mesh mesh({nc, nn, nc, nc}, ne, {pecell.data(), pedge.data(), pecell.data(), pecell.data()},
            {2, 2, 2, 2});
  Problem<SOA> p(std::move(mesh),
                 {{res_calc::RES_DIM, sizeof(double)},
                  {res_calc::X_DIM, sizeof(double)},
                  {res_calc::Q_DIM, sizeof(double)},
                  {res_calc::ADT_DIM, sizeof(double)}},
                 {}, block_size);
  auto fill = [&](unsigned m, const std::vector<double> &aos, MY_SIZE dim) {
    const MY_SIZE n = p.mesh.numPoints(m);
    for (MY_SIZE i = 0; i < n; ++i)
      for (MY_SIZE j = 0; j < dim; ++j)
        p.point_weights[m].template operator[]<double>(index<SOA>(n, i, dim, j)) =
            aos[std::size_t(i) * dim + j];
  };
  fill(0, res, 4);
  fill(1, x, 2);
  fill(2, q, 4);
  fill(3, adt, 1);
  return p;
}

// mine on the 20 x 17 fillEdgeList grid, PointDim 4, CellDim 4, U[0,1)
template <bool SOA> Problem<SOA> mine_grid() {
  const std::uint32_t N = 20, M = 17, np = N * M, ne = (N - 1) * M + N * (M - 1);
  std::vector<std::uint32_t> e(2 * ne);
  fot::check(fot_gen_grid_edges(N, M, e.data()), "fot_gen_grid_edges");
  Problem<SOA> p(Mesh({np}, ne, {e.data()}, {2}), {{4, sizeof(double)}}, {{4, sizeof(double)}},
                 128);
  std::vector<double> pd(4 * np), w(4 * ne);
  fot::check(fot_gen_uniform(pd.size(), 5, 0.0, 1.0, pd.data()), "fot_gen_uniform");
  fot::check(fot_gen_uniform(w.size(), 6, 0.0, 1.0, w.data()), "fot_gen_uniform");
  for (MY_SIZE i = 0; i < np; ++i)
    for (MY_SIZE j = 0; j < 4; ++j)
      p.point_weights[0].template operator[]<double>(index<SOA>(np, i, 4, j)) = pd[4 * i + j];
  std::copy(w.begin(), w.end(), p.cell_weights[0].template begin<double>());
  return p;
}

// hex corner-force scatter on a 5 x 4 x 3 point grid, forces N(0,1)
template <bool SOA> Problem<SOA> hex_grid() {
  const std::uint32_t n1 = 5, n2 = 4, n3 = 3, np = n1 * n2 * n3,
                      nc = (n1 - 1) * (n2 - 1) * (n3 - 1);
  std::vector<std::uint32_t> c(8 * nc);
  fot::check(fot_gen_hex_grid(n1, n2, n3, c.data()), "fot_gen_hex_grid");
  Problem<SOA> p(Mesh({np}, nc, {c.data()}, {8}), {{3, sizeof(double)}}, {{24, sizeof(double)}},
                 128);
  std::vector<double> F(24 * nc);
  fot::check(fot_gen_normal(F.size(), 9, 0.0, 1.0, F.data()), "fot_gen_normal");
  std::copy(F.begin(), F.end(), p.cell_weights[0].template begin<double>());
  return p;
}

// lulesh on a 5 x 4 x 3 node lattice: node (l, r, c) at (r, c, l) +
// U(-0.2, 0.2) (seed 21), forces N(0,1) (seed 22), sigma N(0,1) (seed 23),
// volumes zero; both mappings the hexes (lulesh.cu:10-24)
template <bool SOA> Problem<SOA> lulesh_grid() {
  const std::uint32_t n1 = 5, n2 = 4, n3 = 3, np = n1 * n2 * n3,
                      nc = (n1 - 1) * (n2 - 1) * (n3 - 1);
  std::vector<std::uint32_t> c(8 * nc);
  fot::check(fot_gen_hex_grid(n1, n2, n3, c.data()), "fot_gen_hex_grid");
  Problem<SOA> p(Mesh({np, np}, nc, {c.data(), c.data()}, {8, 8}),
                 {{3, sizeof(double)}, {3, sizeof(double)}},
                 {{3, sizeof(double)}, {1, sizeof(double)}}, 128);
  std::vector<double> jit(3 * np), f(3 * np), sig(3 * nc);
  fot::check(fot_gen_uniform(jit.size(), 21, -0.2, 0.2, jit.data()), "fot_gen_uniform");
  fot::check(fot_gen_normal(f.size(), 22, 0.0, 1.0, f.data()), "fot_gen_normal");
  fot::check(fot_gen_normal(sig.size(), 23, 0.0, 1.0, sig.data()), "fot_gen_normal");
  for (std::uint32_t l = 0; l < n3; ++l)
    for (std::uint32_t r = 0; r < n1; ++r)
      for (std::uint32_t cc = 0; cc < n2; ++cc) {
        const MY_SIZE i = l * n1 * n2 + r * n2 + cc;
        const double base[3] = {double(r), double(cc), double(l)};
        for (MY_SIZE d = 0; d < 3; ++d) {
          p.point_weights[0].template operator[]<double>(index<SOA>(np, i, 3, d)) = f[3 * i + d];
          p.point_weights[1].template operator[]<double>(index<SOA>(np, i, 3, d)) =
              base[d] + jit[3 * i + d];
        }
      }
  std::copy(sig.begin(), sig.end(), p.cell_weights[0].template begin<double>());
  std::fill(p.cell_weights[1].template begin<double>(),
            p.cell_weights[1].template begin<double>() + nc, 0.0);
  return p;
}

// volna on a 14 x 9 triangle mesh (seed 4), mini_aero on 5 x 4 x 3 hexes (seed 2)
template <bool SOA> Problem<SOA> volna_mesh() {
  const std::uint32_t nx = 14, ny = 9, nc = 2 * nx * ny;
  const std::uint32_t ne = nx * (ny + 1) + (nx + 1) * ny + nx * ny;
  std::vector<std::uint32_t> ec(2 * ne);
  std::vector<double> bnd(ne), nrm(2 * ne), fl(3 * ne), ba(2 * ne), vol(nc), st(4 * nc);
  fot::check(fot_gen_tri_mesh(nx, ny, 4, 6, ec.data(), bnd.data(), nrm.data(), fl.data(),
                              ba.data(), vol.data(), st.data()),
             "fot_gen_tri_mesh");
  Problem<SOA> p(Mesh({nc, nc}, ne, {ec.data(), ec.data()}, {2, 2}),
                 {{4, sizeof(double)}, {1, sizeof(double)}},
                 {{3, sizeof(double)}, {2, sizeof(double)}, {2, sizeof(double)},
                  {1, sizeof(double)}},
                 128);
  for (MY_SIZE i = 0; i < nc; ++i) {
    for (MY_SIZE j = 0; j < 4; ++j)
      p.point_weights[0].template operator[]<double>(index<SOA>(nc, i, 4, j)) = st[4 * i + j];
    p.point_weights[1].template operator[]<double>(i) = vol[i];
  }
  std::copy(fl.begin(), fl.end(), p.cell_weights[0].template begin<double>());
  std::copy(ba.begin(), ba.end(), p.cell_weights[1].template begin<double>());
  std::copy(nrm.begin(), nrm.end(), p.cell_weights[2].template begin<double>());
  std::copy(bnd.begin(), bnd.end(), p.cell_weights[3].template begin<double>());
  return p;
}

template <bool SOA> Problem<SOA> mini_aero_mesh() {
  const std::uint32_t n1 = 5, n2 = 4, n3 = 3, nc = n1 * n2 * n3;
  const std::uint32_t nf = (n1 - 1) * n2 * n3 + n1 * (n2 - 1) * n3 + n1 * n2 * (n3 - 1);
  std::vector<std::uint32_t> fc(2 * nf);
  std::vector<double> face(12 * nf), cells(28 * nc);
  fot::check(fot_gen_hex_faces(n1, n2, n3, 2, 6, fc.data(), face.data(), cells.data()),
             "fot_gen_hex_faces");
  Problem<SOA> p(Mesh({nc, nc}, nf, {fc.data(), fc.data()}, {2, 2}),
                 {{5, sizeof(double)}, {28, sizeof(double)}}, {{12, sizeof(double)}}, 128);
  for (MY_SIZE i = 0; i < nc; ++i)
    for (MY_SIZE j = 0; j < 28; ++j)
      p.point_weights[1].template operator[]<double>(index<SOA>(nc, i, 28, j)) =
          cells[28 * i + j];
  std::copy(face.begin(), face.end(), p.cell_weights[0].template begin<double>());
  return p;
}

void io_round_trips() {
  // Mesh text format and writeCellList (mesh.hpp:58-98, 184-197)
  std::istringstream m0("4 3\n0 1\n1 2\n2 3\n"), m1("5 3\n4 0\n0 1\n1 3\n");
  Mesh mesh({&m0, &m1}, {2, 2});
  CHECK(mesh.numCells() == 3 && mesh.numPoints(0) == 4 && mesh.numPoints(1) == 5);
  std::ostringstream os;
  mesh.writeCellList(os, 1);
  CHECK(os.str() == "5 3\n4 0\n0 1\n1 3\n");
  // out-of-range ids and short files are InvalidInputFile (mesh.hpp:84-96)
  bool thrown = false;
  try {
    std::istringstream bad("3 2\n0 1\n1 3\n");
    Mesh b({&bad}, {2});
  } catch (const InvalidInputFile &e) {
    thrown = e.line == 2;
  }
  CHECK(thrown);
  thrown = false;
  try {
    std::istringstream bad("3 2\n0 1\n");
    Mesh b({&bad}, {2});
  } catch (const InvalidInputFile &) {
    thrown = true;
  }
  CHECK(thrown);
  // point / cell data and partition streams (problem.hpp:553-632)
  std::istringstream mi("4 3\n0 1\n1 2\n2 3\n");
  Problem<true> p({&mi}, {2}, {{2, sizeof(double)}}, {{1, sizeof(double)}}, 64);
  std::istringstream pd("1 2\n3 4\n5 6\n7 8\n"), cd("0.5\n0.25\n0.125\n");
  p.readPointData<double>(pd, 0);
  p.readCellData<double>(cd, 0);
  CHECK(p.point_weights[0].operator[]<double>(index<true>(4, 2, 2, 1)) == 6.0);
  CHECK(p.cell_weights[0].operator[]<double>(2) == 0.125);
  std::ostringstream wp;
  p.writePointData<double>(wp);
  CHECK(wp.str() == "1 2\n3 4\n5 6\n7 8\n");
  p.partition_vector = {2, 0, 1};
  std::ostringstream part;
  p.writePartition(part);
  std::istringstream part_in(part.str());
  p.partition_vector.clear();
  p.readPartition(part_in);
  CHECK((p.partition_vector == std::vector<MY_SIZE>{2, 0, 1}));
  // colourCells (mesh.hpp:117-162): path 0-1-2-3 needs two colours
  const auto colours = p.mesh.colourCells();
  CHECK(colours.size() == 2 && (colours[0] == std::vector<MY_SIZE>{0, 2}));
  std::printf("io round trips ok\n");
}

template <bool SOA> void cpu_cases(const char *tag) {
  {
    auto p = airfoil<SOA>(128);
    p.template loopCPUCellCentred<res_calc::StepSeq>(3);
    dump(p, std::string("res_serial_") + tag);
  }
  {
    auto p = airfoil<SOA>(128);
    p.template loopCPUCellCentredOMP<res_calc::StepOMP>(3);
    dump(p, std::string("res_omp_") + tag);
  }
  {
    // the pre-passes of res_calc.cu:129-181: results must not depend on them
    auto p = airfoil<SOA>(128);
    p.reorder();
    p.partition(1.001f);
    p.reorderToPartition();
    p.renumberPoints();
    p.template loopCPUCellCentred<res_calc::StepSeq>(3);
    dump(p, std::string("res_reordered_") + tag);
  }
  {
    auto p = airfoil<SOA>(64);
    p.partitionOurs(1.001f, 4);
    p.reorderToPartition();
    p.renumberPoints();
    p.template loopCPUCellCentredOMP<res_calc::StepOMP>(3);
    dump(p, std::string("res_ours_") + tag);
  }
  {
    auto p = mine_grid<SOA>();
    p.template loopCPUCellCentred<mine::StepSeq<2, 4, 4>>(3);
    dump(p, std::string("mine_serial_") + tag);
  }
  {
    auto p = hex_grid<SOA>();
    p.template loopCPUCellCentred<hex_scatter::StepSeq>(2);
    dump(p, std::string("hex_serial_") + tag);
  }
  {
    auto p = volna_mesh<SOA>();
    p.template loopCPUCellCentred<volna::StepSeq>(2);
    dump(p, std::string("volna_serial_") + tag);
  }
  {
    auto p = mini_aero_mesh<SOA>();
    p.template loopCPUCellCentred<mini_aero::StepSeq>(2);
    dump(p, std::string("ma_serial_") + tag);
  }
  {
    auto p = lulesh_grid<SOA>();
    p.template loopCPUCellCentred<lulesh::StepSeq>(2);
    dump(p, std::string("lulesh_serial_") + tag);
  }
  {
    auto p = lulesh_grid<SOA>();
    p.template loopCPUCellCentredOMP<lulesh::StepOMP>(2);
    dump(p, std::string("lulesh_omp_") + tag);
  }
}

template <bool SOA> void gpu_cases(const char *tag) {
  const std::string t(tag);
  {
    auto p = airfoil<SOA>(128);
    p.template loopGPUCellCentred<res_calc::StepGPUGlobal>(3);
    dump(p, "res_global_" + t);
  }
  {
    auto p = airfoil<SOA>(128);
    p.template loopGPUEdgeCentred<res_calc::StepGPUAtomic>(3);
    dump(p, "res_atomic_" + t);
  }
  {
    auto p = airfoil<SOA>(128);
    p.template loopGPUHierarchical<res_calc::StepGPUHierarchical>(3);
    dump(p, "res_owner_" + t);
  }
  {
    auto p = airfoil<SOA>(128);
    p.hierarchical_flags = 0;  // the reference's per-block-colour plan
    p.partitionOurs(1.001f, 4);
    p.reorderToPartition();
    p.renumberPoints();
    p.template loopGPUHierarchical<res_calc::StepGPUHierarchical>(3);
    dump(p, "res_hier_" + t);
  }
  {
    auto p = mine_grid<SOA>();
    p.template loopGPUHierarchical<mine::StepGPUHierarchical<2, 4, 4>>(3);
    dump(p, "mine_owner_" + t);
  }
  {
    auto p = hex_grid<SOA>();
    p.template loopGPUEdgeCentred<hex_scatter::StepGPUAtomic>(2);
    dump(p, "hex_atomic_" + t);
  }
  {
    auto p = hex_grid<SOA>();
    p.template loopGPUCellCentred<hex_scatter::StepGPUGlobal>(2);
    dump(p, "hex_global_" + t);
  }
  {
    auto p = volna_mesh<SOA>();
    p.template loopGPUHierarchical<volna::StepGPUHierarchical>(2);
    dump(p, "volna_owner_" + t);
  }
  {
    auto p = mini_aero_mesh<SOA>();
    p.template loopGPUHierarchical<mini_aero::StepGPUHierarchical>(2);
    dump(p, "ma_owner_" + t);
  }
  {
    auto p = mini_aero_mesh<SOA>();
    p.template loopGPUCellCentred<mini_aero::StepGPUGlobal>(2);
    dump(p, "ma_global_" + t);
  }
  {
    auto p = mini_aero_mesh<SOA>();
    p.template loopGPUEdgeCentred<mini_aero::StepGPUAtomic>(2);
    dump(p, "ma_atomic_" + t);
  }
  {
    auto p = lulesh_grid<SOA>();
    p.template loopGPUCellCentred<lulesh::StepGPUGlobal>(2);
    dump(p, "lulesh_global_" + t);
  }
  {
    auto p = lulesh_grid<SOA>();
    p.template loopGPUEdgeCentred<lulesh::StepGPUAtomic>(2);
    dump(p, "lulesh_atomic_" + t);
  }
  {
    auto p = lulesh_grid<SOA>();
    p.template loopGPUHierarchical<lulesh::StepGPUHierarchical>(2);
    dump(p, "lulesh_owner_" + t);
  }
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s cpu|gpu <outdir>\n", argv[0]);
    return 2;
  }
  g_out = argv[2];
  const std::string mode = argv[1];
  try {
    if (mode == "cpu") {
      io_round_trips();
      cpu_cases<false>("aos");
      cpu_cases<true>("soa");
    } else if (mode == "gpu") {
      gpu_cases<false>("aos");
      gpu_cases<true>("soa");
    } else {
      return 2;
    }
  } catch (const fot::Error &e) {
    std::fprintf(stderr, "fot::Error: %s\n", e.what());
    return 1;
  } catch (const InvalidInputFile &e) {
    std::fprintf(stderr, "InvalidInputFile: %s %u %u\n", e.input_type.c_str(), e.mapping_ind,
                 e.line);
    return 1;
  }
  std::printf("done\n");
  return 0;
}
