// Drives the C++ drop-in (include/fot/problem.hpp) the way the reference's
// drivers drive Problem<SOA> (res_calc.cu, tests.hpp): builds problems,
// runs the CPU or GPU loops through the kernel structs and writes each
// result, in original point order, as raw little-endian doubles into
// <outdir>/<name>.bin for the Python side to check against the oracle.
//   dropin_test cpu <outdir>   CPU loops, pre-passes and IO round trips
//   dropin_test gpu <outdir>   GPU loops (needs a device)
//   dropin_test step <outdir>  the StepGPU*::call<SOA> functors driven per
//                              colour as the reference's Problem drives them
// Exit status 0 on success; a failed self-check prints and exits 1.
#include <algorithm>
#include <cstdio>
#include <cstring>
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
template <bool SOA> Problem<SOA> hex_grid(MY_SIZE block_size = 128) {
  const std::uint32_t n1 = 5, n2 = 4, n3 = 3, np = n1 * n2 * n3,
                      nc = (n1 - 1) * (n2 - 1) * (n3 - 1);
  std::vector<std::uint32_t> c(8 * nc);
  fot::check(fot_gen_hex_grid(n1, n2, n3, c.data()), "fot_gen_hex_grid");
  Problem<SOA> p(Mesh({np}, nc, {c.data()}, {8}), {{3, sizeof(double)}}, {{24, sizeof(double)}},
                 block_size);
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
  // the reference's per-block-colour plan on a reordered, bisected mesh, and
  // the same plan run by the two-phase write-back: the same bits
  for (std::uint32_t flags : {0u, std::uint32_t(FOT_LOOP_TWO_PHASE)}) {
    auto p = hex_grid<SOA>(7);  // 24 elements in 4 blocks: most points shared
    p.hierarchical_flags = flags;
    p.reorder();
    p.partition(1.001f, true);
    p.reorderToPartition();
    p.renumberPoints();
    p.template loopGPUHierarchical<hex_scatter::StepGPUHierarchical>(2);
    dump(p, (flags ? "hex_2phase_" : "hex_hier_") + t);
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


// ---- the functor-level step API driven as the reference's Problem drives it
// A device buffer that frees itself (the reference's device_data_t,
// data_t.cu:58-74).
struct DevBuf {
  void *p = nullptr;
  DevBuf() = default;
  DevBuf(const void *host, std::size_t bytes) {
    if (!bytes) return;
    fot::check(fot_malloc(&p, bytes), "fot_malloc");
    fot::check(fot_memcpy_htod(p, host, bytes), "fot_memcpy_htod");
  }
  DevBuf(DevBuf &&o) noexcept : p(o.p) { o.p = nullptr; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    std::swap(p, o.p);
    return *this;
  }
  DevBuf(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) fot_free(p);
  }
};
template <class T> DevBuf dev_vec(const std::vector<T> &v) {
  return DevBuf(v.data(), v.size() * sizeof(T));
}
// a device array of device pointers
DevBuf dev_ptrs(const std::vector<const void *> &v) { return dev_vec(v); }

// cell data entry cw gathered for `cells` (SOA, stride cells.size()), as
// problem.hpp:124-143 and copyCellWeights (colouring.hpp:310-335) lay it out
template <bool SOA>
std::vector<char> gather_cells(const Problem<SOA> &p, unsigned cw, const MY_SIZE *cells,
                               MY_SIZE n) {
  const data_t &w = p.cell_weights[cw];
  const MY_SIZE dim = w.getDim(), ts = w.getTypeSize(), nc = p.mesh.numCells();
  std::vector<char> out(std::size_t(n) * dim * ts);
  for (MY_SIZE i = 0; i < n; ++i)
    for (MY_SIZE d = 0; d < dim; ++d)
      std::memcpy(out.data() + std::size_t(index<true>(n, i, dim, d)) * ts,
                  w.cbegin() + std::size_t(index<true>(nc, cells[i], dim, d)) * ts, ts);
  return out;
}

// The point data's device mirrors and strides; `out` starts as a copy of
// point data 0 (problem.hpp:92-95, 203-209).
template <bool SOA> struct StepPoints {
  std::vector<const void *> host_ptrs;
  DevBuf ptrs, stride, out;
  std::size_t out_bytes = 0;
  explicit StepPoints(Problem<SOA> &p) {
    for (data_t &pw : p.point_weights) {
      pw.initDeviceMemory();
      host_ptrs.push_back(pw.getDeviceData());
    }
    ptrs = dev_ptrs(host_ptrs);
    stride = dev_vec(p.mesh.numPoints());
    out_bytes = p.point_weights[0].bytes();
    out = DevBuf(p.point_weights[0].cbegin(), out_bytes);
  }
  // after each iteration point data 0 = out (problem.hpp:179-185, 320-327)
  void copy_back(Problem<SOA> &p) {
    fot::check(fot_memcpy_dtod(p.point_weights[0].getDeviceData(), out.p, out_bytes, nullptr),
               "fot_memcpy_dtod");
    fot::check(fot_device_synchronize(), "fot_device_synchronize");
  }
};

// Problem::loopGPUCellCentred (problem.hpp:90-194) with the launches made
// by Step::call<SOA> itself: per colour, compacted AOS cell lists and SOA
// cell data of the colour's cells, device arrays of device pointers, a
// synchronisation after each launch.  `atomic` runs every cell as one
// colour through the atomics functor (the launch behind loopGPUEdgeCentred).
template <class Step, bool SOA>
void step_loop_global(Problem<SOA> &p, MY_SIZE num, MY_SIZE block_size, bool atomic) {
  std::vector<std::vector<MY_SIZE>> colours;
  if (atomic) {
    colours.emplace_back(p.mesh.numCells());
    for (MY_SIZE i = 0; i < p.mesh.numCells(); ++i) colours[0][i] = i;
  } else {
    colours = p.mesh.colourCells();
  }
  std::vector<std::vector<DevBuf>> lists(colours.size()), weights(colours.size());
  std::vector<DevBuf> list_ptrs, weight_ptrs;
  for (std::size_t c = 0; c < colours.size(); ++c) {
    const std::vector<MY_SIZE> &cells = colours[c];
    std::vector<const void *> lp, wp;
    for (unsigned m = 0; m < p.mesh.numMappings(); ++m) {
      const MY_SIZE dim = p.mesh.cell_to_node[m].getDim();
      std::vector<MY_SIZE> l(std::size_t(cells.size()) * dim);
      for (std::size_t i = 0; i < cells.size(); ++i)
        std::copy_n(p.mesh.cell_to_node[m].template cbegin<MY_SIZE>() + dim * cells[i], dim,
                    l.begin() + dim * i);
      lists[c].push_back(dev_vec(l));
      lp.push_back(lists[c].back().p);
    }
    for (unsigned cw = 0; cw < p.cell_weights.size(); ++cw) {
      weights[c].push_back(dev_vec(gather_cells(p, cw, cells.data(), MY_SIZE(cells.size()))));
      wp.push_back(weights[c].back().p);
    }
    list_ptrs.push_back(dev_ptrs(lp));
    weight_ptrs.push_back(dev_ptrs(wp.empty() ? std::vector<const void *>{nullptr} : wp));
  }
  StepPoints<SOA> pts(p);
  for (MY_SIZE it = 0; it < num; ++it) {
    for (std::size_t c = 0; c < colours.size(); ++c) {
      const MY_SIZE n = MY_SIZE(colours[c].size());
      const MY_SIZE nb = (n + block_size - 1) / block_size;
      Step::template call<SOA>(static_cast<const void **>(pts.ptrs.p), pts.out.p,
                               static_cast<void **>(weight_ptrs[c].p),
                               static_cast<const MY_SIZE **>(list_ptrs[c].p), n,
                               static_cast<MY_SIZE *>(pts.stride.p), n, nb, block_size);
      fot::check(fot_device_synchronize(), "fot_device_synchronize");
    }
    pts.copy_back(p);
  }
  p.point_weights[0].flushToHost();
}

// One block colour of a HierarchicalColourMemory on the device
// (getDeviceMemoryOfOneColour, colouring.hpp:341-393).
struct DevColour {
  MY_SIZE num_cells = 0, num_blocks = 0, shared_size = 0;
  DevBuf ptbc, ptbc_off, ncc, cc, bo, list_ptrs, weight_ptrs;
  std::vector<DevBuf> lists, weights;
};

// Problem::loopGPUHierarchical (problem.hpp:198-339) with the launches made
// by Step::call<SOA>: the plan of fot_hplan_build, one launch per block
// colour with the reference's cache size (shared_size + 32) * Dim *
// sizeof(double), a synchronisation after each.
template <class Step, bool SOA>
void step_loop_hier(Problem<SOA> &p, MY_SIZE num, MY_SIZE block_size) {
  std::vector<const std::uint32_t *> maps;
  std::vector<std::uint32_t> dims;
  for (unsigned m = 0; m < p.mesh.numMappings(); ++m) {
    maps.push_back(p.mesh.cell_to_node[m].template cbegin<MY_SIZE>());
    dims.push_back(p.mesh.cell_to_node[m].getDim());
  }
  fot_hplan *plan = nullptr;
  fot::check(fot_hplan_build(maps.data(), dims.data(), MY_SIZE(maps.size()), p.mesh.numCells(),
                             p.mesh.numPoints(0), nullptr, block_size, &plan),
             "fot_hplan_build");
  fot_hplan_info info{};
  fot::check(fot_hplan_info_get(plan, &info), "fot_hplan_info_get");
  std::vector<DevColour> cols(info.num_colours);
  for (MY_SIZE c = 0; c < info.num_colours; ++c) {
    fot_hplan_colour v{};
    fot::check(fot_hplan_colour_get(plan, c, &v), "fot_hplan_colour_get");
    DevColour &d = cols[c];
    d.num_cells = v.num_cells, d.num_blocks = v.num_blocks, d.shared_size = v.shared_size;
    d.ptbc = DevBuf(v.points_to_be_cached, sizeof(MY_SIZE) * v.num_cached);
    d.ptbc_off = DevBuf(v.points_to_be_cached_offsets, sizeof(MY_SIZE) * (v.num_blocks + 1));
    d.ncc = DevBuf(v.num_cell_colours, v.num_blocks);
    d.cc = DevBuf(v.cell_colours, v.num_cells);
    d.bo = DevBuf(v.block_offsets, sizeof(MY_SIZE) * (v.num_blocks + 1));
    std::vector<const void *> lp, wp;
    for (unsigned m = 0; m < maps.size(); ++m) {
      d.lists.emplace_back(v.cell_list[m], sizeof(MY_SIZE) * v.num_cells * dims[m]);
      lp.push_back(d.lists.back().p);
    }
    for (unsigned cw = 0; cw < p.cell_weights.size(); ++cw) {
      d.weights.push_back(dev_vec(gather_cells(p, cw, v.cell_inv_permutation, v.num_cells)));
      wp.push_back(d.weights.back().p);
    }
    d.list_ptrs = dev_ptrs(lp);
    d.weight_ptrs = dev_ptrs(wp.empty() ? std::vector<const void *>{nullptr} : wp);
  }
  fot_hplan_free(plan);
  StepPoints<SOA> pts(p);
  const MY_SIZE mult = MY_SIZE(sizeof(double)) * p.point_weights[0].getDim();
  std::uint32_t flag = 0;
  fot::check(fot_lds_overflow_flag(&flag, 1), "fot_lds_overflow_flag");
  for (MY_SIZE it = 0; it < num; ++it) {
    for (const DevColour &d : cols) {
      Step::template call<SOA>(
          static_cast<const void **>(pts.ptrs.p), pts.out.p, static_cast<MY_SIZE *>(d.ptbc.p),
          static_cast<MY_SIZE *>(d.ptbc_off.p), static_cast<void **>(d.weight_ptrs.p),
          static_cast<const MY_SIZE **>(d.list_ptrs.p), static_cast<std::uint8_t *>(d.ncc.p),
          static_cast<std::uint8_t *>(d.cc.p), static_cast<MY_SIZE *>(d.bo.p), d.num_cells,
          static_cast<MY_SIZE *>(pts.stride.p), d.num_cells, d.num_blocks, block_size,
          (d.shared_size + 32) * mult);
      fot::check(fot_device_synchronize(), "fot_device_synchronize");
    }
    pts.copy_back(p);
  }
  fot::check(fot_lds_overflow_flag(&flag, 0), "fot_lds_overflow_flag");
  CHECK(flag == 0);
  // an 8-byte cache holds no block's points (every target has three or more
  // values): every block skips its work and raises the flag, every block
  // exits, and the output is untouched
  const DevColour &d0 = cols[0];
  std::vector<char> before(pts.out_bytes), after(pts.out_bytes);
  fot::check(fot_memcpy_dtoh(before.data(), pts.out.p, pts.out_bytes), "fot_memcpy_dtoh");
  Step::template call<SOA>(
      static_cast<const void **>(pts.ptrs.p), pts.out.p, static_cast<MY_SIZE *>(d0.ptbc.p),
      static_cast<MY_SIZE *>(d0.ptbc_off.p), static_cast<void **>(d0.weight_ptrs.p),
      static_cast<const MY_SIZE **>(d0.list_ptrs.p), static_cast<std::uint8_t *>(d0.ncc.p),
      static_cast<std::uint8_t *>(d0.cc.p), static_cast<MY_SIZE *>(d0.bo.p), d0.num_cells,
      static_cast<MY_SIZE *>(pts.stride.p), d0.num_cells, d0.num_blocks, block_size, 8);
  fot::check(fot_device_synchronize(), "fot_device_synchronize");
  fot::check(fot_memcpy_dtoh(after.data(), pts.out.p, pts.out_bytes), "fot_memcpy_dtoh");
  fot::check(fot_lds_overflow_flag(&flag, 1), "fot_lds_overflow_flag");
  CHECK(flag == 1);
  CHECK(after == before);
  fot::check(fot_lds_overflow_flag(&flag, 0), "fot_lds_overflow_flag");
  CHECK(flag == 0);
  p.point_weights[0].flushToHost();
}

template <class Glob, class Atom, class Hier, bool SOA, class Make>
void step_cases(const std::string &name, Make make, MY_SIZE iters, const std::string &tag) {
  for (MY_SIZE bs : {128u, 256u, 512u}) {
    const std::string b = "_b" + std::to_string(bs) + "_" + tag;
    {
      auto p = make();
      step_loop_global<Glob>(p, iters, bs, false);
      dump(p, "step_" + name + "_global" + b);
    }
    {
      auto p = make();
      step_loop_global<Atom>(p, iters, bs, true);
      dump(p, "step_" + name + "_atomic" + b);
    }
    {
      auto p = make();
      step_loop_hier<Hier>(p, iters, bs);
      dump(p, "step_" + name + "_hier" + b);
    }
  }
}

template <bool SOA> void step_api_cases(const char *tag) {
  const std::string t(tag);
  step_cases<res_calc::StepGPUGlobal, res_calc::StepGPUAtomic, res_calc::StepGPUHierarchical,
             SOA>("res", [] { return airfoil<SOA>(128); }, 3, t);
  step_cases<mine::StepGPUGlobal<2, 4, 4>, mine::StepGPUAtomic<2, 4, 4>,
             mine::StepGPUHierarchical<2, 4, 4>, SOA>("mine", [] { return mine_grid<SOA>(); },
                                                      3, t);
  step_cases<hex_scatter::StepGPUGlobal, hex_scatter::StepGPUAtomic,
             hex_scatter::StepGPUHierarchical, SOA>("hex", [] { return hex_grid<SOA>(); }, 2, t);
  step_cases<mini_aero::StepGPUGlobal, mini_aero::StepGPUAtomic,
             mini_aero::StepGPUHierarchical, SOA>("ma", [] { return mini_aero_mesh<SOA>(); }, 2,
                                                  t);
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s cpu|gpu|step <outdir>\n", argv[0]);
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
    } else if (mode == "step") {
      step_api_cases<false>("aos");
      step_api_cases<true>("soa");
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
