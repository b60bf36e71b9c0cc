// Problem<SOA> of the C++ drop-in (problem.hpp:29-652 of the reference): the
// same fields, constructors, loops and pre-passes.  CPU loops call the user
// struct's `call<SOA>` on the host; GPU loops run libfot's loop engine
// (fot_loop_*) with the struct's functor:
//   loopGPUCellCentred<F>   global colouring, one launch per colour
//   loopGPUEdgeCentred<F>   FP64 global atomics, one launch
//   loopGPUHierarchical<F>  LDS-blocked hierarchical loop; by default the
//                           owner-compute plan (one launch, no block colours),
//                           or the reference's per-block-colour plan when
//                           `hierarchical_flags` is 0
// Reorder / partition pre-passes run libfot's native planners.
#ifndef FOT_PROBLEM_HPP
#define FOT_PROBLEM_HPP

#include <cmath>
#include <iostream>
#include <memory>

#include "data.hpp"
#include "kernels.hpp"

constexpr MY_SIZE DEFAULT_BLOCK_SIZE = 128;  // problem.hpp:16

template <bool SOA = false> struct Problem {
  Mesh mesh;
  std::vector<data_t> point_weights;
  std::vector<data_t> cell_weights;
  const MY_SIZE block_size;
  std::vector<MY_SIZE> partition_vector;
  std::vector<MY_SIZE> applied_permutation;
  // loopGPUHierarchical's plan: FOT_LOOP_* flags (default: owner compute, two
  // cells per thread, and for res_calc the gather accumulation); 0 gives the
  // reference's per-block-colour launches, and FOT_LOOP_TWO_PHASE the same
  // plan and bits in two launches (every block at once, shared points merged
  // in block-colour order), the faster choice for 3D element loops on a mesh
  // reorder()ed and partition(tolerance, true)ed into compact blocks
  std::uint32_t hierarchical_flags =
      FOT_LOOP_OWNER_COMPUTE | FOT_LOOP_TWO_PER_THREAD | FOT_LOOP_OWNER_GATHER;

  // problem.hpp:59-79
  Problem(const std::vector<std::istream *> &mesh_is, const std::vector<MY_SIZE> &mesh_dim,
          const std::vector<std::pair<MY_SIZE, unsigned>> &point_data_params,
          const std::vector<std::pair<MY_SIZE, unsigned>> &cell_data_params,
          MY_SIZE _block_size = DEFAULT_BLOCK_SIZE)
      : mesh(mesh_is, mesh_dim), block_size(_block_size) {
    init(point_data_params, cell_data_params);
  }
  // problem.hpp:39-56 (protected there; public here for generated meshes)
  Problem(Mesh &&mesh_, const std::vector<std::pair<MY_SIZE, unsigned>> &point_data_params,
          const std::vector<std::pair<MY_SIZE, unsigned>> &cell_data_params,
          MY_SIZE _block_size = DEFAULT_BLOCK_SIZE)
      : mesh(std::move(mesh_)), block_size(_block_size) {
    init(point_data_params, cell_data_params);
  }
  Problem(Problem &&) = default;

  // ---- CPU loops (problem.hpp:342-437) ----------------------------------
  template <class UserFunc> void stepCPUCellCentred(char *temp) {
    Ptrs p(*this);
    for (MY_SIZE c = 0; c < mesh.numCells(); ++c)
      UserFunc::template call<SOA>(p.point.data(), temp, p.cell.data(), p.maps.data(), c,
                                   p.stride.data(), mesh.numCells());
  }
  template <class UserFunc> void loopCPUCellCentred(MY_SIZE num) {
    data_t &pw = point_weights[0];
    std::vector<char> temp(pw.cbegin(), pw.cend());
    for (MY_SIZE i = 0; i < num; ++i) {
      stepCPUCellCentred<UserFunc>(temp.data());
      std::copy(temp.begin(), temp.end(), pw.begin());
    }
  }
  template <class UserFunc>
  void stepCPUCellCentredOMP(const std::vector<MY_SIZE> &inds, char *out) {
    Ptrs p(*this);
    const long n = long(inds.size());
#pragma omp parallel for
    for (long i = 0; i < n; ++i)
      UserFunc::template call<SOA>(p.point.data(), out, p.cell.data(), p.maps.data(),
                                   inds[std::size_t(i)], p.stride.data(), mesh.numCells());
  }
  template <class UserFunc> void loopCPUCellCentredOMP(MY_SIZE num) {
    data_t &pw = point_weights[0];
    std::vector<char> temp(pw.cbegin(), pw.cend());
    const std::vector<std::vector<MY_SIZE>> colours = mesh.colourCells();
    for (MY_SIZE i = 0; i < num; ++i) {
      for (const auto &c : colours) stepCPUCellCentredOMP<UserFunc>(c, temp.data());
      std::copy(temp.begin(), temp.end(), pw.begin());
    }
  }

  // ---- GPU loops (problem.hpp:90-339) ------------------------------------
  template <class UserFunc> void loopGPUCellCentred(MY_SIZE num) {
    runEngine(UserFunc::functor(), FOT_STRATEGY_GLOBAL_COLOUR, 0, num, nullptr);
  }
  // No reference counterpart: the plain FP64-atomics loop.
  template <class UserFunc> void loopGPUEdgeCentred(MY_SIZE num) {
    runEngine(UserFunc::functor(), FOT_STRATEGY_ATOMIC, 0, num, nullptr);
  }
  template <class UserFunc, unsigned SharedMemoryMultiplier = 0>
  void loopGPUHierarchical(MY_SIZE num) {
    const std::vector<MY_SIZE> *part = partition_vector.empty() ? nullptr : &partition_vector;
    runEngine(UserFunc::functor(), FOT_STRATEGY_HIERARCHICAL, hierarchical_flags, num, part);
  }

  // ---- pre-passes (problem.hpp:439-551) ----------------------------------
  // reorder<SameMapping>: bandwidth-reducing point order of mapping 0 (GPS
  // style, libfot's stand-in for Scotch), cells sorted by their point tuple,
  // the other mappings renumbered with the same permutation (SameMapping) or
  // by first touch.  A mapping identical to mapping 0 (res_calc's q and adt)
  // always takes mapping 0's permutation, so the three stay one map.
  template <bool SameMapping = false> void reorder() {
    std::vector<bool> same(mesh.numMappings(), SameMapping);
    for (unsigned m = 1; m < mesh.numMappings(); ++m)
      same[m] = same[m] || (mesh.numPoints(m) == mesh.numPoints(0) &&
                            mesh.cell_to_node[m].getDim() == mesh.cell_to_node[0].getDim() &&
                            std::equal(mesh.cell_to_node[m].cbegin(), mesh.cell_to_node[m].cend(),
                                       mesh.cell_to_node[0].cbegin()));
    std::vector<MY_SIZE> ind, end;
    pointGraph(0, ind, end);
    std::vector<MY_SIZE> perm(mesh.numPoints(0));
    fot::check(fot_gps_order(ind.data(), end.data(), mesh.numPoints(0), perm.data()),
               "fot_gps_order");
    const std::vector<MY_SIZE> inv = mesh.reorder(perm);
    reorderData<SOA>(point_weights[0], perm);
    for (unsigned m = 1; m < mesh.numMappings(); ++m) {
      const std::vector<MY_SIZE> p = same[m] ? perm : mesh.getPointRenumberingPermutation(m);
      reorderData<SOA>(point_weights[m], mesh.renumberPoints(p, m));
    }
    for (data_t &cw : cell_weights) reorderDataInverse<true>(cw, inv);
    compose(perm);
  }
  // partition: k-way partition of the cell-to-cell graph into parts of at
  // most floor(block_size / tolerance) cells (partitionMetisEnh,
  // partition.hpp:23-29, with libfot's partitioner in place of METIS).
  // bisect (no reference counterpart): recursive bisection into equal,
  // compact parts instead of greedy growing
  void partition(float tolerance, bool bisect = false) {
    std::vector<MY_SIZE> ind, end;
    cellGraph(ind, end);
    const MY_SIZE max_part = std::max<MY_SIZE>(1, MY_SIZE(std::floor(block_size / tolerance)));
    partition_vector.assign(mesh.numCells(), 0);
    std::uint32_t nparts = 0;
    if (bisect)
      fot::check(fot_bisect_partition(ind.data(), end.data(), mesh.numCells(), max_part,
                                      partition_vector.data(), &nparts),
                 "fot_bisect_partition");
    else
      fot::check(fot_kway_partition(ind.data(), end.data(), mesh.numCells(), max_part,
                                    partition_vector.data(), &nparts),
                 "fot_kway_partition");
  }
  // partitionOurs (partition.cu:42-215): coarse parts of block_size *
  // num_blocks_in_coarse cells, then the heuristic partition inside each
  void partitionOurs(float tolerance, MY_SIZE num_blocks_in_coarse) {
    std::vector<MY_SIZE> ind, end;
    cellGraph(ind, end);
    // METIS's ubvec = tolerance (partition.cu:55-61) lets a coarse part
    // exceed the average of block_size * num_blocks_in_coarse cells by that
    // factor: the stand-in's parts are bounded by the same size
    const MY_SIZE coarse_size = std::max<MY_SIZE>(
        1, MY_SIZE(std::floor(double(block_size) * num_blocks_in_coarse * tolerance)));
    std::vector<MY_SIZE> coarse(mesh.numCells());
    std::uint32_t nparts = 0;
    fot::check(fot_kway_partition(ind.data(), end.data(), mesh.numCells(), coarse_size,
                                  coarse.data(), &nparts),
               "fot_kway_partition");
    partition_vector.assign(mesh.numCells(), 0);
    std::uint32_t nblocks = 0;
    fot::check(fot_heuristic_partition(ind.data(), end.data(), mesh.numCells(), coarse.data(),
                                       block_size, partition_vector.data(), &nblocks),
               "fot_heuristic_partition");
  }
  void reorderToPartition() {
    const std::vector<MY_SIZE> inv = mesh.reorderToPartition(partition_vector);
    for (data_t &cw : cell_weights) reorderDataInverse<true>(cw, inv);
  }
  void renumberPoints() {
    for (unsigned m = 0; m < mesh.numMappings(); ++m) {
      const std::vector<MY_SIZE> perm = mesh.getPointRenumberingByPartition(m, partition_vector);
      mesh.renumberPoints(perm, m);
      reorderData<SOA>(point_weights[m], perm);
      if (m == 0) compose(perm);
    }
  }

  // ---- IO (problem.hpp:553-632) -------------------------------------------
  void writePartition(std::ostream &os) const {
    os << block_size << "\n";
    for (MY_SIZE p : partition_vector) os << p << "\n";
  }
  void readPartition(std::istream &is) {
    if (!is) throw InvalidInputFile{"partition input", 0, 0};
    MY_SIZE bs = 0;
    is >> bs;
    if (!is) throw InvalidInputFile{"partition input", 0, 1};
    if (bs != block_size)
      std::cerr << "Warning: block size in file (" << bs << ") doesn't equal to used block size ("
                << block_size << ")" << std::endl;
    partition_vector.resize(mesh.numCells());
    for (MY_SIZE i = 0; i < mesh.numCells(); ++i) {
      is >> partition_vector[i];
      if (!is) throw InvalidInputFile{"partition input", 0, i + 1};
    }
  }
  template <class DataType> void readPointData(std::istream &is, unsigned m) {
    if (!is) throw InvalidInputFile{"point data input", m, 0};
    data_t &d = point_weights[m];
    for (MY_SIZE i = 0; i < mesh.numPoints(m); ++i) {
      for (MY_SIZE j = 0; j < d.getDim(); ++j)
        is >> d.template operator[]<DataType>(index<SOA>(d.getSize(), i, d.getDim(), j));
      if (!is && !is.eof()) throw InvalidInputFile{"point data input", m, i};
    }
  }
  template <class DataType> void writePointData(std::ostream &os) const {
    const MY_SIZE n = mesh.numPoints(0), dim = point_weights[0].getDim();
    for (MY_SIZE i = 0; i < n; ++i) {
      const MY_SIZE p = applied_permutation.empty() ? i : applied_permutation[i];
      for (MY_SIZE j = 0; j < dim; ++j)
        os << (j ? " " : "")
           << point_weights[0].template operator[]<DataType>(index<SOA>(n, p, dim, j));
      os << "\n";
    }
  }
  template <class DataType> void readCellData(std::istream &is, unsigned k) {
    if (!is) throw InvalidInputFile{"cell data input", k, 0};
    data_t &d = cell_weights[k];
    for (MY_SIZE i = 0; i < mesh.numCells(); ++i) {
      for (MY_SIZE j = 0; j < d.getDim(); ++j)
        is >> d.template operator[]<DataType>(j * d.getSize() + i);
      if (!is) throw InvalidInputFile{"cell data input", k, i};
    }
  }
  // problem.hpp:634-652: bytes of all point data (target twice) and cell
  // data plus the mappings
  double calcDataSize() const {
    double r = 0;
    for (unsigned m = 0; m < mesh.numMappings(); ++m)
      r += double(point_weights[m].bytes()) +
           double(mesh.numCells()) * mesh.cell_to_node[m].getDim() * sizeof(MY_SIZE);
    r += double(point_weights[0].bytes());
    for (const data_t &c : cell_weights) r += double(c.bytes());
    return r;
  }

 private:
  void init(const std::vector<std::pair<MY_SIZE, unsigned>> &pd,
            const std::vector<std::pair<MY_SIZE, unsigned>> &cd) {
    if (pd.size() != mesh.numMappings())
      throw std::invalid_argument("one point data description per mapping");
    for (unsigned m = 0; m < mesh.numMappings(); ++m)
      point_weights.emplace_back(mesh.numPoints(m), pd[m].first, pd[m].second);
    for (const auto &c : cd) cell_weights.emplace_back(mesh.numCells(), c.first, c.second);
  }

  // host pointer tables for the CPU call sites (problem.hpp:343-355)
  struct Ptrs {
    std::vector<const void *> point;
    std::vector<void *> cell;
    std::vector<const MY_SIZE *> maps;
    std::vector<unsigned> stride;
    explicit Ptrs(Problem &pr) {
      for (data_t &d : pr.point_weights) point.push_back(d.cbegin());
      for (data_t &d : pr.cell_weights) cell.push_back(d.begin());
      for (data_t &d : pr.mesh.cell_to_node) maps.push_back(d.cbegin<MY_SIZE>());
      for (unsigned m = 0; m < pr.mesh.numMappings(); ++m) stride.push_back(pr.mesh.numPoints(m));
    }
  };

  void pointGraph(unsigned m, std::vector<MY_SIZE> &ind, std::vector<MY_SIZE> &end) const {
    const MY_SIZE md = mesh.cell_to_node[m].getDim();
    ind.assign(mesh.numPoints(m) + 1, 0);
    end.assign(std::max<std::size_t>(1, std::size_t(mesh.numCells()) * md * (md - 1)), 0);
    std::uint32_t narcs = 0;
    fot::check(fot_graph_csr(mesh.cell_to_node[m].cbegin<MY_SIZE>(), mesh.numCells(), md,
                             mesh.numPoints(m), ind.data(), end.data(), &narcs),
               "fot_graph_csr");
    end.resize(narcs);
  }
  void cellGraph(std::vector<MY_SIZE> &ind, std::vector<MY_SIZE> &end) const {
    std::vector<const MY_SIZE *> maps;
    std::vector<MY_SIZE> dims, np = mesh.numPoints();
    for (const data_t &d : mesh.cell_to_node) {
      maps.push_back(d.cbegin<MY_SIZE>());
      dims.push_back(d.getDim());
    }
    std::uint64_t narcs = 0;
    ind.assign(mesh.numCells() + 1, 0);
    fot::check(fot_cell_to_cell_csr(maps.data(), dims.data(), mesh.numMappings(),
                                    mesh.numCells(), np.data(), ind.data(), nullptr, &narcs),
               "fot_cell_to_cell_csr");
    end.assign(std::max<std::uint64_t>(1, narcs), 0);
    fot::check(fot_cell_to_cell_csr(maps.data(), dims.data(), mesh.numMappings(),
                                    mesh.numCells(), np.data(), ind.data(), end.data(), &narcs),
               "fot_cell_to_cell_csr");
  }
  // applied_permutation[i]: current id of original point i
  void compose(const std::vector<MY_SIZE> &perm) {
    if (applied_permutation.empty()) {
      applied_permutation = perm;
    } else {
      for (MY_SIZE &a : applied_permutation) a = perm[a];
    }
  }

  // the engine with the reference's loop semantics: out starts as a copy of
  // point data 0, and after every iteration point data 0 = out
  void runEngine(fot_functor f, std::uint32_t strategy, std::uint32_t flags, MY_SIZE num,
                 const std::vector<MY_SIZE> *part) {
    fot_problem_desc d{};
    d.num_cells = mesh.numCells();
    d.num_mappings = mesh.numMappings();
    d.soa = SOA;
    for (unsigned m = 0; m < mesh.numMappings(); ++m) {
      point_weights[m].initDeviceMemory();
      d.cell_to_node[m] = mesh.cell_to_node[m].cbegin<MY_SIZE>();
      d.mesh_dim[m] = mesh.cell_to_node[m].getDim();
      d.num_points[m] = mesh.numPoints(m);
      d.point_data[m] = point_weights[m].template getDeviceData<double>();
      d.point_dim[m] = point_weights[m].getDim();
    }
    d.num_cell_data = std::uint32_t(cell_weights.size());
    for (std::size_t k = 0; k < cell_weights.size(); ++k) {
      cell_weights[k].initDeviceMemory();
      d.cell_data[k] = cell_weights[k].template getDeviceData<double>();
      d.cell_dim[k] = cell_weights[k].getDim();
    }
    data_t out(point_weights[0].getSize(), point_weights[0].getDim(),
               point_weights[0].getTypeSize());
    out.initDeviceMemory();
    d.point_data_out = out.template getDeviceData<double>();
    fot_loop_config cfg{};
    cfg.strategy = strategy;
    cfg.block_size = block_size;
    cfg.partition = part ? part->data() : nullptr;
    cfg.flags = flags;
    fot_loop *loop = nullptr;
    fot::check(fot_loop_create(&f, &d, &cfg, &loop), "fot_loop_create");
    std::unique_ptr<fot_loop, fot_status (*)(fot_loop *)> guard(loop, fot_loop_destroy);
    fot::check(fot_loop_run(loop, num, nullptr), "fot_loop_run");
    point_weights[0].flushToHost();
  }
};

#endif  // FOT_PROBLEM_HPP
