// Host buffers with a device mirror, the mesh, and the error types of the
// C++ drop-in.  Mirrors data_t (data_t.hpp:10-53), index<SOA>
// (data_t.hpp:105-114), Mesh (mesh.hpp:23-388) and InvalidInputFile
// (mesh.hpp:18-21); all device work and planning goes through libfot's C-ABI
// (fot_capi.h).  Header-only, C++17, needs no HIP headers.
#ifndef FOT_DATA_HPP
#define FOT_DATA_HPP

#include <algorithm>
#include <cassert>
#include <cstdint>
#include <cstring>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../fot_capi.h"

#ifndef MY_SIZE
#define MY_SIZE std::uint32_t
#endif

namespace fot {

// A libfot call failed: the status and libfot's last error message.
struct Error : std::runtime_error {
  fot_status status;
  Error(fot_status s, const std::string &where)
      : std::runtime_error(where + ": " + fot_status_string(s) + ": " +
                           fot_last_error_message()),
        status(s) {}
};

}  // namespace fot

// mesh.hpp:18-21
struct InvalidInputFile {
  std::string input_type;
  MY_SIZE mapping_ind, line;
};

namespace fot {

inline void check(fot_status s, const char *where) {
  if (s == FOT_OK) return;
  if (s == FOT_ERR_INVALID_INPUT) throw InvalidInputFile{where, 0, 0};
  throw Error(s, where);
}

}  // namespace fot

// data_t.hpp:105-114
template <bool SOA = false>
inline MY_SIZE index(MY_SIZE num_points, MY_SIZE node_ind, MY_SIZE Dim, MY_SIZE dim) {
  return SOA ? dim * num_points + node_ind : node_ind * Dim + dim;
}

// data_t (data_t.hpp:10-53): `size` elements of `dim` values of `type_size`
// bytes on the host, plus a lazily allocated device copy.
class data_t {
  MY_SIZE size_, dim_;
  unsigned type_size_;
  std::vector<char> host_;
  void *dev_ = nullptr;

 public:
  data_t(MY_SIZE set_size, MY_SIZE set_dim, unsigned type_size)
      : size_(set_size), dim_(set_dim), type_size_(type_size),
        host_(std::size_t(set_size) * set_dim * type_size) {}
  template <typename T> static data_t create(MY_SIZE size, MY_SIZE dim) {
    return data_t(size, dim, sizeof(T));
  }
  data_t(const data_t &) = delete;
  data_t &operator=(const data_t &) = delete;
  data_t(data_t &&o) noexcept
      : size_(o.size_), dim_(o.dim_), type_size_(o.type_size_), host_(std::move(o.host_)),
        dev_(o.dev_) {
    o.dev_ = nullptr;
  }
  data_t &operator=(data_t &&o) noexcept {
    std::swap(size_, o.size_);
    std::swap(dim_, o.dim_);
    std::swap(type_size_, o.type_size_);
    std::swap(host_, o.host_);
    std::swap(dev_, o.dev_);
    return *this;
  }
  ~data_t() {
    if (dev_) fot_free(dev_);
  }

  template <typename T = char> T &operator[](MY_SIZE i) {
    return reinterpret_cast<T *>(host_.data())[i];
  }
  template <typename T = char> const T &operator[](MY_SIZE i) const {
    return reinterpret_cast<const T *>(host_.data())[i];
  }
  template <typename T = char> T *begin() { return reinterpret_cast<T *>(host_.data()); }
  template <typename T = char> T *end() {
    return reinterpret_cast<T *>(host_.data() + host_.size());
  }
  template <typename T = char> const T *cbegin() const {
    return reinterpret_cast<const T *>(host_.data());
  }
  template <typename T = char> const T *cend() const {
    return reinterpret_cast<const T *>(host_.data() + host_.size());
  }

  MY_SIZE getSize() const { return size_; }
  unsigned getTypeSize() const { return type_size_; }
  MY_SIZE getDim() const { return dim_; }
  std::size_t bytes() const { return host_.size(); }

  // device mirror (data_t.cu:3-56)
  void initDeviceMemory() {
    if (!dev_ && !host_.empty()) fot::check(fot_malloc(&dev_, host_.size()), "fot_malloc");
    flushToDevice();
  }
  void flushToDevice() {
    if (dev_) fot::check(fot_memcpy_htod(dev_, host_.data(), host_.size()), "fot_memcpy_htod");
  }
  void flushToHost() {
    if (dev_) fot::check(fot_memcpy_dtoh(host_.data(), dev_, host_.size()), "fot_memcpy_dtoh");
  }
  template <typename T = char> T *getDeviceData() { return static_cast<T *>(dev_); }
  template <typename T = char> const T *getDeviceData() const {
    return static_cast<const T *>(dev_);
  }
};

// reorderData<SOA> (data_t.hpp:117-146): new[perm[i]] = old[i] per element.
template <bool SOA, typename UnsignedType>
inline void reorderData(data_t &d, const std::vector<UnsignedType> &perm) {
  const MY_SIZE n = d.getSize(), dim = d.getDim();
  const unsigned ts = d.getTypeSize();
  std::vector<char> old(d.cbegin(), d.cend());
  for (MY_SIZE i = 0; i < n; ++i)
    for (MY_SIZE k = 0; k < dim; ++k)
      std::memcpy(d.begin() + std::size_t(index<SOA>(n, MY_SIZE(perm[i]), dim, k)) * ts,
                  old.data() + std::size_t(index<SOA>(n, i, dim, k)) * ts, ts);
}

// reorderDataInverse<SOA> (data_t.hpp:148-167): new[i] = old[inv[i]].
template <bool SOA>
inline void reorderDataInverse(data_t &d, const std::vector<MY_SIZE> &inv) {
  const MY_SIZE n = d.getSize(), dim = d.getDim();
  const unsigned ts = d.getTypeSize();
  std::vector<char> old(d.cbegin(), d.cend());
  for (MY_SIZE i = 0; i < n; ++i)
    for (MY_SIZE k = 0; k < dim; ++k)
      std::memcpy(d.begin() + std::size_t(index<SOA>(n, i, dim, k)) * ts,
                  old.data() + std::size_t(index<SOA>(n, inv[i], dim, k)) * ts, ts);
}

// Mesh (mesh.hpp:23-388): mappings cell -> points, AOS `dim * cell + j`.
class Mesh {
  std::vector<MY_SIZE> num_points_;

 public:
  std::vector<data_t> cell_to_node;

  // text format (mesh.hpp:58-98): "num_points num_cells" then dim ids per cell
  Mesh(const std::vector<std::istream *> &iis, const std::vector<MY_SIZE> &mesh_dim) {
    assert(iis.size() == mesh_dim.size());
    MY_SIZE num_cells = 0;
    for (std::size_t m = 0; m < iis.size(); ++m) {
      std::istream &is = *iis[m];
      if (!is) throw InvalidInputFile{"graph input", MY_SIZE(m), 0};
      MY_SIZE np = 0, nc = 0;
      is >> np >> nc;
      if (!is) throw InvalidInputFile{"graph input", MY_SIZE(m), 0};
      if (m > 0 && nc != num_cells) throw InvalidInputFile{"graph input", MY_SIZE(m), 0};
      num_cells = nc;
      num_points_.push_back(np);
      cell_to_node.emplace_back(data_t::create<MY_SIZE>(nc, mesh_dim[m]));
      MY_SIZE *c = cell_to_node.back().begin<MY_SIZE>();
      for (MY_SIZE i = 0; i < nc; ++i) {
        for (MY_SIZE j = 0; j < mesh_dim[m]; ++j) is >> c[std::size_t(mesh_dim[m]) * i + j];
        if (!is) throw InvalidInputFile{"graph input", MY_SIZE(m), i + 1};
        for (MY_SIZE j = 0; j < mesh_dim[m]; ++j)
          if (c[std::size_t(mesh_dim[m]) * i + j] >= np)
            throw InvalidInputFile{"graph input", MY_SIZE(m), i + 1};
      }
    }
  }
  // in-memory mappings (the reference's protected constructor, mesh.hpp:36-55)
  Mesh(const std::vector<MY_SIZE> &num_points, MY_SIZE num_cells,
       const std::vector<const MY_SIZE *> &maps, const std::vector<MY_SIZE> &mesh_dim)
      : num_points_(num_points) {
    assert(maps.size() == mesh_dim.size() && maps.size() == num_points.size());
    for (std::size_t m = 0; m < maps.size(); ++m) {
      cell_to_node.emplace_back(data_t::create<MY_SIZE>(num_cells, mesh_dim[m]));
      if (maps[m])
        std::copy_n(maps[m], std::size_t(num_cells) * mesh_dim[m],
                    cell_to_node.back().begin<MY_SIZE>());
    }
  }
  Mesh(Mesh &&) = default;
  Mesh &operator=(Mesh &&) = default;

  MY_SIZE numCells() const { return cell_to_node.empty() ? 0 : cell_to_node[0].getSize(); }
  MY_SIZE numPoints(unsigned m) const { return num_points_[m]; }
  const std::vector<MY_SIZE> numPoints() const { return num_points_; }
  unsigned numMappings() const { return unsigned(cell_to_node.size()); }

  // mesh.hpp:184-197
  void writeCellList(std::ostream &os, unsigned m) const {
    os << numPoints(m) << " " << numCells() << "\n";
    const MY_SIZE dim = cell_to_node[m].getDim();
    const MY_SIZE *c = cell_to_node[m].cbegin<MY_SIZE>();
    for (MY_SIZE i = 0; i < numCells(); ++i) {
      for (MY_SIZE j = 0; j < dim; ++j) os << (j ? " " : "") << c[std::size_t(dim) * i + j];
      os << "\n";
    }
  }

  // mesh.hpp:117-162: cells per colour, greedy least-used colouring on
  // mapping 0 (cells [from, to) only)
  std::vector<std::vector<MY_SIZE>> colourCells(MY_SIZE from = 0,
                                                MY_SIZE to = MY_SIZE(-1)) const {
    to = std::min(to, numCells());
    from = std::min(from, to);
    const MY_SIZE dim = cell_to_node[0].getDim();
    std::vector<std::uint8_t> colour(to - from);
    std::uint32_t nc = 0;
    fot::check(fot_colour_cells(cell_to_node[0].cbegin<MY_SIZE>() + std::size_t(dim) * from,
                                to - from, dim, numPoints(0), colour.data(), &nc),
               "fot_colour_cells");
    std::vector<std::vector<MY_SIZE>> out(nc);
    for (MY_SIZE i = 0; i < to - from; ++i) out[colour[i]].push_back(from + i);
    return out;
  }

  // mesh.hpp:272-285
  template <class UnsignedType>
  const std::vector<UnsignedType> &renumberPoints(const std::vector<UnsignedType> &perm,
                                                  unsigned m) {
    for (MY_SIZE *a = cell_to_node[m].begin<MY_SIZE>(); a != cell_to_node[m].end<MY_SIZE>(); ++a)
      *a = MY_SIZE(perm[*a]);
    return perm;
  }

  // mesh.hpp:199-225: renumber mapping 0, sort cells by sorted point tuple;
  // returns the inverse cell permutation and applies it to every mapping
  template <typename UnsignedType>
  std::vector<MY_SIZE> reorder(const std::vector<UnsignedType> &point_perm) {
    std::vector<MY_SIZE> perm(point_perm.begin(), point_perm.end());
    std::vector<MY_SIZE> inv(numCells());
    fot::check(fot_mesh_reorder_cells(cell_to_node[0].begin<MY_SIZE>(), numCells(),
                                      cell_to_node[0].getDim(), perm.data(), inv.data()),
               "fot_mesh_reorder_cells");
    for (unsigned m = 1; m < numMappings(); ++m) reorderDataInverse<false>(cell_to_node[m], inv);
    return inv;
  }

  // mesh.hpp:227-245: sorts the partition, permutes every mapping
  std::vector<MY_SIZE> reorderToPartition(std::vector<MY_SIZE> &partition_vector) {
    std::vector<MY_SIZE> inv(numCells());
    fot::check(fot_reorder_to_partition(partition_vector.data(), numCells(), inv.data()),
               "fot_reorder_to_partition");
    for (data_t &c : cell_to_node) reorderDataInverse<false>(c, inv);
    return inv;
  }

  // mesh.hpp:247-270: first-touch renumbering of mapping m
  std::vector<MY_SIZE> getPointRenumberingPermutation(unsigned m) const {
    std::vector<MY_SIZE> perm(numPoints(m));
    fot::check(fot_point_renumbering_first_touch(cell_to_node[m].cbegin<MY_SIZE>(), numCells(),
                                                 cell_to_node[m].getDim(), numPoints(m),
                                                 perm.data()),
               "fot_point_renumbering_first_touch");
    return perm;
  }

  // mesh.hpp:346-388 (getPointToPartition + getPointRenumberingPermutation2)
  std::vector<MY_SIZE> getPointRenumberingByPartition(unsigned m,
                                                      const std::vector<MY_SIZE> &part) const {
    std::vector<MY_SIZE> perm(numPoints(m));
    fot::check(fot_point_renumbering_by_partition(cell_to_node[m].cbegin<MY_SIZE>(), numCells(),
                                                  cell_to_node[m].getDim(), numPoints(m),
                                                  part.data(), perm.data()),
               "fot_point_renumbering_by_partition");
    return perm;
  }
};

#endif  // FOT_DATA_HPP
