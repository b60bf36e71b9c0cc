// fot_features.cpp -- cheap features of a loop for the learned cost model:
// sizes, shapes, the point-degree distribution, the global colour count and
// the index locality of mapping 0 (cell spans and jumps between consecutive
// cells).  No plan is built.  No reference counterpart.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fot_internal.hpp"

using fot::fail;

static fot_status fot_loop_features_impl(const fot_functor *f, const fot_problem_desc *d,
                                        float *out) {
  if (!f || !d || !out) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  if (d->num_mappings == 0 || d->num_cells == 0 || !d->cell_to_node[0])
    return fail(FOT_ERR_INVALID_ARGUMENT, "empty problem");
  const uint32_t nc = d->num_cells, md = d->mesh_dim[0], np = std::max(1u, d->num_points[0]);
  const uint32_t *m0 = d->cell_to_node[0];
  std::vector<uint32_t> deg(np, 0);
  for (size_t i = 0; i < size_t(nc) * md; ++i) {
    if (m0[i] >= np) return fail(FOT_ERR_INVALID_INPUT, "mapping 0 entry out of range");
    ++deg[m0[i]];
  }
  double mean = double(nc) * md / np, var = 0, mx = 0;
  double hist[4] = {0, 0, 0, 0};
  for (uint32_t v : deg) {
    var += (v - mean) * (v - mean);
    mx = std::max(mx, double(v));
    hist[v <= 2 ? 0 : v <= 4 ? 1 : v <= 8 ? 2 : 3] += 1;
  }
  var /= np;
  std::vector<uint8_t> col(nc);
  uint32_t ncol = 0;
  const fot_status cs = fot_colour_cells(m0, nc, md, np, col.data(), &ncol);
  if (cs == FOT_ERR_COLOUR_OVERFLOW) ncol = 65; // more than the 64 colours
  else if (cs != FOT_OK) return cs;
  double span = 0, span256 = 0, span4k = 0, jump = 0, jump256 = 0, jump4k = 0;
  uint32_t prev_min = 0;
  for (uint32_t c = 0; c < nc; ++c) {
    uint32_t lo = m0[size_t(c) * md], hi = lo;
    for (uint32_t k = 1; k < md; ++k) {
      lo = std::min(lo, m0[size_t(c) * md + k]);
      hi = std::max(hi, m0[size_t(c) * md + k]);
    }
    const double s = hi - lo;
    span += s / np, span256 += s < 256, span4k += s < 4096;
    if (c > 0) {
      const double j = std::fabs(double(lo) - double(prev_min));
      jump += j / np, jump256 += j < 256, jump4k += j < 4096;
    }
    prev_min = lo;
  }
  double cell_doubles = 0, gather_doubles = 0, bytes = 0;
  for (uint32_t k = 0; k < d->num_cell_data; ++k) cell_doubles += d->cell_dim[k];
  for (uint32_t m = 1; m < d->num_mappings; ++m)
    gather_doubles += double(d->mesh_dim[m]) * d->point_dim[m];
  bytes += 16.0 * np * d->point_dim[0];                 // target read + write
  for (uint32_t m = 1; m < d->num_mappings; ++m)         // other point data once
    bytes += 8.0 * d->num_points[m] * d->point_dim[m];
  for (uint32_t m = 0; m < d->num_mappings; ++m) bytes += 4.0 * nc * d->mesh_dim[m];
  bytes += 8.0 * nc * cell_doubles;
  const double n1 = std::max(1.0, double(nc - 1));
  const float v[24] = {
      float(std::log2(double(nc))),
      float(std::log2(double(np))),
      float(mean),
      float(md),
      float(d->point_dim[0]),
      float(cell_doubles),
      float(gather_doubles),
      float(std::sqrt(var) / std::max(mean, 1e-9)),
      float(mx / std::max(mean, 1e-9)),
      float(hist[0] / np),
      float(hist[1] / np),
      float(hist[2] / np),
      float(hist[3] / np),
      float(ncol),
      float(span / nc),
      float(span256 / nc),
      float(span4k / nc),
      float(jump / n1),
      float(jump256 / n1),
      float(jump4k / n1),
      float(std::log2(std::max(bytes, 1.0))),
      float(f->kernel == FOT_KERNEL_RES_CALC ? 1 : 0),
      float(f->kernel == FOT_KERNEL_MINE ? 1 : 0),
      float(std::log2(8.0 * np * d->point_dim[0])),
  };
  std::copy(v, v + 24, out);
  return FOT_OK;
}
extern "C" fot_status fot_loop_features(const fot_functor *f, const fot_problem_desc *d,
                                        float *out) {
  return fot::guarded("fot_loop_features", [&] { return fot_loop_features_impl(f, d, out); });
}
