/*
 * fot_oracle.h -- CPU oracle for the edge-centred indirect-increment loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libfot, the Problem
 * header library) links, calls or includes this file.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only as
 * the checker or as the timed CPU baseline.
 *
 * It restates the reference's serial loop, its OpenMP colour-parallel loop,
 * the user functions and the planning algorithms (global colouring,
 * hierarchical block/thread colouring, partition reordering and point
 * renumbering) in plain C.  Each function cites the reference file:line it
 * follows.  Pinning of each function (golden vectors, invariants, closed
 * forms) is done by tests/test_oracle_*.py; a function without such a check
 * would say "parity unpinned" here -- every function below is pinned.
 */
#ifndef FOT_ORACLE_H
#define FOT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_MAPPINGS 8
#define ORACLE_MAX_CELL_DATA 8

/* kernels (user functions) restated by the oracle */
enum {
  ORACLE_RES_CALC = 0,  /* kernels/res_calc_func.hpp:11-57            */
  ORACLE_MINE2 = 1,     /* kernels/mine_func.hpp:11-29, MeshDim 2     */
  ORACLE_MINE4 = 2,     /* kernels/mine_func.hpp:31-56, MeshDim 4     */
  ORACLE_MINE8 = 3,     /* kernels/mine_func.hpp:58-99, MeshDim 8     */
  ORACLE_HEX_SCATTER = 4, /* lulesh struct API, scatter-only functor  */
  ORACLE_MINI_AERO = 5, /* kernels/mini_aero_func.hpp:12-85           */
  ORACLE_VOLNA = 6,     /* kernels/volna_func.hpp:16-36 in FP64       */
  ORACLE_GETACC = 7,    /* kernels/getacc_scatter_func.hpp:17-94      */
  ORACLE_LULESH = 8     /* kernels/lulesh_func.hpp:13-38              */
};

/*
 * A problem as the reference's Problem<SOA> holds it (problem.hpp:29-35):
 * mapping k has cell_to_node[k] (AOS, mesh_dim[k] ids per cell, u32) into a
 * set of num_points[k] points carrying point_data[k] (point_dim[k] doubles,
 * AOS or SOA by `soa`).  point_data[0] is the increment target.  Cell data is
 * always SOA with stride num_cells.
 */
typedef struct {
  int kernel;
  int soa;
  uint32_t num_cells;
  uint32_t num_mappings;
  const uint32_t *cell_to_node[ORACLE_MAX_MAPPINGS];
  uint32_t mesh_dim[ORACLE_MAX_MAPPINGS];
  uint32_t num_points[ORACLE_MAX_MAPPINGS];
  double *point_data[ORACLE_MAX_MAPPINGS];
  uint32_t point_dim[ORACLE_MAX_MAPPINGS];
  const double *cell_data[ORACLE_MAX_CELL_DATA];
  uint32_t cell_dim[ORACLE_MAX_CELL_DATA];
  uint32_t num_cell_data;
} oracle_problem;

/* ---- user functions (exposed for closed-form tests) ------------------- */
void oracle_res_calc_func(const double *x1, const double *x2, const double *q1,
                          const double *q2, const double *adt1,
                          const double *adt2, double *res1, double *res2,
                          unsigned x_stride, unsigned q_stride);

/* lulesh's element function (kernels/lulesh_func.hpp:13-38): corner
 * coordinates x, y, z [8] and the diagonal stresses give the corner forces
 * f[i * 8 + j] (component i, corner j) and the element volume *determ. */
void oracle_lulesh_func(const double *x, const double *y, const double *z, double sxx,
                        double syy, double szz, double *f, double *determ);

/* ---- loops ------------------------------------------------------------- */
/* Serial loop, cells in index order, Jacobi double buffer
 * (problem.hpp:342-380).  point_data[0] is updated in place.  If abs_sum is
 * non-NULL (same layout as point_data[0]) it receives S = |initial| + sum of
 * |increment| over all iterations: the error scale of the parity bound. */
int oracle_loop_serial(const oracle_problem *p, uint32_t iters,
                       double *abs_sum, double *loop_seconds);

/* Colour-major loop (problem.hpp:382-437): global colouring on mapping 0,
 * then cells colour by colour.  nthreads > 1 runs each colour with OpenMP
 * (the CPU baseline); the result is identical for any thread count.
 * loop_seconds (nullable) receives the iteration loop's time, colouring
 * excluded, as the reference times it. */
int oracle_loop_colour(const oracle_problem *p, uint32_t iters, int nthreads,
                       double *loop_seconds);

/* Replays the summation order of the hierarchical GPU loop on the CPU for a
 * plan built by oracle_hplan_build: per block, a zeroed cache accumulates the
 * increments thread colour by thread colour, then the cache is added to the
 * output; blocks run in plan order (block colour major). */
struct oracle_hplan;
int oracle_loop_hier_replay(const oracle_problem *p,
                            const struct oracle_hplan *plan, uint32_t iters);

/* Owner-compute plan as libfot exports it (fot_oplan_view): blocks own
 * mapping-0 points; every cell touching an owned point runs in the block. */
typedef struct {
  uint32_t mesh_dim, num_blocks;
  const uint32_t *blk_cell_off, *blk_cache_off, *blk_owned, *cached, *cells, *loc;
  const uint8_t *colour, *blk_ncol;
} oracle_oplan;

/* Replays the summation order of an owner-compute plan: per block a cache
 * holding its owned points' current values takes the increments of its
 * cells thread colour by thread colour (halo slots dropped), then replaces
 * those values. */
int oracle_loop_owner_replay(const oracle_problem *p, const oracle_oplan *plan,
                             uint32_t iters);

/* ---- planning ---------------------------------------------------------- */
/* Mesh::colourCells (mesh.hpp:117-162): greedy colouring on mapping 0,
 * least-used available colour among 64.  colour_out[num_cells].  Returns the
 * number of colours, or -1 if 64 colours do not suffice. */
int oracle_colour_cells(const uint32_t *cell_to_node, uint32_t num_cells,
                        uint32_t mesh_dim, uint32_t num_points,
                        uint8_t *colour_out);

/* colourCellsWithOrdering (colouring.cu:94-102, partition.cu:281-301):
 * in-block cell graph, smallest-last order, greedy first-fit colouring.
 * mapping has num_cells*mapping_dim local point ids.  Returns the number of
 * colours (<= 128) or -1. */
int oracle_colour_cells_with_ordering(const uint32_t *mapping,
                                      uint32_t num_cells, uint32_t mapping_dim,
                                      uint8_t *colours_out);

/* HierarchicalColourMemory<SOA> (colouring.hpp:44-335), host side. */
typedef struct oracle_hplan {
  uint32_t num_colours;   /* block colours */
  uint32_t num_mappings;
  uint32_t mesh_dim[ORACLE_MAX_MAPPINGS];
  /* per block colour c, arrays are concatenated with offsets: */
  uint32_t *colour_cell_offset;  /* [num_colours+1] cells per colour       */
  uint32_t *colour_block_offset; /* [num_colours+1] blocks per colour      */
  uint32_t *colour_cache_offset; /* [num_colours+1] cached points per col. */
  /* per colour, SOA with stride = #cells in colour, like cellListSOA */
  uint32_t *cell_list[ORACLE_MAX_MAPPINGS]; /* [mesh_dim*num_cells]         */
  uint8_t *cell_colours;         /* [num_cells]                           */
  uint8_t *num_cell_colours;     /* [num_blocks]                          */
  uint32_t *block_offsets;       /* [num_blocks + num_colours] per colour:
                                    nb_c+1 entries, relative to colour    */
  uint32_t *points_to_be_cached; /* [total cached]                        */
  uint32_t *points_to_be_cached_offsets; /* [num_blocks + num_colours]    */
  uint32_t *cell_inv_permutation; /* [num_cells] plan position -> cell    */
  uint32_t num_cells, num_blocks, num_cached;
} oracle_hplan;

int oracle_hplan_build(const uint32_t *const *cell_to_node,
                       const uint32_t *mesh_dim, uint32_t num_mappings,
                       uint32_t num_cells, uint32_t num_points0,
                       const uint32_t *partition /* NULL: chunks */,
                       uint32_t block_size, oracle_hplan *out);
void oracle_hplan_free(oracle_hplan *plan);

/* Mesh::reorderToPartition (mesh.hpp:227-245): returns the inverse cell
 * permutation (new cell i = old cell inv[i]) and sorts partition in place. */
void oracle_reorder_to_partition(uint32_t *partition, uint32_t num_cells,
                                 uint32_t *inv_perm_out);

/* Mesh::getPointToPartition + getPointRenumberingPermutation2
 * (mesh.hpp:346-388): point permutation (old -> new). */
void oracle_point_renumbering_by_partition(const uint32_t *cell_to_node,
                                           uint32_t num_cells,
                                           uint32_t mesh_dim,
                                           uint32_t num_points,
                                           const uint32_t *partition,
                                           uint32_t *perm_out);

/* Mesh::getPointRenumberingPermutation (mesh.hpp:247-270): first touch. */
void oracle_point_renumbering_first_touch(const uint32_t *cell_to_node,
                                          uint32_t num_cells,
                                          uint32_t mesh_dim,
                                          uint32_t num_points,
                                          uint32_t *perm_out);

/* Mesh::reorder (mesh.hpp:199-225) given the mapping-0 point permutation:
 * renumbers mapping 0 in place and returns the inverse cell permutation
 * (cells sorted by their sorted point tuple). */
void oracle_mesh_reorder_cells(uint32_t *cell_to_node0, uint32_t num_cells,
                               uint32_t mesh_dim, const uint32_t *point_perm,
                               uint32_t *inv_perm_out);

/* GraphCSR<T>(mesh) (reorder.hpp:39-83): point-to-point CSR.  indices has
 * num_points+1 entries; endpoints must hold at least num_cells*mesh_dim*
 * (mesh_dim-1) entries.  Returns the number of arcs. */
uint32_t oracle_graph_csr(const uint32_t *cell_to_node, uint32_t num_cells,
                          uint32_t mesh_dim, uint32_t num_points,
                          uint32_t *indices, uint32_t *endpoints);

/* Mesh::getCellToCellGraph (mesh.hpp:317-344): sorted unique (i, j) pairs,
 * i < j, of cells sharing a point in any mapping.  pairs_out holds up to
 * max_pairs pairs; returns the number of pairs (may exceed max_pairs). */
uint64_t oracle_cell_to_cell_graph(const uint32_t *const *cell_to_node,
                                   const uint32_t *mesh_dim,
                                   uint32_t num_mappings, uint32_t num_cells,
                                   const uint32_t *num_points,
                                   uint32_t *pairs_out, uint64_t max_pairs);

/* HeuristicalPartition (partition.cu:128-275) given the coarse partition of
 * the cell-to-cell graph (the METIS step it cannot run here).  Implemented in
 * oracle_partition.cpp because its tie-breaking is that of libstdc++'s
 * std::multimap.  csr is the cell-to-cell GraphCSR. */
int oracle_heuristic_partition(const uint32_t *csr_indices,
                               const uint32_t *csr_endpoints,
                               uint32_t num_vertices,
                               const uint32_t *coarse_partition,
                               uint32_t block_size, uint32_t *result_out);

/* PriorityQueue (priority_queue.hpp:7-58) over int priorities, for the
 * reference's own priority-queue tests.  ops: 0 push(i,p), 1 popMax -> out
 * (index, priority), 2 modify(i,p), 3 getPriority(i) -> out (found, prio).
 * Returns number of outputs written. */
int oracle_pq_run(const int *ops, uint32_t num_ops, int *out);

#ifdef __cplusplus
}
#endif

#endif
