/*
 * fot_capi.h -- C-ABI of libfot, the MI355X-native edge-centred
 * indirect-increment loop engine.
 *
 * Plain C: pointers, sizes and status codes only.  Everything here is exported
 * by cuda_fot_amd/lib/libfot.so.  Each entry point names the reference
 * interface of bgd54/cuda_fot it replaces (file:line); entry points with no
 * reference counterpart say so.
 *
 * Conventions
 *  - Every call returns fot_status; FOT_OK is 0.  The reference aborts the
 *    process on any CUDA error (helper_cuda.h checkCudaErrors -> exit); this
 *    library never exits: it returns a status and keeps a per-thread message
 *    (fot_last_error_message).
 *  - "device" pointers are HIP device memory; "host" pointers are CPU memory.
 *  - Mapping arrays are u32 (the reference's MY_SIZE, Makefile:75), AOS
 *    (mesh_dim ids per cell) unless stated.  Point data is FP64, AOS or SOA as
 *    the `soa` flag says (data_t.hpp:104-115 index<SOA>); cell data is SOA
 *    with stride num_cells.
 *  - Streams are hipStream_t passed as void*; NULL is the default stream.
 */
#ifndef FOT_CAPI_H
#define FOT_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FOT_MAX_MAPPINGS 8
#define FOT_MAX_CELL_DATA 8

typedef enum {
  FOT_OK = 0,
  FOT_ERR_INVALID_ARGUMENT = 1,
  FOT_ERR_HIP = 2,
  FOT_ERR_OUT_OF_MEMORY = 3,
  /* more colours than the reference's colour sets hold (mesh.hpp:25-29:
   * 64 global colours, 128 block / thread colours) */
  FOT_ERR_COLOUR_OVERFLOW = 4,
  FOT_ERR_UNSUPPORTED = 5,
  /* malformed mesh / partition / data stream (mesh.hpp:18-21
   * InvalidInputFile) */
  FOT_ERR_INVALID_INPUT = 6,
  FOT_ERR_NO_DEVICE = 7,
  /* a bounded device-side wait ran out (work-queue kernels) */
  FOT_ERR_TIMEOUT = 8
} fot_status;

/* ------------------------------------------------------------------ */
/* Library, errors                                                      */
/* ------------------------------------------------------------------ */
/* No reference counterpart (the reference exits on error). */
const char *fot_status_string(fot_status status);
const char *fot_last_error_message(void);
int fot_version(void);

/* ------------------------------------------------------------------ */
/* Device memory, streams, timers                                       */
/* Replace data_t / device_data_t device mirrors (data_t.hpp:10-74,     */
/* data_t.cu:1-93) and CudaTimer (timer.hpp:70-102).                    */
/* ------------------------------------------------------------------ */
fot_status fot_device_count(int *count);
fot_status fot_set_device(int device);
fot_status fot_malloc(void **device_ptr, size_t bytes);
fot_status fot_free(void *device_ptr);
fot_status fot_memcpy_htod(void *device_dst, const void *host_src, size_t bytes);
fot_status fot_memcpy_dtoh(void *host_dst, const void *device_src, size_t bytes);
fot_status fot_memcpy_dtod(void *device_dst, const void *device_src, size_t bytes,
                           void *stream);
fot_status fot_device_synchronize(void);
fot_status fot_stream_create(void **stream);
fot_status fot_stream_destroy(void *stream);
fot_status fot_stream_synchronize(void *stream);
/* copyKernel (data_t.cu:76-93): out -> in after each iteration, in
 * 16-byte vectors. */
fot_status fot_copy(void *device_dst, const void *device_src, size_t bytes,
                    void *stream);

typedef struct fot_timer fot_timer;
fot_status fot_timer_create(fot_timer **timer);
fot_status fot_timer_destroy(fot_timer *timer);
fot_status fot_timer_start(fot_timer *timer, void *stream);
fot_status fot_timer_stop(fot_timer *timer, void *stream);
/* waits for the stop event, returns start->stop in milliseconds */
fot_status fot_timer_elapsed_ms(fot_timer *timer, float *ms);

/* ------------------------------------------------------------------ */
/* Kernel functors                                                       */
/* ------------------------------------------------------------------ */
typedef enum {
  /* Airfoil edge flux, kernels/res_calc.hpp + res_calc_func.hpp.
   * mappings: 0 res (dim 4, target), 1 x (dim 2), 2 q (dim 4), 3 adt (1) */
  FOT_KERNEL_RES_CALC = 0,
  /* synthetic increment, kernels/mine.hpp + mine_func.hpp; mesh_dim 2/4/8,
   * point_dim = PointDim, cell_dim = CellDim (1 or PointDim) */
  FOT_KERNEL_MINE = 1,
  /* scatter-only hex functor behind the lulesh struct API
   * (kernels/lulesh.hpp:7-10 shapes): 8 nodes, dim 3, 24 corner forces as
   * cell data [i * 8 + j] */
  FOT_KERNEL_HEX_SCATTER = 2,
  /* face flux, kernels/mini_aero.hpp + mini_aero_func.hpp: mapping 0 flux
   * (dim 5, target), mapping 1 cell data (dim 28), 12 face doubles */
  FOT_KERNEL_MINI_AERO = 3,
  /* shallow-water edge increment in FP64, kernels/volna.hpp +
   * volna_func.hpp: mappings 0 and 1 the same edge->cell map, point data 0
   * state (dim 4, target), point data 1 cell volume (dim 1); cell data
   * fluxes (3), bathymetry source (2), edge normals (2), right-boundary flag
   * (1, stored as 0.0 / 1.0 where the reference has an int) */
  FOT_KERNEL_VOLNA = 4,
  /* corner scatter of a 2D quad mesh, kernels/getacc_scatter.hpp +
   * getacc_scatter_func.hpp: 4 points of dim 4 (ndmass, ndarea, ndub, ndvb);
   * cell data cnmass (4), rho (1), cnwt (4), cnfx (4), cnfy (4) */
  FOT_KERNEL_GETACC = 5,
  /* lulesh element forces, kernels/lulesh.hpp + lulesh_func.hpp:13-38:
   * 8 points of dim 3 (the forces, target); point data 1 the node
   * coordinates (dim 3, same point set, read through mapping 0); cell data 0
   * the diagonal stresses (3), cell data 1 the element volume (1), written */
  FOT_KERNEL_LULESH = 6
} fot_kernel;

typedef struct {
  uint32_t kernel;    /* fot_kernel */
  uint32_t mesh_dim;  /* MeshDim of mapping 0 */
  uint32_t point_dim; /* dim of the increment target */
  uint32_t cell_dim;  /* dim of cell data 0 (mine: 1 or point_dim) */
} fot_functor;

/* Fills the functor description of a kernel with its fixed shapes
 * (mine: pass mesh_dim/point_dim/cell_dim; others ignore them). */
fot_status fot_functor_init(fot_functor *f, fot_kernel kernel, uint32_t mesh_dim,
                            uint32_t point_dim, uint32_t cell_dim);

/*
 * One launch of the global-colouring step over the cells of one colour.
 * Replaces <app>::StepGPUGlobal::call<SOA> (kernels/res_calc.hpp:164-177,
 * kernels/mine.hpp:197-212, kernels/lulesh.hpp:71-83,
 * kernels/mini_aero.hpp:124-137).  point_data, cell_data, cell_to_node and
 * point_stride are DEVICE arrays (of device pointers / sizes), exactly as the
 * reference's Problem::loopGPUCellCentred builds them (problem.hpp:158-176).
 * cell_to_node[k] is the colour's compacted AOS list; cell data SOA with
 * stride cell_stride.  Increments are added with plain read-modify-write.
 */
fot_status fot_step_gpu_global(const fot_functor *f, int soa,
                               const void *const *point_data, void *point_data_out,
                               void *const *cell_data,
                               const uint32_t *const *cell_to_node,
                               uint32_t num_cells, const uint32_t *point_stride,
                               uint32_t cell_stride, uint32_t num_blocks,
                               uint32_t block_size, void *stream);

/*
 * Same arguments, over ALL cells in one launch, increments added with FP64
 * global atomics (hardware global atomic add, 64-lane coalesced map reads).
 * No reference counterpart: the reference has no atomics; this is the launch
 * behind Problem::loopGPUEdgeCentred.
 */
fot_status fot_step_gpu_atomic(const fot_functor *f, int soa,
                               const void *const *point_data, void *point_data_out,
                               void *const *cell_data,
                               const uint32_t *const *cell_to_node,
                               uint32_t num_cells, const uint32_t *point_stride,
                               uint32_t cell_stride, uint32_t num_blocks,
                               uint32_t block_size, void *stream);

/*
 * One launch of the hierarchical step over the blocks of one block colour.
 * Replaces <app>::StepGPUHierarchical::call<SOA> (kernels/res_calc.hpp:257-276,
 * kernels/mine.hpp:460-481, kernels/lulesh.hpp:140-160,
 * kernels/mini_aero.hpp:225-245).  Arguments as the reference's
 * Problem::loopGPUHierarchical passes them (problem.hpp:289-301): plan arrays
 * of HierarchicalColourMemory for one colour, cell_to_node[k] SOA with stride
 * num_cells, cell_to_node[0] holding block-local point indices.  cache_size is
 * the dynamic LDS size in bytes.
 */
fot_status fot_step_gpu_hierarchical(
    const fot_functor *f, int soa, const void *const *point_data,
    void *point_data_out, const uint32_t *points_to_be_cached,
    const uint32_t *points_to_be_cached_offsets, void *const *cell_data,
    const uint32_t *const *cell_to_node, const uint8_t *num_cell_colours,
    const uint8_t *cell_colours, const uint32_t *block_offsets,
    uint32_t num_cells, const uint32_t *point_stride, uint32_t cell_stride,
    uint32_t num_blocks, uint32_t block_size, uint32_t cache_size, void *stream);

/* mini_aero's user function (kernels/mini_aero_func.hpp:11-85) on the host,
 * the same code the kernels run: flux[5] of a face from the left / right
 * cells' 28 values (5 conservatives, 3 coordinates, 15 gradients,
 * 5 limiters; element stride cell_stride) and the face's 12 values
 * (coordinates, normal, tangent, binormal; stride face_stride). */
fot_status fot_mini_aero_face_flux(const double *left, const double *right,
                                   uint32_t cell_stride, const double *face,
                                   uint32_t face_stride, double *flux);

/* lulesh's element function (kernels/lulesh_func.hpp:13-38, the reference's
 * user_func_host) on the host, the same code the kernels run: corner
 * coordinates x, y, z [8], stresses sig[0], sig[sig_stride],
 * sig[2 * sig_stride] -> forces f[i * 8 + j] (component i, corner j) and the
 * element volume. */
fot_status fot_lulesh_element(const double *x, const double *y, const double *z,
                              const double *sig, uint32_t sig_stride, double *f,
                              double *determ);

/* Reads (and with clear != 0 resets) the flag a hierarchical launch sets
 * when a block's points do not fit cache_size; such blocks are skipped.  No
 * reference counterpart (the reference writes past its shared memory). */
fot_status fot_lds_overflow_flag(uint32_t *flag, int clear);

/* ------------------------------------------------------------------ */
/* Host planning (pure CPU, no device needed)                            */
/* ------------------------------------------------------------------ */

/* Mesh::colourCells (mesh.hpp:117-162): greedy colouring on one mapping,
 * least-used available colour among 64.  colour_out[num_cells].
 * FOT_ERR_COLOUR_OVERFLOW if 64 colours do not suffice. */
fot_status fot_colour_cells(const uint32_t *cell_to_node, uint32_t num_cells,
                            uint32_t mesh_dim, uint32_t num_points,
                            uint8_t *colour_out, uint32_t *num_colours);

/* colourCellsWithOrdering (colouring.cu:94-102): in-block colouring of
 * cells given block-local point ids; smallest-last order then first fit. */
fot_status fot_colour_cells_with_ordering(const uint32_t *mapping,
                                          uint32_t num_cells,
                                          uint32_t mapping_dim,
                                          uint8_t *colours_out,
                                          uint32_t *num_colours);

/* HierarchicalColourMemory<SOA> (colouring.hpp:22-404), host part. */
typedef struct fot_hplan fot_hplan;
fot_status fot_hplan_build(const uint32_t *const *cell_to_node,
                           const uint32_t *mesh_dim, uint32_t num_mappings,
                           uint32_t num_cells, uint32_t num_points0,
                           const uint32_t *partition /* nullable */,
                           uint32_t block_size, fot_hplan **plan);
void fot_hplan_free(fot_hplan *plan);

typedef struct {
  uint32_t num_colours; /* block colours */
  uint32_t num_blocks;
  uint32_t num_cells;
  uint32_t num_cached;  /* sum over blocks of points_to_be_cached */
  uint32_t max_cached;  /* largest per-block cache */
  uint32_t max_block;   /* largest block (cells) */
  uint32_t max_thread_colours;
  double avg_thread_colours;
} fot_hplan_info;
fot_status fot_hplan_info_get(const fot_hplan *plan, fot_hplan_info *info);

/* One block colour of the plan (MemoryOfOneColour, colouring.hpp:24-41).
 * Pointers stay owned by the plan.  cell_list[k] is SOA with stride
 * num_cells (cellListSOA, colouring.hpp:300-308); cell_inv_permutation[i] is
 * the original cell at plan position i of this colour (copyCellWeights,
 * colouring.hpp:310-335). */
typedef struct {
  uint32_t num_blocks, num_cells, num_cached, shared_size;
  const uint32_t *points_to_be_cached;
  const uint32_t *points_to_be_cached_offsets; /* num_blocks + 1 */
  const uint32_t *cell_list[FOT_MAX_MAPPINGS];
  const uint8_t *cell_colours;                 /* num_cells */
  const uint8_t *num_cell_colours;             /* num_blocks */
  const uint32_t *block_offsets;               /* num_blocks + 1 */
  const uint32_t *cell_inv_permutation;        /* num_cells */
} fot_hplan_colour;
fot_status fot_hplan_colour_get(const fot_hplan *plan, uint32_t colour,
                                fot_hplan_colour *out);

/* Owner-compute plan (no reference counterpart): each mapping-0 point is
 * owned by one block -- its part in point_partition when given; else, with a
 * per-cell partition, the first block of the hierarchical cell split whose
 * cells touch it; with neither, its part in a partition of mapping 0's point
 * graph into parts of at most block_size points.  Blocks running more than
 * 1024 cells split their owned points.  A block runs every cell touching an owned
 * point, in thread colours that separate cells sharing an owned point, and
 * writes back only its owned points: its LDS cache starts from the owned
 * points' current values, takes the increments colour by colour and is
 * stored back.  Blocks never share a written point, so one launch needs no
 * block colouring and the result does not depend on block order. */
typedef struct fot_oplan fot_oplan;
typedef struct {
  uint32_t mesh_dim, num_blocks, num_positions, num_cached;
  const uint32_t *blk_cell_off;  /* num_blocks + 1 */
  const uint32_t *blk_cache_off; /* num_blocks + 1 */
  const uint32_t *blk_owned;     /* num_blocks: owned points (prefix) */
  const uint32_t *cached;        /* owned ascending, then halo ascending */
  const uint32_t *cells;         /* original cell per plan position */
  const uint32_t *loc;           /* num_positions * mesh_dim local ids */
  const uint8_t *colour;         /* num_positions */
  const uint8_t *blk_ncol;       /* num_blocks */
} fot_oplan_view;
fot_status fot_oplan_build(const uint32_t *cell_to_node0, uint32_t num_cells,
                           uint32_t mesh_dim, uint32_t num_points,
                           const uint32_t *partition /* nullable, per cell */,
                           uint32_t block_size,
                           const uint32_t *point_partition /* nullable, per point */,
                           fot_oplan **plan);
fot_status fot_oplan_get(const fot_oplan *plan, fot_oplan_view *view);
void fot_oplan_free(fot_oplan *plan);

/* countCacheLinesForBlock (problem.hpp:654-672) with a line size parameter
 * (the reference's 32-byte sectors, or CDNA's 128-byte lines): distinct
 * lines holding the data of `points` (dim values of type_size bytes each,
 * AOS, or SOA where each component counts its own lines). */
fot_status fot_count_cache_lines(const uint32_t *points, uint32_t n, uint32_t dim,
                                 uint32_t type_size, uint32_t line_bytes, int soa,
                                 uint64_t *lines);

/* Mesh::reorderToPartition (mesh.hpp:227-245): sorts partition in place,
 * writes the inverse cell permutation (new cell i = old inv_perm[i]). */
fot_status fot_reorder_to_partition(uint32_t *partition, uint32_t num_cells,
                                    uint32_t *inv_perm);

/* Mesh::getPointToPartition + getPointRenumberingPermutation2
 * (mesh.hpp:346-388): point permutation old -> new. */
fot_status fot_point_renumbering_by_partition(const uint32_t *cell_to_node,
                                              uint32_t num_cells,
                                              uint32_t mesh_dim,
                                              uint32_t num_points,
                                              const uint32_t *partition,
                                              uint32_t *perm);

/* Mesh::getPointRenumberingPermutation (mesh.hpp:247-270), first touch. */
fot_status fot_point_renumbering_first_touch(const uint32_t *cell_to_node,
                                             uint32_t num_cells,
                                             uint32_t mesh_dim,
                                             uint32_t num_points,
                                             uint32_t *perm);

/* Mesh::reorder (mesh.hpp:199-225), mapping-0 part: renumber with
 * point_perm, sort cells by their sorted point tuple.  Rewrites
 * cell_to_node0 and returns the inverse cell permutation. */
fot_status fot_mesh_reorder_cells(uint32_t *cell_to_node0, uint32_t num_cells,
                                  uint32_t mesh_dim, const uint32_t *point_perm,
                                  uint32_t *inv_perm);

/* GraphCSR<T>(mesh) (reorder.hpp:39-83): point-to-point CSR.  endpoints
 * holds num_cells*mesh_dim*(mesh_dim-1) entries; *num_arcs gets the count. */
fot_status fot_graph_csr(const uint32_t *cell_to_node, uint32_t num_cells,
                         uint32_t mesh_dim, uint32_t num_points,
                         uint32_t *indices, uint32_t *endpoints,
                         uint32_t *num_arcs);

/* Mesh::getCellToCellGraph (mesh.hpp:317-344) as a CSR over cells with
 * ascending neighbour rows (the GraphCSR of the pair mesh).  Two calls:
 * endpoints == NULL returns *num_arcs only. */
fot_status fot_cell_to_cell_csr(const uint32_t *const *cell_to_node,
                                const uint32_t *mesh_dim, uint32_t num_mappings,
                                uint32_t num_cells, const uint32_t *num_points,
                                uint32_t *indices, uint32_t *endpoints,
                                uint64_t *num_arcs);

/* HeuristicalPartition (partition.cu:128-275) given a coarse partition of
 * the cell-to-cell CSR.  result[num_cells] block ids; *num_blocks. */
fot_status fot_heuristic_partition(const uint32_t *csr_indices,
                                   const uint32_t *csr_endpoints,
                                   uint32_t num_vertices,
                                   const uint32_t *coarse_partition,
                                   uint32_t block_size, uint32_t *result,
                                   uint32_t *num_blocks);

/* k-way partition of a CSR graph into parts of at most max_part vertices
 * (stand-in for METIS_PartGraphKway, partition.cu:11-31, which is absent:
 * greedy graph growing from a BFS-ordered seed sequence).  No reference
 * counterpart for the algorithm; the interface follows partitionMetisEnh
 * (partition.hpp:23-29): part sizes <= floor(block_size / tolerance). */
fot_status fot_kway_partition(const uint32_t *csr_indices,
                              const uint32_t *csr_endpoints,
                              uint32_t num_vertices, uint32_t max_part,
                              uint32_t *part, uint32_t *num_parts);

/* Recursive bisection of a CSR graph into ceil(n / max_part) parts of equal
 * size (within one vertex): each piece is ordered breadth first from the far
 * end of a sweep from its first vertex and cut at the share of its parts, so
 * parts are compact and parts with nearby ids are neighbours.  A second
 * stand-in for METIS (partition.cu:11-31, METIS's own recursive-bisection
 * mode); the interface is fot_kway_partition's. */
fot_status fot_bisect_partition(const uint32_t *csr_indices, const uint32_t *csr_endpoints,
                                uint32_t num_vertices, uint32_t max_part, uint32_t *part,
                                uint32_t *num_parts);

/* scipy.sparse.csgraph.reverse_cuthill_mckee of the CSR matrix's pattern,
 * the ordering scripts/rcm.py:16-25 applies to an edge list: perm old->new
 * (rcm.py's np.argsort of scipy's result).  The pattern is symmetrised
 * (A + A^T); degrees count a diagonal entry twice; components are seeded in
 * seed_order (scipy: numpy's argsort of the degrees, whose order among equal
 * degrees depends on the host's numpy; NULL: ascending degree, ties to the
 * lower id).  Given the same seed order the permutation is scipy's. */
fot_status fot_rcm_order(const uint32_t *csr_indices, const uint32_t *csr_endpoints,
                         uint32_t num_vertices, const uint32_t *seed_order /* nullable */,
                         uint32_t *perm);
/* Bandwidth-reducing level-structure numbering from a pseudo-peripheral
 * vertex, permutation old->new: the stand-in for Scotch strategy "g"
 * (reorder.hpp:135-175, Scotch is absent here). */
fot_status fot_gps_order(const uint32_t *csr_indices, const uint32_t *csr_endpoints,
                         uint32_t num_vertices, uint32_t *perm);

/* ------------------------------------------------------------------ */
/* Synthetic mesh generators (no reference counterpart except where     */
/* noted).  All randomness comes from SplitMix64(seed).                 */
/* ------------------------------------------------------------------ */

/* Grid::fillEdgeList (grid.hpp:77-104): N x M points, right and down edges;
 * edges[2*((N-1)*M + N*(M-1))]. */
fot_status fot_gen_grid_edges(uint32_t N, uint32_t M, uint32_t *edges);
/* Grid::fillCellList3D (grid.hpp:415-460): N1 x N2 x N3 points, hexes with
 * 8 corners each; cells[8*(N1-1)*(N2-1)*(N3-1)]. */
fot_status fot_gen_hex_grid(uint32_t N1, uint32_t N2, uint32_t N3, uint32_t *cells);
/* Grid::fillCellList (grid.hpp:386-411): N x M points, quads with corners
 * (top_left, top_left + 1, bottom_left, bottom_left + 1); cells[4(N-1)(M-1)]. */
fot_status fot_gen_quad_grid(uint32_t N, uint32_t M, uint32_t *cells);

/* Airfoil-like quad mesh of nx x ny cells: (nx+1)(ny+1) nodes perturbed by
 * U(-0.2,0.2), interior edges only, pedge/pecell oriented so that (dy,-dx)
 * points from cell 1 to cell 2.  shuffle: bit 0 nodes, bit 1 cells, bit 2
 * edges.  Sizes: num_nodes=(nx+1)(ny+1), num_cells=nx*ny,
 * num_edges=nx*(ny-1)+(nx-1)*ny.  Any output may be NULL.  q and adt are
 * AOS (q: 4 per cell). */
fot_status fot_gen_airfoil(uint32_t nx, uint32_t ny, uint64_t seed,
                           uint32_t shuffle, uint32_t *pedge, uint32_t *pecell,
                           double *x, double *q, double *adt);

/* Random permutation (Fisher-Yates on SplitMix64). */
fot_status fot_gen_permutation(uint32_t n, uint64_t seed, uint32_t *perm);
/* n doubles uniform in [lo, hi) / normal (mean, sd) (Box-Muller). */
fot_status fot_gen_uniform(uint64_t n, uint64_t seed, double lo, double hi,
                           double *out);
fot_status fot_gen_normal(uint64_t n, uint64_t seed, double mean, double sd,
                          double *out);

/* Triangle mesh of the shallow-water edge loop (kernels/volna.hpp shapes):
 * an (nx+1) x (ny+1) node grid perturbed by U(-0.2,0.2), each quad split on a
 * random diagonal.  C = 2 nx ny cells, E = nx(ny+1) + (nx+1)ny + nx ny edges
 * (boundary edges included).  edge_cells[2E]: (left, right), (cell, cell) on
 * the boundary; is_boundary[E] 0.0 / 1.0; normals SOA [2][E], unit, left to
 * right (outward on the boundary); fluxes SOA [3][E] and bathy SOA [2][E]
 * N(0, 0.1); volume[C] triangle areas; state AOS [C][4]: h U[0.5,2),
 * hu, hv N(0, 0.1), zb U(-1,0].  shuffle: bit 1 cells, bit 2 edges.  Any
 * output may be NULL.  No reference counterpart (the reference reads Volna
 * meshes from files, volna.cu:10-36). */
/* mini_aero-style 3D hex mesh of n1 x n2 x n3 unit cells: interior faces
 * only, F = (n1-1)n2n3 + n1(n2-1)n3 + n1n2(n3-1), each mapping to its two
 * cells (lower index side first).  face_cells[2F]; face SOA [12][F]:
 * centre, normal (unit, lower -> upper), tangent, binormal; cells AOS
 * [C][28]: conservatives freestream (rho 1, u 50, v 10, w 5, T 300) times
 * (1 + 1e-2 N(0,1)) per component, centre, gradients N(0, 1e-3),
 * limiters U[0,1).  shuffle: bit 1 cells, bit 2 faces.  Any output may be
 * NULL.  No reference counterpart (mini_aero.cu reads its mesh). */
fot_status fot_gen_hex_faces(uint32_t n1, uint32_t n2, uint32_t n3, uint64_t seed,
                             uint32_t shuffle, uint32_t *face_cells, double *face,
                             double *cells);
fot_status fot_gen_tri_mesh(uint32_t nx, uint32_t ny, uint64_t seed, uint32_t shuffle,
                            uint32_t *edge_cells, double *is_boundary, double *normals,
                            double *fluxes, double *bathy, double *volume, double *state);

/* ------------------------------------------------------------------ */
/* Loop engine (MI355X-native fast path)                                 */
/* Replaces Problem<SOA>::loopGPUCellCentred / loopGPUHierarchical      */
/* (problem.hpp:89-339) and adds loopGPUEdgeCentred (atomics).           */
/* ------------------------------------------------------------------ */
typedef enum {
  FOT_STRATEGY_ATOMIC = 0,        /* one launch, FP64 global atomics      */
  FOT_STRATEGY_GLOBAL_COLOUR = 1, /* one launch per global colour        */
  FOT_STRATEGY_HIERARCHICAL = 2   /* LDS blocks, thread + block colours  */
} fot_strategy;

typedef struct {
  uint32_t num_cells;
  uint32_t num_mappings;
  const uint32_t *cell_to_node[FOT_MAX_MAPPINGS]; /* HOST, AOS */
  uint32_t mesh_dim[FOT_MAX_MAPPINGS];
  uint32_t num_points[FOT_MAX_MAPPINGS];
  const double *point_data[FOT_MAX_MAPPINGS];   /* DEVICE; [0] = iterate */
  uint32_t point_dim[FOT_MAX_MAPPINGS];
  const double *cell_data[FOT_MAX_CELL_DATA];   /* DEVICE, SOA */
  uint32_t cell_dim[FOT_MAX_CELL_DATA];
  uint32_t num_cell_data;
  double *point_data_out;                       /* DEVICE accumulation */
  int soa;
} fot_problem_desc;

typedef struct {
  uint32_t strategy;      /* fot_strategy */
  uint32_t block_size;    /* threads (atomic/global) or cells per block */
  const uint32_t *partition; /* HOST, nullable: block id per cell */
  uint32_t use_graph;     /* capture one iteration in a hipGraph */
  uint32_t flags;         /* FOT_LOOP_* */
} fot_loop_config;

/* hierarchical: stage mapping-0 input point data in LDS as well (res_calc:
 * q and adt; mini_aero: the cell values; lulesh: the node coordinates;
 * volna: the cell volumes; mine: the iterate).  Functors other than res_calc
 * gather per cell as without the flag when the staged values and the cache
 * do not fit in LDS together. */
#define FOT_LOOP_STAGE_INPUTS 1u
/* hierarchical: single launch, blocks ordered by per-block dependency
 * flags instead of one launch per block colour */
#define FOT_LOOP_DEPENDENCY_ORDER 2u
/* hierarchical: owner-compute plan (fot_oplan_build), one launch, no block
 * colours.  With FOT_LOOP_POINT_PARTITION, `partition` holds a part id per
 * mapping-0 point; ownership as fot_oplan_build defines it. */
#define FOT_LOOP_OWNER_COMPUTE 4u
#define FOT_LOOP_POINT_PARTITION 8u
/* owner compute: two cells per thread (more loads in flight per wave) */
#define FOT_LOOP_TWO_PER_THREAD 16u
/* owner compute, res_calc: gather mapping-1 (node) data straight from HBM per
 * cell instead of staging each block's nodes in LDS first */
#define FOT_LOOP_GATHER_NODES 32u
/* owner compute, res_calc: 1/rho and the pressure computed once per staged
 * cell (LDS) instead of once per edge side; the same operations, the same
 * bits */
#define FOT_LOOP_POINT_TERMS 64u
/* owner compute, res_calc: each owned cell gathers its edges' fluxes in
 * thread-colour order from LDS (three barriers, no per-colour barriers);
 * the same sums in the same order, the same bits.  Falls back to the
 * colour loop when a cell has more than 8 edges. */
#define FOT_LOOP_OWNER_GATHER 128u
/* owner compute, res_calc gather: persistent workgroups (as many as are
 * resident at once), each running a sequence of blocks and loading the next
 * block's plan entries while it computes the current one; the same bits.
 * Ignored by the other kernels. */
#define FOT_LOOP_PERSISTENT 256u
/* owner compute, res_calc gather: load each owned cell's edge list (8 bytes
 * per cell) from the plan instead of building it in LDS from the edges'
 * positions (the default when every cell has at most one edge per thread
 * colour); the same bits */
#define FOT_LOOP_LISTS_IN_HBM 512u
/* hierarchical (block colours, not owner compute): every block in ONE
 * launch, none waiting for another.  A point only one block touches gets
 * that block's sum added in place; a point several blocks touch gets each
 * block's sum stored in a partial slot, and a second launch adds a point's
 * partials in block-colour order.  The additions of the per-colour launches
 * in their order: the same bits.  No reference counterpart (the reference
 * launches once per block colour, problem.hpp:276-303). */
#define FOT_LOOP_TWO_PHASE 1024u
/* Nontemporal hints on the data a step reads or writes once: the owned
 * output values of the owner-compute res_calc gather kernel (one block per
 * workgroup, lists built in LDS, no per-cell terms), and the cell data of
 * the two-phase hex corner scatter, mini_aero face flux and lulesh element
 * forces (not staged).
 * The same values: the same bits.  By default this is on when the step's
 * data (point arrays, output, cell data and plan) exceed the 256 MB Infinity
 * Cache, where it keeps the read-once data from evicting the point data
 * other blocks touch again, and off below, where it slows the step (every
 * array is read again the next step); FOT_LOOP_NONTEMPORAL forces it on and
 * FOT_LOOP_CACHED off where it exists.  fot_loop_stats.nontemporal reports
 * the choice.  No reference counterpart. */
#define FOT_LOOP_NONTEMPORAL 2048u
#define FOT_LOOP_CACHED 4096u

typedef struct fot_loop fot_loop;
/* Builds the plan (host), uploads it, permutes cell data. */
fot_status fot_loop_create(const fot_functor *f, const fot_problem_desc *desc,
                           const fot_loop_config *cfg, fot_loop **loop);
fot_status fot_loop_destroy(fot_loop *loop);
/* One iteration's increments into point_data_out (no copy back). */
fot_status fot_loop_step(fot_loop *loop, void *stream);
/* `iters` iterations with the reference's semantics: out starts as a copy of
 * point_data[0]; after each iteration point_data[0] = out. */
fot_status fot_loop_run(fot_loop *loop, uint32_t iters, void *stream);
/* Synchronises and reports FOT_ERR_TIMEOUT if a dependency-ordered launch
 * ran out of its bounded wait (no reference counterpart). */
fot_status fot_loop_check(const fot_loop *loop);

typedef struct {
  uint32_t strategy, num_launches_per_iter, num_colours, num_blocks;
  uint32_t block_size, max_cached, lds_bytes, max_thread_colours;
  uint64_t plan_device_bytes;
  double avg_thread_colours, reuse_factor;
  double redundancy; /* owner compute: cell runs per cell (>= 1) */
  /* hierarchical: 128-byte lines of the target data per block, the
   * reference's "average cache_line / block" (problem.hpp:321-334) */
  double avg_cache_lines;
  /* hierarchical launches: workgroup size, and (owner res_calc) the staging
   * rounds each thread runs over the block's staged points */
  uint32_t threads_per_block;
  uint32_t stage_per_thread;
  /* single-launch strategies: workgroups per launch */
  uint32_t launch_grid;
  /* 1 when the launch uses nontemporal hints (FOT_LOOP_NONTEMPORAL) */
  uint32_t nontemporal;
} fot_loop_stats;
fot_status fot_loop_stats_get(const fot_loop *loop, fot_loop_stats *stats);

/* ------------------------------------------------------------------ */
/* Batched loops (no reference counterpart): many problems of one        */
/* functor, each planned as an owner-compute hierarchical loop with its  */
/* own config, run together by ONE persistent launch per step whose      */
/* workgroups take (problem, block) items from a device work queue.      */
/* ------------------------------------------------------------------ */
typedef struct fot_batch fot_batch;
/* descs/cfgs: one per problem (same functor and layout).  Each cfg's
 * block_size / partition (+ FOT_LOOP_POINT_PARTITION) shape that problem's
 * plan; the strategy is always owner-compute hierarchical.  use_graph != 0
 * captures one timestep in a hipGraph that fot_batch_run replays: the
 * batched increments, then point_data[0] = out for every problem when the
 * functor reads its own iterate (mine); otherwise that copy runs once,
 * after the last step, as in the eager run.  A batch with no work items
 * (every mesh empty) and no per-step copy captures no graph. */
fot_status fot_batch_create(const fot_functor *f, uint32_t num_problems,
                            const fot_problem_desc *descs, const fot_loop_config *cfgs,
                            uint32_t use_graph, fot_batch **batch);
fot_status fot_batch_destroy(fot_batch *batch);
/* One step: every problem's increments into its point_data_out. */
fot_status fot_batch_step(fot_batch *batch, void *stream);
/* `iters` iterations with fot_loop_run's semantics for every problem. */
fot_status fot_batch_run(fot_batch *batch, uint32_t iters, void *stream);
typedef struct {
  uint32_t num_problems, num_work_items, grid, threads, lds_bytes, uses_graph;
} fot_batch_stats;
fot_status fot_batch_stats_get(const fot_batch *batch, fot_batch_stats *stats);

/* ------------------------------------------------------------------ */
/* Context layer: a device and one stream, the buffers allocated through */
/* it, plans, one-shot loops, graphs and batches run on that stream.     */
/* The reference runs on one host thread and the default stream          */
/* (problem.hpp:177, 302), with Problem owning its device mirrors         */
/* (data_t.cu:3-14, 50-56).  A context is not thread-safe; use one per    */
/* host thread.  Every call below runs on the context's stream and       */
/* returns once its work is done, unless it says otherwise.               */
/* ------------------------------------------------------------------ */
typedef struct fot_ctx fot_ctx;
/* Makes `device` current and creates the context's stream (default
 * priority). */
fot_status fot_ctx_create(fot_ctx **ctx, int device);
/* Waits for the stream, frees every buffer the context still owns, destroys
 * the stream.  Plans, graphs and batches are freed by their own calls. */
fot_status fot_ctx_destroy(fot_ctx *ctx);
/* The context's hipStream_t, for the entry points that take a stream. */
fot_status fot_ctx_stream(const fot_ctx *ctx, void **stream);
fot_status fot_ctx_synchronize(fot_ctx *ctx);

/* device_data_t (data_t.hpp:10-74): a device buffer owned by the context;
 * upload / download are initDeviceMemory / flushToHost (data_t.cu:3-56),
 * complete on return. */
fot_status fot_buf_alloc(fot_ctx *ctx, size_t bytes, void **device_ptr);
fot_status fot_buf_upload(fot_ctx *ctx, void *device_dst, const void *host_src, size_t bytes);
fot_status fot_buf_download(fot_ctx *ctx, void *host_dst, const void *device_src,
                            size_t bytes);
/* FOT_ERR_INVALID_ARGUMENT for a pointer this context did not allocate. */
fot_status fot_buf_free(fot_ctx *ctx, void *device_ptr);

/* A plan on the device for one problem and config: the colourings and the
 * block plan the reference builds and uploads at the start of every loop
 * call (problem.hpp:106-140, 203, 258-265), built once and kept.  The same
 * object as fot_loop_create's; fot_loop_step / fot_loop_run take it. */
typedef struct fot_loop fot_plan;
fot_status fot_plan_upload(fot_ctx *ctx, const fot_functor *f, const fot_problem_desc *desc,
                           const fot_loop_config *cfg, fot_plan **plan);
fot_status fot_plan_free(fot_plan *plan);

/* Problem::loopGPUCellCentred / loopGPUHierarchical (problem.hpp:90-339) and
 * loopGPUEdgeCentred (atomics), as the reference calls them: plans, runs
 * `iters` iterations with fot_loop_run's semantics on the context's stream,
 * waits, and frees the plan. */
fot_status fot_loop_launch(fot_ctx *ctx, const fot_functor *f, const fot_problem_desc *desc,
                           const fot_loop_config *cfg, uint32_t iters);

/* One iteration's launches of a plan (fot_loop_step: the increments into
 * point_data_out, no copy back) captured in a hipGraph (no reference
 * counterpart).  A plan over no cells captures nothing and its graph's launch
 * does nothing. */
typedef struct fot_graph fot_graph;
fot_status fot_graph_capture(fot_ctx *ctx, fot_plan *plan, fot_graph **graph);
/* Replays the graph on the context's stream; returns without waiting. */
fot_status fot_graph_launch(fot_ctx *ctx, fot_graph *graph);
fot_status fot_graph_free(fot_graph *graph);

/* fot_batch_run on the context's stream (no reference counterpart). */
fot_status fot_batch_launch(fot_ctx *ctx, fot_batch *batch, uint32_t iters);

/* ------------------------------------------------------------------ */
/* Autotuner support (no reference counterpart)                          */
/* ------------------------------------------------------------------ */
#define FOT_NUM_FEATURES 24
/* Cheap features of a loop (no plan is built), out[FOT_NUM_FEATURES]:
 *  0 log2 cells, 1 log2 points of mapping 0, 2 mean point degree,
 *  3 mesh_dim, 4 point_dim, 5 cell-data doubles per cell, 6 point-data
 *  doubles gathered per cell through mappings >= 1, 7 degree std / mean,
 *  8 max degree / mean, 9-12 fractions of points with degree <= 2, 3-4,
 *  5-8, > 8, 13 global colours (65 if 64 do not suffice), 14 mean cell
 *  span / points, 15-16 fractions of cells with span < 256 / < 4096,
 *  17 mean jump between consecutive cells' lowest ids / points, 18-19
 *  fractions of jumps < 256 / < 4096, 20 log2 algorithmic bytes per
 *  iteration, 21 res_calc, 22 reads its iterate (mine), 23 log2 target
 *  bytes.  Host arrays only (point_data may be NULL). */
fot_status fot_loop_features(const fot_functor *f, const fot_problem_desc *desc, float *out);
/* Plans each candidate config and times `steps` steps after `warmup` on a
 * private stream with HIP events: times_ms[i] = ms per step, or -1 when the
 * config cannot be planned for this mesh (colour overflow, block too big). */
fot_status fot_autotune_measure(const fot_functor *f, const fot_problem_desc *desc,
                                const fot_loop_config *candidates, uint32_t n,
                                uint32_t warmup, uint32_t steps, float *times_ms);

/* ------------------------------------------------------------------ */
/* Learned cost model (no reference counterpart): an FP32 MLP            */
/* in -> hidden -> hidden -> out with ReLU, trained and evaluated on the */
/* GPU by hand-written kernels (matrix products on MFMA).                */
/* Parameters, flat: W1[hidden][in], b1, W2[hidden][hidden], b2,          */
/* W3[out][hidden], b3.  Host arrays are row-major float.                 */
/* ------------------------------------------------------------------ */
typedef struct fot_mlp fot_mlp;
fot_status fot_mlp_create(uint32_t in, uint32_t hidden, uint32_t out, uint64_t seed,
                          fot_mlp **mlp);
fot_status fot_mlp_destroy(fot_mlp *mlp);
fot_status fot_mlp_num_params(const fot_mlp *mlp, uint64_t *n);
fot_status fot_mlp_get_params(const fot_mlp *mlp, float *params, uint64_t n);
/* also resets the optimiser state */
fot_status fot_mlp_set_params(fot_mlp *mlp, const float *params, uint64_t n);
/* y[rows][out] = MLP(x[rows][in]) */
fot_status fot_mlp_forward(fot_mlp *mlp, const float *x, uint32_t rows, float *y);
/* fot_mlp_forward on DEVICE buffers x[rows][in] and y[rows][out], enqueued on
 * `stream` (returns without waiting), allocating nothing: no hipMalloc, no
 * hipFree, no allocating library call.  No reference counterpart; it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid it).  The two hidden layers'
 * activations go to a fixed 4 MiB device array (one per device, held until
 * the process ends), so 2 * rows * hidden <= 1048576 (rows <= 8192 at
 * hidden 64); larger calls return FOT_ERR_INVALID_ARGUMENT.  The current
 * device must be the one the MLP was created on.  Every call on a device
 * shares that device's array: calls that overlap on one device (different
 * streams, or host threads) race; calls on one stream are ordered. */
fot_status fot_mlp_forward_noalloc(const fot_mlp *mlp, const float *x, uint32_t rows, float *y,
                                   void *stream);
/* loss = sum(mask (y - t)^2) / sum(mask) (mask NULL: all ones) and its
 * gradient with respect to the flat parameters */
fot_status fot_mlp_loss_grad(fot_mlp *mlp, const float *x, const float *t, const float *mask,
                             uint32_t rows, float *loss, float *grad);
/* `epochs` full-batch Adam steps (beta 0.9 / 0.999, eps 1e-8, L2 weight
 * decay); loss_per_epoch (nullable) gets the loss before each step */
fot_status fot_mlp_train(fot_mlp *mlp, const float *x, const float *t, const float *mask,
                         uint32_t rows, uint32_t epochs, float lr, float weight_decay,
                         float *loss_per_epoch);
/* The MFMA GEMM on DEVICE buffers: C[M][N] = op(A) op(B), A row-major M x K
 * (K x M when trans_a), B row-major K x N (N x K when trans_b). */
fot_status fot_gemm_f32(const float *A, const float *B, float *C, uint32_t M, uint32_t N,
                        uint32_t K, int trans_a, int trans_b, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FOT_CAPI_H */
