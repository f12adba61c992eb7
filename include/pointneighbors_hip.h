/*
 * pointneighbors_hip.h — C-ABI of the MI355X (gfx950) neighbourhood-search engine.
 *
 * This is the drop-in boundary under PointNeighbors.jl's GridNeighborhoodSearch +
 * FullGridCellList + PrecomputedNeighborhoodSearch path.  Every entry point is plain C:
 * device pointers, sizes and status codes; no language runtime types cross it.
 * A Julia binding calls these through `ccall`; the Python host mirror in
 * pointneighbors.jl_amd/pointneighbors calls them through ctypes (same convention).
 *
 * Conventions
 *  - Coordinates are column-major D x N matrices (the reference's layout,
 *    src/neighborhood_search.jl:24-26): point i occupies elements [D*i, D*i + D).
 *  - Point indices crossing the ABI are 0-based unless an `index_base` argument says
 *    otherwise (a Julia binding passes index_base = 1).
 *  - All device work is enqueued on the grid's stream (pn_grid_set_stream).  Functions
 *    documented as "synchronizes" wait for that stream before returning, like the
 *    reference's KernelAbstractions.synchronize after every @threaded loop
 *    (src/util.jl:170).
 *  - Status codes: 0 = PN_OK; on failure the thread-local message is pn_last_error().
 *  - A grid is not reentrant: calls on one grid are serialized by the caller.
 */
#ifndef POINTNEIGHBORS_HIP_H
#define POINTNEIGHBORS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PN_ABI_VERSION 1

enum pn_status {
    PN_OK = 0,
    PN_ERR_ARGUMENT = 1,       /* ArgumentError in the reference constructors */
    PN_ERR_OUT_OF_BOUNDS = 2,  /* "particle coordinates are NaN or outside the domain
                                  bounds of the cell list" (src/cell_lists/full_grid.jl:187-213) */
    PN_ERR_CELL_FULL = 3,      /* "cell list is full. Use a larger `max_points_per_cell`."
                                  (src/vector_of_vectors.jl:114-121) */
    PN_ERR_QUERY_BOUNDS = 4,   /* BoundsError: a query point's neighbour cells leave the grid
                                  (LinearIndices getindex, src/cell_lists/full_grid.jl:157-161) */
    PN_ERR_NOT_INITIALIZED = 5,
    PN_ERR_HIP = 6,            /* HIP runtime failure */
    PN_ERR_LIST_FULL = 7,      /* a precomputed neighbour list exceeds max_neighbors
                                  (pushat! overflow, src/vector_of_vectors.jl:79) */
    PN_ERR_UNSUPPORTED = 8
};

enum pn_dtype { PN_FLOAT32 = 0, PN_FLOAT64 = 1 };

/* Which points a traversal iterates over (the `x` / system_coords side). */
enum pn_queries {
    PN_QUERIES_GRID = 0,   /* x is y: the grid's own binned points */
    PN_QUERIES_BINNED = 1  /* the point set last passed to pn_grid_bin_queries */
};

typedef struct pn_grid* pn_grid_t;

/* Library identity.  No reference counterpart. */
int pn_abi_version(void);
const char* pn_last_error(void);

/* ------------------------------------------------------------------------------------
 * Grid lifetime.  Replaces FullGridCellList(; min_corner, max_corner, search_radius,
 * max_points_per_cell) (src/cell_lists/full_grid.jl:48-82) together with
 * GridNeighborhoodSearch{D}(; search_radius, cell_list) (src/nhs_grid.jl:77-129) and
 * Adapt.adapt(backend, nhs) (src/gpu.jl:10-12).
 *
 * The caller computes the padded min corner and n_cells exactly as the reference
 * constructor does and passes them here:
 *   coord_dtype   element type of the coordinate matrices x and y
 *   radius_dtype  type of search_radius (determines the distance type, src/nhs_grid.jl:547)
 *   cell_dtype    promote(coord type, padded-corner type): the type of
 *                 (x .- min_corner) ./ cell_size (src/cell_lists/full_grid.jl:93)
 *   min_corner    padded min corner, values exact in cell_dtype
 *   n_cells       cells per dimension (including the padding layer)
 * ---------------------------------------------------------------------------------- */
int pn_grid_create(int ndims, int coord_dtype, int radius_dtype, int cell_dtype,
                   const double* min_corner, const int64_t* n_cells, double search_radius,
                   int32_t max_points_per_cell, pn_grid_t* out);
int pn_grid_destroy(pn_grid_t grid);

/* GridNeighborhoodSearch with a SpatialHashingCellList (src/cell_lists/spatial_hashing.jl:1-183;
 * its traversal with the collision checks, src/nhs_grid.jl:479-575): an unbounded domain.  A
 * point's cell is floor_to_int.(x ./ search_radius); its bucket is the reference's spatial_hash
 * (Ihmsen et al. 2011) mod list_size.  pn_grid_update counting-sorts the points by bucket
 * (max_points_per_cell bounds a bucket, as the reference's DynamicVectorOfVectors does:
 * PN_ERR_CELL_FULL), and a cell coordinate outside Int32 (NaN, Inf included) is
 * PN_ERR_OUT_OF_BOUNDS, the reference's InexactError.  Every traversal entry point works on the
 * handle; a traversal scans the 3^D neighbouring cells' buckets and keeps a candidate only if
 * its own cell is the scanned cell, so neighbour sets equal those of a FullGridCellList.  The
 * cell type is the promotion of the coordinate and radius types.  pn_grid_set_periodic applies:
 * cells are then wrapped (periodic_cell_index) and distances take the minimum image.  No
 * batching. */
/* (Both grid constructors accept max_points_per_cell up to INT32_MAX / 27 = 79536431: a 3^D
 * stencil's candidates are counted in int32.) */
int pn_grid_create_spatial_hash(int ndims, int coord_dtype, int radius_dtype, double search_radius,
                                int64_t list_size, int32_t max_points_per_cell, pn_grid_t* out);

/* GridNeighborhoodSearch with a PeriodicBox (src/nhs_grid.jl:102-125, PeriodicBox
 * src/neighborhood_search.jl:129-141).  The caller sizes the periodic grid as the
 * reference constructor does:
 *   n_periodic  floor.(Int, (box size .+ 10eps()) / search_radius), each >= 3
 *   cell_size   box size ./ n_periodic, values exact in the radius type
 *   box_size    max_corner - min_corner, values exact in the radius type
 * Cell coordinates are then wrapped to mod.(cell .- 2, n_periodic) .+ 2
 * (periodic_cell_index, src/nhs_grid.jl:609-620), neighbour cells wrap, and a candidate
 * with d2 > r^2 is tested again at its minimum image (compute_periodic_distance,
 * src/neighborhood_search.jl:428-436).  The cell list needs n_periodic + 2 cells per
 * dimension.  Invalidates earlier builds; not available for batched clouds. */
int pn_grid_set_periodic(pn_grid_t grid, const double* cell_size, const int64_t* n_periodic,
                         const double* box_size);
/* stream: a hipStream_t (NULL = default stream). */
int pn_grid_set_stream(pn_grid_t grid, void* stream);

/* ------------------------------------------------------------------------------------
 * initialize!/update! with the default ParallelUpdate strategy
 * (src/nhs_grid.jl:220-225, :255-292, :470-477): bins the points of y selected by
 * eachindex_y into cells with a counting sort.
 *   y         device pointer, D x n_y of coord_dtype
 *   sel_list  NULL: the selection is the range [sel_first, sel_first + sel_count);
 *             else a device int32 array of sel_count 0-based indices (eachindex_y)
 * pn_grid_update waits for the binning's status and reports the reference's errors:
 *   PN_ERR_OUT_OF_BOUNDS (first offending point via pn_grid_error), PN_ERR_CELL_FULL.
 *   The status reaches pinned host memory from the scan, before the last build kernels end:
 *   the call returns while they still run on the grid's stream (later work on that stream
 *   is ordered after them; other streams or the host synchronize with it first).
 * pn_grid_update_async only enqueues; pn_grid_check later waits and reports.
 * ---------------------------------------------------------------------------------- */
int pn_grid_update(pn_grid_t grid, const void* y, int64_t n_y, int64_t sel_first,
                   int64_t sel_count, const int32_t* sel_list);
int pn_grid_update_async(pn_grid_t grid, const void* y, int64_t n_y, int64_t sel_first,
                         int64_t sel_count, const int32_t* sel_list);
int pn_grid_check(pn_grid_t grid);
/* After PN_ERR_OUT_OF_BOUNDS / PN_ERR_QUERY_BOUNDS: the offending 0-based point index and
 * its 1-based cell coordinates (ndims entries, Int64 wrap as in src/util.jl:19-34). */
int pn_grid_error(pn_grid_t grid, int64_t* point, int64_t* cell);

/* Re-gathers the grid's cell-sorted copy of y from y (cells unchanged).  Needed when y
 * changed in place without update!: the reference reads neighbour coordinates at
 * traversal time (src/nhs_grid.jl:543-545) while cell membership stays as built. */
int pn_grid_refresh(pn_grid_t grid, const void* y, int64_t n_y);

/* Bins a query point set x (the `system_coords` of foreach_point_neighbor,
 * src/neighborhood_search.jl:183-201) with the grid's geometry, restricted to `points`
 * (same selection convention as pn_grid_update).  Synchronizes; PN_ERR_QUERY_BOUNDS when
 * a query cell is outside 2:(n_d - 1). */
int pn_grid_bin_queries(pn_grid_t grid, const void* x, int64_t n_x, int64_t sel_first,
                        int64_t sel_count, const int32_t* sel_list);

/* Counters of the last build (host values; synchronizes): number of binned points,
 * number of non-empty cells, largest cell population. */
int pn_grid_stats(pn_grid_t grid, int64_t* n_binned, int64_t* n_nonempty, int64_t* max_count);
/* Copies the cell table of the grid (device -> device, enqueued): cell_start has
 * prod(n_cells) + 1 entries (0-based offsets into the sorted order), sorted_index the
 * original 0-based point index of every binned point in cell order (cells ascending,
 * original index ascending within a cell).  Either pointer may be NULL. */
int pn_grid_copy_cells(pn_grid_t grid, int32_t* cell_start, int32_t* sorted_index);

/* ------------------------------------------------------------------------------------
 * foreach_point_neighbor with a fixed functor menu (src/neighborhood_search.jl:183-201,
 * hot loop src/nhs_grid.jl:519-575).  A pair (i, j) is visited iff j is a binned point
 * in one of the 3^D cells around i's cell and d2 <= r*r with
 * d2 = ((p1*p1 + p2*p2) + p3*p3), p = convert(T_r, x_i - y_j).
 * Outputs are indexed by the original query index.  accumulate != 0 adds to the
 * output (the `out[i] += ...` closure form); 0 overwrites.  Points outside the query
 * selection are not touched.  All enqueue only.
 * ---------------------------------------------------------------------------------- */
/* Closure `n_neighbors[i] += 1` (benchmarks/count_neighbors.jl:24-27). */
int pn_count_neighbors(pn_grid_t grid, int queries, int32_t* counts, int accumulate);
/* mapreduce_neighbor(+, ...) of (j + index_base) (test/neighborhood_search.jl:389-399). */
int pn_sum_neighbor_indices(pn_grid_t grid, int queries, int64_t* sums, int index_base,
                            int accumulate);
/* mapreduce_neighbor(f, op, ...; init) of every query point (src/neighborhood_search.jl:340-383)
 * in double: value[i] = op over the neighbours of f(i, j, pos_diff, d), count[i] = number of
 * neighbours (an empty neighbourhood leaves value[i] at op's identity: the caller returns
 * init).  fn: 0 one, 1 neighbour index j + index_base, 2 distance, 3 squared distance,
 * 4 + k component k of pos_diff; op: 0 +, 1 min, 2 max.  Overwrites value and count. */
int pn_mapreduce_neighbors(pn_grid_t grid, int queries, int fn, int op, int index_base, double* value,
                           int32_t* count);
/* The n-body closure of benchmarks/n_body.jl:34-51: for every neighbour j at distance
 * d >= sqrt(eps), dv[:, i] += -G mass[j] pos_diff / d^3.  dv: D x N (coordinates' layout);
 * mass per grid point.  Coordinates and radius of one type, 2D or 3D. */
int pn_nbody_accel(pn_grid_t grid, int queries, double G, const void* mass, void* dv, int accumulate);
/* TrixiParticles' fluid-fluid interact! as timed by benchmark_wcsph
 * (benchmarks/smoothed_particle_hydrodynamics.jl:44-102): WCSPH with ContinuityDensity, the
 * pressure `pressure` precomputed, ArtificialViscosityMonaghan(alpha, beta, epsilon) and
 * DensityDiffusionMolteniColagrossi(delta), smoothing length h, sound speed c.  v and dv are
 * (D+1) x N matrices: velocity, then density.  dv accumulates when accumulate != 0 (interact!
 * does).  Queries are the grid's own points (PN_QUERIES_GRID). */
int pn_wcsph_interact(pn_grid_t grid, int kernel, double h, double sound_speed, double alpha,
                      double beta, double epsilon, double delta, const void* mass, const void* v,
                      const void* pressure, void* dv, int accumulate);

/* ------------------------------------------------------------------------------------
 * PrecomputedNeighborhoodSearch lists as CSR (src/nhs_precomputed.jl:187-206):
 *   pn_csr_count: row_ptr[n_rows + 1] (int32, device) from neighbour counts of the query
 *     rows; returns nnz and the longest row (synchronizes).  PN_ERR_LIST_FULL when a
 *     row exceeds max_neighbors (> 0), the reference's pushat! overflow.
 *   pn_csr_fill: col_idx[nnz] with every row sorted ascending (sorteach!,
 *     src/vector_of_vectors.jl:177-212), values j + index_base.
 * n_rows is the number of columns of x (rows of points outside the selection are empty).
 * ---------------------------------------------------------------------------------- */
int pn_csr_count(pn_grid_t grid, int queries, int64_t n_rows, int32_t* row_ptr,
                 int32_t max_neighbors, int64_t* nnz, int64_t* max_row);
int pn_csr_fill(pn_grid_t grid, int queries, int64_t n_rows, const int32_t* row_ptr,
                int32_t* col_idx, int index_base);

/* ------------------------------------------------------------------------------------
 * foreach_point_neighbor over precomputed lists (src/nhs_precomputed.jl:210-247): every
 * stored neighbour is visited, without a radius test.  These take the CSR arrays and a
 * stream directly, so they work after freeze_neighborhood_search dropped the grid
 * (src/nhs_precomputed.jl:268-277).  Rows are selected like pn_grid_update's points.
 * stored_base is the index base the lists were filled with.  Enqueue only.
 * ---------------------------------------------------------------------------------- */
int pn_list_count_neighbors(const int32_t* row_ptr, int64_t n_rows, int64_t sel_first,
                            int64_t sel_count, const int32_t* sel_list, int32_t* counts,
                            int accumulate, void* stream);
int pn_list_sum_neighbor_indices(const int32_t* row_ptr, const int32_t* col_idx, int64_t n_rows,
                                 int stored_base, int64_t sel_first, int64_t sel_count,
                                 const int32_t* sel_list, int64_t* sums, int index_base,
                                 int accumulate, void* stream);

/* ------------------------------------------------------------------------------------
 * Interaction functors and their adjoints.  No reference counterpart in PointNeighbors.jl:
 * these are the closures a WCSPH step passes to foreach_point_neighbor (the workload of
 * benchmarks/smoothed_particle_hydrodynamics.jl:45-102, physics from TrixiParticles) and a
 * Wendland C2 pair energy, plus hand-written reverse-mode kernels that recompute the
 * traversal instead of storing pair lists.  Fields are arrays of the radius type indexed
 * by original point index; vector fields (dv, x_bar, dv_bar) are D x N like coordinates.
 * kernel: PN_KERNEL_CUBIC_SPLINE (Monaghan) or PN_KERNEL_WENDLAND_C2, support 2h.
 * Functions without a `queries` argument need x == y (the grid's own points).  Enqueue only.
 *   pn_sph_density            rho_i = sum_j m_j W(r_ij)  (self included)
 *   pn_sph_pressure_accel     dv_i = -sum_{j != i} m_j (P_i + P_j) grad W_ij,  P = p / rho^2
 *   pn_pair_energy            e_i = 1/2 sum_{j != i} eps (1 - r/rc)^4 (1 + 4 r/rc), r < rc
 *   *_vjp                     vector-Jacobian products for upstream gradients of the outputs
 *   pn_tait_eos               p = B ((rho/rho0)^gamma - 1), B = rho0 c^2 / gamma
 *   pn_tait_pressure_term     P = p / rho^2 with the same p: the per-point term of the pressure
 *                             acceleration, in one pass
 * ---------------------------------------------------------------------------------- */
enum pn_kernel { PN_KERNEL_CUBIC_SPLINE = 0, PN_KERNEL_WENDLAND_C2 = 1 };

int pn_sph_density(pn_grid_t grid, int queries, int kernel, double h, const void* mass,
                   void* rho, int accumulate);
int pn_sph_density_vjp(pn_grid_t grid, int kernel, double h, const void* mass,
                       const void* rho_bar, void* x_bar, void* mass_bar, int accumulate);
int pn_sph_pressure_accel(pn_grid_t grid, int kernel, double h, const void* mass, const void* P,
                          void* dv, int accumulate);
int pn_sph_pressure_accel_vjp(pn_grid_t grid, int kernel, double h, const void* mass,
                              const void* P, const void* dv_bar, void* x_bar, void* P_bar,
                              void* mass_bar, int accumulate);
int pn_pair_energy(pn_grid_t grid, int queries, double eps, double cutoff, void* e,
                   int accumulate);
int pn_pair_energy_vjp(pn_grid_t grid, double eps, double cutoff, const void* e_bar,
                       void* x_bar, int accumulate);
int pn_tait_eos(const void* rho, void* p, int64_t n, int dtype, double rho0, double c,
                double gamma, void* stream);
int pn_tait_pressure_term(const void* rho, void* P, int64_t n, int dtype, double rho0, double c,
                          double gamma, void* stream);

/* ------------------------------------------------------------------------------------
 * Batches of independent clouds.  No reference counterpart: the nearest concept is one
 * GridNeighborhoodSearch per point set made with copy_neighborhood_search
 * (src/neighborhood_search.jl:107-110, src/nhs_grid.jl:640-649), launched per cloud.
 * pn_batch_create returns a grid whose cell table concatenates the clouds' grids (same
 * radius, own padded min corner and n_cells each, computed as FullGridCellList does);
 * points are the concatenation of the clouds (cloud c owns [point_offsets[c],
 * point_offsets[c+1])).  One pn_grid_update(_async) bins every cloud, one traversal call
 * serves every cloud (PN_QUERIES_GRID only).  Launch shapes do not depend on the data, so
 * an update+query step can be captured in a hipGraph (use pn_grid_update_async inside).
 *   min_corners, n_cells: n_clouds x ndims, row-major per cloud
 * ---------------------------------------------------------------------------------- */
int pn_batch_create(int ndims, int coord_dtype, int radius_dtype, int cell_dtype, int64_t n_clouds,
                    const double* min_corners, const int64_t* n_cells, double search_radius,
                    int32_t max_points_per_cell, const int64_t* point_offsets, pn_grid_t* out);

/* Tuning knobs (the autotuner's choices).  No reference counterpart.  None changes a
 * neighbour set; "algorithm" and "chunk" change the order of floating-point sums.
 *   "algorithm"        0: candidate-major cells (default), 1: per-query candidate slices
 *   "persist_blocks"   workgroups of the persistent traversal grid (multiple of 8)
 *   "chunk"            candidates per chunk: -1 per functor, 0 small, 1 large
 *   "canonical_order"  1: in-cell order by original index; 0: atomic order
 *   "pair_cache"       1 (grids without a periodic box): the first SPH density / pressure /
 *                      pair-energy traversal of
 *                      the grid's own points after a build records each cell's accepted pairs
 *                      (about 2 bytes per pair in device memory); later ones on the same build
 *                      replay them instead of testing candidates, with identical results.
 *                      0 (default): always test.  The adjoints always recompute the traversal.
 *   "pair_cache_per_point"  cache capacity in 2-byte entries per point (0: 108 in 3D, 36 in
 *                      2D, 12 in 1D); cells that do not fit are traversed in full
 *   "time_launches"    1: record a HIP event pair on the grid's stream around every traversal
 *                      kernel from now on (at most 4096 until read), 0: stop; either clears
 *                      the records */
int pn_set_tuning(pn_grid_t grid, const char* key, int64_t value);

/* Measurement hook, no reference counterpart: waits for the last recorded launch, writes the
 * durations in ms of the traversal kernels recorded since "time_launches" was set (at most
 * max, in launch order), their number to *n, and clears the records. */
int pn_grid_launch_times(pn_grid_t grid, float* ms, int64_t max, int64_t* n);

/* Adds U[-amplitude, amplitude] to each of the n values of x (a pure function of seed,
 * *step and the index), then increments *step on the device: a replayable per-step
 * perturbation like benchmarks/update.jl:28-49.  Enqueue only. */
int pn_displace_uniform(void* x, int64_t n, int dtype, double amplitude, uint64_t seed,
                        int64_t* step, void* stream);

/* SPH density over precomputed lists (a TLSPH-style sum; mapreduce_neighbor_inner of
 * PrecomputedNeighborhoodSearch, src/nhs_precomputed.jl:210-247: every stored neighbour,
 * distance recomputed, no radius test).  x, y: D x N (original order); mass per y point;
 * rho per x point; rows selected like pn_grid_update's points.  Enqueue only. */
int pn_list_sph_density(const int32_t* row_ptr, const int32_t* col_idx, int64_t n_rows,
                        int stored_base, int ndims, int coord_dtype, int radius_dtype, int kernel,
                        double h, const void* x, const void* y, const void* mass, int64_t sel_first,
                        int64_t sel_count, const int32_t* sel_list, void* rho, int accumulate,
                        void* stream);

/* CSR lists (pn_csr_fill) -> the reference's DynamicVectorOfVectors{Int32} backend
 * (src/vector_of_vectors.jl:3-31; PrecomputedNeighborhoodSearch's `backend` and
 * `transpose_backend`, src/nhs_precomputed.jl:40-66).  lengths[i] = length of row i.  Entry k of
 * row i (k < max_inner) goes to backend[k + max_inner * i] (Array{Int32,2}(max_inner, N)), or
 * with transpose != 0 to backend[i + ld * k] (the row-major array behind the reference's
 * PermutedDimsArray; ld = max_outer_length >= n_rows).  Slots past a row's length are not
 * written.  The caller checks max_row <= max_inner (pn_csr_count reports max_row): longer rows
 * are the reference's list overflow.  Enqueue only. */
int pn_csr_to_dvov(const int32_t* row_ptr, const int32_t* col_idx, int64_t n_rows, int max_inner,
                   int transpose, int64_t ld, int32_t* backend, int32_t* lengths, void* stream);

/* TLSPH deformation gradient over the stored lists of a PrecomputedNeighborhoodSearch built on
 * the initial coordinates X (TrixiParticles 0.5 calc_deformation_grad! and the kernel-gradient
 * correction matrix, run by benchmarks/smoothed_particle_hydrodynamics.jl:127-146 through
 * foreach_point_neighbor):
 *     G_a = sum_{b in list(a), |X_a - X_b| >= sqrt(eps)} V_b (x_b - x_a) (grad_a W(X_a - X_b))^T
 *     out_a = G_a L_a  (L_a given), G_a^-1  (invert != 0, L == NULL), or G_a
 * With x = X, G is the symmetric correction matrix M, and invert gives L = M^-1 (closed-form
 * adjugate over determinant); with that L an affine deformation x = A X + c gives out_a = A
 * exactly.  X, x: D x N; volume (m / rho0) per point;
 * L, out: D x D per point, column-major (Julia's Array{T,3}); one floating type for all
 * arrays (dtype).  Rows selected like pn_grid_update's points.  Enqueue only. */
int pn_list_tlsph_deformation_grad(const int32_t* row_ptr, const int32_t* col_idx, int64_t n_rows,
                                   int stored_base, int ndims, int dtype, int kernel, double h,
                                   const void* X, const void* x, const void* volume, const void* L,
                                   int invert, int64_t sel_first, int64_t sel_count,
                                   const int32_t* sel_list, void* out, void* stream);

/* TLSPH interaction forces over the same stored lists (TrixiParticles 0.5
 * interact_structure_structure!, the loop the reference times in benchmark_tlsph,
 * benchmarks/smoothed_particle_hydrodynamics.jl:115-121 with the system of :146-187):
 *     dv_a (+)= sum_{b in list(a), |X_ab| >= sqrt(eps)} m_b (Pr_a + Pr_b) grad_a W(X_ab) + f_ab / m_a
 *     f_ab = (alpha / 2) V_a V_b W(|X_ab|) / |X_ab|^2 E delta_ab x_ab / |x_ab|,
 *     delta_ab = ((F_a + F_b) X_ab - 2 x_ab) . x_ab / |x_ab|,  V = m / rho
 * (the penalty force of Ganzenmueller 2015, PenaltyForceGanzenmueller; skipped when
 * penalty_alpha == 0, then x, rho and F may be NULL).  X (initial), x (current): D x N;
 * mass and rho (material density) per point; pk1_rho2 = pk1_corrected / rho^2 and F: D x D per
 * point, column-major.  dv: D x N.  The lists are X's own (n_rows = N, columns index X).  One
 * floating type for all arrays.  workspace: NULL, or
 * pn_tlsph_interaction_workspace_bytes(ndims, dtype, n_rows) bytes of 16-byte aligned device
 * memory; with the penalty on, the call then packs each point's fields into one record there
 * and gathers those (same results bit for bit, fewer loads).  Enqueue only. */
int pn_list_tlsph_interaction(const int32_t* row_ptr, const int32_t* col_idx, int64_t n_rows,
                              int stored_base, int ndims, int dtype, int kernel, double h,
                              const void* X, const void* x, const void* mass, const void* rho,
                              const void* pk1_rho2, const void* F, double young_modulus,
                              double penalty_alpha, int64_t sel_first, int64_t sel_count,
                              const int32_t* sel_list, void* dv, int accumulate, void* workspace,
                              void* stream);

/* Size in bytes of pn_list_tlsph_interaction's workspace (-1 for an unsupported ndims/dtype):
 * n_rows records of 2 (D + 1) + 2 D^2 values rounded up to 16 bytes.  No counterpart in the
 * reference (an implementation buffer). */
int64_t pn_tlsph_interaction_workspace_bytes(int ndims, int dtype, int64_t n_rows);

/* Corrected first Piola-Kirchhoff stress over rho^2 of a Saint Venant-Kirchhoff material
 * (TrixiParticles compute_pk1_corrected!, run once by the set-up of benchmark_tlsph,
 * benchmarks/smoothed_particle_hydrodynamics.jl:186): E = (F^T F - I) / 2,
 * S = lame_lambda tr(E) I + 2 lame_mu E, out = F S L / rho^2 (L == NULL: identity).  F, L, out:
 * D x D per point, column-major; rho per point.  Enqueue only. */
int pn_tlsph_pk1_rho2(const void* F, const void* L, const void* rho, int64_t n, int ndims, int dtype,
                      double lame_lambda, double lame_mu, void* out, void* stream);

/* pn_list_sph_density of a PrecomputedNeighborhoodSearch with a PeriodicBox: a stored pair
 * with d2 > search_radius^2 uses its minimum image (compute_periodic_distance,
 * src/neighborhood_search.jl:428-436, called at src/nhs_precomputed.jl:236-238).
 * box_size: max_corner - min_corner, values exact in the radius type.  Enqueue only. */
int pn_list_sph_density_periodic(const int32_t* row_ptr, const int32_t* col_idx, int64_t n_rows,
                                 int stored_base, int ndims, int coord_dtype, int radius_dtype,
                                 int kernel, double h, const void* x, const void* y,
                                 const void* mass, int64_t sel_first, int64_t sel_count,
                                 const int32_t* sel_list, void* rho, int accumulate,
                                 double search_radius, const double* box_size, void* stream);

/* The n-body closure (benchmarks/n_body.jl:34-51) over precomputed lists
 * (src/nhs_precomputed.jl:210-247: stored neighbours, distance recomputed, no radius test):
 * dv[:, i] (+)= sum over the stored j with d >= sqrt(eps) of -G mass[j] pos_diff / d^3.
 * x, y, dv: D x N of `dtype`; lists as pn_list_sph_density.  Enqueue only. */
int pn_list_nbody_accel(const int32_t* row_ptr, const int32_t* col_idx, int64_t n_rows, int stored_base,
                        int ndims, int dtype, double G, const void* x, const void* y, const void* mass,
                        int64_t sel_first, int64_t sel_count, const int32_t* sel_list, void* dv,
                        int accumulate, void* stream);

/* Allocation-free initialize! + neighbour count.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Equals pn_grid_create + pn_grid_update +
 * pn_count_neighbors(PN_QUERIES_GRID, accumulate = 0) on the same arguments (geometry as
 * pn_grid_create, y: D x N, counts: N int32), enqueued on `stream`, then synchronizes and
 * returns the reference's errors (PN_ERR_OUT_OF_BOUNDS, PN_ERR_CELL_FULL).  Every buffer
 * beyond y and counts lives in a static device array of pn_noalloc_scratch_bytes() bytes per
 * GPU (64 MiB: about 28 bytes per Float32 3D point plus 12 bytes per cell); larger problems
 * return PN_ERR_UNSUPPORTED before launching anything.  No hipMalloc or hipFree is called.
 * Calls that overlap on one GPU (several threads, or concurrent streams) share that array
 * and race. */
int pn_count_neighbors_noalloc(int ndims, int coord_dtype, int radius_dtype, int cell_dtype,
                               const double* min_corner, const int64_t* n_cells, double search_radius,
                               int32_t max_points_per_cell, const void* y, int64_t n, int32_t* counts,
                               void* stream);
int64_t pn_noalloc_scratch_bytes(void);

#ifdef __cplusplus
}
#endif

#endif /* POINTNEIGHBORS_HIP_H */
