// C-ABI of the MI355X neighbourhood-search engine: object lifetime, binning, batches,
// tuning and list traversals.  Declarations and the reference interfaces each entry
// point replaces: include/pointneighbors_hip.h.  Traversal launches live in
// pn_api_trav.hip, pn_api_sph.hip and pn_api_vjp.hip.
#include "pn_host.h"

namespace pn_host {
thread_local std::string g_err;
}  // namespace pn_host

extern "C" {

int pn_abi_version(void) { return PN_ABI_VERSION; }

const char* pn_last_error(void) { return g_err.c_str(); }

int pn_grid_create(int ndims, int coord_dtype, int radius_dtype, int cell_dtype, const double* min_corner,
                   const int64_t* n_cells, double search_radius, int32_t max_points_per_cell, pn_grid_t* out) {
    if (!out) return fail(PN_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (ndims < 1 || ndims > 3) return fail(PN_ERR_ARGUMENT, "ndims must be 1, 2 or 3");
    for (int t : {coord_dtype, radius_dtype, cell_dtype})
        if (t != PN_FLOAT32 && t != PN_FLOAT64) return fail(PN_ERR_ARGUMENT, "dtype must be PN_FLOAT32 or PN_FLOAT64");
    if (cell_dtype == PN_FLOAT32 && (coord_dtype != PN_FLOAT32 || radius_dtype != PN_FLOAT32))
        return fail(PN_ERR_ARGUMENT, "cell_dtype must be the promotion of the coordinate and corner types");
    if (!(search_radius > 0.0) || !std::isfinite(search_radius))
        return fail(PN_ERR_ARGUMENT, "search_radius must be positive and finite");
    if (max_points_per_cell < 1) return fail(PN_ERR_ARGUMENT, "max_points_per_cell must be positive");
    if (max_points_per_cell > INT32_MAX / 27)  // a 3^D stencil's candidates are counted in int32
        return fail(PN_ERR_ARGUMENT, "max_points_per_cell must be at most 79536431");
    int64_t C = 1;
    for (int d = 0; d < ndims; ++d) {
        if (n_cells[d] < 3) return fail(PN_ERR_ARGUMENT, "every dimension needs at least 3 cells");
        if (n_cells[d] > INT32_MAX) return fail(PN_ERR_ARGUMENT, "too many cells");
        C *= n_cells[d];
        if (C > INT32_MAX - 1) return fail(PN_ERR_ARGUMENT, "more than 2^31 - 2 cells are not supported");
    }
    pn_grid* g = new pn_grid();
    g->nd = ndims;
    g->tc = coord_dtype;
    g->tr = radius_dtype;
    g->tcell = cell_dtype;
    g->radius = search_radius;
    g->max_ppc = max_points_per_cell;
    g->geom.nd = ndims;
    for (int d = 0; d < 3; ++d) {
        g->geom.n[d] = d < ndims ? (int)n_cells[d] : 1;
        g->geom.minc[d] = d < ndims ? min_corner[d] : 0.0;
    }
    g->geom.ncells = (int)C;
    g->geom.cs = search_radius;
    for (int d = 0; d < 3; ++d) {
        g->geom.csd[d] = search_radius;
        g->geom.per[d] = 1;
        g->geom.box[d] = 0.0;
    }
    g->geom.periodic = 0;
    *out = g;
    return PN_OK;
}

int pn_grid_create_spatial_hash(int ndims, int coord_dtype, int radius_dtype, double search_radius,
                                int64_t list_size, int32_t max_points_per_cell, pn_grid_t* out) {
    if (!out) return fail(PN_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (ndims < 1 || ndims > 3) return fail(PN_ERR_ARGUMENT, "ndims must be 1, 2 or 3");
    for (int t : {coord_dtype, radius_dtype})
        if (t != PN_FLOAT32 && t != PN_FLOAT64) return fail(PN_ERR_ARGUMENT, "dtype must be PN_FLOAT32 or PN_FLOAT64");
    if (!(search_radius > 0.0) || !std::isfinite(search_radius))
        return fail(PN_ERR_ARGUMENT, "search_radius must be positive and finite");
    if (max_points_per_cell < 1) return fail(PN_ERR_ARGUMENT, "max_points_per_cell must be positive");
    if (max_points_per_cell > INT32_MAX / 27)  // a 3^D stencil's candidates are counted in int32
        return fail(PN_ERR_ARGUMENT, "max_points_per_cell must be at most 79536431");
    if (list_size < 1 || list_size > INT32_MAX - 1) return fail(PN_ERR_ARGUMENT, "list_size must be in 1:2^31-2");
    pn_grid* g = new pn_grid();
    g->nd = ndims;
    g->tc = coord_dtype;
    g->tr = radius_dtype;
    g->tcell = (coord_dtype == PN_FLOAT64 || radius_dtype == PN_FLOAT64) ? PN_FLOAT64 : PN_FLOAT32;
    g->radius = search_radius;
    g->max_ppc = max_points_per_cell;
    g->geom.nd = ndims;
    for (int d = 0; d < 3; ++d) {
        g->geom.n[d] = 1;
        g->geom.minc[d] = 0.0;
        g->geom.csd[d] = search_radius;
        g->geom.per[d] = 1;
        g->geom.box[d] = 0.0;
    }
    g->geom.ncells = (int)list_size;
    g->geom.cs = search_radius;
    g->geom.periodic = 0;
    g->geom.hash = 1;
    g->geom.list_size = (int)list_size;
    g->geom.list_magic = ~0ull / (unsigned long long)list_size;
    *out = g;
    return PN_OK;
}

int pn_grid_set_periodic(pn_grid_t grid, const double* cell_size, const int64_t* n_periodic,
                         const double* box_size) {
    if (!grid || !cell_size || !n_periodic || !box_size) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    pn_grid& g = *grid;
    if (g.geom.table) return fail(PN_ERR_UNSUPPORTED, "periodic boxes are not supported for batched clouds");
    for (int d = 0; d < g.nd; ++d) {
        if (n_periodic[d] < 3)
            return fail(PN_ERR_ARGUMENT,
                        "the `GridNeighborhoodSearch` needs at least 3 cells in each dimension when used with "
                        "periodicity. Please use no NHS for very small problems.");
        if (!g.geom.hash && n_periodic[d] > g.geom.n[d] - 2)
            return fail(PN_ERR_ARGUMENT, "the cell list has fewer than n_periodic + 2 cells in a dimension");
        if (!(cell_size[d] > 0.0) || !(box_size[d] > 0.0)) return fail(PN_ERR_ARGUMENT, "box and cell sizes must be positive");
    }
    for (int d = 0; d < g.nd; ++d) {
        g.geom.csd[d] = cell_size[d];
        g.geom.per[d] = (int)n_periodic[d];
        g.geom.box[d] = box_size[d];
    }
    g.geom.periodic = 1;
    g.grid.valid = false;
    g.query.valid = false;
    return PN_OK;
}

int pn_grid_destroy(pn_grid_t grid) {
    if (!grid) return PN_OK;
    if (grid->stream) (void)hipStreamSynchronize(grid->stream);
    else (void)hipDeviceSynchronize();
    delete grid;
    return PN_OK;
}

int pn_grid_set_stream(pn_grid_t grid, void* stream) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    grid->stream = (hipStream_t)stream;
    return PN_OK;
}

int pn_grid_update_async(pn_grid_t grid, const void* y, int64_t n_y, int64_t sel_first, int64_t sel_count,
                         const int32_t* sel_list) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    if (grid->geom.table && n_y != grid->host_point_offsets.back())
        return fail(PN_ERR_ARGUMENT, "a batch grid is updated with the concatenated points of all its clouds");
    grid->err_src = y;
    return enqueue_bin(*grid, grid->grid, y, n_y, sel_first, sel_count, sel_list);
}

int pn_grid_check(pn_grid_t grid) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    return check_set(*grid, grid->grid, nullptr, 0, false);
}

int pn_grid_update(pn_grid_t grid, const void* y, int64_t n_y, int64_t sel_first, int64_t sel_count,
                   const int32_t* sel_list) {
    int rc = pn_grid_update_async(grid, y, n_y, sel_first, sel_count, sel_list);
    if (rc) {
        if (grid) grid->grid.valid = false;
        return rc;
    }
    return check_set(*grid, grid->grid, sel_list, sel_first, false);
}

int pn_grid_error(pn_grid_t grid, int64_t* point, int64_t* cell) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    if (point) *point = grid->err_point;
    if (cell)
        for (int d = 0; d < grid->nd; ++d) cell[d] = grid->err_cell[d];
    return PN_OK;
}

int pn_grid_refresh(pn_grid_t grid, const void* y, int64_t n_y) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    BinSet& s = grid->grid;
    if (!s.valid) return fail(PN_ERR_NOT_INITIALIZED, "the neighborhood search is not initialized");
    if (n_y != s.n_src) return fail(PN_ERR_ARGUMENT, "coordinate matrix size differs from the last update!");
    ++grid->gen;  // coordinates changed: recorded pair lists are stale
    s.fresh = false;  // and so are a hash grid's per-bucket cells
    if (s.n_sel == 0) return PN_OK;
    return dispatch_build(*grid, [&](auto dc, auto tcv, auto) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        hipLaunchKernelGGL((pn::k_regather<D, TC>), dim3(blocks_for(s.n_sel, 256)), dim3(256), 0, grid->stream,
                           s.sorted_idx.as<int32_t>(), s.status.as<pn::DevStatus>(), (const TC*)y,
                           s.coords.as<TC>(), s.cap);
        PN_CHECK_LAUNCH();
        s.src = y;
        return PN_OK;
    });
}

int pn_grid_bin_queries(pn_grid_t grid, const void* x, int64_t n_x, int64_t sel_first, int64_t sel_count,
                        const int32_t* sel_list) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    if (grid->geom.table) return fail(PN_ERR_UNSUPPORTED, "a batch grid traverses its own points only");
    int rc = enqueue_bin(*grid, grid->query, x, n_x, sel_first, sel_count, sel_list);
    if (rc) {
        grid->query.valid = false;
        return rc;
    }
    return check_set(*grid, grid->query, sel_list, sel_first, true);
}

int pn_grid_stats(pn_grid_t grid, int64_t* n_binned, int64_t* n_nonempty, int64_t* max_count) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    BinSet& s = grid->grid;
    if (!s.valid) return fail(PN_ERR_NOT_INITIALIZED, "the neighborhood search is not initialized");
    if (!s.stats_known) {
        PN_HIP(hipMemcpyAsync(s.host_status, s.status.p, sizeof(pn::DevStatus), hipMemcpyDeviceToHost,
                              grid->stream));
        PN_HIP(hipStreamSynchronize(grid->stream));
        s.stats = *s.host_status;
        s.stats_known = true;
    }
    if (n_binned) *n_binned = s.stats.n_binned;
    if (n_nonempty) *n_nonempty = s.stats.n_nonempty;
    if (max_count) *max_count = s.stats.max_count;
    return PN_OK;
}

int pn_grid_copy_cells(pn_grid_t grid, int32_t* cell_start, int32_t* sorted_index) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    BinSet& s = grid->grid;
    if (!s.valid) return fail(PN_ERR_NOT_INITIALIZED, "the neighborhood search is not initialized");
    if (cell_start)
        PN_HIP(hipMemcpyAsync(cell_start, s.start.p, (size_t)(grid->geom.ncells + 1) * 4, hipMemcpyDeviceToDevice,
                              grid->stream));
    if (sorted_index && s.n_sel > 0)
        PN_HIP(hipMemcpyAsync(sorted_index, s.sorted_idx.p, (size_t)s.n_sel * 4, hipMemcpyDeviceToDevice,
                              grid->stream));
    return PN_OK;
}

}  // extern "C"

extern "C" {

int pn_list_count_neighbors(const int32_t* row_ptr, int64_t n_rows, int64_t sel_first, int64_t sel_count,
                            const int32_t* sel_list, int32_t* counts, int accumulate, void* stream) {
    if (!row_ptr || !counts) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (!sel_list && (sel_first < 0 || sel_first + sel_count > n_rows))
        return fail(PN_ERR_ARGUMENT, "selection range out of bounds of the neighbor lists");
    if (sel_count <= 0) return PN_OK;
    hipLaunchKernelGGL(pn::k_list_count, dim3(blocks_for(sel_count, 256)), dim3(256), 0, (hipStream_t)stream,
                       row_ptr, sel_first, sel_list, sel_count, counts, accumulate);
    PN_CHECK_LAUNCH();
    return PN_OK;
}

int pn_list_sum_neighbor_indices(const int32_t* row_ptr, const int32_t* col_idx, int64_t n_rows,
                                 int stored_base, int64_t sel_first, int64_t sel_count, const int32_t* sel_list,
                                 int64_t* sums, int index_base, int accumulate, void* stream) {
    if (!row_ptr || !sums) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (!sel_list && (sel_first < 0 || sel_first + sel_count > n_rows))
        return fail(PN_ERR_ARGUMENT, "selection range out of bounds of the neighbor lists");
    if (sel_count <= 0) return PN_OK;
    const int nblocks = (int)round_up8(sel_count);
    hipLaunchKernelGGL(pn::k_list_index_sum, dim3(nblocks), dim3(pn::WAVE), 0, (hipStream_t)stream, row_ptr,
                       col_idx, sel_first, sel_list, sel_count, (long long)index_base - stored_base, sums,
                       accumulate, nblocks);
    PN_CHECK_LAUNCH();
    return PN_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------
// Batches of independent clouds, tuning, synthetic displacement.
// ------------------------------------------------------------------------------------
extern "C" {

int pn_batch_create(int ndims, int coord_dtype, int radius_dtype, int cell_dtype, int64_t n_clouds,
                    const double* min_corners, const int64_t* n_cells, double search_radius,
                    int32_t max_points_per_cell, const int64_t* point_offsets, pn_grid_t* out) {
    if (!out) return fail(PN_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (n_clouds < 1 || n_clouds > INT32_MAX / 2) return fail(PN_ERR_ARGUMENT, "n_clouds must be positive");
    if (!min_corners || !n_cells || !point_offsets) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (point_offsets[0] != 0) return fail(PN_ERR_ARGUMENT, "point_offsets must start at 0");
    std::vector<pn::GeomEntry> table((size_t)n_clouds);
    std::vector<int32_t> cell_offsets((size_t)n_clouds + 1);
    int64_t C = 0;
    for (int64_t c = 0; c < n_clouds; ++c) {
        if (point_offsets[c + 1] < point_offsets[c]) return fail(PN_ERR_ARGUMENT, "point_offsets must not decrease");
        int64_t cc = 1;
        for (int d = 0; d < 3; ++d) {
            const int64_t nd_ = d < ndims ? n_cells[c * ndims + d] : 1;
            if (d < ndims && nd_ < 3) return fail(PN_ERR_ARGUMENT, "every dimension needs at least 3 cells");
            table[c].n[d] = (int)nd_;
            table[c].minc[d] = d < ndims ? min_corners[c * ndims + d] : 0.0;
            cc *= nd_;
        }
        table[c].cell_off = (int)C;
        cell_offsets[c] = (int32_t)C;
        C += cc;
        if (C > INT32_MAX - 1) return fail(PN_ERR_ARGUMENT, "more than 2^31 - 2 cells are not supported");
    }
    cell_offsets[n_clouds] = (int32_t)C;
    const int64_t n_points = point_offsets[n_clouds];
    if (n_points > INT32_MAX - 1) return fail(PN_ERR_ARGUMENT, "more than 2^31 - 2 points are not supported");
    // a 3-cell dummy geometry satisfies the single-grid checks; the table overrides it
    int64_t dummy[3] = {3, 3, 3};
    double zero[3] = {0, 0, 0};
    pn_grid_t g;
    int rc = pn_grid_create(ndims, coord_dtype, radius_dtype, cell_dtype, zero, dummy, search_radius,
                            max_points_per_cell, &g);
    if (rc) return rc;
    g->geom.ncells = (int)C;
    std::vector<int32_t> cloud_of((size_t)std::max<int64_t>(n_points, 1));
    for (int64_t c = 0; c < n_clouds; ++c)
        for (int64_t p = point_offsets[c]; p < point_offsets[c + 1]; ++p) cloud_of[p] = (int32_t)c;
    auto upload = [&](Buf& b, const void* src, size_t bytes) -> int {
        int r = b.ensure(bytes, nullptr);
        if (r) return r;
        PN_HIP(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
        return PN_OK;
    };
    if ((rc = upload(g->geom_table, table.data(), table.size() * sizeof(pn::GeomEntry))) ||
        (rc = upload(g->cell_offs, cell_offsets.data(), cell_offsets.size() * 4)) ||
        (rc = upload(g->cloud_of, cloud_of.data(), cloud_of.size() * 4))) {
        delete g;
        return rc;
    }
    g->geom.table = g->geom_table.as<pn::GeomEntry>();
    g->geom.cell_offsets = g->cell_offs.as<int32_t>();
    g->geom.cloud_of_point = g->cloud_of.as<int32_t>();
    g->geom.n_clouds = (int)n_clouds;
    g->host_table = table;
    g->host_point_offsets.assign(point_offsets, point_offsets + n_clouds + 1);
    *out = g;
    return PN_OK;
}

int pn_grid_launch_times(pn_grid_t grid, float* ms, int64_t max, int64_t* n) {
    if (!grid || !n || (max > 0 && !ms)) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    const int64_t pairs = (int64_t)(grid->timing_used / 2);
    if (pairs > 0) PN_HIP(hipEventSynchronize(grid->timing_events[grid->timing_used - 1]));
    for (int64_t i = 0; i < pairs && i < max; ++i)
        PN_HIP(hipEventElapsedTime(&ms[i], grid->timing_events[2 * i], grid->timing_events[2 * i + 1]));
    *n = pairs;
    grid->timing_used = 0;
    return PN_OK;
}

int pn_set_tuning(pn_grid_t grid, const char* key, int64_t value) {
    if (!grid || !key) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    const std::string k(key);
    if (k == "persist_blocks") {
        if (value < 8 || value % 8) return fail(PN_ERR_ARGUMENT, "persist_blocks must be a positive multiple of 8");
        grid->persist_blocks = value;
        return PN_OK;
    }
    if (k == "chunk") {
        if (value < -1 || value > 1) return fail(PN_ERR_ARGUMENT, "chunk must be -1, 0 or 1");
        grid->chunk = (int)value;
        grid->pc_gen = ~0ull;
        return PN_OK;
    }
    if (k == "algorithm") {
        if (value != 0 && value != 1) return fail(PN_ERR_ARGUMENT, "algorithm must be 0 or 1");
        grid->algorithm = (int)value;
        grid->pc_gen = ~0ull;
        return PN_OK;
    }
    if (k == "time_launches") {
        if (value != 0 && value != 1) return fail(PN_ERR_ARGUMENT, "time_launches must be 0 or 1");
        grid->time_launches = (int)value;
        grid->timing_used = 0;
        return PN_OK;
    }
    if (k == "pair_cache") {
        if (value != 0 && value != 1) return fail(PN_ERR_ARGUMENT, "pair_cache must be 0 or 1");
        grid->pair_cache = (int)value;
        grid->pc_gen = ~0ull;
        return PN_OK;
    }
    if (k == "pair_cache_per_point") {
        if (value < 0) return fail(PN_ERR_ARGUMENT, "pair_cache_per_point must be >= 0");
        grid->pc_per_point = value;
        grid->pc_gen = ~0ull;
        return PN_OK;
    }
    if (k == "canonical_order") {
        if (value != 0 && value != 1) return fail(PN_ERR_ARGUMENT, "canonical_order must be 0 or 1");
        grid->canonical_order = (int)value;
        return PN_OK;
    }
    return fail(PN_ERR_ARGUMENT, "unknown tuning key " + k);
}

int pn_displace_uniform(void* x, int64_t n, int dtype, double amplitude, uint64_t seed, int64_t* step,
                        void* stream) {
    if (!x || !step) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (n <= 0) return PN_OK;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == PN_FLOAT32)
        hipLaunchKernelGGL(pn::k_displace<float>, dim3(blocks_for(n, 256)), dim3(256), 0, s, (float*)x, n,
                           (float)amplitude, seed, (const int64_t*)step);
    else if (dtype == PN_FLOAT64)
        hipLaunchKernelGGL(pn::k_displace<double>, dim3(blocks_for(n, 256)), dim3(256), 0, s, (double*)x, n,
                           amplitude, seed, (const int64_t*)step);
    else
        return fail(PN_ERR_ARGUMENT, "dtype must be PN_FLOAT32 or PN_FLOAT64");
    PN_CHECK_LAUNCH();
    hipLaunchKernelGGL(pn::k_step_increment, dim3(1), dim3(1), 0, s, step);
    PN_CHECK_LAUNCH();
    return PN_OK;
}

}  // extern "C"

extern "C" {

static int list_sph_density(const int32_t* row_ptr, const int32_t* col_idx, int64_t n_rows, int stored_base,
                            int ndims, int coord_dtype, int radius_dtype, int kernel, double h, const void* x,
                            const void* y, const void* mass, int64_t sel_first, int64_t sel_count,
                            const int32_t* sel_list, void* rho, int accumulate, void* stream, double search_radius,
                            const double* box_size) {
    if (!row_ptr || !x || !y || !mass || !rho) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (ndims != 2 && ndims != 3) return fail(PN_ERR_UNSUPPORTED, "the SPH kernels are defined for 2D and 3D");
    if (kernel != pn::KERNEL_CUBIC_SPLINE && kernel != pn::KERNEL_WENDLAND_C2)
        return fail(PN_ERR_ARGUMENT, "unknown smoothing kernel");
    if (!(h > 0.0)) return fail(PN_ERR_ARGUMENT, "smoothing length must be positive");
    if (!sel_list && (sel_first < 0 || sel_first + sel_count > n_rows))
        return fail(PN_ERR_ARGUMENT, "selection range out of bounds of the neighbor lists");
    if (box_size)
        for (int d = 0; d < ndims; ++d)
            if (!(box_size[d] > 0.0)) return fail(PN_ERR_ARGUMENT, "box sizes must be positive");
    if (sel_count <= 0) return PN_OK;
    pn_grid tmp;  // dispatch only
    tmp.nd = ndims;
    tmp.tc = coord_dtype;
    tmp.tr = radius_dtype;
    return dispatch_phys(tmp, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        constexpr int rows = pn::ListShape<TR, false>::rows;
        const int nblocks = (int)round_up8((sel_count + rows - 1) / rows);
        auto kern = pn::SmoothingKernel<TR, D>::make(kernel, h);
        pn::ListPeriodic<TR> per{};
        per.on = box_size ? 1 : 0;
        const TR r = (TR)search_radius;
        per.r2 = r * r;
        for (int d = 0; d < 3; ++d) per.box[d] = (box_size && d < D) ? (TR)box_size[d] : (TR)1;
        hipLaunchKernelGGL((pn::k_list_density<D, TC, TR, pn::SmoothingKernel<TR, D>>), dim3(nblocks), dim3(256),
                           0, (hipStream_t)stream, row_ptr, col_idx, stored_base, sel_first, sel_list, sel_count,
                           (const TC*)x, (const TC*)y, (const TR*)mass, kern, (TR*)rho, accumulate, nblocks, per);
        PN_CHECK_LAUNCH();
        return PN_OK;
    });
}

int pn_csr_to_dvov(const int32_t* row_ptr, const int32_t* col_idx, int64_t n_rows, int max_inner, int transpose,
                   int64_t ld, int32_t* backend, int32_t* lengths, void* stream) {
    if (!row_ptr || !backend || !lengths) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (max_inner <= 0) return fail(PN_ERR_ARGUMENT, "max_inner must be positive");
    if (transpose && ld < n_rows) return fail(PN_ERR_ARGUMENT, "ld must be at least n_rows");
    if (n_rows <= 0) return PN_OK;
    if (transpose) {
        hipLaunchKernelGGL(pn::k_csr_to_dvov_rows, dim3(blocks_for(n_rows, 256)), dim3(256), 0, (hipStream_t)stream,
                           row_ptr, col_idx, n_rows, max_inner, ld, backend, lengths);
    } else {
        const int nblocks = (int)std::min<int64_t>(n_rows, 65536);
        hipLaunchKernelGGL(pn::k_csr_to_dvov_cols, dim3(nblocks), dim3(pn::WAVE), 0, (hipStream_t)stream, row_ptr,
                           col_idx, n_rows, max_inner, backend, lengths, nblocks);
    }
    PN_CHECK_LAUNCH();
    return PN_OK;
}

int pn_list_tlsph_deformation_grad(const int32_t* row_ptr, const int32_t* col_idx, int64_t n_rows, int stored_base,
                                   int ndims, int dtype, int kernel, double h, const void* X, const void* x,
                                   const void* volume, const void* L, int invert, int64_t sel_first,
                                   int64_t sel_count, const int32_t* sel_list, void* out, void* stream) {
    if (!row_ptr || !X || !x || !volume || !out) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (ndims != 2 && ndims != 3) return fail(PN_ERR_UNSUPPORTED, "the TLSPH kernels are defined for 2D and 3D");
    if (kernel != pn::KERNEL_CUBIC_SPLINE && kernel != pn::KERNEL_WENDLAND_C2)
        return fail(PN_ERR_ARGUMENT, "unknown smoothing kernel");
    if (!(h > 0.0)) return fail(PN_ERR_ARGUMENT, "smoothing length must be positive");
    if (dtype != PN_FLOAT32 && dtype != PN_FLOAT64) return fail(PN_ERR_ARGUMENT, "dtype must be PN_FLOAT32 or PN_FLOAT64");
    if (invert && L) return fail(PN_ERR_ARGUMENT, "invert needs L == NULL");
    if (!sel_list && (sel_first < 0 || sel_first + sel_count > n_rows))
        return fail(PN_ERR_ARGUMENT, "selection range out of bounds of the neighbor lists");
    if (sel_count <= 0) return PN_OK;
    auto run = [&](auto dc, auto tv) -> int {
        constexpr int D = decltype(dc)::value;
        using T = decltype(tv);
        constexpr int rows = pn::ListShape<T, true>::rows;
        const int nblocks = (int)round_up8((sel_count + rows - 1) / rows);
        auto kern = pn::SmoothingKernel<T, D>::make(kernel, h);
        hipLaunchKernelGGL((pn::k_list_tlsph_grad<D, T, pn::SmoothingKernel<T, D>>), dim3(nblocks), dim3(256), 0,
                           (hipStream_t)stream, row_ptr, col_idx, stored_base, sel_first, sel_list, sel_count,
                           (const T*)X, (const T*)x, (const T*)volume, kern, (const T*)L, invert ? 1 : 0, (T*)out,
                           nblocks);
        PN_CHECK_LAUNCH();
        return PN_OK;
    };
    if (ndims == 2) return dtype == PN_FLOAT32 ? run(pn::IntC<2>{}, 0.0f) : run(pn::IntC<2>{}, 0.0);
    return dtype == PN_FLOAT32 ? run(pn::IntC<3>{}, 0.0f) : run(pn::IntC<3>{}, 0.0);
}

int pn_list_tlsph_interaction(const int32_t* row_ptr, const int32_t* col_idx, int64_t n_rows, int stored_base,
                              int ndims, int dtype, int kernel, double h, const void* X, const void* x,
                              const void* mass, const void* rho, const void* pk1_rho2, const void* F,
                              double young_modulus, double penalty_alpha, int64_t sel_first, int64_t sel_count,
                              const int32_t* sel_list, void* dv, int accumulate, void* workspace, void* stream) {
    if (!row_ptr || !X || !mass || !pk1_rho2 || !dv) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (ndims != 2 && ndims != 3) return fail(PN_ERR_UNSUPPORTED, "the TLSPH kernels are defined for 2D and 3D");
    if (kernel != pn::KERNEL_CUBIC_SPLINE && kernel != pn::KERNEL_WENDLAND_C2)
        return fail(PN_ERR_ARGUMENT, "unknown smoothing kernel");
    if (!(h > 0.0)) return fail(PN_ERR_ARGUMENT, "smoothing length must be positive");
    if (dtype != PN_FLOAT32 && dtype != PN_FLOAT64) return fail(PN_ERR_ARGUMENT, "dtype must be PN_FLOAT32 or PN_FLOAT64");
    if (!(penalty_alpha >= 0.0)) return fail(PN_ERR_ARGUMENT, "penalty_alpha must be non-negative");
    const bool pen = penalty_alpha > 0.0;
    if (pen && (!x || !rho || !F)) return fail(PN_ERR_ARGUMENT, "the penalty force needs x, rho and F");
    if (workspace && ((uintptr_t)workspace & 15u) != 0) return fail(PN_ERR_ARGUMENT, "workspace must be 16-byte aligned");
    if (!sel_list && (sel_first < 0 || sel_first + sel_count > n_rows))
        return fail(PN_ERR_ARGUMENT, "selection range out of bounds of the neighbor lists");
    if (sel_count <= 0) return PN_OK;
    auto run = [&](auto dc, auto tv) -> int {
        constexpr int D = decltype(dc)::value;
        using T = decltype(tv);
        using K = pn::SmoothingKernel<T, D>;
        constexpr int rows = pn::ListShape<T, true>::rows;
        const int nblocks = (int)round_up8((sel_count + rows - 1) / rows);
        auto kern = K::make(kernel, h);
        const T hae = (T)(0.5 * penalty_alpha * young_modulus);
        if (pen && workspace) {
            uint4* rec = (uint4*)workspace;
            hipLaunchKernelGGL((pn::k_tlsph_pack<D, T>), dim3(blocks_for(n_rows, 256)), dim3(256), 0,
                               (hipStream_t)stream, (const T*)X, (const T*)x, (const T*)mass, (const T*)rho,
                               (const T*)pk1_rho2, (const T*)F, n_rows, rec);
            PN_CHECK_LAUNCH();
            hipLaunchKernelGGL((pn::k_list_tlsph_force_rec<D, T, K>), dim3(nblocks), dim3(256), 0, (hipStream_t)stream,
                               row_ptr, col_idx, stored_base, sel_first, sel_list, sel_count, rec, (const T*)rho, hae,
                               kern, (T*)dv, accumulate ? 1 : 0, nblocks);
        } else if (pen)
            hipLaunchKernelGGL((pn::k_list_tlsph_force<D, T, K, true>), dim3(nblocks), dim3(256), 0,
                               (hipStream_t)stream, row_ptr, col_idx, stored_base, sel_first, sel_list, sel_count,
                               (const T*)X, (const T*)x, (const T*)mass, (const T*)rho, (const T*)pk1_rho2,
                               (const T*)F, hae, kern, (T*)dv, accumulate ? 1 : 0, nblocks);
        else
            hipLaunchKernelGGL((pn::k_list_tlsph_force<D, T, K, false>), dim3(nblocks), dim3(256), 0,
                               (hipStream_t)stream, row_ptr, col_idx, stored_base, sel_first, sel_list, sel_count,
                               (const T*)X, (const T*)x, (const T*)mass, (const T*)rho, (const T*)pk1_rho2,
                               (const T*)F, hae, kern, (T*)dv, accumulate ? 1 : 0, nblocks);
        PN_CHECK_LAUNCH();
        return PN_OK;
    };
    if (ndims == 2) return dtype == PN_FLOAT32 ? run(pn::IntC<2>{}, 0.0f) : run(pn::IntC<2>{}, 0.0);
    return dtype == PN_FLOAT32 ? run(pn::IntC<3>{}, 0.0f) : run(pn::IntC<3>{}, 0.0);
}

int64_t pn_tlsph_interaction_workspace_bytes(int ndims, int dtype, int64_t n_rows) {
    if ((ndims != 2 && ndims != 3) || (dtype != PN_FLOAT32 && dtype != PN_FLOAT64) || n_rows < 0) return -1;
    auto per = [&](auto dc, auto tv) -> int64_t {
        return 16 * (int64_t)pn::TlsphRec<decltype(dc)::value, decltype(tv)>::NV;
    };
    const int64_t b = ndims == 2 ? (dtype == PN_FLOAT32 ? per(pn::IntC<2>{}, 0.0f) : per(pn::IntC<2>{}, 0.0))
                                 : (dtype == PN_FLOAT32 ? per(pn::IntC<3>{}, 0.0f) : per(pn::IntC<3>{}, 0.0));
    return b * n_rows;
}

int pn_tlsph_pk1_rho2(const void* F, const void* L, const void* rho, int64_t n, int ndims, int dtype,
                      double lame_lambda, double lame_mu, void* out, void* stream) {
    if (!F || !rho || !out) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (ndims != 2 && ndims != 3) return fail(PN_ERR_UNSUPPORTED, "the TLSPH kernels are defined for 2D and 3D");
    if (dtype != PN_FLOAT32 && dtype != PN_FLOAT64) return fail(PN_ERR_ARGUMENT, "dtype must be PN_FLOAT32 or PN_FLOAT64");
    if (n < 0) return fail(PN_ERR_ARGUMENT, "n must be non-negative");
    if (n == 0) return PN_OK;
    auto run = [&](auto dc, auto tv) -> int {
        constexpr int D = decltype(dc)::value;
        using T = decltype(tv);
        hipLaunchKernelGGL((pn::k_tlsph_pk1_rho2<D, T>), dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                           (const T*)F, (const T*)L, (const T*)rho, n, (T)lame_lambda, (T)lame_mu, (T*)out);
        PN_CHECK_LAUNCH();
        return PN_OK;
    };
    if (ndims == 2) return dtype == PN_FLOAT32 ? run(pn::IntC<2>{}, 0.0f) : run(pn::IntC<2>{}, 0.0);
    return dtype == PN_FLOAT32 ? run(pn::IntC<3>{}, 0.0f) : run(pn::IntC<3>{}, 0.0);
}

int pn_list_sph_density(const int32_t* row_ptr, const int32_t* col_idx, int64_t n_rows, int stored_base, int ndims,
                        int coord_dtype, int radius_dtype, int kernel, double h, const void* x, const void* y,
                        const void* mass, int64_t sel_first, int64_t sel_count, const int32_t* sel_list, void* rho,
                        int accumulate, void* stream) {
    return list_sph_density(row_ptr, col_idx, n_rows, stored_base, ndims, coord_dtype, radius_dtype, kernel, h, x, y,
                            mass, sel_first, sel_count, sel_list, rho, accumulate, stream, 0.0, nullptr);
}

int pn_list_sph_density_periodic(const int32_t* row_ptr, const int32_t* col_idx, int64_t n_rows, int stored_base,
                                 int ndims, int coord_dtype, int radius_dtype, int kernel, double h, const void* x,
                                 const void* y, const void* mass, int64_t sel_first, int64_t sel_count,
                                 const int32_t* sel_list, void* rho, int accumulate, double search_radius,
                                 const double* box_size, void* stream) {
    if (!box_size) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    return list_sph_density(row_ptr, col_idx, n_rows, stored_base, ndims, coord_dtype, radius_dtype, kernel, h, x, y,
                            mass, sel_first, sel_count, sel_list, rho, accumulate, stream, search_radius, box_size);
}

}  // extern "C"
