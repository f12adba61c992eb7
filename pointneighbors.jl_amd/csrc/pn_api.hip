// C-ABI of the MI355X neighbourhood-search engine: object lifetime, device buffers,
// kernel launches and error reporting.  Declarations and the reference interfaces
// each entry point replaces: include/pointneighbors_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pointneighbors_hip.h"
#include "pn_build.h"
#include "pn_sph.h"
#include "pn_traverse.h"
#include "pn_cells.h"
#include "pn_lists.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define PN_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(PN_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define PN_CHECK_LAUNCH() PN_HIP(hipGetLastError())

size_t dtype_size(int t) { return t == PN_FLOAT64 ? 8 : 4; }

int64_t round_up8(int64_t n) { return (n + 7) / 8 * 8; }

// Device buffer that only grows.
struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need, hipStream_t stream) {
        if (need <= bytes) return PN_OK;
        if (p) {
            PN_HIP(hipStreamSynchronize(stream));
            PN_HIP(hipFree(p));
            p = nullptr;
            bytes = 0;
        }
        size_t alloc = std::max<size_t>(need, 256);
        PN_HIP(hipMalloc(&p, alloc));
        bytes = alloc;
        return PN_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const {
        return (T*)p;
    }
};

// Exclusive scan over n int32 values (cell histograms and CSR row counts).
struct Scanner {
    Buf block_sums;
    int run(const int32_t* counts, int64_t n, int32_t* start, int32_t* nonempty, pn::DevStatus* st,
            hipStream_t stream) {
        const int64_t nb = std::max<int64_t>(1, (n + pn::SCAN_TILE - 1) / pn::SCAN_TILE);
        int rc = block_sums.ensure((size_t)nb * 8, stream);
        if (rc) return rc;
        auto* bs = block_sums.as<unsigned long long>();
        hipLaunchKernelGGL(pn::k_scan_reduce, dim3((unsigned)nb), dim3(pn::SCAN_THREADS), 0, stream, counts,
                           n, bs, st);
        PN_CHECK_LAUNCH();
        hipLaunchKernelGGL(pn::k_scan_top, dim3(1), dim3(pn::SCAN_THREADS), 0, stream, bs, nb, start, n, st);
        PN_CHECK_LAUNCH();
        hipLaunchKernelGGL(pn::k_scan_down, dim3((unsigned)nb), dim3(pn::SCAN_THREADS), 0, stream, counts, n,
                           (const unsigned long long*)bs, start, nonempty);
        PN_CHECK_LAUNCH();
        return PN_OK;
    }
};

// One binned point set: the grid's y or a query set x.
struct BinSet {
    Buf counts, start, nonempty, keys, slots, perm_tmp, sorted_idx, coords, status;
    pn::DevStatus* host_status = nullptr;  // pinned
    Scanner scan;
    int64_t cap = 0;
    int64_t n_sel = 0;        // selected points of the last binning
    int64_t n_src = 0;        // columns of the source matrix
    const void* src = nullptr;
    bool valid = false;       // last binning succeeded
    bool stats_known = false;
    pn::DevStatus stats{};

    ~BinSet() {
        counts.release();
        start.release();
        nonempty.release();
        keys.release();
        slots.release();
        perm_tmp.release();
        sorted_idx.release();
        coords.release();
        status.release();
        scan.block_sums.release();
        if (host_status) (void)hipHostFree(host_status);
    }

    pn::SetView view() const {
        pn::SetView v;
        v.start = start.as<int32_t>();
        v.nonempty = nonempty.as<int32_t>();
        v.sorted_idx = sorted_idx.as<int32_t>();
        v.coords = coords.p;
        v.cap = cap;
        v.st = status.as<pn::DevStatus>();
        return v;
    }
};

}  // namespace

struct pn_grid {
    int nd = 0, tc = 0, tr = 0, tcell = 0;
    pn::Geom geom{};
    double radius = 0.0;
    int32_t max_ppc = 100;
    hipStream_t stream = nullptr;
    BinSet grid;
    BinSet query;
    Buf fields;  // per-neighbour fields gathered into the grid's sorted order
    // batch of independent clouds: per-cloud geometry, point -> cloud, cell offsets
    Buf geom_table, cloud_of, cell_offs;
    std::vector<pn::GeomEntry> host_table;
    std::vector<int64_t> host_point_offsets;
    int64_t persist_blocks = 1 << 20;  // workgroups of the persistent traversal (multiple of 8)
    int chunk = 0;                  // candidate chunk: -1 per functor, 0 small, 1 large
    int algorithm = 0;              // traversal: 0 candidate-major cells, 1 per-query slices
    int canonical_order = 1;        // 1: in-cell order = ascending original index
    Buf csr_counts;
    Buf csr_status;
    Scanner csr_scan;
    pn::DevStatus* csr_host_status = nullptr;
    int64_t err_point = -1;
    int64_t err_cell[3] = {0, 0, 0};
    const void* err_src = nullptr;

    ~pn_grid() {
        fields.release();
        geom_table.release();
        cloud_of.release();
        cell_offs.release();
        csr_counts.release();
        csr_status.release();
        csr_scan.block_sums.release();
        if (csr_host_status) (void)hipHostFree(csr_host_status);
    }
};

namespace {

// Template dispatch over (ndims, coordinate type, radius type, cell type).
template <class Fn>
int dispatch_build(const pn_grid& g, Fn&& fn) {
    auto by_d = [&](auto tcv, auto tcellv) -> int {
        using TC = decltype(tcv);
        using TCELL = decltype(tcellv);
        switch (g.nd) {
            case 1: return fn(std::integral_constant<int, 1>{}, TC{}, TCELL{});
            case 2: return fn(std::integral_constant<int, 2>{}, TC{}, TCELL{});
            case 3: return fn(std::integral_constant<int, 3>{}, TC{}, TCELL{});
        }
        return fail(PN_ERR_ARGUMENT, "ndims must be 1, 2 or 3");
    };
    if (g.tc == PN_FLOAT32 && g.tcell == PN_FLOAT32) return by_d(float{}, float{});
    if (g.tc == PN_FLOAT32) return by_d(float{}, double{});
    return by_d(double{}, double{});
}

template <class Fn>
int dispatch_trav(const pn_grid& g, Fn&& fn) {
    auto by_d = [&](auto tcv, auto trv) -> int {
        using TC = decltype(tcv);
        using TR = decltype(trv);
        switch (g.nd) {
            case 1: return fn(std::integral_constant<int, 1>{}, TC{}, TR{});
            case 2: return fn(std::integral_constant<int, 2>{}, TC{}, TR{});
            case 3: return fn(std::integral_constant<int, 3>{}, TC{}, TR{});
        }
        return fail(PN_ERR_ARGUMENT, "ndims must be 1, 2 or 3");
    };
    if (g.tc == PN_FLOAT32 && g.tr == PN_FLOAT32) return by_d(float{}, float{});
    if (g.tc == PN_FLOAT32) return by_d(float{}, double{});
    if (g.tr == PN_FLOAT32) return by_d(double{}, float{});
    return by_d(double{}, double{});
}

int ensure_set(pn_grid& g, BinSet& s, int64_t n_sel) {
    const int64_t C = g.geom.ncells;
    int rc;
    if ((rc = s.counts.ensure((size_t)C * 4, g.stream))) return rc;
    if ((rc = s.start.ensure((size_t)(C + 1) * 4, g.stream))) return rc;
    if ((rc = s.nonempty.ensure((size_t)C * 4, g.stream))) return rc;
    if ((rc = s.status.ensure(sizeof(pn::DevStatus), g.stream))) return rc;
    if (!s.host_status) PN_HIP(hipHostMalloc((void**)&s.host_status, sizeof(pn::DevStatus), hipHostMallocDefault));
    if (n_sel > s.cap) {
        int64_t cap = std::max<int64_t>(n_sel, 64);
        if ((rc = s.keys.ensure((size_t)cap * 4, g.stream))) return rc;
        if ((rc = s.slots.ensure((size_t)cap * 4, g.stream))) return rc;
        if ((rc = s.perm_tmp.ensure((size_t)cap * 4, g.stream))) return rc;
        if ((rc = s.sorted_idx.ensure((size_t)cap * 4, g.stream))) return rc;
        // planes are strided by cap: reallocate exactly
        PN_HIP(hipStreamSynchronize(g.stream));
        s.coords.release();
        if ((rc = s.coords.ensure((size_t)cap * g.nd * dtype_size(g.tc), g.stream))) return rc;
        s.cap = cap;
    }
    return PN_OK;
}

unsigned blocks_for(int64_t n, int threads) { return (unsigned)std::max<int64_t>(1, (n + threads - 1) / threads); }

// Enqueues the counting-sort binning of the selected points of src into s.
int enqueue_bin(pn_grid& g, BinSet& s, const void* src, int64_t n_src, int64_t sel_first, int64_t sel_count,
                const int32_t* sel_list) {
    if (!src && sel_count > 0) return fail(PN_ERR_ARGUMENT, "coordinate pointer is NULL");
    if (sel_count < 0 || n_src < 0) return fail(PN_ERR_ARGUMENT, "negative size");
    if (n_src > INT32_MAX - 1 || sel_count > INT32_MAX - 1)
        return fail(PN_ERR_ARGUMENT, "more than 2^31 - 2 points are not supported");
    if (!sel_list && (sel_first < 0 || sel_first + sel_count > n_src))
        return fail(PN_ERR_ARGUMENT, "selection range out of bounds of the coordinate matrix");
    int rc = ensure_set(g, s, sel_count);
    if (rc) return rc;
    const int64_t C = g.geom.ncells;
    auto* st = s.status.as<pn::DevStatus>();
    hipLaunchKernelGGL(pn::k_reset_status, dim3(1), dim3(1), 0, g.stream, st);
    PN_CHECK_LAUNCH();
    hipLaunchKernelGGL(pn::k_fill_i32, dim3(blocks_for(std::min<int64_t>(C, 1 << 20), 256)), dim3(256), 0,
                       g.stream, s.counts.as<int32_t>(), C, 0);
    PN_CHECK_LAUNCH();
    rc = dispatch_build(g, [&](auto dc, auto tcv, auto tcellv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TCELL = decltype(tcellv);
        if (sel_count > 0) {
            hipLaunchKernelGGL((pn::k_bin<D, TC, TCELL>), dim3(blocks_for(sel_count, 256)), dim3(256), 0, g.stream,
                               g.geom, (const TC*)src, sel_first, sel_list, sel_count, s.keys.as<int32_t>(),
                               s.slots.as<int32_t>(), s.counts.as<int32_t>(), st);
            PN_CHECK_LAUNCH();
        }
        int rc2 = s.scan.run(s.counts.as<int32_t>(), C, s.start.as<int32_t>(), s.nonempty.as<int32_t>(), st,
                             g.stream);
        if (rc2) return rc2;
        if (sel_count > 0) {
            hipLaunchKernelGGL(pn::k_scatter, dim3(blocks_for(sel_count, 256)), dim3(256), 0, g.stream,
                               s.keys.as<int32_t>(), s.slots.as<int32_t>(), sel_first, sel_list, sel_count,
                               s.start.as<int32_t>(), s.perm_tmp.as<int32_t>());
            PN_CHECK_LAUNCH();
            hipLaunchKernelGGL((pn::k_canon<D, TC>), dim3(blocks_for(sel_count, 256)), dim3(256), 0, g.stream,
                               s.keys.as<int32_t>(), s.slots.as<int32_t>(), sel_first, sel_list, sel_count,
                               s.start.as<int32_t>(), s.perm_tmp.as<int32_t>(), (const TC*)src,
                               s.sorted_idx.as<int32_t>(), s.coords.as<TC>(), s.cap, g.canonical_order);
            PN_CHECK_LAUNCH();
        }
        return PN_OK;
    });
    if (rc) return rc;
    s.n_sel = sel_count;
    s.n_src = n_src;
    s.src = src;
    s.valid = true;  // provisional until checked
    s.stats_known = false;
    return PN_OK;
}

// Host coordinates and 1-based cell of one point (error path only).
int describe_point(pn_grid& g, const void* src, int64_t p) {
    g.err_point = p;
    double xs[3] = {0, 0, 0};
    if (g.tc == PN_FLOAT32) {
        float f[3];
        PN_HIP(hipMemcpy(f, (const float*)src + g.nd * p, g.nd * 4, hipMemcpyDeviceToHost));
        for (int d = 0; d < g.nd; ++d) xs[d] = f[d];
    } else {
        PN_HIP(hipMemcpy(xs, (const double*)src + g.nd * p, g.nd * 8, hipMemcpyDeviceToHost));
    }
    const double* minc = g.geom.minc;
    if (!g.host_table.empty()) {
        auto it = std::upper_bound(g.host_point_offsets.begin(), g.host_point_offsets.end(), p);
        const size_t cloud = (size_t)(it - g.host_point_offsets.begin()) - 1;
        minc = g.host_table[std::min(cloud, g.host_table.size() - 1)].minc;
    }
    for (int d = 0; d < g.nd; ++d) {
        double q;
        if (g.tcell == PN_FLOAT32)
            q = (double)(((float)xs[d] - (float)minc[d]) / (float)g.geom.csd[d]);
        else
            q = (xs[d] - minc[d]) / g.geom.csd[d];
        double r = std::floor(q);
        int64_t c;
        if (std::isnan(r) || r >= 9223372036854775808.0) c = INT64_MAX;
        else if (r <= -9223372036854775808.0) c = INT64_MIN;
        else c = (int64_t)r;
        c = (int64_t)((uint64_t)c + 1u);
        if (g.geom.periodic) {
            int64_t m = (int64_t)((uint64_t)c - 2u) % g.geom.per[d];
            if (m < 0) m += g.geom.per[d];
            c = m + 2;
        }
        g.err_cell[d] = c;
    }
    return PN_OK;
}

// Synchronizes and turns the device status of a binning pass into the reference's errors.
int check_set(pn_grid& g, BinSet& s, const int32_t* sel_list, int64_t sel_first, bool is_query) {
    PN_HIP(hipMemcpyAsync(s.host_status, s.status.p, sizeof(pn::DevStatus), hipMemcpyDeviceToHost, g.stream));
    PN_HIP(hipStreamSynchronize(g.stream));
    s.stats = *s.host_status;
    s.stats_known = true;
    if (s.stats.first_bad != ~0ull) {
        s.valid = false;
        int64_t a = (int64_t)s.stats.first_bad;
        int64_t p = a + sel_first;
        if (sel_list) {
            int32_t pi;
            PN_HIP(hipMemcpy(&pi, sel_list + a, 4, hipMemcpyDeviceToHost));
            p = pi;
        }
        int rc = describe_point(g, s.src, p);
        if (rc) return rc;
        if (is_query)
            return fail(PN_ERR_QUERY_BOUNDS,
                        "BoundsError: the neighbor cells of a query point are out of bounds of the cell grid");
        return fail(PN_ERR_OUT_OF_BOUNDS,
                    "particle coordinates are NaN or outside the domain bounds of the cell list");
    }
    if (!is_query && s.stats.max_count > g.max_ppc) {
        s.valid = false;
        return fail(PN_ERR_CELL_FULL, "cell list is full. Use a larger `max_points_per_cell`.");
    }
    return PN_OK;
}

}  // namespace

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
        if (n_periodic[d] > g.geom.n[d] - 2)
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

namespace {

const BinSet* query_set(pn_grid& g, int queries) {
    if (queries == PN_QUERIES_GRID) return g.grid.valid ? &g.grid : nullptr;
    if (queries == PN_QUERIES_BINNED) return (g.query.valid && g.grid.valid) ? &g.query : nullptr;
    return nullptr;
}

// Candidate chunk staged in LDS per wavefront.  Functors with per-neighbour fields and
// pair queues use smaller chunks: LDS per wavefront bounds the waves per CU.
template <class TC, class F>
struct Chunk {
    static constexpr int large = sizeof(TC) == 4 ? 512 : 256;
    static constexpr int small = sizeof(TC) == 4 ? 256 : 128;
};

int check_queries(pn_grid& g, int queries, const BinSet*& Q) {
    Q = query_set(g, queries);
    if (!Q) {
        if (queries != PN_QUERIES_GRID && queries != PN_QUERIES_BINNED)
            return fail(PN_ERR_ARGUMENT, "queries must be PN_QUERIES_GRID or PN_QUERIES_BINNED");
        return fail(PN_ERR_NOT_INITIALIZED, "the neighborhood search is not initialized");
    }
    return PN_OK;
}

// Gathers the functor's per-neighbour fields into sorted order, then launches the
// wave-per-cell traversal for the query set.
template <int D, class TC, class TR, class F>
int launch_traverse_t(pn_grid& g, const BinSet* Q, F f, const pn::FieldSrc<TR>* src) {
    if (Q->n_sel == 0) return PN_OK;
    if constexpr (F::kFields > 0) {
        const BinSet& G = g.grid;
        const int64_t stride = std::max<int64_t>(G.cap, 1);
        int rc = g.fields.ensure((size_t)F::kFields * stride * sizeof(TR), g.stream);
        if (rc) return rc;
        if (G.n_sel > 0) {
            hipLaunchKernelGGL(pn::k_gather_fields<TR>, dim3(blocks_for(G.n_sel, 256)), dim3(256), 0, g.stream,
                               G.sorted_idx.as<int32_t>(), G.status.as<pn::DevStatus>(), *src, g.fields.as<TR>(),
                               stride);
            PN_CHECK_LAUNCH();
        }
        f.fsorted = g.fields.p;
        f.fstride = stride;
    }
    // Upper bound on non-empty query cells; the kernel reads the exact count on device.
    int64_t nwork = std::min<int64_t>(g.geom.ncells, Q->n_sel);
    if (Q->stats_known) nwork = Q->stats.n_nonempty;
    if (nwork == 0) return PN_OK;
    const int nblocks = (int)std::min<int64_t>(round_up8(nwork), g.persist_blocks);
    const TR r = (TR)g.radius;
    const TR r2 = r * r;
    if (g.algorithm == 0 || g.geom.periodic) {  // only the cell kernel handles periodic boxes
        auto launch = [&](auto gc, auto perc) {
            constexpr int GG = decltype(gc)::value;
            constexpr bool PER = decltype(perc)::value;
            hipLaunchKernelGGL((pn::k_cells<D, TC, TR, F, GG, PER>), dim3(nblocks), dim3(pn::WAVE), 0, g.stream,
                               g.geom, Q->view(), g.grid.view(), r2, f, nblocks);
        };
        using Large = pn::IntC<pn::CellGroups<TC, F>::large>;
        using Small = pn::IntC<pn::CellGroups<TC, F>::small>;
        if (g.geom.periodic) {
            if (g.chunk != 0) launch(Large{}, std::true_type{});
            else launch(Small{}, std::true_type{});
        } else {
            if (g.chunk != 0) launch(Large{}, std::false_type{});
            else launch(Small{}, std::false_type{});
        }
        PN_CHECK_LAUNCH();
        return PN_OK;
    }
    const bool large = g.chunk < 0 ? !F::kQueue : g.chunk == 1;
    if (large)
        hipLaunchKernelGGL((pn::k_traverse<D, TC, TR, F, Chunk<TC, F>::large>), dim3(nblocks), dim3(pn::WAVE), 0,
                           g.stream, g.geom, Q->view(), g.grid.view(), r2, f, nblocks);
    else
        hipLaunchKernelGGL((pn::k_traverse<D, TC, TR, F, Chunk<TC, F>::small>), dim3(nblocks), dim3(pn::WAVE), 0,
                           g.stream, g.geom, Q->view(), g.grid.view(), r2, f, nblocks);
    PN_CHECK_LAUNCH();
    return PN_OK;
}

template <class F>
int launch_traverse(pn_grid& g, int queries, const F& f) {
    const BinSet* Q;
    int rc = check_queries(g, queries, Q);
    if (rc) return rc;
    return dispatch_trav(g, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        return launch_traverse_t<D, TC, TR, F>(g, Q, f, (const pn::FieldSrc<TR>*)nullptr);
    });
}

template <class TR>
pn::FieldSrc<TR> fields(std::initializer_list<std::pair<const void*, int>> l) {
    pn::FieldSrc<TR> s{};
    s.n = 0;
    for (auto& e : l) {
        s.p[s.n] = (const TR*)e.first;
        s.step[s.n] = e.second;
        ++s.n;
    }
    return s;
}

}  // namespace

extern "C" {

int pn_count_neighbors(pn_grid_t grid, int queries, int32_t* counts, int accumulate) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    if (!counts) return fail(PN_ERR_ARGUMENT, "counts is NULL");
    pn::CountF f;
    f.out = counts;
    f.accumulate = accumulate;
    return launch_traverse(*grid, queries, f);
}

int pn_sum_neighbor_indices(pn_grid_t grid, int queries, int64_t* sums, int index_base, int accumulate) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    if (!sums) return fail(PN_ERR_ARGUMENT, "sums is NULL");
    pn::IndexSumF f;
    f.out = sums;
    f.accumulate = accumulate;
    f.base = index_base;
    return launch_traverse(*grid, queries, f);
}

int pn_csr_count(pn_grid_t grid, int queries, int64_t n_rows, int32_t* row_ptr, int32_t max_neighbors,
                 int64_t* nnz, int64_t* max_row) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    if (!row_ptr) return fail(PN_ERR_ARGUMENT, "row_ptr is NULL");
    pn_grid& g = *grid;
    const BinSet* Q = query_set(g, queries);
    if (!Q) return fail(PN_ERR_NOT_INITIALIZED, "the neighborhood search is not initialized");
    if (n_rows < Q->n_src) return fail(PN_ERR_ARGUMENT, "n_rows is smaller than the number of query points");
    if (n_rows > INT32_MAX - 1) return fail(PN_ERR_ARGUMENT, "too many rows");
    int rc;
    if ((rc = g.csr_counts.ensure((size_t)std::max<int64_t>(n_rows, 1) * 4, g.stream))) return rc;
    if ((rc = g.csr_status.ensure(sizeof(pn::DevStatus), g.stream))) return rc;
    if (!g.csr_host_status)
        PN_HIP(hipHostMalloc((void**)&g.csr_host_status, sizeof(pn::DevStatus), hipHostMallocDefault));
    auto* st = g.csr_status.as<pn::DevStatus>();
    hipLaunchKernelGGL(pn::k_reset_status, dim3(1), dim3(1), 0, g.stream, st);
    PN_CHECK_LAUNCH();
    hipLaunchKernelGGL(pn::k_fill_i32, dim3(blocks_for(std::min<int64_t>(n_rows, 1 << 20), 256)), dim3(256), 0,
                       g.stream, g.csr_counts.as<int32_t>(), n_rows, 0);
    PN_CHECK_LAUNCH();
    pn::CountF f;
    f.out = g.csr_counts.as<int32_t>();
    f.accumulate = 0;
    if ((rc = launch_traverse(g, queries, f))) return rc;
    if ((rc = g.csr_scan.run(g.csr_counts.as<int32_t>(), n_rows, row_ptr, nullptr, st, g.stream))) return rc;
    PN_HIP(hipMemcpyAsync(g.csr_host_status, st, sizeof(pn::DevStatus), hipMemcpyDeviceToHost, g.stream));
    PN_HIP(hipStreamSynchronize(g.stream));
    const pn::DevStatus hs = *g.csr_host_status;
    if (nnz) *nnz = (int64_t)hs.total64;
    if (max_row) *max_row = hs.max_count;
    if (hs.total64 > (unsigned long long)INT32_MAX)
        return fail(PN_ERR_UNSUPPORTED, "more than 2^31 - 1 neighbor pairs do not fit int32 row pointers");
    if (max_neighbors > 0 && hs.max_count > max_neighbors)
        return fail(PN_ERR_LIST_FULL, "cell list is full. Use a larger `max_points_per_cell`.");
    return PN_OK;
}

int pn_csr_fill(pn_grid_t grid, int queries, int64_t n_rows, const int32_t* row_ptr, int32_t* col_idx,
                int index_base) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    if (!row_ptr) return fail(PN_ERR_ARGUMENT, "row_ptr is NULL");
    pn_grid& g = *grid;
    const BinSet* Q = query_set(g, queries);
    if (!Q) return fail(PN_ERR_NOT_INITIALIZED, "the neighborhood search is not initialized");
    if (n_rows < Q->n_src) return fail(PN_ERR_ARGUMENT, "n_rows is smaller than the number of query points");
    if (g.csr_host_status && g.csr_host_status->max_count > pn::SORT_CAP)
        return fail(PN_ERR_UNSUPPORTED, "neighbor lists longer than 4096 entries are not supported");
    pn::FillF f;
    f.row_ptr = row_ptr;
    f.col = col_idx;
    f.base = index_base;
    f.row_end = 0;
    int rc = launch_traverse(g, queries, f);
    if (rc) return rc;
    if (n_rows == 0) return PN_OK;
    const int nblocks = (int)round_up8(n_rows);
    hipLaunchKernelGGL(pn::k_sort_rows, dim3(nblocks), dim3(pn::WAVE), 0, g.stream, row_ptr, n_rows, col_idx,
                       nblocks);
    PN_CHECK_LAUNCH();
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
// Interaction functors: SPH density / pressure acceleration, pair energy, adjoints.
// ------------------------------------------------------------------------------------
namespace {

int sph_precheck(pn_grid_t grid, int kernel, double h) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    if (grid->nd < 2) return fail(PN_ERR_UNSUPPORTED, "the SPH kernels are defined for 2D and 3D");
    if (kernel != pn::KERNEL_CUBIC_SPLINE && kernel != pn::KERNEL_WENDLAND_C2)
        return fail(PN_ERR_ARGUMENT, "unknown smoothing kernel");
    if (!(h > 0.0)) return fail(PN_ERR_ARGUMENT, "smoothing length must be positive");
    return PN_OK;
}

// Dispatch over (D in {2,3}, TC, TR) for the physics functors.
template <class Fn>
int dispatch_phys(const pn_grid& g, Fn&& fn) {
    return dispatch_trav(g, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        if constexpr (D == 1) {
            return fail(PN_ERR_UNSUPPORTED, "the interaction functors are defined for 2D and 3D");
        } else {
            return fn(dc, tcv, trv);
        }
    });
}

int require_self(pn_grid& g, int queries) {
    if (queries != PN_QUERIES_GRID)
        return fail(PN_ERR_UNSUPPORTED, "this functor needs x == y (the grid's own points)");
    if (!g.grid.valid) return fail(PN_ERR_NOT_INITIALIZED, "the neighborhood search is not initialized");
    return PN_OK;
}

}  // namespace

extern "C" {

int pn_sph_density(pn_grid_t grid, int queries, int kernel, double h, const void* mass, void* rho, int accumulate) {
    int rc = sph_precheck(grid, kernel, h);
    if (rc) return rc;
    if (!mass || !rho) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    pn_grid& g = *grid;
    const BinSet* Q;
    if ((rc = check_queries(g, queries, Q))) return rc;
    return dispatch_phys(g, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        pn::DensityF<TR, D> f;
        f.kern = pn::SmoothingKernel<TR, D>::make(kernel, h);
        f.rho = (TR*)rho;
        f.accumulate = accumulate;
        auto src = fields<TR>({{mass, 1}});
        return launch_traverse_t<D, TC, TR>(g, Q, f, &src);
    });
}

int pn_sph_pressure_accel(pn_grid_t grid, int kernel, double h, const void* mass, const void* P, void* dv,
                          int accumulate) {
    int rc = sph_precheck(grid, kernel, h);
    if (rc) return rc;
    if (!mass || !P || !dv) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    pn_grid& g = *grid;
    if ((rc = require_self(g, PN_QUERIES_GRID))) return rc;
    const BinSet* Q = &g.grid;
    return dispatch_phys(g, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        pn::PressureAccelF<TR, D> f;
        f.kern = pn::SmoothingKernel<TR, D>::make(kernel, h);
        f.P = (const TR*)P;
        f.dv = (TR*)dv;
        f.accumulate = accumulate;
        auto src = fields<TR>({{mass, 1}, {P, 1}});
        return launch_traverse_t<D, TC, TR>(g, Q, f, &src);
    });
}

int pn_sph_density_vjp(pn_grid_t grid, int kernel, double h, const void* mass, const void* rho_bar, void* x_bar,
                       void* mass_bar, int accumulate) {
    int rc = sph_precheck(grid, kernel, h);
    if (rc) return rc;
    if (!mass || !rho_bar || !x_bar) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    pn_grid& g = *grid;
    if ((rc = require_self(g, PN_QUERIES_GRID))) return rc;
    const BinSet* Q = &g.grid;
    return dispatch_phys(g, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        pn::DensityVjpF<TR, D> f;
        f.kern = pn::SmoothingKernel<TR, D>::make(kernel, h);
        f.mass = (const TR*)mass;
        f.g = (const TR*)rho_bar;
        f.xbar = (TR*)x_bar;
        f.mbar = (TR*)mass_bar;
        f.accumulate = accumulate;
        auto src = fields<TR>({{mass, 1}, {rho_bar, 1}});
        return launch_traverse_t<D, TC, TR>(g, Q, f, &src);
    });
}

int pn_sph_pressure_accel_vjp(pn_grid_t grid, int kernel, double h, const void* mass, const void* P,
                              const void* dv_bar, void* x_bar, void* P_bar, void* mass_bar, int accumulate) {
    int rc = sph_precheck(grid, kernel, h);
    if (rc) return rc;
    if (!mass || !P || !dv_bar || !x_bar) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    pn_grid& g = *grid;
    if ((rc = require_self(g, PN_QUERIES_GRID))) return rc;
    const BinSet* Q = &g.grid;
    return dispatch_phys(g, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        pn::PressureVjpF<TR, D> f;
        f.kern = pn::SmoothingKernel<TR, D>::make(kernel, h);
        f.mass = (const TR*)mass;
        f.P = (const TR*)P;
        f.abar = (const TR*)dv_bar;
        f.xbar = (TR*)x_bar;
        f.Pbar = (TR*)P_bar;
        f.mbar = (TR*)mass_bar;
        f.accumulate = accumulate;
        const TR* ab = (const TR*)dv_bar;
        auto src = D == 3 ? fields<TR>({{mass, 1}, {P, 1}, {ab, 3}, {ab + 1, 3}, {ab + 2, 3}})
                          : fields<TR>({{mass, 1}, {P, 1}, {ab, 2}, {ab + 1, 2}});
        return launch_traverse_t<D, TC, TR>(g, Q, f, &src);
    });
}

int pn_pair_energy(pn_grid_t grid, int queries, double eps, double cutoff, void* e, int accumulate) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    if (!e) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (!(cutoff > 0.0)) return fail(PN_ERR_ARGUMENT, "cutoff must be positive");
    pn_grid& g = *grid;
    const BinSet* Q;
    int rc = check_queries(g, queries, Q);
    if (rc) return rc;
    return dispatch_phys(g, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        pn::EnergyF<TR, D> f;
        f.eps = (TR)eps;
        f.inv_rc = (TR)(1.0 / cutoff);
        f.e = (TR*)e;
        f.accumulate = accumulate;
        return launch_traverse_t<D, TC, TR>(g, Q, f, (const pn::FieldSrc<TR>*)nullptr);
    });
}

int pn_pair_energy_vjp(pn_grid_t grid, double eps, double cutoff, const void* e_bar, void* x_bar, int accumulate) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    if (!e_bar || !x_bar) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (!(cutoff > 0.0)) return fail(PN_ERR_ARGUMENT, "cutoff must be positive");
    pn_grid& g = *grid;
    int rc;
    if ((rc = require_self(g, PN_QUERIES_GRID))) return rc;
    const BinSet* Q = &g.grid;
    return dispatch_phys(g, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        pn::EnergyVjpF<TR, D> f;
        f.eps = (TR)eps;
        f.inv_rc = (TR)(1.0 / cutoff);
        f.wq = (const TR*)e_bar;
        f.xbar = (TR*)x_bar;
        f.accumulate = accumulate;
        auto src = fields<TR>({{e_bar, 1}});
        return launch_traverse_t<D, TC, TR>(g, Q, f, &src);
    });
}

int pn_tait_eos(const void* rho, void* p, int64_t n, int dtype, double rho0, double c, double gamma, void* stream) {
    if (!rho || !p) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (n <= 0) return PN_OK;
    if (dtype == PN_FLOAT32) {
        hipLaunchKernelGGL(pn::k_tait_eos<float>, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                           (const float*)rho, (float*)p, n, (float)rho0, (float)c, (float)gamma);
    } else if (dtype == PN_FLOAT64) {
        hipLaunchKernelGGL(pn::k_tait_eos<double>, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                           (const double*)rho, (double*)p, n, rho0, c, gamma);
    } else {
        return fail(PN_ERR_ARGUMENT, "dtype must be PN_FLOAT32 or PN_FLOAT64");
    }
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
        return PN_OK;
    }
    if (k == "algorithm") {
        if (value != 0 && value != 1) return fail(PN_ERR_ARGUMENT, "algorithm must be 0 or 1");
        grid->algorithm = (int)value;
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
    const int nblocks = (int)round_up8(sel_count);
    pn_grid tmp;  // dispatch only
    tmp.nd = ndims;
    tmp.tc = coord_dtype;
    tmp.tr = radius_dtype;
    return dispatch_phys(tmp, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        auto kern = pn::SmoothingKernel<TR, D>::make(kernel, h);
        pn::ListPeriodic<TR> per{};
        per.on = box_size ? 1 : 0;
        const TR r = (TR)search_radius;
        per.r2 = r * r;
        for (int d = 0; d < 3; ++d) per.box[d] = (box_size && d < D) ? (TR)box_size[d] : (TR)1;
        hipLaunchKernelGGL((pn::k_list_density<D, TC, TR, pn::SmoothingKernel<TR, D>>), dim3(nblocks), dim3(pn::WAVE),
                           0, (hipStream_t)stream, row_ptr, col_idx, stored_base, sel_first, sel_list, sel_count,
                           (const TC*)x, (const TC*)y, (const TR*)mass, kern, (TR*)rho, accumulate, nblocks, per);
        PN_CHECK_LAUNCH();
        return PN_OK;
    });
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
