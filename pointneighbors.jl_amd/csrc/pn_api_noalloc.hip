// Allocation-free build + neighbour count: initialize! of a FullGridCellList grid and
// foreach_point_neighbor with the counting closure in one call, for callers that cannot
// allocate device memory (code under graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Every buffer the kernels need beyond the caller's coordinates and
// counts lives in `scratch`, one file-scope device array per GPU of NOALLOC_SCRATCH_BYTES,
// held until the process ends.  Calls on one GPU share that array: calls that overlap on
// one GPU (several host threads, or streams that run concurrently) race.
#include "pn_host.h"

#include <mutex>

namespace {

constexpr int64_t NOALLOC_SCRATCH_BYTES = 64ll << 20;  // 64 MiB per GPU
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of this GPU's scratch array (looked up once per device, with that device
// current).  The array's name is a host-side stand-in: only this pointer may reach kernels.
int scratch_ptr(void** out) {
    int dev = 0;
    PN_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(PN_ERR_UNSUPPORTED, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PN_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PN_OK;
}

int64_t align256(int64_t b) { return (b + 255) & ~int64_t(255); }

// Layout of the build and traversal buffers for n points and C cells.
struct Layout {
    int64_t status, counts, start, nonempty, block_sums, keys, slots, perm_tmp, sorted_idx, coords, total;
};

Layout layout(int64_t n, int64_t C, int nd, int tc_size) {
    Layout L{};
    const int64_t nb = std::max<int64_t>(1, (C + pn::SCAN_TILE - 1) / pn::SCAN_TILE);
    const int64_t cap = std::max<int64_t>(n, 1);
    int64_t o = 0;
    L.status = o, o += align256(sizeof(pn::DevStatus));
    L.counts = o, o += align256(C * 4);
    L.start = o, o += align256((C + 1) * 4);
    L.nonempty = o, o += align256(C * 4);
    L.block_sums = o, o += align256(nb * 16);
    L.keys = o, o += align256(cap * 4);
    L.slots = o, o += align256(cap * 4);
    L.perm_tmp = o, o += align256(cap * 4);
    L.sorted_idx = o, o += align256(cap * 4);
    L.coords = o, o += align256(cap * nd * tc_size);
    L.total = o;
    return L;
}

}  // namespace

extern "C" {

int pn_count_neighbors_noalloc(int ndims, int coord_dtype, int radius_dtype, int cell_dtype, const double* min_corner,
                               const int64_t* n_cells, double search_radius, int32_t max_points_per_cell,
                               const void* y, int64_t n, int32_t* counts, void* stream) {
    // argument checks of pn_grid_create and pn_grid_update
    if (ndims < 1 || ndims > 3) return fail(PN_ERR_ARGUMENT, "ndims must be 1, 2 or 3");
    for (int t : {coord_dtype, radius_dtype, cell_dtype})
        if (t != PN_FLOAT32 && t != PN_FLOAT64) return fail(PN_ERR_ARGUMENT, "dtype must be PN_FLOAT32 or PN_FLOAT64");
    if (cell_dtype == PN_FLOAT32 && (coord_dtype != PN_FLOAT32 || radius_dtype != PN_FLOAT32))
        return fail(PN_ERR_ARGUMENT, "cell_dtype must be the promotion of the coordinate and corner types");
    if (!(search_radius > 0.0) || !std::isfinite(search_radius))
        return fail(PN_ERR_ARGUMENT, "search_radius must be positive and finite");
    if (max_points_per_cell < 1 || max_points_per_cell > INT32_MAX / 27)
        return fail(PN_ERR_ARGUMENT, "max_points_per_cell must be in 1:79536431");
    if (!min_corner || !n_cells) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (n < 0) return fail(PN_ERR_ARGUMENT, "negative size");
    if (n > 0 && (!y || !counts)) return fail(PN_ERR_ARGUMENT, "coordinate or count pointer is NULL");
    int64_t C = 1;
    for (int d = 0; d < ndims; ++d) {
        if (n_cells[d] < 3) return fail(PN_ERR_ARGUMENT, "every dimension needs at least 3 cells");
        if (n_cells[d] > INT32_MAX) return fail(PN_ERR_ARGUMENT, "too many cells");
        C *= n_cells[d];
        if (C > INT32_MAX - 1) return fail(PN_ERR_ARGUMENT, "more than 2^31 - 2 cells are not supported");
    }
    const int tc_size = coord_dtype == PN_FLOAT32 ? 4 : 8;
    const Layout L = layout(n, C, ndims, tc_size);
    if (L.total > NOALLOC_SCRATCH_BYTES)
        return fail(PN_ERR_UNSUPPORTED,
                    "pn_count_neighbors_noalloc: " + std::to_string(n) + " points and " + std::to_string(C) +
                        " cells need " + std::to_string(L.total) + " bytes of scratch, more than its " +
                        std::to_string(NOALLOC_SCRATCH_BYTES) + " (about 28 bytes per Float32 3D point plus 12 " +
                        "bytes per cell); use pn_grid_create / pn_grid_update / pn_count_neighbors");
    if (n == 0) return PN_OK;
    void* base = nullptr;
    int rc = scratch_ptr(&base);
    if (rc) return rc;
    unsigned char* sb = (unsigned char*)base;
    auto* st = (pn::DevStatus*)(sb + L.status);
    auto* cnt = (int32_t*)(sb + L.counts);
    auto* start = (int32_t*)(sb + L.start);
    auto* nonempty = (int32_t*)(sb + L.nonempty);
    auto* bs = (unsigned long long*)(sb + L.block_sums);
    auto* keys = (int32_t*)(sb + L.keys);
    auto* slots = (int32_t*)(sb + L.slots);
    auto* perm_tmp = (int32_t*)(sb + L.perm_tmp);
    auto* sorted_idx = (int32_t*)(sb + L.sorted_idx);
    void* coords = sb + L.coords;
    const hipStream_t s = (hipStream_t)stream;

    // geometry as pn_grid_create sets it
    pn_grid g;
    g.nd = ndims;
    g.tc = coord_dtype;
    g.tr = radius_dtype;
    g.tcell = cell_dtype;
    g.radius = search_radius;
    g.geom.nd = ndims;
    for (int d = 0; d < 3; ++d) {
        g.geom.n[d] = d < ndims ? (int)n_cells[d] : 1;
        g.geom.minc[d] = d < ndims ? min_corner[d] : 0.0;
        g.geom.csd[d] = search_radius;
        g.geom.per[d] = 1;
        g.geom.box[d] = 0.0;
    }
    g.geom.ncells = (int)C;
    g.geom.cs = search_radius;
    g.geom.periodic = 0;

    // build: the counting sort of pn_grid_update (every scratch buffer that must start at
    // zero is cleared by a kernel here; the array keeps the previous call's contents)
    hipLaunchKernelGGL(pn::k_reset_status, dim3(1), dim3(1), 0, s, st);
    PN_CHECK_LAUNCH();
    hipLaunchKernelGGL(pn::k_fill_i32, dim3(blocks_for(std::min<int64_t>(C, 1 << 20), 256)), dim3(256), 0, s, cnt, C,
                       0);
    PN_CHECK_LAUNCH();
    const int64_t nb = std::max<int64_t>(1, (C + pn::SCAN_TILE - 1) / pn::SCAN_TILE);
    rc = dispatch_build(g, [&](auto dc, auto tcv, auto tcellv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TCELL = decltype(tcellv);
        hipLaunchKernelGGL((pn::k_bin<D, TC, TCELL>), dim3(blocks_for(n, 256)), dim3(256), 0, s, g.geom, (const TC*)y,
                           (int64_t)0, (const int32_t*)nullptr, n, keys, slots, cnt, st);
        PN_CHECK_LAUNCH();
        hipLaunchKernelGGL(pn::k_scan_reduce, dim3((unsigned)nb), dim3(pn::SCAN_THREADS), 0, s, (const int32_t*)cnt,
                           C, bs, st);
        PN_CHECK_LAUNCH();
        hipLaunchKernelGGL(pn::k_scan_top, dim3(1), dim3(pn::SCAN_THREADS), 0, s, bs, nb, start, C, st,
                           (pn::StatusMailbox*)nullptr, 0ull);
        PN_CHECK_LAUNCH();
        hipLaunchKernelGGL(pn::k_scan_down, dim3((unsigned)nb), dim3(pn::SCAN_THREADS), 0, s, (const int32_t*)cnt, C,
                           (const unsigned long long*)bs, start, nonempty, (int32_t*)nullptr);
        PN_CHECK_LAUNCH();
        hipLaunchKernelGGL(pn::k_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, s, (const int32_t*)keys,
                           (const int32_t*)slots, n, (const int32_t*)start, perm_tmp, (int32_t*)nullptr);
        PN_CHECK_LAUNCH();
        hipLaunchKernelGGL((pn::k_canon<D, TC>), dim3(blocks_for(n, 256)), dim3(256), 0, s, (const int32_t*)keys,
                           (int64_t)0, (const int32_t*)nullptr, (const pn::DevStatus*)st, (const int32_t*)start,
                           (const int32_t*)perm_tmp, (const TC*)y, sorted_idx, (TC*)coords, std::max<int64_t>(n, 1),
                           1);
        PN_CHECK_LAUNCH();
        return PN_OK;
    });
    if (rc) return rc;

    // count: the cell traversal of pn_count_neighbors over the grid's own points
    pn::SetView v;
    v.start = start;
    v.nonempty = nonempty;
    v.sorted_idx = sorted_idx;
    v.coords = coords;
    v.cap = std::max<int64_t>(n, 1);
    v.st = st;
    pn::CountF f;
    f.out = counts;
    f.accumulate = 0;
    const int nblocks = (int)std::min<int64_t>(round_up8(std::min<int64_t>(C, n)), int64_t(1) << 20);
    rc = dispatch_trav(g, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        constexpr int G = pn::CellGroups<TC, pn::CountF>::large;
        const TR r = (TR)search_radius;
        hipLaunchKernelGGL((pn::k_cells<D, TC, TR, pn::CountF, G, false, 0>), dim3(nblocks), dim3(pn::WAVE), 0, s,
                           g.geom, v, v, r * r, f, nblocks, pn::PairCache{});
        PN_CHECK_LAUNCH();
        return PN_OK;
    });
    if (rc) return rc;

    // the reference's errors (initialize! checks its points, src/cell_lists/full_grid.jl:187-213,
    // src/vector_of_vectors.jl:114-121)
    pn::DevStatus hs;
    PN_HIP(hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s));
    PN_HIP(hipStreamSynchronize(s));
    if (hs.first_bad != ~0ull)
        return fail(PN_ERR_OUT_OF_BOUNDS, "particle coordinates are NaN or outside the domain bounds of the cell list "
                                          "(point " + std::to_string(hs.first_bad) + ")");
    if (hs.max_count > max_points_per_cell)
        return fail(PN_ERR_CELL_FULL, "cell list is full. Use a larger `max_points_per_cell`.");
    return PN_OK;
}

int64_t pn_noalloc_scratch_bytes(void) { return NOALLOC_SCRATCH_BYTES; }

}  // extern "C"
