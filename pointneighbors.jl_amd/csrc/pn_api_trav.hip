// C-ABI traversal entry points: neighbour counts, index sums and CSR neighbour lists.
#include "pn_host.h"
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
    // persistent grids; the LDS sort only runs when some row is longer than 512 entries
    {
        const int nblocks = (int)std::min<int64_t>(round_up8(n_rows), 65536);
        hipLaunchKernelGGL(pn::k_sort_rows, dim3(nblocks), dim3(pn::WAVE), 0, g.stream, row_ptr, n_rows, col_idx,
                           nblocks);
        PN_CHECK_LAUNCH();
    }
    if (!g.csr_host_status || g.csr_host_status->max_count > pn::SORT_REGS_MAX) {
        const int nb_lds = (int)std::min<int64_t>(n_rows, 4096);
        hipLaunchKernelGGL(pn::k_sort_rows_lds, dim3(nb_lds), dim3(pn::WAVE), 0, g.stream, row_ptr, n_rows, col_idx,
                           nb_lds);
        PN_CHECK_LAUNCH();
    }
    return PN_OK;
}

}  // extern "C"
