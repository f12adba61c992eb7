// Precomputed neighbour lists (src/nhs_precomputed.jl:187-247): CSR row sort and the
// list traversals that work without a grid.
#pragma once

#include "pn_build.h"

namespace pn {

// Bitonic sort of each CSR row in LDS, one wavefront per row, rows up to SORT_CAP.
constexpr int SORT_CAP = 4096;

static __global__ __launch_bounds__(WAVE) void k_sort_rows(const int32_t* __restrict__ row_ptr, int64_t n_rows,
                                                    int32_t* __restrict__ col, int nblocks) {
    __shared__ int32_t s[SORT_CAP];
    const int lane = threadIdx.x;
    const int64_t row = xcd_remap(blockIdx.x, nblocks);
    if (row >= n_rows) return;
    const int b = row_ptr[row], e = row_ptr[row + 1];
    const int n = e - b;
    if (n <= 1 || n > SORT_CAP) return;
    int P = 1;
    while (P < n) P <<= 1;
    for (int k = lane; k < P; k += WAVE) s[k] = k < n ? col[b + k] : INT32_MAX;
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = lane; t < (P >> 1); t += WAVE) {
                const int lo = 2 * t - (t & (stride - 1));
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                const int32_t a = s[lo], c = s[hi];
                if ((a > c) == up) {
                    s[lo] = c;
                    s[hi] = a;
                }
            }
            __syncthreads();
        }
    }
    for (int k = lane; k < n; k += WAVE) col[b + k] = s[k];
}

// ------------------------------------------------------------------------------------
// Precomputed-list traversal (src/nhs_precomputed.jl:210-247): iterate the stored list
// of each selected row; no radius test.
// ------------------------------------------------------------------------------------
static __global__ __launch_bounds__(256) void k_list_count(const int32_t* __restrict__ row_ptr, int64_t sel_first,
                                                    const int32_t* __restrict__ sel_list, int64_t n_sel,
                                                    int32_t* __restrict__ out, int accumulate) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_sel) return;
    const int64_t i = sel_point(a, sel_first, sel_list);
    const int32_t c = row_ptr[i + 1] - row_ptr[i];
    out[i] = accumulate ? out[i] + c : c;
}

// Sum of (stored value - stored_base + index_base) over each selected row; one
// wavefront per row, coalesced over the row.
static __global__ __launch_bounds__(WAVE) void k_list_index_sum(const int32_t* __restrict__ row_ptr,
                                                         const int32_t* __restrict__ col, int64_t sel_first,
                                                         const int32_t* __restrict__ sel_list, int64_t n_sel,
                                                         long long shift, int64_t* __restrict__ out,
                                                         int accumulate, int nblocks) {
    const int64_t a = xcd_remap(blockIdx.x, nblocks);
    if (a >= n_sel) return;
    const int64_t i = sel_point(a, sel_first, sel_list);
    const int b = row_ptr[i], e = row_ptr[i + 1];
    long long s = 0;
    for (int k = b + (int)threadIdx.x; k < e; k += WAVE) s += (long long)col[k] + shift;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, WAVE);
    if (threadIdx.x == 0) out[i] = accumulate ? out[i] + s : s;
}

// Periodic box of a list traversal: compute_periodic_distance
// (src/neighborhood_search.jl:428-436) applies when on != 0.
template <class TR>
struct ListPeriodic {
    int on;
    TR r2;
    TR box[3];
};

// SPH density over stored lists (the TLSPH-style sum of a PrecomputedNeighborhoodSearch,
// src/nhs_precomputed.jl:210-247: stored neighbours, distance recomputed, no radius
// test).  One wavefront per selected row, lanes stride the row (coalesced column reads),
// wave reduction in a fixed order.  x and y are D x N coordinate matrices (original
// order), mass per y point, rho per x point.
template <int D, class TC, class TR, class K>
__global__ __launch_bounds__(WAVE) void k_list_density(const int32_t* __restrict__ row_ptr,
                                                       const int32_t* __restrict__ col, int stored_base,
                                                       int64_t sel_first, const int32_t* __restrict__ sel_list,
                                                       int64_t n_sel, const TC* __restrict__ x,
                                                       const TC* __restrict__ y, const TR* __restrict__ mass,
                                                       K kern, TR* __restrict__ rho, int accumulate, int nblocks,
                                                       ListPeriodic<TR> per) {
    const int64_t a = xcd_remap(blockIdx.x, nblocks);
    if (a >= n_sel) return;
    const int64_t i = sel_point(a, sel_first, sel_list);
    const int b = row_ptr[i], e = row_ptr[i + 1];
    TC xi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = x[(int64_t)D * i + d];
    TR s = (TR)0;
    for (int k = b + (int)threadIdx.x; k < e; k += WAVE) {
        const int64_t j = (int64_t)col[k] - stored_base;
        TR pd[D];
#pragma unroll
        for (int d = 0; d < D; ++d) pd[d] = (TR)(xi[d] - y[(int64_t)D * j + d]);
        TR d2 = pd[0] * pd[0];
#pragma unroll
        for (int d = 1; d < D; ++d) {
            const TR p2 = pd[d] * pd[d];
            d2 = d2 + p2;
        }
        if (per.on && d2 > per.r2) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const TR kk = rint(pd[d] / per.box[d]);
                const TR sh = per.box[d] * kk;
                pd[d] = pd[d] - sh;
            }
            d2 = pd[0] * pd[0];
#pragma unroll
            for (int d = 1; d < D; ++d) {
                const TR p2 = pd[d] * pd[d];
                d2 = d2 + p2;
            }
        }
        s += mass[j] * kern.w(kern.sqrt_(d2));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, WAVE);
    if (threadIdx.x == 0) rho[i] = accumulate ? rho[i] + s : s;
}

}  // namespace pn
