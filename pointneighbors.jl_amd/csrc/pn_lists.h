// Precomputed neighbour lists (src/nhs_precomputed.jl:187-247): CSR row sort and the
// list traversals that work without a grid.
#pragma once

#include "pn_build.h"

namespace pn {

// Sorting of CSR rows (sorteach!, src/vector_of_vectors.jl:177-212), rows up to SORT_CAP.
constexpr int SORT_CAP = 4096;

// x from lane (lane ^ m), m in 1..32.  ds_swizzle's bit mode covers xor masks inside
// 32-lane halves; 32 goes through ds_bpermute.
template <int M>
__device__ __forceinline__ int lane_xor(int x) {
    if constexpr (M < 32) return __builtin_amdgcn_ds_swizzle(x, (M << 10) | 0x1F);
    else return __builtin_amdgcn_ds_bpermute((int)((threadIdx.x ^ M) << 2), x);
}

// One bitonic stage (sequence length SIZE, stride STRIDE) over the 64*K keys held as
// v[j] = key (lane + 64 j): ascending overall.
template <int K, int SIZE, int STRIDE>
__device__ __forceinline__ void bitonic_stage(int (&v)[K], int lane) {
    if constexpr (STRIDE >= WAVE) {
        constexpr int JS = STRIDE / WAVE;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if ((j & JS) == 0) {
                const bool asc = (((lane + WAVE * j) & SIZE) == 0);
                const int a = v[j], b = v[j + JS];
                v[j] = asc ? min(a, b) : max(a, b);
                v[j + JS] = asc ? max(a, b) : min(a, b);
            }
        }
    } else {
        const bool lower = (lane & STRIDE) == 0;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int p = lane_xor<STRIDE>(v[j]);
            const bool asc = (((lane + WAVE * j) & SIZE) == 0);
            v[j] = (lower == asc) ? min(v[j], p) : max(v[j], p);
        }
    }
}

template <int K, int SIZE, int STRIDE>
__device__ __forceinline__ void bitonic_strides(int (&v)[K], int lane) {
    bitonic_stage<K, SIZE, STRIDE>(v, lane);
    if constexpr (STRIDE > 1) bitonic_strides<K, SIZE, STRIDE / 2>(v, lane);
}

template <int K, int SIZE>
__device__ __forceinline__ void bitonic_sizes(int (&v)[K], int lane) {
    bitonic_strides<K, SIZE, SIZE / 2>(v, lane);
    if constexpr (SIZE < WAVE * K) bitonic_sizes<K, SIZE * 2>(v, lane);
}

// Sorts row [b, b + n), n <= 64 K, in registers (padding +INT32_MAX sorts last).
template <int K>
__device__ __forceinline__ void sort_row_regs(int32_t* __restrict__ col, int b, int n, int lane) {
    int v[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int k = lane + WAVE * j;
        v[j] = k < n ? col[b + k] : INT32_MAX;
    }
    bitonic_sizes<K, 2>(v, lane);
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int k = lane + WAVE * j;
        if (k < n) col[b + k] = v[j];
    }
}

// A persistent grid of wavefronts walks the rows (XCD-contiguous ranges) and sorts rows of
// up to 512 entries in registers; k_sort_rows_lds sorts the longer ones (up to SORT_CAP).
constexpr int SORT_REGS_MAX = 8 * WAVE;

static __global__ __launch_bounds__(WAVE) void k_sort_rows(const int32_t* __restrict__ row_ptr, int64_t n_rows,
                                                    int32_t* __restrict__ col, int nblocks) {
    const int lane = threadIdx.x;
    const int xcd = blockIdx.x & 7, per_xcd = nblocks >> 3;
    const int64_t chunk = (n_rows + 7) >> 3;
    const int64_t r_end = min(n_rows, (int64_t)(xcd + 1) * chunk);
    for (int64_t row = (int64_t)xcd * chunk + (blockIdx.x >> 3); row < r_end; row += per_xcd) {
        const int b = row_ptr[row], e = row_ptr[row + 1];
        const int n = e - b;
        if (n <= 1 || n > SORT_REGS_MAX) continue;
        if (n <= WAVE) sort_row_regs<1>(col, b, n, lane);
        else if (n <= 2 * WAVE) sort_row_regs<2>(col, b, n, lane);
        else if (n <= 4 * WAVE) sort_row_regs<4>(col, b, n, lane);
        else sort_row_regs<8>(col, b, n, lane);
    }
}

// Rows of SORT_REGS_MAX + 1 .. SORT_CAP entries: bitonic sort in LDS, one row at a time.
static __global__ __launch_bounds__(WAVE) void k_sort_rows_lds(const int32_t* __restrict__ row_ptr, int64_t n_rows,
                                                        int32_t* __restrict__ col, int nblocks) {
    __shared__ int32_t s[SORT_CAP];
    const int lane = threadIdx.x;
    for (int64_t row = blockIdx.x; row < n_rows; row += nblocks) {
        const int b = row_ptr[row], e = row_ptr[row + 1];
        const int n = e - b;
        if (n <= SORT_REGS_MAX || n > SORT_CAP) continue;
        int P = 1;
        while (P < n) P <<= 1;
        __syncthreads();
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
