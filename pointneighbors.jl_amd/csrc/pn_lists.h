// Precomputed neighbour lists (src/nhs_precomputed.jl:187-247): CSR row sort and the
// list traversals that work without a grid.
#pragma once

#include "pn_build.h"
#include "pn_sph.h"

namespace pn {

// Sorting of CSR rows (sorteach!, src/vector_of_vectors.jl:177-212), rows up to SORT_CAP.
constexpr int SORT_CAP = 4096;

// x from lane (lane ^ m), m in 1..32: DPP quad permutations for 1 and 2 (a VALU move),
// ds_swizzle's bit mode for xor masks inside 32-lane halves, ds_bpermute for 32.
template <int M>
__device__ __forceinline__ int lane_xor(int x) {
    if constexpr (M == 1) return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
    else if constexpr (M == 2) return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
    else if constexpr (M < 32) return __builtin_amdgcn_ds_swizzle(x, (M << 10) | 0x1F);
    else return __builtin_amdgcn_ds_bpermute((int)((threadIdx.x ^ M) << 2), x);
}

// Median of three: with c = INT32_MIN it is min(a, b), with c = INT32_MAX it is max(a, b).
__device__ __forceinline__ int med3_i32(int a, int b, int c) {
    int r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// One bitonic stage (sequence length SIZE, stride STRIDE) over R rows of 64*K keys each,
// v[r][j] = key (lane + 64 j) of row r: ascending overall.  A lane that keeps the smaller
// key of its pair takes med3(own, partner, INT32_MIN), the other med3(.., INT32_MAX); the
// selector depends on the lane only, so it stays in a register across rows.
template <int K, int R, int SIZE, int STRIDE>
__device__ __forceinline__ void bitonic_stage(int (&v)[R][K], int lane) {
    if constexpr (STRIDE >= WAVE) {
        constexpr int JS = STRIDE / WAVE;
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if ((j & JS) == 0) {
                    const bool asc = ((WAVE * j) & SIZE) == 0;
                    const int a = v[r][j], b = v[r][j + JS];
                    v[r][j] = asc ? min(a, b) : max(a, b);
                    v[r][j + JS] = asc ? max(a, b) : min(a, b);
                }
            }
        }
    } else {
        const bool lower = (lane & STRIDE) == 0;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const bool asc = (((lane + WAVE * j) & SIZE) == 0);
            const int sel = (lower == asc) ? INT32_MIN : INT32_MAX;
#pragma unroll
            for (int r = 0; r < R; ++r) v[r][j] = med3_i32(v[r][j], lane_xor<STRIDE>(v[r][j]), sel);
        }
    }
}

template <int K, int R, int SIZE, int STRIDE>
__device__ __forceinline__ void bitonic_strides(int (&v)[R][K], int lane) {
    bitonic_stage<K, R, SIZE, STRIDE>(v, lane);
    if constexpr (STRIDE > 1) bitonic_strides<K, R, SIZE, STRIDE / 2>(v, lane);
}

template <int K, int R, int SIZE>
__device__ __forceinline__ void bitonic_sizes(int (&v)[R][K], int lane) {
    bitonic_strides<K, R, SIZE, SIZE / 2>(v, lane);
    if constexpr (SIZE < WAVE * K) bitonic_sizes<K, R, SIZE * 2>(v, lane);
}

// Sorts R rows [b[r], b[r] + n[r]), each n[r] <= 64 K, in registers (padding INT32_MAX sorts
// last); the R rows' networks are independent, so their latencies overlap.
template <int K, int R>
__device__ __forceinline__ void sort_rows_regs(int32_t* __restrict__ col, const int (&b)[R], const int (&n)[R],
                                               int lane) {
    int v[R][K];
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int k = lane + WAVE * j;
            v[r][j] = k < n[r] ? col[b[r] + k] : INT32_MAX;
        }
    }
    bitonic_sizes<K, R, 2>(v, lane);
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int k = lane + WAVE * j;
            if (k < n[r]) col[b[r] + k] = v[r][j];
        }
    }
}

template <int K>
__device__ __forceinline__ void sort_row_regs(int32_t* __restrict__ col, int b, int n, int lane) {
    const int bb[1] = {b}, nn[1] = {n};
    sort_rows_regs<K, 1>(col, bb, nn, lane);
}

// A persistent grid of wavefronts walks the rows (XCD-contiguous ranges, two rows per step)
// and sorts rows of up to 512 entries in registers; k_sort_rows_lds sorts the longer ones
// (up to SORT_CAP).
constexpr int SORT_REGS_MAX = 8 * WAVE;

// Sorts rows [b0, b0 + n0) and [b1, b1 + n1) of col (wave-uniform arguments): together when
// both fit two registers per lane, else one at a time; rows longer than SORT_REGS_MAX are left
// to k_sort_rows_lds.
__device__ __forceinline__ void sort_row_pair(int32_t* __restrict__ col, int b0, int n0, int b1, int n1, int lane) {
    const int m = max(n0, n1);
    if (m <= 2 * WAVE) {
        const int bb[2] = {b0, b1}, nn[2] = {n0 > 1 ? n0 : 0, n1 > 1 ? n1 : 0};
        if (m <= WAVE) sort_rows_regs<1, 2>(col, bb, nn, lane);
        else sort_rows_regs<2, 2>(col, bb, nn, lane);
        return;
    }
    const int bs[2] = {b0, b1}, ns[2] = {n0, n1};
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int n = ns[r], b = bs[r];
        if (n <= 1 || n > SORT_REGS_MAX) continue;
        if (n <= WAVE) sort_row_regs<1>(col, b, n, lane);
        else if (n <= 2 * WAVE) sort_row_regs<2>(col, b, n, lane);
        else if (n <= 4 * WAVE) sort_row_regs<4>(col, b, n, lane);
        else sort_row_regs<8>(col, b, n, lane);
    }
}

static __global__ __launch_bounds__(WAVE) void k_sort_rows(const int32_t* __restrict__ row_ptr, int64_t n_rows,
                                                    int32_t* __restrict__ col, int nblocks) {
    const int lane = threadIdx.x;
    const int xcd = blockIdx.x & 7, per_xcd = nblocks >> 3;
    const int64_t chunk = (((n_rows + 7) >> 3) + 1) & ~int64_t(1);  // even: row pairs stay in one range
    const int64_t r_end = min(n_rows, (int64_t)(xcd + 1) * chunk);
    for (int64_t row = (int64_t)xcd * chunk + 2 * (int64_t)(blockIdx.x >> 3); row < r_end; row += 2 * per_xcd) {
        const bool two = row + 1 < r_end;
        const int b0 = row_ptr[row], e0 = row_ptr[row + 1];
        const int e1 = two ? row_ptr[row + 2] : e0;
        sort_row_pair(col, b0, e0 - b0, e0, e1 - e0, lane);
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

// CSR lists -> the reference's DynamicVectorOfVectors backend (src/vector_of_vectors.jl:3-31,
// src/nhs_precomputed.jl:40-66): lengths[i] = row length; entry k of row i (k < max_inner) at
// out[k + max_inner * i] (columns per point), or with transpose at out[i + ld * k] (the k-th
// neighbours of all points contiguous, transpose_backend = true; ld = max_outer_length).
// Slots past a row's length are not written (the reference leaves them undefined).
// Non-transposed: one wavefront per row, coalesced over the row; transposed: one thread per
// row, coalesced over rows for every k.
static __global__ __launch_bounds__(WAVE) void k_csr_to_dvov_cols(const int32_t* __restrict__ row_ptr,
                                                           const int32_t* __restrict__ col, int64_t n_rows,
                                                           int max_inner, int32_t* __restrict__ out,
                                                           int32_t* __restrict__ lengths, int nblocks) {
    for (int64_t i = blockIdx.x; i < n_rows; i += nblocks) {
        const int b = row_ptr[i], n = row_ptr[i + 1] - b;
        if (threadIdx.x == 0) lengths[i] = n;
        const int m = min(n, max_inner);
        for (int k = threadIdx.x; k < m; k += WAVE) out[(int64_t)max_inner * i + k] = col[b + k];
    }
}

static __global__ __launch_bounds__(256) void k_csr_to_dvov_rows(const int32_t* __restrict__ row_ptr,
                                                          const int32_t* __restrict__ col, int64_t n_rows,
                                                          int max_inner, int64_t ld, int32_t* __restrict__ out,
                                                          int32_t* __restrict__ lengths) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_rows) return;
    const int b = row_ptr[i], n = row_ptr[i + 1] - b;
    lengths[i] = n;
    const int m = min(n, max_inner);
    for (int k = 0; k < m; ++k) out[i + ld * k] = col[b + k];
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
// test).  ListShape lanes per selected row (256 / lanes rows per 256-thread block): each
// lane first loads up to LIST_UNROLL of its row's column indices, then all their
// coordinates and masses, so a row's loads are in flight together; the lanes' partial sums
// are combined in a fixed butterfly order.  x and y are D x N coordinate matrices (original
// order), mass per y point, rho per x point.
constexpr int LIST_UNROLL = 4;  // measured best of 1, 2, 4, 8 for the density sum

// Lanes per row of the list kernels (a power of two dividing 64) and rows per 256-thread
// block.  Fewer lanes per row put more rows in flight and shorten the closing butterfly: 8
// measured faster for the density sum (0.555 vs 0.588 ms on the 128^3 lists) and for every
// Float64 TLSPH kernel (correction matrix 0.51 vs 0.57 ms, deformation gradient 0.58 vs 0.60,
// interaction 1.745 vs 1.79), 16 for the Float32 TLSPH kernels (deformation gradient 0.384
// vs 0.409 ms, interaction 0.952 vs 0.965).
template <class T, bool TLSPH>
struct ListShape {
    static constexpr int lanes = (TLSPH && sizeof(T) == 4) ? 16 : 8;
    static constexpr int rows = 256 / lanes;
};

template <int D, class TC, class TR, class K>
__global__ __launch_bounds__(256) void k_list_density(const int32_t* __restrict__ row_ptr,
                                                      const int32_t* __restrict__ col, int stored_base,
                                                      int64_t sel_first, const int32_t* __restrict__ sel_list,
                                                      int64_t n_sel, const TC* __restrict__ x,
                                                      const TC* __restrict__ y, const TR* __restrict__ mass,
                                                      K kern, TR* __restrict__ rho, int accumulate, int nblocks,
                                                      ListPeriodic<TR> per) {
    constexpr int LIST_LANES = ListShape<TR, false>::lanes;
    constexpr int LIST_ROWS_PER_BLOCK = ListShape<TR, false>::rows;
    const int sub = threadIdx.x & (LIST_LANES - 1);
    const int64_t a = (int64_t)xcd_remap(blockIdx.x, nblocks) * LIST_ROWS_PER_BLOCK + (threadIdx.x / LIST_LANES);
    if (a >= n_sel) return;  // a row's lanes leave together
    const int64_t i = sel_point(a, sel_first, sel_list);
    const int b = row_ptr[i], e = row_ptr[i + 1];
    TC xi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = x[(int64_t)D * i + d];
    TR s = (TR)0;
    for (int k0 = b + sub; k0 < e; k0 += LIST_LANES * LIST_UNROLL) {
        int64_t j[LIST_UNROLL];
#pragma unroll
        for (int u = 0; u < LIST_UNROLL; ++u) {
            const int k = k0 + LIST_LANES * u;
            j[u] = k < e ? (int64_t)col[k] - stored_base : -1;
        }
        TC yv[LIST_UNROLL][D];
        TR mv[LIST_UNROLL];
#pragma unroll
        for (int u = 0; u < LIST_UNROLL; ++u) {
            const int64_t jj = j[u] >= 0 ? j[u] : 0;
#pragma unroll
            for (int d = 0; d < D; ++d) yv[u][d] = y[(int64_t)D * jj + d];
            mv[u] = mass[jj];
        }
#pragma unroll
        for (int u = 0; u < LIST_UNROLL; ++u) {
            TR pd[D];
#pragma unroll
            for (int d = 0; d < D; ++d) pd[d] = (TR)(xi[d] - yv[u][d]);
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
            const TR t = mv[u] * kern.w(kern.sqrt_(d2));
            s += j[u] >= 0 ? t : (TR)0;
        }
    }
#pragma unroll
    for (int o = LIST_LANES / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, LIST_LANES);
    if (sub == 0) rho[i] = accumulate ? rho[i] + s : s;
}

// TLSPH deformation gradient over stored lists (TrixiParticles 0.5 calc_deformation_grad!,
// run by benchmarks/smoothed_particle_hydrodynamics.jl:127-146 through
// foreach_point_neighbor over the precomputed lists of the initial coordinates):
//     G_a = sum_{b in list(a), |X_ab| >= sqrt(eps)} V_b (x_b - x_a) (grad_a W(X_ab))^T,
//     grad_a W(X_ab) = W'(|X_ab|) X_ab / |X_ab|,   X_ab = X_a - X_b,
// out_a = G_a L_a (L given), G_a^-1 (invert; closed-form adjugate over determinant) or G_a.
// With x = X this is the (symmetric) kernel-gradient correction matrix M_a, and L = M^-1
// (invert) makes G reproduce affine maps exactly.  ListShape lanes
// per row as in k_list_density; each lane keeps its D x D partial sum, the butterfly over the
// lanes has a fixed order.  X, x: D x N; out and L: D x D per point, column-major.
template <int D, class T, class K>
__global__ __launch_bounds__(256) void k_list_tlsph_grad(const int32_t* __restrict__ row_ptr,
                                                         const int32_t* __restrict__ col, int stored_base,
                                                         int64_t sel_first, const int32_t* __restrict__ sel_list,
                                                         int64_t n_sel, const T* __restrict__ X, const T* __restrict__ x,
                                                         const T* __restrict__ vol, K kern, const T* __restrict__ L,
                                                         int invert, T* __restrict__ out, int nblocks) {
    constexpr int LIST_LANES = ListShape<T, true>::lanes;
    constexpr int LIST_ROWS_PER_BLOCK = ListShape<T, true>::rows;
    const int sub = threadIdx.x & (LIST_LANES - 1);
    const int64_t a = (int64_t)xcd_remap(blockIdx.x, nblocks) * LIST_ROWS_PER_BLOCK + (threadIdx.x / LIST_LANES);
    if (a >= n_sel) return;  // a row's lanes leave together
    const int64_t i = sel_point(a, sel_first, sel_list);
    const int b = row_ptr[i], e = row_ptr[i + 1];
    T Xi[D], xi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        Xi[d] = X[(int64_t)D * i + d];
        xi[d] = x[(int64_t)D * i + d];
    }
    const T tiny = (T)1.4901161193847656e-8;  // sqrt(eps(Float64)): coincident points add nothing
    T g[D][D];
#pragma unroll
    for (int r = 0; r < D; ++r) {
#pragma unroll
        for (int c = 0; c < D; ++c) g[r][c] = (T)0;
    }
    constexpr int UG = 2;  // gathers in flight per lane: 2 measured 1.7x faster than 8 (occupancy)
    for (int k0 = b + sub; k0 < e; k0 += LIST_LANES * UG) {
        int64_t j[UG];
#pragma unroll
        for (int u = 0; u < UG; ++u) {
            const int k = k0 + LIST_LANES * u;
            j[u] = k < e ? (int64_t)col[k] - stored_base : -1;
        }
        T Xj[UG][D], xj[UG][D], vj[UG];
#pragma unroll
        for (int u = 0; u < UG; ++u) {
            const int64_t jj = j[u] >= 0 ? j[u] : 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                Xj[u][d] = X[(int64_t)D * jj + d];
                xj[u][d] = x[(int64_t)D * jj + d];
            }
            vj[u] = vol[jj];
        }
#pragma unroll
        for (int u = 0; u < UG; ++u) {
            T Xab[D];
#pragma unroll
            for (int d = 0; d < D; ++d) Xab[d] = Xi[d] - Xj[u][d];
            T d2 = Xab[0] * Xab[0];
#pragma unroll
            for (int d = 1; d < D; ++d) d2 = d2 + Xab[d] * Xab[d];
            // r and 1/r from one reciprocal square root
            const T inv = prsqrt0(d2);
            const T dist = d2 * inv;
            const bool use = j[u] >= 0 && dist >= tiny;
            // V_b W'(r) / r, zero for pairs that do not contribute
            const T s = use ? vj[u] * kern.dw(dist) * inv : (T)0;
#pragma unroll
            for (int r = 0; r < D; ++r) {
                const T dx = xj[u][r] - xi[r];
#pragma unroll
                for (int c = 0; c < D; ++c) g[r][c] = fma(s * dx, Xab[c], g[r][c]);
            }
        }
    }
#pragma unroll
    for (int o = LIST_LANES / 2; o > 0; o >>= 1) {
#pragma unroll
        for (int r = 0; r < D; ++r) {
#pragma unroll
            for (int c = 0; c < D; ++c) g[r][c] += __shfl_xor(g[r][c], o, LIST_LANES);
        }
    }
    if (sub != 0) return;
    T* o = out + (int64_t)D * D * i;
    if (L) {
        const T* l = L + (int64_t)D * D * i;  // column-major: l[r + D c] = L[r][c]
#pragma unroll
        for (int c = 0; c < D; ++c) {
#pragma unroll
            for (int r = 0; r < D; ++r) {
                T v = g[r][0] * l[0 + D * c];
#pragma unroll
                for (int k = 1; k < D; ++k) v = fma(g[r][k], l[k + D * c], v);
                o[r + D * c] = v;
            }
        }
    } else if (invert) {
        T inv[D][D];
        if constexpr (D == 2) {
            const T rdet = (T)1 / (g[0][0] * g[1][1] - g[0][1] * g[1][0]);
            inv[0][0] = g[1][1] * rdet;
            inv[0][1] = -g[0][1] * rdet;
            inv[1][0] = -g[1][0] * rdet;
            inv[1][1] = g[0][0] * rdet;
        } else {
            // cofactors: inv[r][c] = C[c][r] / det
#pragma unroll
            for (int r = 0; r < 3; ++r) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
                    inv[r][c] = g[r1][c1] * g[r2][c2] - g[r1][c2] * g[r2][c1];
                }
            }
            const T rdet = (T)1 / (g[0][0] * inv[0][0] + g[0][1] * inv[1][0] + g[0][2] * inv[2][0]);
#pragma unroll
            for (int r = 0; r < 3; ++r) {
#pragma unroll
                for (int c = 0; c < 3; ++c) inv[r][c] *= rdet;
            }
        }
#pragma unroll
        for (int c = 0; c < D; ++c) {
#pragma unroll
            for (int r = 0; r < D; ++r) o[r + D * c] = inv[r][c];
        }
    } else {
#pragma unroll
        for (int c = 0; c < D; ++c) {
#pragma unroll
            for (int r = 0; r < D; ++r) o[r + D * c] = g[r][c];
        }
    }
}

// TLSPH interaction over stored lists of the initial coordinates (TrixiParticles'
// interact_structure_structure!, timed by the reference's benchmark_tlsph at
// benchmarks/smoothed_particle_hydrodynamics.jl:115-121, set up at :146-187):
//     dv_a (+)= sum_{b in list(a), |X_ab| >= sqrt(eps)} m_b (Pr_a + Pr_b) grad_a W(X_ab)   [stress]
//              + f_ab / m_a                                                             [penalty]
// Pr = corrected first Piola-Kirchhoff stress / rho^2 (D x D per point, column-major),
// grad_a W = W'(|X_ab|) X_ab / |X_ab|.  The penalty (Ganzenmueller 2015 hourglass control,
// enabled when PEN) with x_ab = x_a - x_b the current difference, F the deformation gradient:
//     delta = ((F_a + F_b) X_ab - 2 x_ab) . x_ab / |x_ab|
//     f_ab  = (alpha / 2) V_a V_b W(|X_ab|) / |X_ab|^2 * E * delta * x_ab / |x_ab|,  V = m / rho.
// Both terms are antisymmetric in (a, b): sum_a m_a dv_a = 0 over symmetric lists.  ListShape
// lanes per row as in k_list_density; dv: D x N.
template <int D, class T, class K, bool PEN>
__global__ __launch_bounds__(256) void k_list_tlsph_force(const int32_t* __restrict__ row_ptr,
                                                          const int32_t* __restrict__ col, int stored_base,
                                                          int64_t sel_first, const int32_t* __restrict__ sel_list,
                                                          int64_t n_sel, const T* __restrict__ X,
                                                          const T* __restrict__ x, const T* __restrict__ mass,
                                                          const T* __restrict__ rho, const T* __restrict__ Pr,
                                                          const T* __restrict__ F, T half_alpha_E, K kern,
                                                          T* __restrict__ dv, int accumulate, int nblocks) {
    constexpr int LIST_LANES = ListShape<T, true>::lanes;
    constexpr int LIST_ROWS_PER_BLOCK = ListShape<T, true>::rows;
    const int sub = threadIdx.x & (LIST_LANES - 1);
    const int64_t a = (int64_t)xcd_remap(blockIdx.x, nblocks) * LIST_ROWS_PER_BLOCK + (threadIdx.x / LIST_LANES);
    if (a >= n_sel) return;  // a row's lanes leave together
    const int64_t i = sel_point(a, sel_first, sel_list);
    const int b = row_ptr[i], e = row_ptr[i + 1];
    T Xi[D], xi[D], Pi[D * D], Fi[PEN ? D * D : 1];
#pragma unroll
    for (int d = 0; d < D; ++d) Xi[d] = X[(int64_t)D * i + d];
#pragma unroll
    for (int k = 0; k < D * D; ++k) Pi[k] = Pr[(int64_t)D * D * i + k];
    T pen_a = (T)0;  // (alpha / 2) E V_a / m_a = (alpha / 2) E / rho_a
    if constexpr (PEN) {
#pragma unroll
        for (int d = 0; d < D; ++d) xi[d] = x[(int64_t)D * i + d];
#pragma unroll
        for (int k = 0; k < D * D; ++k) Fi[k] = F[(int64_t)D * D * i + k];
        pen_a = half_alpha_E / rho[i];
    }
    const T tiny = (T)1.4901161193847656e-8;  // sqrt(eps(Float64)): coincident points add nothing
    T acc[D];
#pragma unroll
    for (int d = 0; d < D; ++d) acc[d] = (T)0;
    constexpr int U = 2;  // gathers in flight per lane (measured best: deeper unrolls cost occupancy)
    for (int k0 = b + sub; k0 < e; k0 += LIST_LANES * U) {
        int64_t j[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + LIST_LANES * u;
            j[u] = k < e ? (int64_t)col[k] - stored_base : -1;
        }
        T Xj[U][D], mj[U], Pj[U][D * D], xj[U][D], rj[U], Fj[U][PEN ? D * D : 1];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t jj = j[u] >= 0 ? j[u] : 0;
#pragma unroll
            for (int d = 0; d < D; ++d) Xj[u][d] = X[(int64_t)D * jj + d];
            mj[u] = mass[jj];
#pragma unroll
            for (int k = 0; k < D * D; ++k) Pj[u][k] = Pr[(int64_t)D * D * jj + k];
            if constexpr (PEN) {
#pragma unroll
                for (int d = 0; d < D; ++d) xj[u][d] = x[(int64_t)D * jj + d];
                rj[u] = rho[jj];
#pragma unroll
                for (int k = 0; k < D * D; ++k) Fj[u][k] = F[(int64_t)D * D * jj + k];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            T Xab[D];
#pragma unroll
            for (int d = 0; d < D; ++d) Xab[d] = Xi[d] - Xj[u][d];
            T d2 = Xab[0] * Xab[0];
#pragma unroll
            for (int d = 1; d < D; ++d) d2 = d2 + Xab[d] * Xab[d];
            const T inv = prsqrt0(d2);
            const T dist = d2 * inv;
            const bool use = j[u] >= 0 && dist >= tiny;
            const T safe = use ? dist : (T)1;
            const T isafe = use ? inv : (T)1;
            // m_b W'(r) / r, so that m_b grad_a W = s X_ab
            const T s = use ? mj[u] * kern.dw(safe) * isafe : (T)0;
#pragma unroll
            for (int r = 0; r < D; ++r) {
                T v = (T)0;
#pragma unroll
                for (int c = 0; c < D; ++c) v = fma(Pi[r + D * c] + Pj[u][r + D * c], Xab[c], v);
                acc[r] = fma(s, v, acc[r]);
            }
            if constexpr (PEN) {
                T xab[D], eps[D];
#pragma unroll
                for (int d = 0; d < D; ++d) xab[d] = xi[d] - xj[u][d];
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    T v = (T)-2 * xab[r];
#pragma unroll
                    for (int c = 0; c < D; ++c) v = fma(Fi[r + D * c] + Fj[u][r + D * c], Xab[c], v);
                    eps[r] = v;
                }
                T c2 = xab[0] * xab[0], dl = eps[0] * xab[0];
#pragma unroll
                for (int d = 1; d < D; ++d) {
                    c2 = c2 + xab[d] * xab[d];
                    dl = fma(eps[d], xab[d], dl);
                }
                // V_b W(r) / r^2 * delta / |x_ab| * (alpha / 2) E / rho_a, times x_ab / |x_ab|
                const T g = use ? pen_a * (mj[u] / rj[u]) * kern.w(safe) * (isafe * isafe) * dl / c2 : (T)0;
#pragma unroll
                for (int d = 0; d < D; ++d) acc[d] = fma(g, xab[d], acc[d]);
            }
        }
    }
#pragma unroll
    for (int o = LIST_LANES / 2; o > 0; o >>= 1) {
#pragma unroll
        for (int d = 0; d < D; ++d) acc[d] += __shfl_xor(acc[d], o, LIST_LANES);
    }
    if (sub != 0) return;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        T* o = dv + (int64_t)D * i + d;
        *o = accumulate ? *o + acc[d] : acc[d];
    }
}

// Per-point record of the penalty interaction, one 16-byte aligned block per point so that a
// neighbour costs TlsphRec::NV vector loads instead of 2 D^2 + 2 D + 2 scalar ones:
//     [X (D), m, x (D), V = m / rho, Pr (D*D, column-major), F (D*D, column-major), pad]
template <int D, class T>
struct TlsphRec {
    static constexpr int USED = 2 * (D + 1) + 2 * D * D;
    static constexpr int PER16 = 16 / (int)sizeof(T);
    static constexpr int R = (USED + PER16 - 1) / PER16 * PER16;  // values per point
    static constexpr int NV = R / PER16;                          // 16-byte loads per point
    static constexpr int X = 0, M = D, XC = D + 1, V = 2 * D + 1, P = 2 * D + 2, F = 2 * D + 2 + D * D;
};

template <int D, class T>
__global__ __launch_bounds__(256) void k_tlsph_pack(const T* __restrict__ X, const T* __restrict__ x,
                                                    const T* __restrict__ mass, const T* __restrict__ rho,
                                                    const T* __restrict__ Pr, const T* __restrict__ F, int64_t n,
                                                    uint4* __restrict__ rec) {
    using RC = TlsphRec<D, T>;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T r[RC::R];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        r[RC::X + d] = X[(int64_t)D * i + d];
        r[RC::XC + d] = x[(int64_t)D * i + d];
    }
    r[RC::M] = mass[i];
    r[RC::V] = mass[i] / rho[i];
#pragma unroll
    for (int k = 0; k < D * D; ++k) {
        r[RC::P + k] = Pr[(int64_t)D * D * i + k];
        r[RC::F + k] = F[(int64_t)D * D * i + k];
    }
#pragma unroll
    for (int k = RC::USED; k < RC::R; ++k) r[k] = (T)0;
#pragma unroll
    for (int q = 0; q < RC::NV; ++q) {
        uint4 v;
        __builtin_memcpy(&v, &r[q * RC::PER16], 16);
        rec[(int64_t)RC::NV * i + q] = v;
    }
}

template <int D, class T>
__device__ __forceinline__ void load_rec(const uint4* __restrict__ rec, int64_t i, T (&r)[TlsphRec<D, T>::R]) {
    using RC = TlsphRec<D, T>;
#pragma unroll
    for (int q = 0; q < RC::NV; ++q) {
        const uint4 v = rec[(int64_t)RC::NV * i + q];
        __builtin_memcpy(&r[q * RC::PER16], &v, 16);
    }
}

// k_list_tlsph_force<D, T, K, true> over packed records (same terms, same operation order, so
// the same bits); one neighbour in flight per lane, which measured fastest with the wide loads.
template <int D, class T, class K>
__global__ __launch_bounds__(256) void k_list_tlsph_force_rec(const int32_t* __restrict__ row_ptr,
                                                              const int32_t* __restrict__ col, int stored_base,
                                                              int64_t sel_first, const int32_t* __restrict__ sel_list,
                                                              int64_t n_sel, const uint4* __restrict__ rec,
                                                              const T* __restrict__ rho, T half_alpha_E, K kern,
                                                              T* __restrict__ dv, int accumulate, int nblocks) {
    using RC = TlsphRec<D, T>;
    constexpr int LIST_LANES = ListShape<T, true>::lanes;
    constexpr int LIST_ROWS_PER_BLOCK = ListShape<T, true>::rows;
    const int sub = threadIdx.x & (LIST_LANES - 1);
    const int64_t a = (int64_t)xcd_remap(blockIdx.x, nblocks) * LIST_ROWS_PER_BLOCK + (threadIdx.x / LIST_LANES);
    if (a >= n_sel) return;  // a row's lanes leave together
    const int64_t i = sel_point(a, sel_first, sel_list);
    const int b = row_ptr[i], e = row_ptr[i + 1];
    T ri[RC::R];
    load_rec<D, T>(rec, i, ri);
    const T pen_a = half_alpha_E / rho[i];
    const T tiny = (T)1.4901161193847656e-8;  // sqrt(eps(Float64)): coincident points add nothing
    T acc[D];
#pragma unroll
    for (int d = 0; d < D; ++d) acc[d] = (T)0;
    for (int k = b + sub; k < e; k += LIST_LANES) {
        T rj[RC::R];
        load_rec<D, T>(rec, (int64_t)col[k] - stored_base, rj);
        T Xab[D];
#pragma unroll
        for (int d = 0; d < D; ++d) Xab[d] = ri[RC::X + d] - rj[RC::X + d];
        T d2 = Xab[0] * Xab[0];
#pragma unroll
        for (int d = 1; d < D; ++d) d2 = d2 + Xab[d] * Xab[d];
        const T inv = prsqrt0(d2);
        const T dist = d2 * inv;
        const bool use = dist >= tiny;
        const T safe = use ? dist : (T)1;
        const T isafe = use ? inv : (T)1;
        const T s = use ? rj[RC::M] * kern.dw(safe) * isafe : (T)0;
#pragma unroll
        for (int r = 0; r < D; ++r) {
            T v = (T)0;
#pragma unroll
            for (int c = 0; c < D; ++c) v = fma(ri[RC::P + r + D * c] + rj[RC::P + r + D * c], Xab[c], v);
            acc[r] = fma(s, v, acc[r]);
        }
        T xab[D], eps[D];
#pragma unroll
        for (int d = 0; d < D; ++d) xab[d] = ri[RC::XC + d] - rj[RC::XC + d];
#pragma unroll
        for (int r = 0; r < D; ++r) {
            T v = (T)-2 * xab[r];
#pragma unroll
            for (int c = 0; c < D; ++c) v = fma(ri[RC::F + r + D * c] + rj[RC::F + r + D * c], Xab[c], v);
            eps[r] = v;
        }
        T c2 = xab[0] * xab[0], dl = eps[0] * xab[0];
#pragma unroll
        for (int d = 1; d < D; ++d) {
            c2 = c2 + xab[d] * xab[d];
            dl = fma(eps[d], xab[d], dl);
        }
        const T g = use ? pen_a * rj[RC::V] * kern.w(safe) * (isafe * isafe) * dl / c2 : (T)0;
#pragma unroll
        for (int d = 0; d < D; ++d) acc[d] = fma(g, xab[d], acc[d]);
    }
#pragma unroll
    for (int o = LIST_LANES / 2; o > 0; o >>= 1) {
#pragma unroll
        for (int d = 0; d < D; ++d) acc[d] += __shfl_xor(acc[d], o, LIST_LANES);
    }
    if (sub != 0) return;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        T* o = dv + (int64_t)D * i + d;
        *o = accumulate ? *o + acc[d] : acc[d];
    }
}

// Corrected first Piola-Kirchhoff stress over rho^2 of a Saint Venant-Kirchhoff material
// (TrixiParticles compute_pk1_corrected!): E = (F^T F - I) / 2, S = lambda tr(E) I + 2 mu E,
// out = F S L / rho^2.  F, L, out: D x D per point, column-major; L == nullptr means identity.
template <int D, class T>
__global__ __launch_bounds__(256) void k_tlsph_pk1_rho2(const T* __restrict__ F, const T* __restrict__ L,
                                                        const T* __restrict__ rho, int64_t n, T lambda, T mu,
                                                        T* __restrict__ out) {
    // every thread reaches the barrier below; past the end it recomputes point n - 1 (n >= 1),
    // whose copy lies outside the block's valid range and is not stored
    const int64_t i = min((int64_t)blockIdx.x * blockDim.x + threadIdx.x, n - 1);
    T f[D * D], E[D * D], S[D * D], P[D * D];
#pragma unroll
    for (int k = 0; k < D * D; ++k) f[k] = F[(int64_t)D * D * i + k];
    T tr = (T)0;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            T v = (T)0;  // (F^T F)_rc = sum_k F_kr F_kc
#pragma unroll
            for (int k = 0; k < D; ++k) v = fma(f[k + D * r], f[k + D * c], v);
            E[r + D * c] = (T)0.5 * (v - (r == c ? (T)1 : (T)0));
        }
#pragma unroll
    for (int d = 0; d < D; ++d) tr += E[d + D * d];
#pragma unroll
    for (int k = 0; k < D * D; ++k) S[k] = (T)2 * mu * E[k];
#pragma unroll
    for (int d = 0; d < D; ++d) S[d + D * d] += lambda * tr;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            T v = (T)0;
#pragma unroll
            for (int k = 0; k < D; ++k) v = fma(f[r + D * k], S[k + D * c], v);
            P[r + D * c] = v;
        }
    const T ri = rho[i];
    const T s = (T)1 / (ri * ri);
    // the block's D*D*256 outputs are staged in LDS and stored contiguously: per-thread stores
    // D*D values apart left partial cache lines and wrote 2.7-4x the output bytes to HBM
    __shared__ T s_out[256 * D * D];
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            T v;
            if (L) {
                v = (T)0;
#pragma unroll
                for (int k = 0; k < D; ++k) v = fma(P[r + D * k], L[(int64_t)D * D * i + k + D * c], v);
            } else {
                v = P[r + D * c];
            }
            s_out[D * D * threadIdx.x + r + D * c] = v * s;
        }
    __syncthreads();
    const int64_t m = min((int64_t)blockDim.x, n - i0) * D * D;  // valid outputs of this block
    for (int64_t t = threadIdx.x; t < m; t += blockDim.x) out[D * D * i0 + t] = s_out[t];
}

}  // namespace pn
