// Device kernels of the MI355X neighbourhood-search engine (gfx950, wave64).
//
//  Build (initialize!/update!, ParallelUpdate, src/nhs_grid.jl:255-281):
//    counting sort of the selected points by linear cell index
//      k_bin       cell key per point + bounds check + wave-aggregated histogram
//      k_scan_*    exclusive scan of the histogram -> cell_start, non-empty cell list
//      k_scatter   point -> provisional slot in its cell segment
//      k_canon     deterministic order inside each cell (original index ascending)
//                  and the cell-sorted coordinate planes
//  Traversal (foreach_point_neighbor, hot loop src/nhs_grid.jl:519-575):
//      k_traverse  one wavefront per query cell; the 3^D stencil is 3^(D-1) contiguous
//                  x-rows of the sorted order, staged in LDS with coalesced loads; lanes
//                  split as (query, candidate slice)
//  Precomputed lists (src/nhs_precomputed.jl:187-206): count, scan, fill, sort rows.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pn {

constexpr int WAVE = 64;

// Grid geometry.  minc and cs are exact in the cell-computation type.
struct Geom {
    int nd;
    int n[3];      // cells per dimension, 1 for unused dimensions
    int ncells;
    double minc[3];
    double cs;
};

// Device-side status of one binning pass.
struct DevStatus {
    unsigned long long first_bad;  // smallest selection position out of bounds
    unsigned long long total64;    // sum of all counts, 64-bit (overflow guard)
    int max_count;                 // largest cell population
    int n_nonempty;                // number of non-empty cells
    int n_binned;                  // number of binned points (low 32 bits of the sum)
    int pad;
};

// Read-only view of a binned point set used by the traversal.
struct SetView {
    const int32_t* start;       // C + 1 cell offsets
    const int32_t* nonempty;    // non-empty cell ids
    const int32_t* sorted_idx;  // original index per sorted position
    const void* coords;         // D planes of cap entries, cell-sorted
    int64_t cap;
    const DevStatus* st;
};

// Workgroups are dispatched round-robin over the 8 XCDs.  Remap so that each XCD
// works on a contiguous range of cells: neighbouring cells share stencil rows, which
// then hit the same L2.  The launch size must be a multiple of 8.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    const int per = nblocks >> 3;
    return (bid & 7) * per + (bid >> 3);
}

// ------------------------------------------------------------------------------------
// Build kernels
// ------------------------------------------------------------------------------------

__global__ void k_reset_status(DevStatus* st) {
    st->first_bad = ~0ull;
    st->total64 = 0ull;
    st->max_count = 0;
    st->n_nonempty = 0;
    st->n_binned = 0;
    st->pad = 0;
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// Cell key of point p: (x - min) / cs in TCELL, floor, 1-based cell must lie in
// 2:(n_d - 1) (src/cell_lists/full_grid.jl:84-94, :187-213).  With f = floor(q) the
// condition reads 1 <= f <= n_d - 2; NaN and +-Inf fail both comparisons, matching the
// saturating floor_to_int (src/util.jl:19-34) whose saturated values are out of range.
template <int D, class TC, class TCELL>
__device__ __forceinline__ int cell_key(const Geom& g, const TC* __restrict__ xp, bool& ok) {
    int key = 0, stride = 1;
    ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        TCELL q = ((TCELL)xp[d] - (TCELL)g.minc[d]) / (TCELL)g.cs;
        TCELL f = floor(q);
        bool in = (f >= (TCELL)1) && (f <= (TCELL)(g.n[d] - 2));
        ok = ok && in;
        key += (in ? (int)f : 0) * stride;
        stride *= g.n[d];
    }
    return key;
}

__device__ __forceinline__ int64_t sel_point(int64_t a, int64_t sel_first, const int32_t* sel_list) {
    return sel_list ? (int64_t)sel_list[a] : sel_first + a;
}

// One thread per selected point.  Lanes of a wavefront whose keys form a run (common:
// inputs are usually close to cell order) share one atomic on the histogram; the slot
// of each point inside its cell segment is the atomic's old value plus its run offset.
template <int D, class TC, class TCELL>
__global__ __launch_bounds__(256) void k_bin(Geom g, const TC* __restrict__ y, int64_t sel_first,
                                             const int32_t* __restrict__ sel_list, int64_t n_sel,
                                             int32_t* __restrict__ keys, int32_t* __restrict__ slots,
                                             int32_t* __restrict__ counts, DevStatus* st) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (WAVE - 1);
    int key = -1;
    if (a < n_sel) {
        int64_t p = sel_point(a, sel_first, sel_list);
        bool ok;
        int k = cell_key<D, TC, TCELL>(g, y + (int64_t)D * p, ok);
        if (ok) key = k;
        else atomicMin(&st->first_bad, (unsigned long long)a);
        keys[a] = key;
    }
    const int prev = __shfl_up(key, 1, WAVE);
    const bool head = (lane == 0) || (key != prev);
    const unsigned long long hm = __ballot(head);
    const unsigned long long above = lane == 63 ? 0ull : (hm >> (lane + 1)) << (lane + 1);
    const int next = above ? __ffsll((long long)above) - 1 : WAVE;
    int base = 0;
    if (head && key >= 0) base = atomicAdd(&counts[key], next - lane);
    const unsigned long long upto = hm & ((2ull << lane) - 1ull);
    const int hl = 63 - __clzll((long long)upto);
    base = __shfl(base, hl, WAVE);
    if (a < n_sel && key >= 0) slots[a] = base + (lane - hl);
}

// Scan of a histogram of C int32 counts, 2048 items per block (256 threads x 8).
// Packed 64-bit value per cell: low word = count, high word = (count > 0); the high
// word's prefix is the position in the non-empty cell list.
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ unsigned long long pack_cell(int32_t c) {
    return (unsigned long long)(uint32_t)c | ((unsigned long long)(c > 0) << 32);
}

__device__ __forceinline__ unsigned long long wave_incl_scan(unsigned long long v, int lane) {
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        unsigned long long t = __shfl_up(v, o, WAVE);
        if (lane >= o) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread (256 threads = 4 waves).
__device__ __forceinline__ unsigned long long block_excl_scan(unsigned long long v,
                                                              unsigned long long* sh /*4*/,
                                                              unsigned long long& total) {
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x >> 6;
    unsigned long long inc = wave_incl_scan(v, lane);
    if (lane == WAVE - 1) sh[w] = inc;
    __syncthreads();
    unsigned long long off = 0;
    total = 0;
#pragma unroll
    for (int i = 0; i < SCAN_THREADS / WAVE; ++i) {
        unsigned long long s = sh[i];
        if (i < w) off += s;
        total += s;
    }
    __syncthreads();
    return off + inc - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const int32_t* __restrict__ counts, int64_t C,
                                                              unsigned long long* __restrict__ block_sums,
                                                              DevStatus* st) {
    __shared__ unsigned long long sh[SCAN_THREADS / WAVE];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    unsigned long long s = 0;
    int mx = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        int64_t c = base + i;
        int32_t v = c < C ? counts[c] : 0;
        s += pack_cell(v);
        mx = max(mx, v);
    }
    unsigned long long total;
    block_excl_scan(s, sh, total);
    if (threadIdx.x == 0) {
        block_sums[blockIdx.x] = total;
        atomicAdd(&st->total64, total & 0xffffffffull);
    }
    // wave max of the populations, one atomic per wavefront
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0 && mx > 0) atomicMax(&st->max_count, mx);
}

// Single workgroup: exclusive scan of the block sums, writes the totals.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_top(unsigned long long* __restrict__ block_sums,
                                                           int64_t nblocks, int32_t* __restrict__ start,
                                                           int64_t C, DevStatus* st) {
    __shared__ unsigned long long sh[SCAN_THREADS / WAVE];
    unsigned long long carry = 0;
    for (int64_t b0 = 0; b0 < nblocks; b0 += SCAN_THREADS) {
        int64_t b = b0 + threadIdx.x;
        unsigned long long v = b < nblocks ? block_sums[b] : 0ull;
        unsigned long long total;
        unsigned long long ex = block_excl_scan(v, sh, total);
        if (b < nblocks) block_sums[b] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) {
        start[C] = (int32_t)(uint32_t)(carry & 0xffffffffull);
        st->n_binned = (int)(uint32_t)(carry & 0xffffffffull);
        st->n_nonempty = (int)(uint32_t)(carry >> 32);
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(const int32_t* __restrict__ counts, int64_t C,
                                                            const unsigned long long* __restrict__ block_sums,
                                                            int32_t* __restrict__ start,
                                                            int32_t* __restrict__ nonempty) {
    __shared__ unsigned long long sh[SCAN_THREADS / WAVE];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int32_t v[SCAN_ITEMS];
    unsigned long long s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        int64_t c = base + i;
        v[i] = c < C ? counts[c] : 0;
        s += pack_cell(v[i]);
    }
    unsigned long long total;
    unsigned long long run = block_sums[blockIdx.x] + block_excl_scan(s, sh, total);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        int64_t c = base + i;
        if (c < C) {
            start[c] = (int32_t)(uint32_t)(run & 0xffffffffull);
            if (nonempty && v[i] > 0) nonempty[(uint32_t)(run >> 32)] = (int32_t)c;
        }
        run += pack_cell(v[i]);
    }
}

__global__ __launch_bounds__(256) void k_scatter(const int32_t* __restrict__ keys,
                                                 const int32_t* __restrict__ slots, int64_t sel_first,
                                                 const int32_t* __restrict__ sel_list, int64_t n_sel,
                                                 const int32_t* __restrict__ start,
                                                 int32_t* __restrict__ perm_tmp) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_sel) return;
    const int key = keys[a];
    if (key < 0) return;
    perm_tmp[start[key] + slots[a]] = (int32_t)sel_point(a, sel_first, sel_list);
}

// Cells with more points than this keep the provisional (atomic) order instead of the
// O(n^2) rank: their neighbour sets are unaffected, only the order inside the cell.
constexpr int CANON_RANK_LIMIT = 1024;

// Final position of each point: its cell's start plus the number of points of the same
// cell with a smaller original index.  Writes the permutation and the cell-sorted
// coordinate planes (plane d at coords + d * cap).
template <int D, class TC>
__global__ __launch_bounds__(256) void k_canon(const int32_t* __restrict__ keys,
                                               const int32_t* __restrict__ slots, int64_t sel_first,
                                               const int32_t* __restrict__ sel_list, int64_t n_sel,
                                               const int32_t* __restrict__ start,
                                               const int32_t* __restrict__ perm_tmp,
                                               const TC* __restrict__ y, int32_t* __restrict__ sorted_idx,
                                               TC* __restrict__ coords, int64_t cap) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_sel) return;
    const int key = keys[a];
    if (key < 0) return;
    const int32_t p = (int32_t)sel_point(a, sel_first, sel_list);
    const int s = start[key], e = start[key + 1];
    int pos;
    if (e - s <= CANON_RANK_LIMIT) {
        int rank = 0;
        for (int q = s; q < e; ++q) rank += perm_tmp[q] < p;
        pos = s + rank;
    } else {
        pos = s + slots[a];
    }
    sorted_idx[pos] = p;
#pragma unroll
    for (int d = 0; d < D; ++d) coords[(int64_t)d * cap + pos] = y[(int64_t)D * p + d];
}

// Re-gathers the sorted coordinate planes from y (update-free coordinate refresh).
template <int D, class TC>
__global__ __launch_bounds__(256) void k_regather(const int32_t* __restrict__ sorted_idx, const DevStatus* st,
                                                  const TC* __restrict__ y, TC* __restrict__ coords,
                                                  int64_t cap) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= st->n_binned) return;
    const int64_t p = sorted_idx[k];
#pragma unroll
    for (int d = 0; d < D; ++d) coords[(int64_t)d * cap + k] = y[(int64_t)D * p + d];
}

// ------------------------------------------------------------------------------------
// Traversal
// ------------------------------------------------------------------------------------

template <int D>
struct Rows {
    static constexpr int R = D == 1 ? 1 : (D == 2 ? 3 : 9);
};

// Functor contract (fixed menu; arbitrary closures cannot cross the C-ABI):
//   Acc                 per-query accumulator, combined over candidate slices
//   kIndex              stage original neighbour indices in LDS
//   init() / combine()  accumulator identity and merge
//   pair(acc, pd, d2, j, scratch)   called for every accepted pair (d2 <= r^2)
//   write(i, acc)       store for original query index i
struct CountF {
    using Acc = int;
    static constexpr bool kIndex = false;
    int32_t* out;
    int accumulate;
    __device__ Acc init() const { return 0; }
    template <class TR, int D>
    __device__ void pair(Acc& a, const TR (&)[D], TR, int, int*) const { a += 1; }
    __device__ void combine(Acc& a, const Acc& b) const { a += b; }
    __device__ void write(int64_t i, const Acc& a) const { out[i] = accumulate ? out[i] + a : a; }
};

struct IndexSumF {
    using Acc = long long;
    static constexpr bool kIndex = true;
    int64_t* out;
    int accumulate;
    int base;
    __device__ Acc init() const { return 0; }
    template <class TR, int D>
    __device__ void pair(Acc& a, const TR (&)[D], TR, int j, int*) const { a += (long long)j + base; }
    __device__ void combine(Acc& a, const Acc& b) const { a += b; }
    __device__ void write(int64_t i, const Acc& a) const { out[i] = accumulate ? out[i] + a : a; }
};

// CSR fill: accepted neighbours go to row i at positions claimed with an LDS counter
// per query; the rows are sorted afterwards, so the claim order does not matter.
struct FillF {
    using Acc = int;
    static constexpr bool kIndex = true;
    const int32_t* row_ptr;
    int32_t* col;
    int base;
    int row_end;  // set per query by the traversal; guards the row against a stale count
    __device__ Acc init() const { return 0; }
    __device__ int row_begin(int64_t i) const { return row_ptr[i]; }
    template <class TR, int D>
    __device__ void pair(Acc&, const TR (&)[D], TR, int j, int* qcnt) const {
        int slot = atomicAdd(qcnt, 1);
        if (slot < row_end) col[slot] = j + base;
    }
    __device__ void combine(Acc&, const Acc&) const {}
    __device__ void write(int64_t, const Acc&) const {}
};

template <class F>
struct IsFill {
    static constexpr bool value = false;
};
template <>
struct IsFill<FillF> {
    static constexpr bool value = true;
};

// One wavefront (= one workgroup) per non-empty query cell.
//  1. lanes 0..R-1 read the [begin, end) of their stencil row (3 x-adjacent cells)
//  2. candidates are staged in LDS in chunks of CH, coalesced; a sentinel at +Inf pads
//     the chunk to S*L entries so every lane runs the same trip count
//  3. the nq query points of the cell (up to 64 at a time) take S = 64 / nq slices of
//     the chunk each; slice s covers [s*L, s*L + L) with L odd (conflict-free banks)
//  4. slice partials are combined in LDS in slice order (deterministic) and written
//     to the query's original index.
template <int D, class TC, class TR, class F, int CH>
__global__ __launch_bounds__(WAVE) void k_traverse(Geom g, SetView Q, SetView G, TR r2, F f, int nblocks) {
    constexpr int R = Rows<D>::R;
    // padded = S * L with L = ceil(mc / S) | 1 <= mc + 2S - 1 <= CH + 127
    constexpr int CAP = CH + 2 * WAVE;
    __shared__ TC s_c[D][CAP];
    __shared__ int s_j[F::kIndex ? CAP : 1];
    __shared__ typename F::Acc s_red[WAVE];
    __shared__ int s_cnt[IsFill<F>::value ? WAVE : 1];
    __shared__ int s_end[IsFill<F>::value ? WAVE : 1];

    const int lane = threadIdx.x;
    const int w = xcd_remap(blockIdx.x, nblocks);
    if (w >= Q.st->n_nonempty) return;  // uniform over the workgroup (one wavefront)
    const int qc = Q.nonempty[w];
    const int qs = Q.start[qc], qe = Q.start[qc + 1];

    // stencil rows
    int rb = 0, rlen = 0;
    if (lane < R) {
        int off = 0;
        if (D >= 2) off += ((lane % 3) - 1) * g.n[0];
        if (D >= 3) off += ((lane / 3) - 1) * g.n[0] * g.n[1];
        const int base = qc - 1 + off;
        rb = G.start[base];
        rlen = G.start[base + 3] - rb;
    }
    int rbeg[R], pre[R + 1];
    pre[0] = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        rbeg[r] = __builtin_amdgcn_readlane(rb, r);
        pre[r + 1] = pre[r] + __builtin_amdgcn_readlane(rlen, r);
    }
    const int M = pre[R];

    const TC* __restrict__ gc = (const TC*)G.coords;
    const TC* __restrict__ qcrd = (const TC*)Q.coords;

    for (int q0 = qs; q0 < qe; q0 += WAVE) {
        const int nq = min(WAVE, qe - q0);
        const int S = WAVE / nq;
        const int q = lane % nq, s = lane / nq;
        const bool act = s < S;
        TC xi[D];
#pragma unroll
        for (int d = 0; d < D; ++d) xi[d] = act ? qcrd[(int64_t)d * Q.cap + q0 + q] : (TC)0;
        typename F::Acc acc = f.init();
        const int64_t orig_i = act ? Q.sorted_idx[q0 + q] : 0;
        __syncthreads();
        if constexpr (IsFill<F>::value) {
            if (lane < nq) {
                s_cnt[lane] = f.row_begin(orig_i);
                s_end[lane] = f.row_begin(orig_i + 1);
            }
        }
        int* qcnt = nullptr;

        for (int m0 = 0; m0 < M; m0 += CH) {
            const int mc = min(CH, M - m0);
            int L = (mc + S - 1) / S;
            L |= 1;
            const int padded = S * L;
            __syncthreads();
            for (int k = lane; k < padded; k += WAVE) {
                if (k < mc) {
                    const int gk = m0 + k;
                    int r = 0;
#pragma unroll
                    for (int t = 1; t < R; ++t) r += (gk >= pre[t]) ? 1 : 0;
                    int rsel_beg = rbeg[0], rsel_pre = pre[0];
#pragma unroll
                    for (int t = 1; t < R; ++t)
                        if (r == t) {
                            rsel_beg = rbeg[t];
                            rsel_pre = pre[t];
                        }
                    const int src = rsel_beg + (gk - rsel_pre);
#pragma unroll
                    for (int d = 0; d < D; ++d) s_c[d][k] = gc[(int64_t)d * G.cap + src];
                    if constexpr (F::kIndex) s_j[k] = G.sorted_idx[src];
                } else {
#pragma unroll
                    for (int d = 0; d < D; ++d) s_c[d][k] = (TC)INFINITY;
                    if constexpr (F::kIndex) s_j[k] = 0;
                }
            }
            __syncthreads();
            if (act) {
                if constexpr (IsFill<F>::value) {
                    qcnt = &s_cnt[q];
                    f.row_end = s_end[q];
                }
                const int kb = s * L;
#pragma unroll 4
                for (int t = 0; t < L; ++t) {
                    const int k = kb + t;
                    TR pd[D];
#pragma unroll
                    for (int d = 0; d < D; ++d) pd[d] = (TR)(xi[d] - s_c[d][k]);
                    TR d2 = pd[0] * pd[0];
#pragma unroll
                    for (int d = 1; d < D; ++d) {
                        TR p2 = pd[d] * pd[d];
                        d2 = d2 + p2;
                    }
                    if (d2 <= r2) {
                        int j = 0;
                        if constexpr (F::kIndex) j = s_j[k];
                        f.template pair<TR, D>(acc, pd, d2, j, qcnt);
                    }
                }
            }
        }
        if constexpr (!IsFill<F>::value) {
            __syncthreads();
            s_red[lane] = acc;
            __syncthreads();
            if (act && s == 0) {
                for (int t = 1; t < S; ++t) f.combine(acc, s_red[q + t * nq]);
                f.write(orig_i, acc);
            }
        }
    }
}

// Bitonic sort of each CSR row in LDS, one wavefront per row, rows up to SORT_CAP.
constexpr int SORT_CAP = 4096;

__global__ __launch_bounds__(WAVE) void k_sort_rows(const int32_t* __restrict__ row_ptr, int64_t n_rows,
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
__global__ __launch_bounds__(256) void k_list_count(const int32_t* __restrict__ row_ptr, int64_t sel_first,
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
__global__ __launch_bounds__(WAVE) void k_list_index_sum(const int32_t* __restrict__ row_ptr,
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

}  // namespace pn
