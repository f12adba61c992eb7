// Build kernels of the MI355X neighbourhood-search engine (gfx950, wave64):
// initialize!/update! with ParallelUpdate semantics (src/nhs_grid.jl:255-281) as a
// counting sort of the selected points by linear cell index.
//      k_bin       cell key per point + bounds check + wave-aggregated histogram
//      k_scan_*    exclusive scan of the histogram -> cell_start, non-empty cell list
//      k_scatter   point -> provisional slot in its cell segment
//      k_canon     deterministic order inside each cell (original index ascending),
//                  one thread per sorted slot, and the cell-sorted coordinate planes
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pn {

constexpr int WAVE = 64;

// Geometry of one cloud's grid inside a batch: its cells occupy the linear range
// [cell_off, cell_off + n0*n1*n2) of the concatenated cell table.
struct GeomEntry {
    int n[3];
    int cell_off;
    double minc[3];
};

// Grid geometry.  minc and cs are exact in the cell-computation type.  A batch of
// independent clouds (same radius, own boxes) sets `table`: point p belongs to cloud
// cloud_of_point[p], whose cells start at cell_offsets[cloud].
struct Geom {
    int nd;
    int n[3];      // cells per dimension, 1 for unused dimensions
    int ncells;    // total cells (all clouds)
    double minc[3];
    double cs;      // search radius
    double csd[3];  // cell size per dimension: the search radius, or box size / per with a periodic box
    int periodic;   // a PeriodicBox is set (src/nhs_grid.jl:102-125)
    int per[3];     // periodic cells per dimension
    double box[3];  // periodic box size
    const GeomEntry* table;
    const int32_t* cloud_of_point;
    const int32_t* cell_offsets;  // n_clouds + 1
    int n_clouds;
    int hash;       // SpatialHashingCellList: keys are hash buckets of unbounded cells
    int list_size;  // its number of buckets (== ncells)
    unsigned long long list_magic;  // floor((2^64 - 1) / list_size), for the bucket reduction
};

// Cloud of a linear cell index in a batch (binary search over cell_offsets).
__device__ __forceinline__ int cloud_of_cell(const Geom& g, int c) {
    int lo = 0, hi = g.n_clouds;  // cell_offsets[lo] <= c < cell_offsets[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (g.cell_offsets[mid] <= c) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Cells per dimension (x, y) of the grid holding cell c: the batch entry of its cloud, or
// gn0, gn1 (the single grid's, read once from the kernel arguments by the caller).  The
// entry is read as one 8-byte load, so the compiler cannot fold the two sources into one
// load through a private copy of the kernel arguments.
__device__ __forceinline__ void grid_dims_xy(const Geom& g, int c, int gn0, int gn1, int& n0, int& n1) {
    if (g.table) {
        const int2 e = *(const int2*)&g.table[cloud_of_cell(g, c)].n[0];
        n0 = e.x;
        n1 = e.y;
    } else {
        n0 = gn0;
        n1 = gn1;
    }
}

// Device-side status of one binning pass.
struct DevStatus {
    unsigned long long first_bad;  // smallest selection position out of bounds
    unsigned long long total64;    // sum of all counts, 64-bit (overflow guard)
    int max_count;                 // largest cell population
    int n_nonempty;                // number of non-empty cells
    int n_binned;                  // number of binned points (low 32 bits of the sum)
    int n_ordered;                 // work items appended by k_order_work
};

// Host copy of a binning's status in pinned, coherent host memory.  k_scan_top, the last
// kernel of a binning to change the fields the host checks, stores the status there and
// then raises `seq` with a system-scope release, so the host can raise the binning's
// errors while the scatter and the in-cell ordering still run.
struct StatusMailbox {
    DevStatus status;
    unsigned long long seq;
};

// Read-only view of a binned point set used by the traversal.
struct SetView {
    const int32_t* start;       // C + 1 cell offsets
    const int32_t* nonempty;    // non-empty cell ids
    const int32_t* sorted_idx;  // original index per sorted position
    const void* coords;         // D planes of cap entries, cell-sorted
    int64_t cap;
    const DevStatus* st;
    // SpatialHashingCellList only (else NULL), as of the last binning: per non-empty bucket the
    // cell of one of its points (3 int32), and bucket_mixed[b] == bucket_epoch when its points
    // lie in more than one cell (entries of earlier binnings hold earlier epochs)
    const int32_t* bucket_cell;
    const int32_t* bucket_mixed;
    int bucket_epoch;
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

static __global__ void k_reset_status(DevStatus* st) {
    st->first_bad = ~0ull;
    st->total64 = 0ull;
    st->max_count = 0;
    st->n_nonempty = 0;
    st->n_binned = 0;
    st->n_ordered = 0;
}

static __global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// floor_to_int (src/util.jl:19-34) of an already floored value: NaN and values >= 2^63
// saturate to typemax(Int64), values <= -2^63 to typemin(Int64).
template <class T>
__device__ __forceinline__ int64_t floor_to_int64(T f) {
    if (!(f < (T)9223372036854775808.0)) return INT64_MAX;  // NaN too
    if (f <= (T)-9223372036854775808.0) return INT64_MIN;
    return (int64_t)f;
}

// Cell key of point p: (x - min) / cs in TCELL, floor, 1-based cell must lie in
// 2:(n_d - 1) (src/cell_lists/full_grid.jl:84-94, :187-213).  With f = floor(q) the
// condition reads 1 <= f <= n_d - 2; NaN and +-Inf fail both comparisons, matching the
// saturating floor_to_int (src/util.jl:19-34) whose saturated values are out of range.
// With a periodic box the 1-based cell floor_to_int(q) + 1 is wrapped to
// mod(cell - 2, per) + 2 in Int64 arithmetic (src/nhs_grid.jl:609-628), so every
// coordinate, NaN included, lands in 2:per+1; that must still lie in 2:(n_d - 1).
template <int D, class TC, class TCELL>
__device__ __forceinline__ int cell_key(const int (&n)[3], const double (&minc)[3], const double (&cs)[3],
                                        const Geom& g, const TC* __restrict__ xp, bool& ok) {
    int key = 0, stride = 1;
    ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        TCELL q = ((TCELL)xp[d] - (TCELL)minc[d]) / (TCELL)cs[d];
        TCELL f = floor(q);
        int c0;
        bool in;
        if (!g.periodic) {
            in = (f >= (TCELL)1) && (f <= (TCELL)(n[d] - 2));
            c0 = in ? (int)f : 0;
        } else {
            // (floor_to_int(q) + 1) - 2 with Int64 wrap, floored mod per
            const int64_t a = (int64_t)((uint64_t)floor_to_int64(f) - 1ull);
            int64_t r = a % (int64_t)g.per[d];
            if (r < 0) r += g.per[d];
            c0 = (int)r + 1;  // 0-based index of the 1-based cell r + 2
            in = c0 <= n[d] - 2;
            c0 = in ? c0 : 0;
        }
        ok = ok && in;
        key += c0 * stride;
        stride *= n[d];
    }
    return key;
}

// mod(c - 2, n) + 2 in wrapping Int64 arithmetic (periodic_cell_index, src/nhs_grid.jl:609-620).
__device__ __forceinline__ int64_t hash_wrap(int64_t c, int n) {
    int64_t m = (int64_t)((uint64_t)c - 2ull) % (int64_t)n;
    if (m < 0) m += n;
    return m + 2;
}

// Cell of a point under a SpatialHashingCellList: floor_to_int.(x ./ cell_size), no corner
// and no +1 (nonperiodic_cell_coords, src/nhs_grid.jl:636-638).  ok is false when a cell
// coordinate does not fit in Int32 (NaN and infinities included): the reference's
// coordinates_flattened (src/cell_lists/spatial_hashing.jl:170-183) throws InexactError there.
// With a PeriodicBox the cell is wrapped, mod.(c .- 2, n_periodic) .+ 2 (periodic_cell_index,
// src/nhs_grid.jl:609-620), which always fits: no error then, as in the reference.
template <int D, class TC, class TCELL, bool PER>
__device__ __forceinline__ void hash_cell_p(const Geom& g, const TC* __restrict__ xp, int64_t (&c)[3], bool& ok) {
    ok = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) c[d] = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const TCELL q = (TCELL)xp[d] / (TCELL)g.csd[d];
        const TCELL f = floor(q);
        if constexpr (PER) {
            c[d] = hash_wrap(floor_to_int64(f), g.per[d]);
        } else {
            ok = ok && (f >= (TCELL)-2147483648.0) && (f <= (TCELL)2147483647.0);
            c[d] = ok ? (int64_t)f : 0;
        }
    }
}

template <int D, class TC, class TCELL>
__device__ __forceinline__ void hash_cell(const Geom& g, const TC* __restrict__ xp, int64_t (&c)[3], bool& ok) {
    if (g.periodic) hash_cell_p<D, TC, TCELL, true>(g, xp, c, ok);
    else hash_cell_p<D, TC, TCELL, false>(g, xp, c, ok);
}

// spatial_hash (src/cell_lists/spatial_hashing.jl:150-164, Ihmsen et al. 2011) in Julia's
// wrapping Int64 arithmetic, 0-based: mod(xor(i * 73856093, j * 19349663, k * 83492791), L),
// the floored mod of the signed hash.  For a negative hash h, mod(h, L) = L - 1 - (~h mod L)
// (~h = -h - 1 >= 0).  x mod L for x < 2^63 is a Barrett reduction with
// magic = floor((2^64 - 1) / L): q = mulhi(x, magic) > x / L - 3/2, so r = x - q L lies in
// [0, 3L/2) and one conditional subtraction ends it (a 64-bit division is a long software
// routine here).
template <int D>
__device__ __forceinline__ int spatial_hash(const int64_t (&c)[3], const Geom& g) {
    uint64_t h = (uint64_t)c[0] * 73856093ull;
    if (D >= 2) h ^= (uint64_t)c[1] * 19349663ull;
    if (D >= 3) h ^= (uint64_t)c[2] * 83492791ull;
    const uint64_t L = (uint64_t)g.list_size;
    const bool neg = (int64_t)h < 0;
    const uint64_t x = neg ? ~h : h;
    uint64_t r = x - __umul64hi(x, g.list_magic) * L;
    r = r >= L ? r - L : r;
    return (int)(neg ? L - 1 - r : r);
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
        int k;
        if (g.hash) {
            int64_t c[3];
            hash_cell<D, TC, TCELL>(g, y + (int64_t)D * p, c, ok);
            k = spatial_hash<D>(c, g);
        } else if (g.table) {
            const GeomEntry e = g.table[g.cloud_of_point[p]];
            k = e.cell_off + cell_key<D, TC, TCELL>(e.n, e.minc, g.csd, g, y + (int64_t)D * p, ok);
        } else {
            k = cell_key<D, TC, TCELL>(g.n, g.minc, g.csd, g, y + (int64_t)D * p, ok);
        }
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

static __global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const int32_t* __restrict__ counts, int64_t C,
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
    // block max of the populations; block sums and maxima go to k_scan_top (no same-address
    // atomics: with millions of cells or buckets they serialised the pass)
    __shared__ int shm[SCAN_THREADS / WAVE];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0) shm[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = shm[0];
#pragma unroll
        for (int i = 1; i < SCAN_THREADS / WAVE; ++i) m = max(m, shm[i]);
        block_sums[blockIdx.x] = total;
        block_sums[gridDim.x + blockIdx.x] = (unsigned long long)(uint32_t)m;
    }
}

// Single workgroup: exclusive scan of the block sums (block_sums[0, nblocks)), writes the
// totals, the 64-bit sum of all counts and the largest count (block maxima at
// block_sums[nblocks, 2 nblocks)).
static __global__ __launch_bounds__(SCAN_THREADS) void k_scan_top(unsigned long long* __restrict__ block_sums,
                                                           int64_t nblocks, int32_t* __restrict__ start,
                                                           int64_t C, DevStatus* st, StatusMailbox* mb,
                                                           unsigned long long seq) {
    __shared__ unsigned long long sh[SCAN_THREADS / WAVE];
    __shared__ unsigned long long sh_sum[SCAN_THREADS / WAVE];
    __shared__ int sh_max[SCAN_THREADS / WAVE];
    unsigned long long carry = 0, sum64 = 0;
    int mx = 0;
    for (int64_t b0 = 0; b0 < nblocks; b0 += SCAN_THREADS) {
        int64_t b = b0 + threadIdx.x;
        unsigned long long v = b < nblocks ? block_sums[b] : 0ull;
        if (b < nblocks) {
            sum64 += v & 0xffffffffull;
            mx = max(mx, (int)(uint32_t)block_sums[nblocks + b]);
        }
        unsigned long long total;
        unsigned long long ex = block_excl_scan(v, sh, total);
        if (b < nblocks) block_sums[b] = carry + ex;
        carry += total;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sum64 += __shfl_xor(sum64, o, WAVE);
        mx = max(mx, __shfl_xor(mx, o, WAVE));
    }
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        sh_sum[threadIdx.x >> 6] = sum64;
        sh_max[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        int m = 0;
#pragma unroll
        for (int i = 0; i < SCAN_THREADS / WAVE; ++i) {
            tot += sh_sum[i];
            m = max(m, sh_max[i]);
        }
        st->total64 = tot;
        st->max_count = m;
        start[C] = (int32_t)(uint32_t)(carry & 0xffffffffull);
        st->n_binned = (int)(uint32_t)(carry & 0xffffffffull);
        st->n_nonempty = (int)(uint32_t)(carry >> 32);
        if (mb) {
            // first_bad comes from k_bin, an earlier kernel on the same stream
            mb->status.first_bad = st->first_bad;
            mb->status.total64 = tot;
            mb->status.max_count = m;
            mb->status.n_nonempty = (int)(uint32_t)(carry >> 32);
            mb->status.n_binned = (int)(uint32_t)(carry & 0xffffffffull);
            mb->status.n_ordered = 0;
            __hip_atomic_store(&mb->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

static __global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(const int32_t* counts, int64_t C,
                                                            const unsigned long long* __restrict__ block_sums,
                                                            int32_t* __restrict__ start,
                                                            int32_t* __restrict__ nonempty,
                                                            int32_t* clear) {
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
    // clear = counts: the counts are read once, here, and left zero for the next binning
    if (clear) {
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; ++i)
            if (base + i < C) clear[base + i] = 0;
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

// Provisional placement: selection position a goes to slot start[key] + slots[a] of its
// cell segment (atomic order).  The slot holds a, not the point index, so k_canon can
// recover the key and break ties between repeated selections.  With reset = slots, each
// entry is set to -1 once read (the hash build reuses the array, cleared, for its bucket
// representatives).
static __global__ __launch_bounds__(256) void k_scatter(const int32_t* __restrict__ keys, const int32_t* slots,
                                                 int64_t n_sel, const int32_t* __restrict__ start,
                                                 int32_t* __restrict__ perm_tmp, int32_t* reset) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_sel) return;
    const int key = keys[a];
    const int slot = key < 0 ? 0 : slots[a];
    if (reset) reset[a] = -1;
    if (key < 0) return;
    perm_tmp[start[key] + slot] = (int32_t)a;
}

// Cells with more points than this keep the provisional (atomic) order instead of the
// O(n^2) rank: their neighbour sets are unaffected, only the order inside the cell.
constexpr int CANON_RANK_LIMIT = 1024;

// Output-major: one thread per sorted slot k of the binned points.  The final position of
// the point in slot k is its cell's start plus its rank in the cell segment by (original
// index, selection position) ascending (canonical = 1: the SerialUpdate insertion order,
// and repeated selections of one point get distinct consecutive slots), or k itself
// (canonical = 0: the order of the reference's ParallelUpdate, nondeterministic inside a
// cell; neighbour sets are the same).  A wavefront's stores land inside the cell segments
// of its 64 slots, so the permutation and the cell-sorted coordinate planes (plane d at
// coords + d * cap) are written in whole lines.
// Slot k's canonical position pos, its selection position a, point p, cell key and the first
// slot s of the cell segment.
struct CanonSlot {
    int pos, key, s;
    int32_t a, p;
};
__device__ __forceinline__ CanonSlot canon_slot(const int32_t* __restrict__ keys, int64_t sel_first,
                                                const int32_t* __restrict__ sel_list,
                                                const int32_t* __restrict__ start,
                                                const int32_t* __restrict__ perm_tmp, int64_t k, int canonical) {
    CanonSlot c;
    c.a = perm_tmp[k];
    c.key = keys[c.a];
    c.p = (int32_t)sel_point(c.a, sel_first, sel_list);
    c.s = start[c.key];
    const int e = start[c.key + 1];
    c.pos = (int)k;
    if (canonical && e - c.s <= CANON_RANK_LIMIT) {
        int rank = 0;
        if (!sel_list) {
            // original index = sel_first + selection position: rank by position
            for (int q = c.s; q < e; ++q) rank += perm_tmp[q] < c.a;
        } else {
            for (int q = c.s; q < e; ++q) {
                const int32_t aq = perm_tmp[q];
                const int32_t pq = sel_list[aq];
                rank += (pq < c.p) || (pq == c.p && aq < c.a);
            }
        }
        c.pos = c.s + rank;
    }
    return c;
}

template <int D, class TC>
__global__ __launch_bounds__(256) void k_canon(const int32_t* __restrict__ keys, int64_t sel_first,
                                               const int32_t* __restrict__ sel_list, const DevStatus* st,
                                               const int32_t* __restrict__ start,
                                               const int32_t* __restrict__ perm_tmp,
                                               const TC* __restrict__ y, int32_t* __restrict__ sorted_idx,
                                               TC* __restrict__ coords, int64_t cap, int canonical) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= st->n_binned) return;
    const CanonSlot c = canon_slot(keys, sel_first, sel_list, start, perm_tmp, k, canonical);
    sorted_idx[c.pos] = c.p;
#pragma unroll
    for (int d = 0; d < D; ++d) coords[(int64_t)d * cap + c.pos] = y[(int64_t)D * c.p + d];
}

// k_canon for a SpatialHashingCellList, which also records the bucket data the cell kernel
// needs (the reference's collision flags, src/cell_lists/spatial_hashing.jl:84-116, computed
// once per binning instead of per push): the point in the segment's first provisional slot s
// gives the bucket's cell (written by slot s itself), and a point whose cell differs from it
// marks the bucket as mixed (bucket_mixed[key] = epoch; every writer stores the same value).
// Slot s also names its selection position a as the bucket's representative: rep[a] = key
// (rep holds -1 elsewhere).
template <int D, class TC, class TCELL>
__global__ __launch_bounds__(256) void k_canon_hash(const int32_t* __restrict__ keys, int64_t sel_first,
                                                    const int32_t* __restrict__ sel_list, const DevStatus* st,
                                                    const int32_t* __restrict__ start,
                                                    const int32_t* __restrict__ perm_tmp,
                                                    const TC* __restrict__ y, int32_t* __restrict__ sorted_idx,
                                                    TC* __restrict__ coords, int64_t cap, int canonical, Geom g,
                                                    int32_t* __restrict__ bucket_cell,
                                                    int32_t* __restrict__ bucket_mixed, int epoch,
                                                    int32_t* __restrict__ rep) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= st->n_binned) return;
    const CanonSlot c = canon_slot(keys, sel_first, sel_list, start, perm_tmp, k, canonical);
    TC x[D], xs[D];
    const int64_t ps = sel_point(perm_tmp[c.s], sel_first, sel_list);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        x[d] = y[(int64_t)D * c.p + d];
        xs[d] = y[(int64_t)D * ps + d];
    }
    sorted_idx[c.pos] = c.p;
#pragma unroll
    for (int d = 0; d < D; ++d) coords[(int64_t)d * cap + c.pos] = x[d];
    int64_t ck[3], cs[3];
    bool ok;
    hash_cell<D, TC, TCELL>(g, x, ck, ok);
    hash_cell<D, TC, TCELL>(g, xs, cs, ok);
    if (k == c.s) {
#pragma unroll
        for (int d = 0; d < 3; ++d) bucket_cell[3 * (int64_t)c.key + d] = (int32_t)cs[d];
        rep[c.a] = c.key;  // the bucket's representative selection position (for k_order_work)
    }
    bool same = true;
#pragma unroll
    for (int d = 0; d < D; ++d) same = same && ck[d] == cs[d];
    if (!same) bucket_mixed[c.key] = epoch;
}

// SpatialHashingCellList work order: overwrites the non-empty bucket list with the buckets in
// the order of their representatives' selection positions (rep[a] >= 0, from k_canon_hash),
// roughly: a workgroup compacts ORDER_ITEMS positions per thread in order and appends them with
// one atomic, so the order across workgroups is their completion order.  Buckets sit at hash
// positions, so bucket order scatters spatially adjacent work items over the XCDs; the
// selection order follows the input's order, which is spatially coherent in every benchmark
// (cell-sorted clouds), and keeps a neighbourhood's buckets in one XCD's L2.  Every work item
// writes only its own queries' results, so results do not depend on the order.
constexpr int ORDER_THREADS = 1024, ORDER_ITEMS = 4;
static __global__ __launch_bounds__(ORDER_THREADS) void k_order_work(const int32_t* __restrict__ rep, int64_t n_sel,
                                                                      DevStatus* st,
                                                                      int32_t* __restrict__ nonempty) {
    __shared__ int s_wave[ORDER_THREADS / WAVE];
    __shared__ int s_base;
    const int64_t a0 = ((int64_t)blockIdx.x * ORDER_THREADS + threadIdx.x) * ORDER_ITEMS;
    int key[ORDER_ITEMS];
    int n = 0;
#pragma unroll
    for (int i = 0; i < ORDER_ITEMS; ++i) {
        const int64_t a = a0 + i;
        key[i] = a < n_sel ? rep[a] : -1;
        n += key[i] >= 0;
    }
    // exclusive scan of n over the workgroup
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    int inc = n;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const int t = __shfl_up(inc, o, WAVE);
        if (lane >= o) inc += t;
    }
    if (lane == WAVE - 1) s_wave[wv] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < ORDER_THREADS / WAVE; ++w) {
            const int v = s_wave[w];
            s_wave[w] = tot;
            tot += v;
        }
        s_base = tot > 0 ? atomicAdd(&st->n_ordered, tot) : 0;
    }
    __syncthreads();
    int pos = s_base + s_wave[wv] + inc - n;
#pragma unroll
    for (int i = 0; i < ORDER_ITEMS; ++i)
        if (key[i] >= 0) nonempty[pos++] = key[i];
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

// Synthetic per-step displacement U[-a, a] of every coordinate, a pure function of
// (seed, step, index): SplitMix64 of a counter, top 53 bits.  The step counter lives on
// the device so a captured graph advances it on every replay (the perturbation of
// benchmarks/update.jl:28-49, made replayable).
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <class T>
__global__ __launch_bounds__(256) void k_displace(T* __restrict__ x, int64_t n, T a, unsigned long long seed,
                                                  const int64_t* __restrict__ step) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long key = splitmix64(seed ^ splitmix64((unsigned long long)*step));
    const double u = (double)(splitmix64(key + (unsigned long long)i) >> 11) * (1.0 / 9007199254740992.0);
    x[i] += (T)((2.0 * u - 1.0) * (double)a);
}

static __global__ void k_step_increment(int64_t* step) { *step += 1; }

}  // namespace pn
