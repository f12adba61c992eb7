// Neighbour traversal (foreach_point_neighbor, hot loop src/nhs_grid.jl:519-575) on
// gfx950: one wavefront per non-empty query cell.
//
// The 3^D cells around a cell are 3^(D-1) x-rows of three adjacent cells each, and an
// x-row is one contiguous range of the cell-sorted order (linear indices are first-
// dimension fastest).  The wavefront stages those ranges in LDS with coalesced loads and
// every query point of the cell tests every staged candidate with the reference's
// predicate: p = convert(T_r, x_i - y_j), d2 = ((p1*p1 + p2*p2) + p3*p3), d2 <= r*r
// (src/nhs_grid.jl:547-555), evaluated without FMA contraction (built -ffp-contract=off).
// The candidate order equals the reference's (CartesianIndices, first dimension
// fastest; inside a cell ascending original index).
#pragma once

#include "pn_build.h"

namespace pn {

template <int D>
struct Rows {
    static constexpr int R = D == 1 ? 1 : (D == 2 ? 3 : 9);
};

// Candidate coordinates in LDS, one wide read per candidate (3D padded to 4).
template <class TC, int D>
struct alignas(sizeof(TC) * (D == 3 ? 4 : D)) Pack {
    TC v[D == 3 ? 4 : D];
};

struct Empty {};

template <class TR, int D>
struct Vec {
    TR v[D];
};

// Functor contract (a fixed menu: arbitrary closures cannot cross the C-ABI):
//   Acc, QState              per-query accumulator and per-query inputs
//   kIndex                   stage original neighbour indices
//   kFields                  per-neighbour TR fields staged from `fsorted`, planes of
//                            `fstride` entries in the grid's sorted order
//   query(i)                 per-query inputs of original query index i
//   init()/combine(a, b)     accumulator identity and merge (slices merged in order)
//   pair(acc, qs, pd, d2, j, fj, scratch)   every accepted pair (d2 <= r^2)
//   write(i, qs, acc)        store for original query index i
//   kQueue                   accepted candidates are queued per lane and the pair body
//                            runs in lockstep over the queues (for bodies much costlier
//                            than the distance test; ~1 in 6 candidates is accepted)
//   kLight                   counting only (no pair body)
//   kCache                   forward pair sum whose pair lists may be recorded once per build
//                            and replayed by later traversals (the adjoints recompute)
struct FunctorBase {
    static constexpr bool kIndex = false;
    static constexpr int kFields = 0;
    static constexpr bool kFill = false;
    static constexpr bool kQueue = false;  // heavy pair body: run it on queued pairs only
    static constexpr bool kLight = false;  // counting only: no pair body
    static constexpr bool kCache = false;  // forward pair sum: may replay a recorded pair list
    using QState = Empty;
    const void* fsorted = nullptr;
    int64_t fstride = 0;
    __device__ QState query(int64_t) const { return QState{}; }
};

struct CountF : FunctorBase {
    using Acc = int;
    static constexpr bool kLight = true;
    int32_t* out;
    int accumulate;
    __device__ Acc init() const { return 0; }
    template <class TR, int D>
    __device__ void pair(Acc& a, const QState&, const TR (&)[D], TR, int, const TR*, int*) const { a += 1; }
    __device__ void combine(Acc& a, const Acc& b) const { a += b; }
    __device__ void write(int64_t i, const QState&, const Acc& a) const { out[i] = accumulate ? out[i] + a : a; }
};

struct IndexSumF : FunctorBase {
    using Acc = long long;
    static constexpr bool kIndex = true;
    int64_t* out;
    int accumulate;
    int base;
    __device__ Acc init() const { return 0; }
    template <class TR, int D>
    __device__ void pair(Acc& a, const QState&, const TR (&)[D], TR, int j, const TR*, int*) const {
        a += (long long)j + base;
    }
    __device__ void combine(Acc& a, const Acc& b) const { a += b; }
    __device__ void write(int64_t i, const QState&, const Acc& a) const { out[i] = accumulate ? out[i] + a : a; }
};

// CSR fill: accepted neighbours go to row i at positions claimed with an LDS counter per
// query; the rows are sorted afterwards, so the claim order does not matter.
struct FillF : FunctorBase {
    using Acc = int;
    static constexpr bool kIndex = true;
    static constexpr bool kFill = true;
    const int32_t* row_ptr;
    int32_t* col;
    int base;
    int row_end;  // set per query by the traversal; guards the row against a stale count
    struct QState {
        int rb, re;  // row i of the CSR arrays: [rb, re)
    };
    __device__ QState query(int64_t i) const { return QState{row_ptr[i], row_ptr[i + 1]}; }
    __device__ Acc init() const { return 0; }
    __device__ int row_begin(int64_t i) const { return row_ptr[i]; }
    template <class TR, int D>
    __device__ void pair(Acc&, const QState&, const TR (&)[D], TR, int j, const TR*, int* qcnt) const {
        int slot = atomicAdd(qcnt, 1);
        if (slot < row_end) col[slot] = j + base;
    }
    __device__ void combine(Acc&, const Acc&) const {}
    __device__ void write(int64_t, const QState&, const Acc&) const {}
};

// compute_periodic_distance (src/neighborhood_search.jl:428-436) for a hash grid with a
// PeriodicBox, in the operations of the cell kernel's periodic_fix: when d2 > r^2 and some
// |pos_diff_d| > 0.49 L_d, pos_diff -= L .* rint.(pos_diff ./ L) and d2 is recomputed
// (otherwise every rint is 0 and nothing changes).
template <int D, class TR>
__device__ __forceinline__ TR hash_min_image(const Geom& g, TR (&pd)[D], TR d2, TR r2) {
    if (!(d2 > r2)) return d2;
    bool need = false;
#pragma unroll
    for (int d = 0; d < D; ++d) need = need || (fabs(pd[d]) > (TR)0.49 * (TR)g.box[d]);
    if (!need) return d2;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const TR box = (TR)g.box[d];
        const TR k = rint(pd[d] / box);
        const TR sft = box * k;
        pd[d] = pd[d] - sft;
    }
    d2 = pd[0] * pd[0];
#pragma unroll
    for (int d = 1; d < D; ++d) {
        const TR p2 = pd[d] * pd[d];
        d2 = d2 + p2;
    }
    return d2;
}

// One query of a grid keyed by a SpatialHashingCellList (src/nhs_grid.jl:519-575 with the
// collision checks of :479-510), scanned by one thread.  The 3^D neighbouring cells are
// visited in CartesianIndices order (first dimension fastest); each one's bucket is scanned
// and an accepted candidate counts only if its own cell is that neighbouring cell.  That is
// the reference's check_collision; it applies the check only to buckets that
// check_cell_collision flags, and for every other bucket it holds anyway.  So each
// neighbour is visited exactly once, in the reference's cell order.
template <int D, class TC, class TR, class TCELL, class F, bool PER>
__device__ __forceinline__ void hash_query(const Geom& g, const SetView& Q, const SetView& G, TR r2, F& f,
                                           int64_t k, int* cnt_slot) {
    constexpr int NF = F::kFields;
    constexpr int NS = D == 1 ? 3 : (D == 2 ? 9 : 27);
    const TC* __restrict__ gc = (const TC*)G.coords;
    const TC* __restrict__ qcrd = (const TC*)Q.coords;
    const TR* __restrict__ fs = (const TR*)f.fsorted;
    TC xi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = qcrd[(int64_t)d * Q.cap + k];
    const int64_t orig_i = Q.sorted_idx[k];
    int64_t c[3];
    bool ok;
    hash_cell_p<D, TC, TCELL, PER>(g, xi, c, ok);  // binning rejected every point with !ok
    const typename F::QState qst = f.query(orig_i);
    typename F::Acc acc = f.init();
    int* qcnt = nullptr;
    if constexpr (F::kFill) {
        *cnt_slot = f.row_begin(orig_i);
        f.row_end = f.row_begin(orig_i + 1);
        qcnt = cnt_slot;
    }
    for (int o = 0; o < NS; ++o) {
        int64_t nc[3] = {c[0], c[1], c[2]};
        int t = o;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            nc[d] = c[d] + (t % 3) - 1;
            if constexpr (PER) nc[d] = hash_wrap(nc[d], g.per[d]);
            t /= 3;
        }
        const int b = spatial_hash<D>(nc, g);
        const int s1 = G.start[b + 1];
        for (int src = G.start[b]; src < s1; ++src) {
            TC xj[D];
            TR pd[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                xj[d] = gc[(int64_t)d * G.cap + src];
                pd[d] = (TR)(xi[d] - xj[d]);
            }
            TR d2 = pd[0] * pd[0];
#pragma unroll
            for (int d = 1; d < D; ++d) {
                TR p2 = pd[d] * pd[d];
                d2 = d2 + p2;
            }
            if constexpr (PER) d2 = hash_min_image<D, TR>(g, pd, d2, r2);
            if (!(d2 <= r2)) continue;
            int64_t cj[3];
            bool okj;
            hash_cell_p<D, TC, TCELL, PER>(g, xj, cj, okj);
            bool same = true;
#pragma unroll
            for (int d = 0; d < D; ++d) same = same && cj[d] == nc[d];
            if (!same) continue;
            int j = 0;
            if constexpr (F::kIndex) j = G.sorted_idx[src];
            TR fj[NF > 0 ? NF : 1];
            if constexpr (NF > 0) {
#pragma unroll
                for (int fi = 0; fi < NF; ++fi) fj[fi] = fs[(int64_t)fi * f.fstride + src];
            }
            f.template pair<TR, D>(acc, qst, pd, d2, j, fj, qcnt);
        }
    }
    if constexpr (!F::kFill) f.write(orig_i, qst, acc);
}

// One wavefront (= one workgroup) per non-empty query cell at a time; a persistent
// grid of workgroups walks the non-empty cells (count read on the device, so launch
// shapes do not depend on the data and can be captured in a graph).
//  1. lanes 0..R-1 read the [begin, end) of their stencil row (3 x-adjacent cells)
//  2. candidates are staged in LDS in chunks of CH, coalesced; a sentinel at +Inf pads
//     the chunk to S*L entries so every lane runs the same trip count
//  3. the nq query points of the cell (up to 64 at a time) take S = 64 / nq slices of
//     the chunk each; slice s covers [s*L, s*L + L) with L odd (conflict-free banks)
//  4. slice partials are combined in LDS in slice order (deterministic) and written
//     to the query's original index.
//
// HASH: the grid is keyed by a SpatialHashingCellList and a work item is a non-empty query
// bucket.  When all its queries share one cell, the pieces are that cell's 3^D neighbouring
// cells' buckets, and a staged candidate whose own cell is not its piece's cell becomes a
// sentinel (hash_query's check, applied once per staged candidate).  A bucket whose queries
// lie in several cells (a hash collision) is scanned query by query with hash_query.
template <int D, class TC, class TR, class F, int CH, bool HASH = false, class TCELL = TR, bool HPER = false>
__global__ __launch_bounds__(WAVE) void k_traverse(Geom g, SetView Q, SetView G, TR r2, F f, int nblocks) {
    constexpr int R = HASH ? (D == 1 ? 3 : (D == 2 ? 9 : 27)) : Rows<D>::R;
    constexpr int NF = F::kFields;
    // A chunk holds at most cmax = CAP - 2S + 1 candidates, so the padded length
    // S * L with L = ceil(mc / S) | 1 <= mc + 2S - 1 never exceeds CAP.
    constexpr int CAP = CH + WAVE;
    __shared__ Pack<TC, D> s_c[CAP];
    __shared__ int s_j[F::kIndex ? CAP : 1];
    __shared__ TR s_f[NF > 0 ? NF : 1][NF > 0 ? CAP : 1];
    __shared__ typename F::Acc s_red[WAVE];
    __shared__ int s_cnt[F::kFill ? WAVE : 1];
    __shared__ int s_end[F::kFill ? WAVE : 1];
    // per-lane queues of accepted candidates (LDS indices < CAP < 2^16), lane-minor
    constexpr int QCAP = 16;
    __shared__ uint16_t s_q[F::kQueue ? QCAP : 1][F::kQueue ? WAVE : 1];

    const int lane = threadIdx.x;
    // Persistent grid: XCD x (= blockIdx % 8) walks a contiguous range of the non-empty
    // cells, so the cells that share stencil rows run on one XCD's L2.
    const int n_work = Q.st->n_nonempty;
    const int gn0 = g.n[0], gn1 = g.n[1];
    const int xcd = blockIdx.x & 7, per_xcd = nblocks >> 3;
    const int chunk = (n_work + 7) >> 3;
    const int w_end = min(n_work, (xcd + 1) * chunk);
    for (int w = xcd * chunk + (int)(blockIdx.x >> 3); w < w_end; w += per_xcd) {
    const int qc = Q.nonempty[w];
    const int qs = Q.start[qc], qe = Q.start[qc + 1];
    int64_t qcell[3] = {0, 0, 0};
    bool uniform = true;
    if constexpr (HASH) {
        const TC* __restrict__ qx = (const TC*)Q.coords;
        TC x0[D];
#pragma unroll
        for (int d = 0; d < D; ++d) x0[d] = qx[(int64_t)d * Q.cap + qs];
        bool ok;
        hash_cell_p<D, TC, TCELL, HPER>(g, x0, qcell, ok);
        for (int k = qs + lane; k < qe; k += WAVE) {
            TC xk[D];
#pragma unroll
            for (int d = 0; d < D; ++d) xk[d] = qx[(int64_t)d * Q.cap + k];
            int64_t ck[3];
            hash_cell_p<D, TC, TCELL, HPER>(g, xk, ck, ok);
#pragma unroll
            for (int d = 0; d < D; ++d) uniform = uniform && ck[d] == qcell[d];
        }
        uniform = !__any(!uniform);
    }
    if (HASH && !uniform) {
        if constexpr (HASH) {
            for (int k = qs + lane; k < qe; k += WAVE)
                hash_query<D, TC, TR, TCELL, F, HPER>(g, Q, G, r2, f, (int64_t)k, &s_cnt[F::kFill ? lane : 0]);
        }
    } else {
    int n0, n1;
    grid_dims_xy(g, qc, gn0, gn1, n0, n1);

    // stencil rows
    int rb = 0, rlen = 0;
    if (lane < R) {
        if constexpr (HASH) {
            int64_t nc[3] = {qcell[0], qcell[1], qcell[2]};
            int t = lane;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                nc[d] = qcell[d] + (t % 3) - 1;
                if constexpr (HPER) nc[d] = hash_wrap(nc[d], g.per[d]);
                t /= 3;
            }
            const int b = spatial_hash<D>(nc, g);
            rb = G.start[b];
            rlen = G.start[b + 1] - rb;
        } else {
            int off = 0;
            if (D >= 2) off += ((lane % 3) - 1) * n0;
            if (D >= 3) off += ((lane / 3) - 1) * n0 * n1;
            const int base = qc - 1 + off;
            rb = G.start[base];
            rlen = G.start[base + 3] - rb;
        }
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
    const TR* __restrict__ fs = (const TR*)f.fsorted;

    for (int q0 = qs; q0 < qe; q0 += WAVE) {
        const int nq = min(WAVE, qe - q0);
        const int S = WAVE / nq;
        const int q = lane % nq, s = lane / nq;
        const bool act = s < S;
        TC xi[D];
#pragma unroll
        for (int d = 0; d < D; ++d) xi[d] = act ? qcrd[(int64_t)d * Q.cap + q0 + q] : (TC)0;
        const int64_t orig_i = act ? Q.sorted_idx[q0 + q] : 0;
        typename F::QState qst{};
        if (act) qst = f.query(orig_i);
        typename F::Acc acc = f.init();
        __syncthreads();
        if constexpr (F::kFill) {
            if (lane < nq) {
                s_cnt[lane] = f.row_begin(orig_i);
                s_end[lane] = f.row_begin(orig_i + 1);
            }
        }
        int* qcnt = nullptr;

        const int cmax = CAP - 2 * S + 1;
        for (int m0 = 0; m0 < M; m0 += cmax) {
            const int mc = min(cmax, M - m0);
            int L = (mc + S - 1) / S;
            L |= 1;
            const int padded = S * L;
            __syncthreads();
            for (int k = lane; k < padded; k += WAVE) {
                Pack<TC, D> pk;
                if (k < mc) {
                    const int gk = m0 + k;
                    int rsel_beg = rbeg[0], rsel_pre = pre[0], rsel_t = 0;
#pragma unroll
                    for (int t = 1; t < R; ++t)
                        if (gk >= pre[t]) {
                            rsel_beg = rbeg[t];
                            rsel_pre = pre[t];
                            rsel_t = t;
                        }
                    const int src = rsel_beg + (gk - rsel_pre);
#pragma unroll
                    for (int d = 0; d < D; ++d) pk.v[d] = gc[(int64_t)d * G.cap + src];
                    if constexpr (HASH) {
                        // a candidate of another cell in this bucket never passes the test
                        int64_t cj[3];
                        bool okj;
                        hash_cell_p<D, TC, TCELL, HPER>(g, pk.v, cj, okj);
                        bool same = true;
                        int t = rsel_t;
#pragma unroll
                        for (int d = 0; d < D; ++d) {
                            int64_t want = qcell[d] + (t % 3) - 1;
                            if constexpr (HPER) want = hash_wrap(want, g.per[d]);
                            same = same && cj[d] == want;
                            t /= 3;
                        }
                        if (!same) {
#pragma unroll
                            for (int d = 0; d < D; ++d) pk.v[d] = (TC)INFINITY;
                        }
                    }
                    if constexpr (F::kIndex) s_j[k] = G.sorted_idx[src];
                    if constexpr (NF > 0) {
#pragma unroll
                        for (int fi = 0; fi < NF; ++fi) s_f[fi][k] = fs[(int64_t)fi * f.fstride + src];
                    }
                } else {
#pragma unroll
                    for (int d = 0; d < D; ++d) pk.v[d] = (TC)INFINITY;
                    if constexpr (F::kIndex) s_j[k] = 0;
                    if constexpr (NF > 0) {
#pragma unroll
                        for (int fi = 0; fi < NF; ++fi) s_f[fi][k] = (TR)0;
                    }
                }
                if constexpr (D == 3) pk.v[3] = (TC)0;
                s_c[k] = pk;
            }
            __syncthreads();
            if (act) {
                if constexpr (F::kFill) {
                    qcnt = &s_cnt[q];
                    f.row_end = s_end[q];
                }
                const int kb = s * L;
                // d2 of the query against LDS candidate k, in the reference's arithmetic
                auto test = [&](int k, TR (&pd)[D]) -> TR {
                    const Pack<TC, D> pk = s_c[k];
#pragma unroll
                    for (int d = 0; d < D; ++d) pd[d] = (TR)(xi[d] - pk.v[d]);
                    TR d2 = pd[0] * pd[0];
#pragma unroll
                    for (int d = 1; d < D; ++d) {
                        TR p2 = pd[d] * pd[d];
                        d2 = d2 + p2;
                    }
                    if constexpr (HASH && HPER) d2 = hash_min_image<D, TR>(g, pd, d2, r2);
                    return d2;
                };
                auto body = [&](int k, const TR (&pd)[D], TR d2) {
                    int j = 0;
                    if constexpr (F::kIndex) j = s_j[k];
                    TR fj[NF > 0 ? NF : 1];
                    if constexpr (NF > 0) {
#pragma unroll
                        for (int fi = 0; fi < NF; ++fi) fj[fi] = s_f[fi][k];
                    }
                    f.template pair<TR, D>(acc, qst, pd, d2, j, fj, qcnt);
                };
                if constexpr (!F::kQueue) {
#pragma unroll 4
                    for (int t = 0; t < L; ++t) {
                        TR pd[D];
                        const TR d2 = test(kb + t, pd);
                        if (d2 <= r2) body(kb + t, pd, d2);
                    }
                } else {
                    // The slot at `cnt` is written unconditionally and kept only when the
                    // candidate is accepted: the test loop stays branch-free.
                    // queue slot i of this lane is s_q[i][lane]; `cnt` counts the kept slots
                    int cnt = 0;
                    auto flush = [&]() {
                        // two entries per round: their LDS reads overlap
                        for (int i = 0; __any(i < cnt); i += 2) {
                            const bool h0 = i < cnt, h1 = i + 1 < cnt;
                            const int k0 = h0 ? (int)s_q[i][lane] : 0;
                            const int k1 = h1 ? (int)s_q[i + 1][lane] : 0;
                            TR pd0[D], pd1[D];
                            const TR d20 = test(k0, pd0);
                            const TR d21 = test(k1, pd1);
                            if (h0) body(k0, pd0, d20);
                            if (h1) body(k1, pd1, d21);
                        }
                        cnt = 0;
                    };
                    int t = 0;
                    for (; t + 4 <= L; t += 4) {
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            TR pd[D];
                            const TR d2 = test(kb + t + u, pd);
                            s_q[cnt][lane] = (uint16_t)(kb + t + u);
                            cnt += (d2 <= r2) ? 1 : 0;
                        }
                        if (__any(cnt > QCAP - 4)) flush();
                    }
                    for (; t < L; ++t) {
                        TR pd[D];
                        const TR d2 = test(kb + t, pd);
                        s_q[cnt][lane] = (uint16_t)(kb + t);
                        cnt += (d2 <= r2) ? 1 : 0;
                    }
                    flush();
                }
            }
        }
        if constexpr (!F::kFill) {
            __syncthreads();
            s_red[lane] = acc;
            __syncthreads();
            if (act && s == 0) {
                for (int t = 1; t < S; ++t) f.combine(acc, s_red[q + t * nq]);
                f.write(orig_i, qst, acc);
            }
        }
    }
    }  // uniform work item
    __syncthreads();
    }  // persistent loop over cells
}

// Gathers per-neighbour fields into the grid's sorted order: out[f * stride + k] =
// src_f[sorted_idx[k] * step_f + off_f] (a field may be one component of a D x N array).
constexpr int MAX_FIELDS = 7;

template <class TR>
struct FieldSrc {
    const TR* p[MAX_FIELDS];
    int step[MAX_FIELDS];
    int n;
    unsigned recip;  // bit f set: field f is gathered as 1 / value (an IEEE division)
};

template <class TR>
__global__ __launch_bounds__(256) void k_gather_fields(const int32_t* __restrict__ sorted_idx, const DevStatus* st,
                                                       FieldSrc<TR> src, TR* __restrict__ out, int64_t stride) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= st->n_binned) return;
    const int64_t p = sorted_idx[k];
#pragma unroll
    for (int f = 0; f < MAX_FIELDS; ++f) {
        if (f < src.n) {
            const TR v = src.p[f][p * src.step[f]];
            out[(int64_t)f * stride + k] = ((src.recip >> f) & 1u) ? (TR)1 / v : v;
        }
    }
}

}  // namespace pn
