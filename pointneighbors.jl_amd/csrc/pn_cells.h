// Candidate-major neighbour traversal (foreach_point_neighbor, hot loop
// src/nhs_grid.jl:519-575) on gfx950: one wavefront per non-empty query cell.
//
// The 3^D cells around a query cell are 3^(D-1) x-rows of three adjacent cells, each a
// contiguous range of the cell-sorted order.  The wavefront loads those candidates into
// registers, 64 per group (lane = candidate), and broadcasts the cell's query points one at
// a time through scalar registers.  Every group then tests its 64 candidates against the
// query with the reference's predicate
//     p = convert(T_r, x_i - y_j), d2 = ((p1*p1 + p2*p2) + p3*p3), accept iff d2 <= r*r
// (src/nhs_grid.jl:547-555, built with -ffp-contract=off), nine VALU instructions per group
// and no LDS traffic.
//   * Counting: a group's accepted candidates are one ballot; its population count goes
//     to the query's count in a scalar register.
//   * Every other functor: the accepted (query, candidate) pairs are appended to a pair
//     list in LDS in query-major order (per query in the reference's candidate order:
//     stencil rows first-dimension fastest, then the in-cell order).  When the list is
//     full, or the candidate chunk ends, every query of the list gets the same number of
//     lanes, which run the pair body over its pairs round-robin; their partial sums are
//     combined by a tree of fixed order, so sums are deterministic.
#pragma once

#include "pn_traverse.h"

#include <type_traits>

namespace pn {

// Pair-list capacity (entries) of the candidate-major traversal.
constexpr int PAIR_CAP = 1024;

// Pair cache: the first traversal of a cacheable functor after a build records, per
// non-empty cell, its flushes; later traversals on the same build replay them instead of
// testing candidates.  Each wavefront appends the records of its cells to its own region
// of `data` (no atomics); a cell's records are contiguous:
//     record = [batch, chunk, i0, i1, qstart[0 .. i1-i0], pair[0 .. P-1]]   (P = qstart[i1-i0])
// head[2w] = first entry / 4 of work item w's records (-1: not cached, the replay runs
// the full traversal for it), head[2w+1] = their length.  A replay loads a cell's records into
// LDS with one wide copy and runs the recorded flushes, so results equal a full traversal.
constexpr int PC_CELL_MAX = PAIR_CAP + WAVE;  // entries of one cell's records (the LDS list)
struct PairCache {
    uint16_t* data;
    long long region;  // entries per wavefront (multiple of 4)
    int* head;
};

template <class T>
__device__ __forceinline__ T lane_bcast(T v, int i) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), i));
    } else {
        const long long u = __builtin_bit_cast(long long, v);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(u & 0xffffffffll), i);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(u >> 32), i);
        return __builtin_bit_cast(T, (long long)(((unsigned long long)hi << 32) | lo));
    }
}

// v with lane `lane` set to the wave-uniform value `val` (v_writelane_b32: one VALU
// instruction; no builtin for it here).
__device__ __forceinline__ int write_lane(int v, int val, int lane) {
    // the lane select goes through m0: one instruction may read only one other SGPR
    asm volatile("v_writelane_b32 %0, %1, m0" : "+v"(v) : "s"(val), "{m0}"(lane));
    return v;
}

// v of lane src (mod 64), for any trivially copyable accumulator (word by word).
template <class A>
__device__ __forceinline__ A shfl_acc(const A& v, int src) {
    static_assert(sizeof(A) % 4 == 0, "accumulators are whole words");
    constexpr int W = sizeof(A) / 4;
    struct Words {
        int w[W];
    };
    Words in = __builtin_bit_cast(Words, v), out;
#pragma unroll
    for (int k = 0; k < W; ++k) out.w[k] = __shfl(in.w[k], src & (WAVE - 1), WAVE);
    return __builtin_bit_cast(A, out);
}

// Candidate groups held in registers per chunk.  Counting: 6 (a 3D stencil of ~14 points per
// cell in one chunk).  Pair bodies: 3 groups (192 candidates) for Float32 by default, which
// keeps LDS small enough for 6 wavefronts per SIMD; the large variant holds the stencil in
// one chunk.
template <class TC, class F>
struct CellGroups {
    static constexpr int large = F::kLight ? 6 : (sizeof(TC) == 4 ? 6 : 4);
    static constexpr int small = F::kLight ? 6 : (sizeof(TC) == 4 ? 3 : 2);
};

template <int N>
using IntC = std::integral_constant<int, N>;

// 64 / n for n = 0 .. 64 (0 -> 64): the lanes per query of a flush, read with a scalar load.
struct LanesTable {
    int v[WAVE + 1];  // words: a byte table is read with a vector load
    constexpr LanesTable() : v{} {
        for (int n = 0; n <= WAVE; ++n) v[n] = WAVE / (n > 0 ? n : 1);
    }
};
static __constant__ constexpr LanesTable kLanesTable{};

// Calls fn(IntC<n>{}) for a runtime n in 1..G, so loops over n groups unroll without guards.
template <int G, class Fn>
__device__ __forceinline__ void with_groups(int n, Fn&& fn) {
    if constexpr (G >= 1) {
        if (n == G) {
            fn(IntC<G>{});
            return;
        }
        with_groups<G - 1>(n, fn);
    }
}

// PCM: pair cache mode of this instance (0 off, 1 record, 2 replay); only functors with
// kCache use 1 and 2.
// HASH: the grid is keyed by a SpatialHashingCellList (a work item is a non-empty query
// bucket, the pieces are the 3^D neighbour cells' buckets) and binned from the current
// coordinates, so the per-bucket cell and collision flags of the binning hold.  A query
// bucket that mixes cells is run once per cell, over that cell's queries.
template <int D, class TC, class TR, class F, int G, bool PER, int PCM, bool HASH = false>
__global__ __launch_bounds__(WAVE) void k_cells(Geom g, SetView Q, SetView Gs, TR r2, F f, int nblocks,
                                                PairCache pc) {
    static_assert(!HASH || PCM == 0, "no pair cache over a hash grid");
    constexpr int R = Rows<D>::R;
    // stencil pieces: rows (two per row with a periodic box), or 3^D buckets
    constexpr int RP = HASH ? (D == 1 ? 3 : (D == 2 ? 9 : 27)) : (PER ? 2 * R : R);
    // the cell arithmetic of a hash grid (as the binning's)
    using TCELL = std::conditional_t<(sizeof(TC) == 8 || sizeof(TR) == 8), double, float>;
    constexpr int NF = F::kFields;
    constexpr int CH = G * WAVE;
    constexpr bool kLight = F::kLight;
    using Acc = typename F::Acc;
    using QS = typename F::QState;
    constexpr int LC = kLight ? 1 : CH;
    constexpr int LQ = kLight ? 1 : WAVE;
    __shared__ Pack<TC, D> s_c[LC];
    __shared__ int s_j[(!kLight && F::kIndex) ? CH : 1];
    // 3D: the first per-neighbour field rides in the padding slot of the candidate record
    constexpr bool kPackF = D == 3 && NF > 0 && std::is_same<TC, TR>::value;
    constexpr int FOFF = kPackF ? 1 : 0;
    constexpr int NFL = NF - FOFF;  // fields kept in their own LDS planes
    __shared__ TR s_f[NFL > 0 ? NFL : 1][(!kLight && NFL > 0) ? CH : 1];
    // Pair list: candidate slots of the chunk, one byte each when a chunk holds at most 256 and
    // no pair cache is involved; a replay copies a cell's 16-bit records (up to PC_CELL_MAX
    // entries) here.
    using PairT = typename std::conditional<(PCM == 0 && CH <= 256), uint8_t, uint16_t>::type;
    constexpr int PAIR_WORDS = (PCM == 0 ? (PAIR_CAP + WAVE) * (int)sizeof(PairT) : 2 * PC_CELL_MAX) / 8;
    __shared__ uint64_t s_pairs64[kLight ? 1 : PAIR_WORDS];
    PairT* const s_pairs = (PairT*)s_pairs64;
    __shared__ uint16_t s_qstart[LQ + 1];
    // the query batch's coordinates, lane k holding query k's: read with one broadcast LDS load
    // per query in the test loop and per lane in the pair bodies; in Float64 pair functors they
    // are taken from the query's lane instead (a readlane per query, a shuffle per flush), which
    // keeps 2 KB of LDS free for occupancy (measured faster there, slower in Float32)
    constexpr bool kQcLds = kLight || sizeof(TC) == 4;
    __shared__ Pack<TC, D> s_qc[kQcLds ? WAVE : 1];
    TC qxr[D];
#pragma unroll
    for (int d = 0; d < D; ++d) qxr[d] = (TC)0;
    // Per-query state (f.query): kept in LDS when it is one word or the CSR fill's row bounds
    // (read at almost every pair), else reloaded from global memory where a flush needs it
    // (larger states would cost occupancy).
    constexpr bool kQstLds = sizeof(QS) <= 4 || F::kFill;
    __shared__ QS s_qst[kQstLds ? LQ : 1];
    // the batch's per-query sums: in a register of the query's lane (lane k holds query k's;
    // frees LDS, hence occupancy) for sums of up to three words or in Float64 (measured faster),
    // else in LDS
    constexpr bool kRegAcc = sizeof(Acc) <= 12 || sizeof(TR) == 8;
    // flushes spread all 64 lanes over their queries for sums of more than one value
    constexpr bool kAllLanes = sizeof(Acc) > sizeof(TR);
    Acc qacc = f.init();
    __shared__ Acc s_acc[(!kLight && !kRegAcc) ? LQ : 1];
    auto query_sum = [&]() -> Acc {
        if constexpr (kRegAcc) return qacc;
        else return s_acc[threadIdx.x];
    };
    __shared__ int s_written[F::kFill ? WAVE : 1];

    const int lane = threadIdx.x;

    // compute_periodic_distance (src/neighborhood_search.jl:428-436): if d2 > r^2 the
    // minimum image pos_diff -= L .* round.(pos_diff ./ L) (round half to even) and d2 is
    // recomputed.  round(pos_diff / L) is 0 unless |pos_diff| > L / 2, so lanes with every
    // |pos_diff_d| <= 0.49 L_d keep pos_diff and d2 unchanged, as the reference would.
    TR box[D], half[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        box[d] = (TR)g.box[d];
        half[d] = (TR)0.49 * box[d];
    }
    auto periodic_fix = [&](TR (&pd)[D], TR d2) -> TR {
        bool need = false;
#pragma unroll
        for (int d = 0; d < D; ++d) need = need || (fabs(pd[d]) > half[d]);
        need = need && (d2 > r2);
        if (__builtin_amdgcn_ballot_w64(need)) {
            if (need) {
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const TR k = rint(pd[d] / box[d]);
                    const TR s = box[d] * k;
                    pd[d] = pd[d] - s;
                }
                d2 = pd[0] * pd[0];
#pragma unroll
                for (int d = 1; d < D; ++d) {
                    const TR p2 = pd[d] * pd[d];
                    d2 = d2 + p2;
                }
            }
        }
        return d2;
    };

    // Record mode: write position in this wavefront's region, start of the current cell's
    // records, and whether they all fitted.
    long long rec_pos = (long long)blockIdx.x * pc.region;
    const long long rec_end = rec_pos + pc.region;
    long long cell_base = rec_pos;
    bool cell_ok = true;

    // Runs the pair body over pairs [0, P) of list `prp`, which hold queries i0 .. i1-1
    // (qsp[q - i0] = first pair of query q, qsp[i1 - i0] = P).  The 64 lanes are shared
    // out over the i1 - i0 queries (floor or ceil of 64 / (i1 - i0) each), which take their
    // query's pairs round-robin; a query's partial sums are combined by a tree of fixed order
    // and then added, in flush order, to the query's sum, so sums are deterministic.
    int cur_q0 = 0;  // sorted position of the current query batch
    // query q's coordinates (every lane calls it: the shuffle reads lane q)
    auto qpack = [&](int q) {
        Pack<TC, D> p;
        if constexpr (kQcLds) {
            p = s_qc[q];
        } else {
#pragma unroll
            for (int d = 0; d < (D == 3 ? 4 : D); ++d) p.v[d] = (TC)0;
#pragma unroll
            for (int d = 0; d < D; ++d) p.v[d] = shfl_acc(qxr[d], q);
        }
        return p;
    };
    // ... and its per-query state.
    auto qstate = [&](int q) -> QS {
        if constexpr (kQstLds) return s_qst[q];
        else return f.query((int64_t)Q.sorted_idx[cur_q0 + q]);
    };
    auto flush = [&](int i0, int i1, int P, int batch, int chunk, const uint16_t* qsp, const PairT* prp) {
        const int nf = i1 - i0;
        __syncthreads();
        if constexpr (F::kCache && PCM == 1) {
            if (cell_ok && P > 0) {
                const long long size = 5 + nf + P;
                if (rec_pos + size > rec_end || rec_pos + size - cell_base > PC_CELL_MAX) {
                    cell_ok = false;
                } else {
                    uint16_t* rec = pc.data + rec_pos;
                    if (lane == 0) {
                        rec[0] = (uint16_t)batch;
                        rec[1] = (uint16_t)chunk;
                        rec[2] = (uint16_t)i0;
                        rec[3] = (uint16_t)i1;
                    }
                    for (int t = lane; t <= nf; t += WAVE) rec[4 + t] = qsp[t];
                    for (int k = lane; k < P; k += WAVE) rec[5 + nf + k] = prp[k];
                    rec_pos += size;
                }
            }
        }
        if constexpr (F::kFill) {
            // CSR fill: pair k goes to lane k % 64, so a store instruction writes consecutive
            // slots of one or two rows; each lane's query cursor only moves forward.
            int fq = -1, nxt = 0, row0 = 0;  // row0: slot of pair qsp[fq] of query fq
            QS qst{};
            for (int k = lane; k < P; k += WAVE) {
                if (k >= nxt) {
                    do {
                        ++fq;
                        nxt = qsp[fq + 1];
                    } while (k >= nxt);
                    qst = qstate(i0 + fq);
                    row0 = qst.rb + s_written[i0 + fq] - (int)qsp[fq];
                }
                const int slot = row0 + k;
                if (slot < qst.re) f.col[slot] = s_j[prp[k]] + f.base;
            }
            __syncthreads();
            if (lane < nf) s_written[i0 + lane] += qsp[lane + 1] - qsp[lane];
            __syncthreads();
            return;
        }
        // Every lane takes one query: with L = floor(64 / nf), the first rr = 64 - L nf queries
        // get L + 1 lanes and the others L (query q starts at lane q L + min(q, rr)); sums of
        // one value leave the spare lanes idle (rr = 0: measured 1% faster for the density,
        // where the wider tree costs more than the shorter pair loop saves).  Lane
        // l of query q takes its pairs qsp[q] + sub + Lq t, so no lane changes query inside
        // the loop; the Lq partial sums of a query are then combined by a tree of fixed order.
        const int L = kLanesTable.v[nf];
        const int rr = (kAllLanes && nf > 0) ? WAVE - L * nf : 0;
        const int cut = rr * (L + 1);
        const bool wide = lane < cut;
        const int Lq = wide ? L + 1 : L;
        const int lrel = wide ? lane : lane - cut;
        // lrel / Lq exactly: (lrel + 1/2) / Lq is at least 1/(2 Lq) away from an integer
        const int qrel = (int)(((float)lrel + 0.5f) * __builtin_amdgcn_rcpf((float)Lq));
        const int q = wide ? qrel : rr + qrel, sub = lrel - qrel * Lq;
        Acc a = f.init();
        const Pack<TC, D> qp = qpack(i0 + q);
        if (q < nf) {
            const QS qst = qstate(i0 + q);
            const int ke = qsp[q + 1];
            for (int k = qsp[q] + sub; k < ke; k += Lq) {
                const int m = prp[k];
                const Pack<TC, D> pk = s_c[m];
                TR pd[D];
#pragma unroll
                for (int d = 0; d < D; ++d) pd[d] = (TR)(qp.v[d] - pk.v[d]);
                TR d2 = pd[0] * pd[0];
#pragma unroll
                for (int d = 1; d < D; ++d) {
                    const TR p2 = pd[d] * pd[d];
                    d2 = d2 + p2;
                }
                if constexpr (PER) d2 = periodic_fix(pd, d2);
                int j = 0;
                if constexpr (F::kIndex) j = s_j[m];
                TR fj[NF > 0 ? NF : 1];
                if constexpr (kPackF) fj[0] = pk.v[3];
                if constexpr (NFL > 0) {
#pragma unroll
                    for (int fi = FOFF; fi < NF; ++fi) fj[fi] = s_f[fi - FOFF][m];
                }
                f.template pair<TR, D>(a, qst, pd, d2, j, fj, nullptr);
            }
        }
        // tree over the query's Lq lanes towards its first lane, with P2 the largest power of
        // two <= the widest Lq (uniform): fold the lanes past P2, then halve, skipping
        // partners past the query's own lanes (Lq = P2 - 1 when L + 1 is a power of two)
        const int Lw = rr > 0 ? L + 1 : L;
        int P2 = 1;
        while (2 * P2 <= Lw) P2 *= 2;
        if (Lw > P2) {
            const Acc t = shfl_acc(a, lane + P2);
            if (sub < Lq - P2) f.combine(a, t);
        }
        for (int o = P2 / 2; o > 0; o >>= 1) {
            const Acc t = shfl_acc(a, lane + o);
            if (sub < o && sub + o < Lq) f.combine(a, t);
        }
        if constexpr (kRegAcc) {
            // lane k (query k of the batch) takes its flush sum from its query's first lane
            const int kq = lane - i0;
            const Acc t = shfl_acc(a, kq * L + min(kq, rr));
            if (lane >= i0 && lane < i1) f.combine(qacc, t);
        } else {
            if (q < nf && sub == 0) {
                Acc t = s_acc[i0 + q];
                f.combine(t, a);
                s_acc[i0 + q] = t;
            }
        }
        __syncthreads();
    };

    const TC* __restrict__ gc = (const TC*)Gs.coords;
    const TC* __restrict__ qcrd = (const TC*)Q.coords;
    const TR* __restrict__ fs = (const TR*)f.fsorted;

    // Persistent grid: XCD x (= blockIdx % 8) walks a contiguous range of the non-empty
    // cells, so the cells that share stencil rows run on one XCD's L2.
    const int n_work = Q.st->n_nonempty;
    const int gn0 = g.n[0], gn1 = g.n[1];
    const int xcd = blockIdx.x & 7, per_xcd = nblocks >> 3;
    const int wchunk = (n_work + 7) >> 3;
    const int w_end = min(n_work, (xcd + 1) * wchunk);
    for (int w = xcd * wchunk + (int)(blockIdx.x >> 3); w < w_end; w += per_xcd) {
        cell_base = (rec_pos + 3) & ~3ll;  // 8-byte aligned cell records
        rec_pos = cell_base;
        cell_ok = true;
        const int qc = Q.nonempty[w];
        const int qs = Q.start[qc], qe = Q.start[qc + 1];
        // The piece table in scalar registers: element m of piece u sits at sorted position
        // m + rdel[u]; piece u holds elements [pre[u], pre[u + 1]).
        // A hash grid (which needs each element's piece for its cell check) and more than 9
        // pieces (a periodic box: 18) keep the table in lanes, lane u < RP holding pre[u] and
        // rdel[u]: scalar registers would not hold 18 or 27 pieces, and the select chains over
        // them became scratch-memory table loads.
        constexpr bool kLaneTable = HASH || RP > 9;
        int rdel[kLaneTable ? 1 : RP], pre[kLaneTable ? 1 : RP + 1];
        int pre_l = 0, rdel_l = 0;
        int M = 0;
        // Lane u < RP holds piece u's first sorted position (rb) and length (rlen).
        auto set_pieces = [&](int rb, int rlen) {
            int pre_v = rlen;  // rlen of lanes <= u, summed
#pragma unroll
            for (int o = 1; o < 32; o <<= 1) {
                const int t = __shfl_up(pre_v, o, WAVE);
                if (lane >= o) pre_v += t;
            }
            M = __builtin_amdgcn_readlane(pre_v, RP - 1);
            if constexpr (kLaneTable) {
                pre_l = pre_v - rlen;
                rdel_l = rb - pre_l;
            } else {
                pre[0] = 0;
#pragma unroll
                for (int t = 0; t < RP; ++t) {
                    rdel[t] = __builtin_amdgcn_readlane(rb, t) - pre[t];
                    pre[t + 1] = pre[t] + __builtin_amdgcn_readlane(rlen, t);
                }
            }
        };
        // Lane table: piece tp of element m of the group of 64 elements from mlo, and its offset
        // del (sorted position m + del).  The group spans pieces ulo..uhi (a few: found with two
        // ballots); only those are scanned, with the table read from lanes.  Called with every
        // lane of the wave active.
        auto lane_piece = [&](int mlo, int m, int& del, int& tp) {
            const bool inr = lane < RP;
            const int ulo = __builtin_popcountll(__builtin_amdgcn_ballot_w64(inr && pre_l <= mlo)) - 1;
            const int uhi = __builtin_popcountll(__builtin_amdgcn_ballot_w64(inr && pre_l <= mlo + WAVE - 1)) - 1;
            del = __builtin_amdgcn_readlane(rdel_l, ulo);
            tp = ulo;
            for (int t = ulo + 1; t <= uhi; ++t) {
                const int p = __builtin_amdgcn_readlane(pre_l, t);
                const int r = __builtin_amdgcn_readlane(rdel_l, t);
                del = m >= p ? r : del;
                tp = m >= p ? t : tp;
            }
        };
        // HASH: the cell whose 3^D neighbour cells the pieces are, and the pieces whose bucket
        // holds points of several cells (their elements are checked one by one).
        int64_t hc[3] = {0, 0, 0};
        unsigned long long hchk = 0;
        bool qmixed = false;
        // Pieces: the buckets of the 3^D cells around hc in CartesianIndices order (first
        // dimension fastest), as hash_query visits them.  A bucket whose points all lie in
        // one other cell (a collision) contributes nothing; a bucket that mixes cells is
        // scanned with a per-element cell check.
        auto hash_pieces = [&]() {
            int rb = 0, rlen = 0;
            bool chk = false;
            if (lane < RP) {
                int64_t nc[3] = {hc[0], hc[1], hc[2]};
                int t = lane;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    nc[d] = hc[d] + (t % 3) - 1;
                    if constexpr (PER) nc[d] = hash_wrap(nc[d], g.per[d]);
                    t /= 3;
                }
                const int b = spatial_hash<D>(nc, g);
                rb = Gs.start[b];
                rlen = Gs.start[b + 1] - rb;
                if (rlen > 0) {
                    if (Gs.bucket_mixed[b] == Gs.bucket_epoch) {
                        chk = true;
                    } else {
                        bool same = true;
#pragma unroll
                        for (int d = 0; d < D; ++d) same = same && (int64_t)Gs.bucket_cell[3 * (int64_t)b + d] == nc[d];
                        if (!same) rlen = 0;
                    }
                }
            }
            hchk = __builtin_amdgcn_ballot_w64(chk);
            set_pieces(rb, rlen);
        };
        if constexpr (HASH) {
            qmixed = Q.bucket_mixed[qc] == Q.bucket_epoch;
            if (!qmixed) {
#pragma unroll
                for (int d = 0; d < D; ++d) hc[d] = Q.bucket_cell[3 * (int64_t)qc + d];
                hash_pieces();
            }
        } else {
        int n0, n1;
        grid_dims_xy(g, qc, gn0, gn1, n0, n1);
        // Stencil rows, each one or two pieces of the sorted order: piece u covers
        // [rdel[u] + pre[u], rdel[u] + pre[u + 1]).  A row of 3 x-adjacent cells is one
        // contiguous range; with a periodic box a row whose x-neighbour wraps is two
        // (src/nhs_grid.jl:593-620), kept in the reference's cell order (x offset -1, 0, 1).
        int rb = 0, rlen = 0;
        if (lane < RP) {
            const int t = PER ? lane >> 1 : lane, piece = PER ? lane & 1 : 0;
            const int dy = D >= 2 ? (t % 3) - 1 : 0, dz = D >= 3 ? (t / 3) - 1 : 0;
            if constexpr (!PER) {
                const int base = qc - 1 + dy * n0 + dz * n0 * n1;
                rb = Gs.start[base];
                rlen = Gs.start[base + 3] - rb;
            } else {
                // 0-based periodic cells are 1..per; wrap(c) = ((c - 1) mod per) + 1
                auto wrap = [](int c, int per) { return ((c - 1 + per) % per) + 1; };
                const int cx = qc % n0, cy = (qc / n0) % n1, cz = qc / (n0 * n1);
                const int yy = D >= 2 ? wrap(cy + dy, g.per[1]) : cy;
                const int zz = D >= 3 ? wrap(cz + dz, g.per[2]) : cz;
                const int row = yy * n0 + zz * n0 * n1;
                const int xm = wrap(cx - 1, g.per[0]), xp = wrap(cx + 1, g.per[0]);
                int lo, hi;  // cells [lo, hi) of this piece
                if (xm == cx - 1 && xp == cx + 1) {
                    lo = piece == 0 ? cx - 1 : 0;
                    hi = piece == 0 ? cx + 2 : 0;
                } else if (xm != cx - 1) {
                    lo = piece == 0 ? xm : cx;
                    hi = piece == 0 ? xm + 1 : cx + 2;
                } else {
                    lo = piece == 0 ? cx - 1 : xp;
                    hi = piece == 0 ? cx + 1 : xp + 1;
                }
                if (hi > lo) {
                    rb = Gs.start[row + lo];
                    rlen = Gs.start[row + hi] - rb;
                }
            }
        }
        set_pieces(rb, rlen);
        }

        // HASH, pieces of a bucket that mixes cells: an element whose own cell (from its
        // coordinates x) is not its piece's cell becomes +Inf, never accepted (hash_query's
        // check, once per staged element).
        auto hash_check = [&](int tp, TC* x) {
            if ((hchk >> tp) & 1) {
                int64_t cj[3];
                bool okj;
                hash_cell_p<D, TC, TCELL, PER>(g, x, cj, okj);
                bool same = true;
                int t = tp;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    int64_t want = hc[d] + (t % 3) - 1;
                    if constexpr (PER) want = hash_wrap(want, g.per[d]);
                    same = same && cj[d] == want;
                    t /= 3;
                }
                if (!same) {
#pragma unroll
                    for (int d = 0; d < D; ++d) x[d] = (TC)INFINITY;
                }
            }
        };

        // Counting: loads chunk m0 (mc elements) into registers; group gi holds elements
        // m0 + 64 gi + lane, the piece of element m is the last u with pre[u] <= m.  Lanes
        // past the chunk end hold +Inf (never accepted).
        auto stage = [&](int m0, int mc, TC (&cx)[G][D]) {
            const int ng = (mc + WAVE - 1) / WAVE;
#pragma unroll
            for (int gi = 0; gi < G; ++gi) {
                if (gi < ng) {
                    const int ml = gi * WAVE + lane;
                    const int m = m0 + ml;
                    int del, tp = 0;
                    if constexpr (kLaneTable) {
                        lane_piece(m0 + gi * WAVE, m, del, tp);
                    } else {
                        del = rdel[0];
#pragma unroll
                        for (int t = 1; t < RP; ++t) del = m >= pre[t] ? rdel[t] : del;
                    }
                    if (ml < mc) {
                        const unsigned src = (unsigned)(m + del);
#pragma unroll
                        for (int d = 0; d < D; ++d) cx[gi][d] = gc[(int64_t)d * Gs.cap + src];
                        if constexpr (HASH) {
                            if (hchk) hash_check(tp, cx[gi]);
                        }
                    } else {
#pragma unroll
                        for (int d = 0; d < D; ++d) cx[gi][d] = (TC)INFINITY;
                    }
                }
            }
        };

        // Pair mode: stages chunk m0 (mc elements: coordinates, indices, fields) in LDS.  Lane
        // l gathers elements m0 + 64 gi + l of every group first (all loads in flight), then
        // writes them; the test registers cx (lanes past the chunk: +Inf) come from the same
        // loads.
        auto stage_lds = [&](int m0, int mc, TC (&cx)[G][D]) {
            const int ng = (mc + WAVE - 1) / WAVE;
            Pack<TC, D> pk[G];
            int jv[G];
            TR fv[G][NF > 0 ? NF : 1];
#pragma unroll
            for (int gi = 0; gi < G; ++gi) {
                const int ml = gi * WAVE + lane;
#pragma unroll
                for (int d = 0; d < (D == 3 ? 4 : D); ++d) pk[gi].v[d] = (TC)0;
                jv[gi] = 0;
#pragma unroll
                for (int fi = 0; fi < (NF > 0 ? NF : 1); ++fi) fv[gi][fi] = (TR)0;
                const int m = m0 + ml;
                int hdel = 0, htp = 0;
                if constexpr (kLaneTable) {
                    if (gi < ng) lane_piece(m0 + gi * WAVE, m, hdel, htp);
                }
                if (gi < ng && ml < mc) {
                    int del;
                    if constexpr (kLaneTable) {
                        del = hdel;
                    } else {
                        // a sum of per-piece steps (a select chain here becomes a scratch table load)
                        del = rdel[0];
#pragma unroll
                        for (int t = 1; t < RP; ++t) del += m >= pre[t] ? rdel[t] - rdel[t - 1] : 0;
                    }
                    const unsigned src = (unsigned)(m + del);
#pragma unroll
                    for (int d = 0; d < D; ++d) pk[gi].v[d] = gc[(int64_t)d * Gs.cap + src];
                    if constexpr (HASH) {
                        if (hchk) hash_check(htp, pk[gi].v);
                    }
                    if constexpr (F::kIndex) jv[gi] = Gs.sorted_idx[src];
                    if constexpr (NF > 0) {
#pragma unroll
                        for (int fi = 0; fi < NF; ++fi) fv[gi][fi] = fs[(int64_t)fi * f.fstride + src];
                    }
                    if constexpr (kPackF) pk[gi].v[3] = fv[gi][0];
                }
            }
            __syncthreads();  // the previous chunk's pairs are done
#pragma unroll
            for (int gi = 0; gi < G; ++gi) {
                const int ml = gi * WAVE + lane;
                if (gi < ng) {
                    const bool in = ml < mc;
                    if (in) {
                        s_c[ml] = pk[gi];
                        if constexpr (F::kIndex) s_j[ml] = jv[gi];
                        if constexpr (NFL > 0) {
#pragma unroll
                            for (int fi = FOFF; fi < NF; ++fi) s_f[fi - FOFF][ml] = fv[gi][fi];
                        }
                    }
#pragma unroll
                    for (int d = 0; d < D; ++d) cx[gi][d] = in ? pk[gi].v[d] : (TC)INFINITY;
                }
            }
            __syncthreads();
        };

        // Query batch of up to 64 points starting at sorted position q0.
        auto begin_batch = [&](int q0, int nq, TC (&xq)[D], int64_t& orig_i) {
            cur_q0 = q0;
            const bool qact = lane < nq;
#pragma unroll
            for (int d = 0; d < D; ++d) xq[d] = qact ? qcrd[(int64_t)d * Q.cap + q0 + lane] : (TC)0;
            orig_i = qact ? Q.sorted_idx[q0 + lane] : 0;
            __syncthreads();  // the previous batch is done with s_qc
            if constexpr (kQcLds) {
                Pack<TC, D> pk;
#pragma unroll
                for (int d = 0; d < (D == 3 ? 4 : D); ++d) pk.v[d] = d < D ? xq[d] : (TC)0;
                s_qc[lane] = pk;
            } else {
#pragma unroll
                for (int d = 0; d < D; ++d) qxr[d] = xq[d];
            }
            __syncthreads();
            if constexpr (!kLight) {
                if (qact) {
                    if constexpr (kQstLds) s_qst[lane] = f.query(orig_i);
                    if constexpr (kRegAcc) qacc = f.init();
                    else s_acc[lane] = f.init();
                    if constexpr (F::kFill) s_written[lane] = 0;
                }
            }
        };

        int rbase = -1, rsize = 0;
        if constexpr (F::kCache && PCM == 2) {
            rbase = pc.head[2 * (int64_t)w];
            rsize = pc.head[2 * (int64_t)w + 1];
        }
        if constexpr (PCM == 2) {
        if (rbase >= 0) {
            // replay: copy the cell's records to LDS (8 bytes per lane and step), then run
            // the recorded flushes batch by batch
            __syncthreads();
            const uint64_t* src = (const uint64_t*)pc.data + rbase;
            for (int k = lane; k < (rsize + 3) / 4; k += WAVE) s_pairs64[k] = src[k];
            __syncthreads();
            int pos = 0;
            for (int q0 = qs, b = 0; q0 < qe; q0 += WAVE, ++b) {
                const int nq = min(WAVE, qe - q0);
                TC xq[D];
                int64_t orig_i;
                begin_batch(q0, nq, xq, orig_i);
                int staged = -1;
                while (pos < rsize && (int)s_pairs[pos] == b) {
                    const int chunk = s_pairs[pos + 1], i0 = s_pairs[pos + 2], i1 = s_pairs[pos + 3];
                    const int nf = i1 - i0;
                    const int P = s_pairs[pos + 4 + nf];
                    if (chunk != staged) {
                        const int m0 = chunk * CH;
                        TC unused[G][D];
                        stage_lds(m0, min(CH, M - m0), unused);
                        staged = chunk;
                    }
                    flush(i0, i1, P, b, chunk, s_pairs + pos + 4, s_pairs + pos + 5 + nf);
                    pos += 5 + nf + P;
                }
                __syncthreads();
                if (lane < nq) f.write(orig_i, qstate(lane), query_sum());
            }
            continue;
        }
        }

        for (int q0 = qs, b = 0; q0 < qe; q0 += WAVE, ++b) {
            const int nq = min(WAVE, qe - q0);
            const bool qact = lane < nq;
            TC xq[D];
            int64_t orig_i;
            begin_batch(q0, nq, xq, orig_i);
            int mycount = 0;
            // HASH, a bucket of several query cells: one pass per cell over the queries of
            // the batch in that cell (qsel), the pieces rebuilt for it; else one pass.
            unsigned long long qpend = 0, qsel = ~0ull;
            int64_t qcl[3] = {0, 0, 0};
            if constexpr (HASH) {
                if (qmixed) {
                    bool ok;
                    hash_cell_p<D, TC, TCELL, PER>(g, xq, qcl, ok);
                    qpend = __builtin_amdgcn_ballot_w64(qact);
                }
            }
            do {
            if constexpr (HASH) {
                if (qmixed) {
                    const int ld = __builtin_ctzll(qpend);
                    bool same = qact;
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        hc[d] = lane_bcast(qcl[d], ld);
                        same = same && qcl[d] == hc[d];
                    }
                    qsel = __builtin_amdgcn_ballot_w64(same);
                    qpend &= ~qsel;
                    hash_pieces();
                }
            }

            for (int m0 = 0, chunk = 0; m0 < M; m0 += CH, ++chunk) {
                const int mc = min(CH, M - m0);
                const int ng = (mc + WAVE - 1) / WAVE;
                TC cx[G][D];
                if constexpr (kLight) {
                    stage(m0, mc, cx);
                } else {
                    stage_lds(m0, mc, cx);
                }

                // d2 of broadcast query xs against group gi, in the reference's arithmetic
                auto test = [&](const TC (&xs)[D], int gi) -> bool {
                    TR pd[D];
#pragma unroll
                    for (int d = 0; d < D; ++d) pd[d] = (TR)(xs[d] - cx[gi][d]);
                    TR d2 = pd[0] * pd[0];
#pragma unroll
                    for (int d = 1; d < D; ++d) {
                        const TR p2 = pd[d] * pd[d];
                        d2 = d2 + p2;
                    }
                    if constexpr (PER) d2 = periodic_fix(pd, d2);
                    return d2 <= r2;
                };

                if constexpr (kLight) {
                    with_groups<G>(ng, [&](auto ngc) {
                        constexpr int NG = decltype(ngc)::value;
                        for (int i = 0; i < nq; ++i) {
                            if constexpr (HASH) {
                                if (!((qsel >> i) & 1)) continue;
                            }
                            TC xs[D];
#pragma unroll
                            for (int d = 0; d < D; ++d) xs[d] = s_qc[i].v[d];
                            int c = 0;
#pragma unroll
                            for (int gi = 0; gi < NG; ++gi)
                                c += __builtin_popcountll(__builtin_amdgcn_ballot_w64(test(xs, gi)));
                            // lane i holds query i's count (no earlier chunk: nothing to add)
                            if (chunk != 0) c += __builtin_amdgcn_readlane(mycount, i);
                            mycount = write_lane(mycount, c, i);
                        }
                    });
                } else {
                    int P = 0, i0 = 0;
                    int qfirst = 0;  // lane t: first list entry of query i0 + t
                    // flushes queries i0 .. i1-1 (list starts to LDS first)
                    auto flush_to = [&](int i1) {
                        const int nf = i1 - i0;
                        if (lane < nf) s_qstart[lane] = (uint16_t)qfirst;
                        if (lane == 0) s_qstart[nf] = (uint16_t)P;
                        flush(i0, i1, P, b, chunk, s_qstart, s_pairs);
                    };
                    with_groups<G>(ng, [&](auto ngc) {
                        constexpr int NG = decltype(ngc)::value;
                        for (int i = 0; i < nq; ++i) {
                            if (P + NG * WAVE > PAIR_CAP) {
                                flush_to(i);
                                i0 = i;
                                P = 0;
                            }
                            qfirst = write_lane(qfirst, P, i - i0);
                            if constexpr (HASH) {
                                if (!((qsel >> i) & 1)) continue;  // no pairs in this pass
                            }
                            TC xs[D];
#pragma unroll
                            for (int d = 0; d < D; ++d) xs[d] = kQcLds ? s_qc[i].v[d] : lane_bcast(qxr[d], i);
#pragma unroll
                            for (int gi = 0; gi < NG; ++gi) {
                                // accepted lanes take the next list slots in lane order (the
                                // write runs under their execution mask)
                                const bool acc = test(xs, gi);
                                const unsigned long long mask = __builtin_amdgcn_ballot_w64(acc);
                                if (acc) {
                                    // P + (accepted lanes below this one)
                                    const int slot = __builtin_amdgcn_mbcnt_hi(
                                        (unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, (unsigned)P));
                                    s_pairs[slot] = (PairT)(gi * WAVE + lane);
                                }
                                P += __builtin_popcountll(mask);
                            }
                        }
                    });
                    flush_to(nq);
                }
            }
            } while (qpend);
            if constexpr (kLight) {
                if (qact) f.write(orig_i, QS{}, mycount);
            } else {
                __syncthreads();
                if (qact) f.write(orig_i, qstate(lane), query_sum());
            }

        }
        if constexpr (F::kCache && PCM == 1) {
            {
                if (lane == 0) {
                    pc.head[2 * (int64_t)w] = cell_ok ? (int)(cell_base >> 2) : -1;
                    pc.head[2 * (int64_t)w + 1] = (int)(rec_pos - cell_base);
                }
                if (!cell_ok) rec_pos = cell_base;
            }
        }
    }
}

}  // namespace pn
