// Host-side machinery shared by the C-ABI translation units: error reporting, device
// buffers, binned point sets, the grid object and the traversal launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
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

namespace pn_host {
// message of the last failed call on this thread (pn_last_error)
extern thread_local std::string g_err;
}  // namespace pn_host

namespace {

using pn_host::g_err;

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
            hipStream_t stream, pn::StatusMailbox* mb = nullptr, unsigned long long seq = 0,
            bool clear_counts = false) {
        const int64_t nb = std::max<int64_t>(1, (n + pn::SCAN_TILE - 1) / pn::SCAN_TILE);
        int rc = block_sums.ensure((size_t)nb * 16, stream);  // block sums, then block maxima
        if (rc) return rc;
        auto* bs = block_sums.as<unsigned long long>();
        hipLaunchKernelGGL(pn::k_scan_reduce, dim3((unsigned)nb), dim3(pn::SCAN_THREADS), 0, stream, counts,
                           n, bs, st);
        PN_CHECK_LAUNCH();
        hipLaunchKernelGGL(pn::k_scan_top, dim3(1), dim3(pn::SCAN_THREADS), 0, stream, bs, nb, start, n, st, mb,
                           seq);
        PN_CHECK_LAUNCH();
        hipLaunchKernelGGL(pn::k_scan_down, dim3((unsigned)nb), dim3(pn::SCAN_THREADS), 0, stream, counts, n,
                           (const unsigned long long*)bs, start, nonempty,
                           clear_counts ? const_cast<int32_t*>(counts) : nullptr);
        PN_CHECK_LAUNCH();
        return PN_OK;
    }
};

// One binned point set: the grid's y or a query set x.
struct BinSet {
    Buf counts, start, nonempty, keys, slots, perm_tmp, sorted_idx, coords, status;
    Buf bucket_cell, bucket_mixed;  // SpatialHashingCellList: per-bucket cell and collision flag
    int hash_epoch = 0;             // bucket_mixed value of a mixed bucket of the last binning
    pn::DevStatus* host_status = nullptr;  // pinned
    // pinned coherent host memory k_scan_top publishes the status to; mail_seq is the
    // sequence number of the last binning that asked for it
    pn::StatusMailbox* mailbox = nullptr;
    unsigned long long mail_seq = 0;
    bool mail_pending = false;
    bool mailbox_failed = false;  // its allocation failed once: not tried again
    Scanner scan;
    bool counts_zero = false;  // `counts` holds zeros (left so by the last binning's scan)
    int64_t cap = 0;
    int64_t n_sel = 0;        // selected points of the last binning
    int64_t n_src = 0;        // columns of the source matrix
    const void* src = nullptr;
    bool valid = false;       // last binning succeeded
    bool fresh = false;       // binned from the current coordinates (no refresh since)
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
        bucket_cell.release();
        bucket_mixed.release();
        scan.block_sums.release();
        if (host_status) (void)hipHostFree(host_status);
        if (mailbox) (void)hipHostFree(mailbox);
    }

    pn::SetView view() const {
        pn::SetView v;
        v.start = start.as<int32_t>();
        v.nonempty = nonempty.as<int32_t>();
        v.sorted_idx = sorted_idx.as<int32_t>();
        v.coords = coords.p;
        v.cap = cap;
        v.st = status.as<pn::DevStatus>();
        v.bucket_cell = bucket_cell.as<int32_t>();
        v.bucket_mixed = bucket_mixed.as<int32_t>();
        v.bucket_epoch = hash_epoch;
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
    int pair_cache = 0;             // record / replay forward pair lists (pn::PairCache)
    int64_t pc_per_point = 0;       // cache capacity in entries per point (0: by dimension)
    uint64_t gen = 0;               // build generation: bumped by every binning or refresh
    Buf pc_data, pc_head;           // pair cache storage
    int64_t pc_region = 0;          // entries per wavefront of the recording launch
    uint64_t pc_gen = ~0ull;        // generation the cache was recorded at
    int pc_groups = 0;              // candidate groups of the recording kernel
    Buf csr_counts;
    Buf csr_status;
    Scanner csr_scan;
    pn::DevStatus* csr_host_status = nullptr;
    int64_t err_point = -1;
    int64_t err_cell[3] = {0, 0, 0};
    const void* err_src = nullptr;
    // traversal kernel timing (tuning key "time_launches"): events 2i and 2i + 1 bracket the
    // i-th traversal kernel launched since timing was switched on
    int time_launches = 0;
    std::vector<hipEvent_t> timing_events;
    size_t timing_used = 0;

    ~pn_grid() {
        for (hipEvent_t e : timing_events) (void)hipEventDestroy(e);
        fields.release();
        pc_data.release();
        pc_head.release();
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
    const void* counts_before = s.counts.p;
    if ((rc = s.counts.ensure((size_t)C * 4, g.stream))) return rc;
    if (s.counts.p != counts_before) s.counts_zero = false;
    if ((rc = s.start.ensure((size_t)(C + 1) * 4, g.stream))) return rc;
    if ((rc = s.nonempty.ensure((size_t)C * 4, g.stream))) return rc;
    if ((rc = s.status.ensure(sizeof(pn::DevStatus), g.stream))) return rc;
    if (g.geom.hash) {
        if ((rc = s.bucket_cell.ensure((size_t)C * 12, g.stream))) return rc;
        const size_t before = s.bucket_mixed.bytes;
        if ((rc = s.bucket_mixed.ensure((size_t)C * 4, g.stream))) return rc;
        if (s.bucket_mixed.bytes != before) s.hash_epoch = 0;  // new memory: cleared before use
    }
    if (!s.host_status) PN_HIP(hipHostMalloc((void**)&s.host_status, sizeof(pn::DevStatus), hipHostMallocDefault));
    if (!s.mailbox && !s.mailbox_failed) {
        // optional: without it the status is copied after a stream synchronization
        if (hipHostMalloc((void**)&s.mailbox, sizeof(pn::StatusMailbox),
                          hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
            (void)hipGetLastError();
            s.mailbox = nullptr;
            s.mailbox_failed = true;
        } else {
            std::memset((void*)s.mailbox, 0, sizeof(pn::StatusMailbox));
            s.mail_seq = 0;
        }
    }
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

// Records a HIP event on the grid's stream before and after the traversal kernel launched in
// its scope, when "time_launches" is on (at most MAX_TIMED launches until they are read).
struct LaunchTimer {
    static constexpr size_t MAX_TIMED = 4096;
    pn_grid& g;
    bool on;
    explicit LaunchTimer(pn_grid& grid) : g(grid), on(grid.time_launches != 0) {
        if (!on) return;
        if (g.timing_used + 2 > 2 * MAX_TIMED) {
            on = false;
            return;
        }
        while (g.timing_events.size() < g.timing_used + 2) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) {
                on = false;
                return;
            }
            g.timing_events.push_back(e);
        }
        (void)hipEventRecord(g.timing_events[g.timing_used], g.stream);
    }
    ~LaunchTimer() {
        if (!on) return;
        (void)hipEventRecord(g.timing_events[g.timing_used + 1], g.stream);
        g.timing_used += 2;
    }
};

unsigned blocks_for(int64_t n, int threads) { return (unsigned)std::max<int64_t>(1, (n + threads - 1) / threads); }

// Enqueues the counting-sort binning of the selected points of src into s.
int enqueue_bin(pn_grid& g, BinSet& s, const void* src, int64_t n_src, int64_t sel_first, int64_t sel_count,
                const int32_t* sel_list) {
    ++g.gen;  // recorded pair lists belong to the previous build
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
    ++s.mail_seq;  // a new number for every binning enqueued, also one that fails below
    s.mail_pending = false;
    hipLaunchKernelGGL(pn::k_reset_status, dim3(1), dim3(1), 0, g.stream, st);
    PN_CHECK_LAUNCH();
    if (!s.counts_zero) {
        hipLaunchKernelGGL(pn::k_fill_i32, dim3(blocks_for(std::min<int64_t>(C, 1 << 20), 256)), dim3(256), 0,
                           g.stream, s.counts.as<int32_t>(), C, 0);
        PN_CHECK_LAUNCH();
    }
    s.counts_zero = false;  // until this binning's scan is enqueued
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
                             g.stream, s.mailbox, s.mail_seq, true);
        if (rc2) return rc2;
        s.counts_zero = true;
        if (sel_count > 0) {
            // a hash build reuses `slots`, cleared to -1 here, for its bucket representatives
            hipLaunchKernelGGL(pn::k_scatter, dim3(blocks_for(sel_count, 256)), dim3(256), 0, g.stream,
                               s.keys.as<int32_t>(), s.slots.as<int32_t>(), sel_count, s.start.as<int32_t>(),
                               s.perm_tmp.as<int32_t>(), g.geom.hash ? s.slots.as<int32_t>() : nullptr);
            PN_CHECK_LAUNCH();
            // one thread per binned slot (at most sel_count; the count is read on the device)
            if (!g.geom.hash) {
                hipLaunchKernelGGL((pn::k_canon<D, TC>), dim3(blocks_for(sel_count, 256)), dim3(256), 0, g.stream,
                                   s.keys.as<int32_t>(), sel_first, sel_list, (const pn::DevStatus*)st,
                                   s.start.as<int32_t>(), s.perm_tmp.as<int32_t>(), (const TC*)src,
                                   s.sorted_idx.as<int32_t>(), s.coords.as<TC>(), s.cap, g.canonical_order);
                PN_CHECK_LAUNCH();
            } else {
                // mixed-bucket flags of earlier binnings hold earlier epochs; fresh memory (epoch 0)
                // and a wrapped counter clear the array first
                if (s.hash_epoch <= 0 || s.hash_epoch == INT32_MAX) {
                    hipLaunchKernelGGL(pn::k_fill_i32, dim3(blocks_for(std::min<int64_t>(C, 1 << 20), 256)),
                                       dim3(256), 0, g.stream, s.bucket_mixed.as<int32_t>(), C, 0);
                    PN_CHECK_LAUNCH();
                    s.hash_epoch = 0;
                }
                ++s.hash_epoch;
                // bucket representatives by selection position, in `slots` (-1 after k_scatter)
                int32_t* rep = s.slots.as<int32_t>();
                hipLaunchKernelGGL((pn::k_canon_hash<D, TC, TCELL>), dim3(blocks_for(sel_count, 256)), dim3(256), 0,
                                   g.stream, s.keys.as<int32_t>(), sel_first, sel_list, (const pn::DevStatus*)st,
                                   s.start.as<int32_t>(), s.perm_tmp.as<int32_t>(), (const TC*)src,
                                   s.sorted_idx.as<int32_t>(), s.coords.as<TC>(), s.cap, g.canonical_order, g.geom,
                                   s.bucket_cell.as<int32_t>(), s.bucket_mixed.as<int32_t>(), s.hash_epoch, rep);
                PN_CHECK_LAUNCH();
                const int64_t per_block = (int64_t)pn::ORDER_THREADS * pn::ORDER_ITEMS;
                hipLaunchKernelGGL(pn::k_order_work, dim3((unsigned)((sel_count + per_block - 1) / per_block)),
                                   dim3(pn::ORDER_THREADS), 0, g.stream, (const int32_t*)rep, sel_count, st,
                                   s.nonempty.as<int32_t>());
                PN_CHECK_LAUNCH();
            }
        }
        return PN_OK;
    });
    if (rc) return rc;
    s.mail_pending = true;
    s.n_sel = sel_count;
    s.n_src = n_src;
    s.src = src;
    s.valid = true;  // provisional until checked
    s.fresh = true;
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

// The status of the last binning of s.  The mailbox path spins on the sequence number
// k_scan_top raises (the host returns while the rest of the build still runs; a kernel
// error after it surfaces at the next synchronization, as for any enqueued work); after
// STATUS_SPIN_US without it, or with no mailbox, the status is copied after a stream
// synchronization.
constexpr int64_t STATUS_SPIN_US = 5000;

int wait_status(pn_grid& g, BinSet& s) {
    if (s.mailbox && s.mail_pending) {
        s.mail_pending = false;
        const unsigned long long want = s.mail_seq;
        const auto t0 = std::chrono::steady_clock::now();
        for (unsigned it = 1;; ++it) {
            if (__atomic_load_n(&s.mailbox->seq, __ATOMIC_ACQUIRE) == want) {
                std::memcpy(&s.stats, (const void*)&s.mailbox->status, sizeof(pn::DevStatus));
                return PN_OK;
            }
#if defined(__x86_64__) || defined(__i386__)
            __builtin_ia32_pause();
#endif
            if ((it & 255) == 0 &&
                std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(STATUS_SPIN_US))
                break;
        }
    }
    PN_HIP(hipMemcpyAsync(s.host_status, s.status.p, sizeof(pn::DevStatus), hipMemcpyDeviceToHost, g.stream));
    PN_HIP(hipStreamSynchronize(g.stream));
    s.stats = *s.host_status;
    return PN_OK;
}

// Waits for the status of a binning pass and turns it into the reference's errors.
int check_set(pn_grid& g, BinSet& s, const int32_t* sel_list, int64_t sel_first, bool is_query) {
    int rc0 = wait_status(g, s);
    if (rc0) return rc0;
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
    LaunchTimer timer(g);  // brackets the one traversal kernel launched below
    if (g.geom.hash && (g.algorithm == 1 || !(Q->fresh && g.grid.fresh))) {
        // a SpatialHashingCellList whose coordinates moved since the binning (the per-bucket
        // cells of the binning no longer hold: every candidate's cell is recomputed), or the
        // per-query algorithm asked for
        using TCELL = std::conditional_t<(sizeof(TC) == 8 || sizeof(TR) == 8), double, float>;
        const bool large_h = g.chunk < 0 ? !F::kQueue : g.chunk == 1;
        auto launch_h = [&](auto chunk, auto perc) {
            constexpr int CHK = decltype(chunk)::value;
            constexpr bool PER = decltype(perc)::value;
            hipLaunchKernelGGL((pn::k_traverse<D, TC, TR, F, CHK, true, TCELL, PER>), dim3(nblocks), dim3(pn::WAVE), 0,
                               g.stream, g.geom, Q->view(), g.grid.view(), r2, f, nblocks);
        };
        using LargeC = pn::IntC<Chunk<TC, F>::large>;
        using SmallC = pn::IntC<Chunk<TC, F>::small>;
        if (g.geom.periodic) {
            if (large_h) launch_h(LargeC{}, std::true_type{});
            else launch_h(SmallC{}, std::true_type{});
        } else {
            if (large_h) launch_h(LargeC{}, std::false_type{});
            else launch_h(SmallC{}, std::false_type{});
        }
        PN_CHECK_LAUNCH();
        return PN_OK;
    }
    if (g.geom.hash) {  // a SpatialHashingCellList binned from the current coordinates
        auto launch_hc = [&](auto gc, auto perc) {
            constexpr int GG = decltype(gc)::value;
            constexpr bool PER = decltype(perc)::value;
            hipLaunchKernelGGL((pn::k_cells<D, TC, TR, F, GG, PER, 0, true>), dim3(nblocks), dim3(pn::WAVE), 0,
                               g.stream, g.geom, Q->view(), g.grid.view(), r2, f, nblocks, pn::PairCache{});
        };
        using Large = pn::IntC<pn::CellGroups<TC, F>::large>;
        using Small = pn::IntC<pn::CellGroups<TC, F>::small>;
        if (g.geom.periodic) {
            if (g.chunk != 0) launch_hc(Large{}, std::true_type{});
            else launch_hc(Small{}, std::true_type{});
        } else {
            if (g.chunk != 0) launch_hc(Large{}, std::false_type{});
            else launch_hc(Small{}, std::false_type{});
        }
        PN_CHECK_LAUNCH();
        return PN_OK;
    }
    if (g.algorithm == 0 || g.geom.periodic) {  // only the cell kernel handles periodic boxes
        pn::PairCache pc{};
        int pcm = 0;  // pair cache mode of the launch
        auto launch = [&](auto gc, auto perc) {
            constexpr int GG = decltype(gc)::value;
            constexpr bool PER = decltype(perc)::value;
            if constexpr (F::kCache && !PER) {
                if (pcm == 1) {
                    hipLaunchKernelGGL((pn::k_cells<D, TC, TR, F, GG, PER, 1>), dim3(nblocks), dim3(pn::WAVE), 0,
                                       g.stream, g.geom, Q->view(), g.grid.view(), r2, f, nblocks, pc);
                    return;
                }
                if (pcm == 2) {
                    hipLaunchKernelGGL((pn::k_cells<D, TC, TR, F, GG, PER, 2>), dim3(nblocks), dim3(pn::WAVE), 0,
                                       g.stream, g.geom, Q->view(), g.grid.view(), r2, f, nblocks, pc);
                    return;
                }
            }
            hipLaunchKernelGGL((pn::k_cells<D, TC, TR, F, GG, PER, 0>), dim3(nblocks), dim3(pn::WAVE), 0, g.stream,
                               g.geom, Q->view(), g.grid.view(), r2, f, nblocks, pc);
        };
        using Large = pn::IntC<pn::CellGroups<TC, F>::large>;
        using Small = pn::IntC<pn::CellGroups<TC, F>::small>;
        if constexpr (F::kCache) {
            if (g.pair_cache && Q == &g.grid && !g.geom.periodic) {
                const int groups = g.chunk != 0 ? Large::value : Small::value;
                int rc = PN_OK;
                // per wavefront: its share of about `per_point` entries per point, at least one
                // cell's records
                const int64_t per_point = g.pc_per_point > 0 ? g.pc_per_point : (D == 3 ? 108 : (D == 2 ? 36 : 12));
                int64_t region = (Q->n_sel * per_point + nblocks - 1) / nblocks;
                if (g.pc_per_point == 0) region = std::max<int64_t>(region, pn::PC_CELL_MAX + 8);
                region = (region + 3) & ~int64_t(3);
                const int64_t cap = region * nblocks;
                if (cap / 4 > INT32_MAX) return fail(PN_ERR_UNSUPPORTED, "pair cache larger than 2^33 entries");
                if ((rc = g.pc_head.ensure((size_t)nwork * 8, g.stream)) ||
                    (rc = g.pc_data.ensure((size_t)cap * 2, g.stream)))
                    return rc;
                pc.data = g.pc_data.as<uint16_t>();
                pc.region = (long long)region;
                pc.head = g.pc_head.as<int>();
                if (g.pc_gen == g.gen && g.pc_groups == groups && g.pc_region == region) {
                    pcm = 2;
                } else {
                    pcm = 1;
                    g.pc_gen = g.gen;
                    g.pc_groups = groups;
                    g.pc_region = region;
                }
            }
        }
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

