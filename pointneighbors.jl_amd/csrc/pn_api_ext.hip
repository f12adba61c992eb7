// C-ABI entry points of the benchmark closures and per-point reductions: n-body
// acceleration (grid and precomputed lists), the WCSPH fluid-fluid interaction, and
// mapreduce_neighbor over a menu of value functions and operators.
#include "pn_host.h"
#include "pn_interact.h"

namespace {

// Dispatch over D in {2, 3} with one floating type for coordinates and radius.
template <class Fn>
int dispatch_same(const pn_grid& g, Fn&& fn) {
    if (g.tc != g.tr) return fail(PN_ERR_UNSUPPORTED, "this functor needs coordinates and radius of one type");
    if (g.nd != 2 && g.nd != 3) return fail(PN_ERR_UNSUPPORTED, "this functor is defined for 2D and 3D");
    auto by_t = [&](auto dc) -> int {
        if (g.tc == PN_FLOAT32) return fn(dc, 0.0f);
        return fn(dc, 0.0);
    };
    if (g.nd == 2) return by_t(pn::IntC<2>{});
    return by_t(pn::IntC<3>{});
}

}  // namespace

extern "C" {

int pn_nbody_accel(pn_grid_t grid, int queries, double G, const void* mass, void* dv, int accumulate) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    if (!mass || !dv) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    pn_grid& g = *grid;
    const BinSet* Q;
    int rc = check_queries(g, queries, Q);
    if (rc) return rc;
    return dispatch_same(g, [&](auto dc, auto tv) -> int {
        constexpr int D = decltype(dc)::value;
        using T = decltype(tv);
        pn::NBodyF<T, D> f;
        f.G = (T)G;
        f.dv = (T*)dv;
        f.accumulate = accumulate;
        auto src = fields<T>({{mass, 1}});
        return launch_traverse_t<D, T, T>(g, Q, f, &src);
    });
}

int pn_wcsph_interact(pn_grid_t grid, int kernel, double h, double sound_speed, double alpha, double beta,
                      double epsilon, double delta, const void* mass, const void* v, const void* pressure, void* dv,
                      int accumulate) {
    int rc = sph_precheck(grid, kernel, h);
    if (rc) return rc;
    if (!mass || !v || !pressure || !dv) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (!(sound_speed >= 0.0) || !(epsilon >= 0.0)) return fail(PN_ERR_ARGUMENT, "invalid WCSPH parameters");
    pn_grid& g = *grid;
    if ((rc = require_self(g, PN_QUERIES_GRID))) return rc;
    const BinSet* Q = &g.grid;
    return dispatch_same(g, [&](auto dc, auto tv) -> int {
        constexpr int D = decltype(dc)::value;
        using T = decltype(tv);
        pn::WCSPHInteractF<T, D> f;
        f.kern = pn::SmoothingKernel<T, D>::make(kernel, h);
        f.prm.h = (T)h;
        f.prm.c = (T)sound_speed;
        f.prm.alpha = (T)alpha;
        f.prm.beta = (T)beta;
        f.prm.eps_h2 = (T)(epsilon * h * h);
        f.prm.dhc = (T)(delta * h * sound_speed);
        f.vmat = (const T*)v;
        f.pressure = (const T*)pressure;
        f.dvmat = (T*)dv;
        f.accumulate = accumulate;
        const T* vm = (const T*)v;
        // fields in sorted order: m_b, rho_b, p_b, (Float64: 1 / rho_b,) v_b
        using F = pn::WCSPHInteractF<T, D>;
        pn::FieldSrc<T> src{};
        src.n = F::kFields;
        src.p[0] = (const T*)mass;
        src.step[0] = 1;
        src.p[1] = vm + D;
        src.step[1] = D + 1;
        src.p[2] = (const T*)pressure;
        src.step[2] = 1;
        if (F::kInvRho) {
            src.p[3] = vm + D;
            src.step[3] = D + 1;
            src.recip = 1u << 3;
        }
        for (int d = 0; d < D; ++d) {
            src.p[F::kV + d] = vm + d;
            src.step[F::kV + d] = D + 1;
        }
        return launch_traverse_t<D, T, T>(g, Q, f, &src);
    });
}

int pn_mapreduce_neighbors(pn_grid_t grid, int queries, int fn, int op, int index_base, double* value,
                           int32_t* count) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    if (!value || !count) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (fn < pn::MR_ONE || fn >= pn::MR_POSDIFF + grid->nd) return fail(PN_ERR_ARGUMENT, "unknown value function");
    if (op < pn::MR_ADD || op > pn::MR_MAX) return fail(PN_ERR_ARGUMENT, "unknown operator");
    pn::MapReduceF f;
    f.fn = fn;
    f.op = op;
    f.base = index_base;
    f.value = value;
    f.count = count;
    return launch_traverse(*grid, queries, f);
}

int pn_list_nbody_accel(const int32_t* row_ptr, const int32_t* col_idx, int64_t n_rows, int stored_base, int ndims,
                        int dtype, double G, const void* x, const void* y, const void* mass, int64_t sel_first,
                        int64_t sel_count, const int32_t* sel_list, void* dv, int accumulate, void* stream) {
    if (!row_ptr || !x || !y || !mass || !dv) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (!sel_list && (sel_first < 0 || sel_first + sel_count > n_rows))
        return fail(PN_ERR_ARGUMENT, "selection range out of bounds of the neighbor lists");
    if (dtype != PN_FLOAT32 && dtype != PN_FLOAT64) return fail(PN_ERR_ARGUMENT, "dtype must be PN_FLOAT32 or PN_FLOAT64");
    pn_grid tmp;  // dispatch only
    tmp.nd = ndims;
    tmp.tc = dtype;
    tmp.tr = dtype;
    if (sel_count <= 0) return PN_OK;
    const int nblocks = (int)round_up8((sel_count + 15) / 16);
    return dispatch_same(tmp, [&](auto dc, auto tv) -> int {
        constexpr int D = decltype(dc)::value;
        using T = decltype(tv);
        hipLaunchKernelGGL((pn::k_list_nbody<D, T, T>), dim3(nblocks), dim3(256), 0, (hipStream_t)stream, row_ptr,
                           col_idx, stored_base, sel_first, sel_list, sel_count, (const T*)x, (const T*)y,
                           (const T*)mass, (T)G, (T*)dv, accumulate, nblocks);
        PN_CHECK_LAUNCH();
        return PN_OK;
    });
}

}  // extern "C"
