// C-ABI physics entry points: SPH density, pressure acceleration, pair energy, Tait EOS.
#include "pn_host.h"

extern "C" {
int pn_sph_density(pn_grid_t grid, int queries, int kernel, double h, const void* mass, void* rho, int accumulate) {
    int rc = sph_precheck(grid, kernel, h);
    if (rc) return rc;
    if (!mass || !rho) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    pn_grid& g = *grid;
    const BinSet* Q;
    if ((rc = check_queries(g, queries, Q))) return rc;
    return dispatch_phys(g, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        pn::DensityF<TR, D> f;
        f.kern = pn::SmoothingKernel<TR, D>::make(kernel, h);
        f.rho = (TR*)rho;
        f.accumulate = accumulate;
        auto src = fields<TR>({{mass, 1}});
        return launch_traverse_t<D, TC, TR>(g, Q, f, &src);
    });
}

int pn_sph_pressure_accel(pn_grid_t grid, int kernel, double h, const void* mass, const void* P, void* dv,
                          int accumulate) {
    int rc = sph_precheck(grid, kernel, h);
    if (rc) return rc;
    if (!mass || !P || !dv) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    pn_grid& g = *grid;
    if ((rc = require_self(g, PN_QUERIES_GRID))) return rc;
    const BinSet* Q = &g.grid;
    return dispatch_phys(g, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        pn::PressureAccelF<TR, D> f;
        f.kern = pn::SmoothingKernel<TR, D>::make(kernel, h);
        f.P = (const TR*)P;
        f.dv = (TR*)dv;
        f.accumulate = accumulate;
        auto src = fields<TR>({{mass, 1}, {P, 1}});
        return launch_traverse_t<D, TC, TR>(g, Q, f, &src);
    });
}

int pn_pair_energy(pn_grid_t grid, int queries, double eps, double cutoff, void* e, int accumulate) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    if (!e) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (!(cutoff > 0.0)) return fail(PN_ERR_ARGUMENT, "cutoff must be positive");
    pn_grid& g = *grid;
    const BinSet* Q;
    int rc = check_queries(g, queries, Q);
    if (rc) return rc;
    return dispatch_phys(g, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        pn::EnergyF<TR, D> f;
        f.eps = (TR)eps;
        f.inv_rc = (TR)(1.0 / cutoff);
        f.e = (TR*)e;
        f.accumulate = accumulate;
        return launch_traverse_t<D, TC, TR>(g, Q, f, (const pn::FieldSrc<TR>*)nullptr);
    });
}

namespace {
int tait_eos(const void* rho, void* out, int64_t n, int dtype, double rho0, double c, double gamma, int over_rho2,
             void* stream) {
    if (!rho || !out) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (n <= 0) return PN_OK;
    if (dtype == PN_FLOAT32) {
        hipLaunchKernelGGL(pn::k_tait_eos<float>, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                           (const float*)rho, (float*)out, n, (float)rho0, (float)c, (float)gamma, over_rho2);
    } else if (dtype == PN_FLOAT64) {
        hipLaunchKernelGGL(pn::k_tait_eos<double>, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                           (const double*)rho, (double*)out, n, rho0, c, gamma, over_rho2);
    } else {
        return fail(PN_ERR_ARGUMENT, "dtype must be PN_FLOAT32 or PN_FLOAT64");
    }
    PN_CHECK_LAUNCH();
    return PN_OK;
}
}  // namespace

int pn_tait_eos(const void* rho, void* p, int64_t n, int dtype, double rho0, double c, double gamma, void* stream) {
    return tait_eos(rho, p, n, dtype, rho0, c, gamma, 0, stream);
}

int pn_tait_pressure_term(const void* rho, void* P, int64_t n, int dtype, double rho0, double c, double gamma,
                          void* stream) {
    return tait_eos(rho, P, n, dtype, rho0, c, gamma, 1, stream);
}

}  // extern "C"
