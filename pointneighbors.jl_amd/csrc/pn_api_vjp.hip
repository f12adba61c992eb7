// C-ABI adjoint entry points (reverse mode of the physics pair sums; the traversal is
// recomputed).
#include "pn_host.h"

extern "C" {
int pn_sph_density_vjp(pn_grid_t grid, int kernel, double h, const void* mass, const void* rho_bar, void* x_bar,
                       void* mass_bar, int accumulate) {
    int rc = sph_precheck(grid, kernel, h);
    if (rc) return rc;
    if (!mass || !rho_bar || !x_bar) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    pn_grid& g = *grid;
    if ((rc = require_self(g, PN_QUERIES_GRID))) return rc;
    const BinSet* Q = &g.grid;
    return dispatch_phys(g, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        pn::DensityVjpF<TR, D> f;
        f.kern = pn::SmoothingKernel<TR, D>::make(kernel, h);
        f.mass = (const TR*)mass;
        f.g = (const TR*)rho_bar;
        f.xbar = (TR*)x_bar;
        f.mbar = (TR*)mass_bar;
        f.accumulate = accumulate;
        auto src = fields<TR>({{mass, 1}, {rho_bar, 1}});
        return launch_traverse_t<D, TC, TR>(g, Q, f, &src);
    });
}

int pn_sph_pressure_accel_vjp(pn_grid_t grid, int kernel, double h, const void* mass, const void* P,
                              const void* dv_bar, void* x_bar, void* P_bar, void* mass_bar, int accumulate) {
    int rc = sph_precheck(grid, kernel, h);
    if (rc) return rc;
    if (!mass || !P || !dv_bar || !x_bar) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    pn_grid& g = *grid;
    if ((rc = require_self(g, PN_QUERIES_GRID))) return rc;
    const BinSet* Q = &g.grid;
    return dispatch_phys(g, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        pn::PressureVjpF<TR, D> f;
        f.kern = pn::SmoothingKernel<TR, D>::make(kernel, h);
        f.mass = (const TR*)mass;
        f.P = (const TR*)P;
        f.abar = (const TR*)dv_bar;
        f.xbar = (TR*)x_bar;
        f.Pbar = (TR*)P_bar;
        f.mbar = (TR*)mass_bar;
        f.accumulate = accumulate;
        const TR* ab = (const TR*)dv_bar;
        auto src = D == 3 ? fields<TR>({{mass, 1}, {P, 1}, {ab, 3}, {ab + 1, 3}, {ab + 2, 3}})
                          : fields<TR>({{mass, 1}, {P, 1}, {ab, 2}, {ab + 1, 2}});
        return launch_traverse_t<D, TC, TR>(g, Q, f, &src);
    });
}

int pn_pair_energy_vjp(pn_grid_t grid, double eps, double cutoff, const void* e_bar, void* x_bar, int accumulate) {
    if (!grid) return fail(PN_ERR_ARGUMENT, "grid is NULL");
    if (!e_bar || !x_bar) return fail(PN_ERR_ARGUMENT, "NULL pointer");
    if (!(cutoff > 0.0)) return fail(PN_ERR_ARGUMENT, "cutoff must be positive");
    pn_grid& g = *grid;
    int rc;
    if ((rc = require_self(g, PN_QUERIES_GRID))) return rc;
    const BinSet* Q = &g.grid;
    return dispatch_phys(g, [&](auto dc, auto tcv, auto trv) -> int {
        constexpr int D = decltype(dc)::value;
        using TC = decltype(tcv);
        using TR = decltype(trv);
        pn::EnergyVjpF<TR, D> f;
        f.eps = (TR)eps;
        f.inv_rc = (TR)(1.0 / cutoff);
        f.wq = (const TR*)e_bar;
        f.xbar = (TR*)x_bar;
        f.accumulate = accumulate;
        auto src = fields<TR>({{e_bar, 1}});
        return launch_traverse_t<D, TC, TR>(g, Q, f, &src);
    });
}

}  // extern "C"
