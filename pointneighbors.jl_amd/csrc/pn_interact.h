// Interaction functors of the reference's benchmark closures and per-point reductions:
//
//   NBodyF       benchmarks/n_body.jl:34-51 (benchmark_n_body):
//                    dv_i (+)= sum_{j, d >= sqrt(eps)} -G m_j (x_i - x_j) / d^3
//   WCSPHInteractF  the fluid-fluid interact! of TrixiParticles 0.5 that the reference's
//                benchmark_wcsph times (benchmarks/smoothed_particle_hydrodynamics.jl:44-102):
//                weakly compressible SPH with ContinuityDensity, a precomputed pressure,
//                Monaghan's artificial viscosity and Molteni-Colagrossi density diffusion.
//                Per pair (a, b), pos_diff = x_a - x_b, d = |pos_diff|, h the smoothing
//                length, gradW = W'(d)/d pos_diff (zero for d < sqrt(eps)):
//                    dv_a   += -m_b (p_a + p_b) / (rho_a rho_b) gradW           pressure
//                              -m_b Pi_ab gradW                                  viscosity
//                    drho_a += rho_a / rho_b m_b (v_a - v_b) . gradW             continuity
//                              + delta h c (m_b / rho_b) psi . gradW              diffusion
//                    Pi_ab = (-alpha c mu + beta mu^2) / rho_mean if (v_a - v_b) . pos_diff < 0,
//                    mu = h (v_a - v_b) . pos_diff / (d^2 + eps_v h^2), rho_mean = (rho_a + rho_b) / 2,
//                    psi = 2 (rho_a - rho_b) pos_diff / d^2
//                The velocity and density of a point are column a of v, a (D+1) x N matrix
//                (velocity rows, then the density: TrixiParticles' layout with
//                ContinuityDensity); dv has the same layout.
//   MapReduceF   mapreduce_neighbor (src/neighborhood_search.jl:340-383) for a menu of value
//                functions (1, neighbour index, distance, squared distance, a pos_diff
//                component) and operators (+, min, max), per query point, in double, with
//                the number of neighbours (an empty neighbourhood returns `init`).
#pragma once

#include "pn_sph.h"

namespace pn {

// dv_i (+)= -G sum_{j, d >= sqrt(eps(T))} m_j pos_diff / d^3.  Fields: m_j.
template <class TR, int D>
struct NBodyF : FunctorBase {
    using Acc = Vec<TR, D>;
    static constexpr bool kQueue = true;
    static constexpr int kFields = 1;
    TR G;
    TR* dv;
    int accumulate;
    __device__ Acc init() const {
        Acc a;
#pragma unroll
        for (int d = 0; d < D; ++d) a.v[d] = (TR)0;
        return a;
    }
    template <class T, int DD>
    __device__ void pair(Acc& a, const QState&, const T (&pd)[DD], T d2, int, const T* fj, int*) const {
        // 1/d^3 from the reciprocal square root (no division); d < sqrt(eps) adds nothing
        const T inv_d = prsqrt0(d2);
        const T dist = d2 * inv_d;
        const T c = dist < sqrt_eps<T>() ? (T)0 : -G * fj[0] * (inv_d * inv_d * inv_d);
#pragma unroll
        for (int d = 0; d < D; ++d) a.v[d] = pfma(c, pd[d], a.v[d]);
    }
    __device__ void combine(Acc& a, const Acc& b) const {
#pragma unroll
        for (int d = 0; d < D; ++d) a.v[d] += b.v[d];
    }
    __device__ void write(int64_t i, const QState&, const Acc& a) const {
#pragma unroll
        for (int d = 0; d < D; ++d) dv[D * i + d] = accumulate ? dv[D * i + d] + a.v[d] : a.v[d];
    }
};

// Constants of the WCSPH interaction.
template <class TR>
struct WCSPHParams {
    TR h;           // smoothing length
    TR c;           // sound speed
    TR alpha, beta;  // Monaghan viscosity
    TR eps_h2;      // epsilon * h^2 (Monaghan's singularity guard)
    TR dhc;         // delta * h * c (Molteni-Colagrossi diffusion)
};

// WCSPH fluid-fluid interaction.  Fields (per neighbour, sorted order): m_b, rho_b, p_b,
// in Float64 1 / rho_b (divided once per point by the field gather instead of once per
// pair), v_b (D values).  Query state: v_a, rho_a, p_a (reloaded where a flush needs it).
template <class TR, int D>
struct WCSPHInteractF : FunctorBase {
    struct Acc {
        TR v[D + 1];  // dv_a, then drho_a
    };
    static constexpr bool kQueue = true;
    static constexpr bool kInvRho = sizeof(TR) == 8;
    static constexpr int kV = kInvRho ? 4 : 3;  // field of v_b's first component
    static constexpr int kFields = kV + D;
    struct QState {
        TR v[D];
        TR rho, p, inv_rho;
    };
    SmoothingKernel<TR, D> kern;
    WCSPHParams<TR> prm;
    const TR* vmat;  // (D + 1) x N
    const TR* pressure;
    TR* dvmat;       // (D + 1) x N
    int accumulate;
    __device__ QState query(int64_t i) const {
        QState q;
#pragma unroll
        for (int d = 0; d < D; ++d) q.v[d] = vmat[(D + 1) * i + d];
        q.rho = vmat[(D + 1) * i + D];
        q.p = pressure[i];
        q.inv_rho = prcp(q.rho);
        return q;
    }
    __device__ Acc init() const {
        Acc a;
#pragma unroll
        for (int d = 0; d <= D; ++d) a.v[d] = (TR)0;
        return a;
    }
    template <class T, int DD>
    __device__ void pair(Acc& a, const QState& q, const T (&pd)[DD], T d2, int, const T* fj, int*) const {
        const T mb = fj[0], rhob = fj[1], pb = fj[2];
        // d >= sqrt(eps): inv_d = 1/d; the self pair and coincident points have no gradient
        const T inv_d = prsqrt0(d2);
        const T dist = d2 * inv_d;
        const T gw = dist < sqrt_eps<T>() ? (T)0 : kern.dw(dist) * inv_d;  // gradW = gw * pos_diff
        T vr = (T)0;
#pragma unroll
        for (int d = 0; d < D; ++d) vr = pfma(q.v[d] - fj[kV + d], pd[d], vr);
        const T inv_rhob = kInvRho ? fj[3] : prcp(rhob);
        // pressure and viscosity (the viscosity only for approaching pairs, selected per lane);
        // 1 / rho_mean = 2 / (rho_a + rho_b)
        T coef = -mb * (q.p + pb) * q.inv_rho * inv_rhob;
        // evaluated only when a lane of the wavefront has an approaching pair (a fluid at rest
        // skips it; the Float64 reciprocals are divisions)
        if (__builtin_amdgcn_ballot_w64(vr < (T)0)) {
            const T mu = prm.h * vr * prcp(d2 + prm.eps_h2);
            const T pi = pfma(prm.beta, mu, -prm.alpha * prm.c) * mu * ((T)2 * prcp(q.rho + rhob));
            coef = vr < (T)0 ? pfma(-mb, pi, coef) : coef;
        }
        const T s = coef * gw;
#pragma unroll
        for (int d = 0; d < D; ++d) a.v[d] = pfma(s, pd[d], a.v[d]);
        // continuity and density diffusion: psi . gradW = 2 (rho_a - rho_b) gw
        const T cont = q.rho * inv_rhob * mb * vr;
        const T diff = prm.dhc * mb * inv_rhob * (T)2 * (q.rho - rhob);
        a.v[D] = pfma(cont + diff, gw, a.v[D]);
    }
    __device__ void combine(Acc& a, const Acc& b) const {
#pragma unroll
        for (int d = 0; d <= D; ++d) a.v[d] += b.v[d];
    }
    __device__ void write(int64_t i, const QState&, const Acc& a) const {
#pragma unroll
        for (int d = 0; d <= D; ++d) {
            TR* o = dvmat + (D + 1) * i + d;
            *o = accumulate ? *o + a.v[d] : a.v[d];
        }
    }
};

enum { MR_ONE = 0, MR_INDEX = 1, MR_DISTANCE = 2, MR_DISTANCE2 = 3, MR_POSDIFF = 4 };  // + dim
enum { MR_ADD = 0, MR_MIN = 1, MR_MAX = 2 };

// mapreduce_neighbor per query: value[i] = op over the neighbours of f(i, j, pos_diff, d),
// count[i] = number of neighbours.  Values in double; the index base is added to j.
struct MapReduceF : FunctorBase {
    struct Acc {
        double v;
        int n;
    };
    static constexpr bool kIndex = true;
    int fn, op, base;
    double* value;
    int32_t* count;
    __device__ double ident() const {
        return op == MR_ADD ? 0.0 : (op == MR_MIN ? __builtin_huge_val() : -__builtin_huge_val());
    }
    __device__ double apply(double a, double b) const {
        return op == MR_ADD ? a + b : (op == MR_MIN ? fmin(a, b) : fmax(a, b));
    }
    __device__ Acc init() const { return Acc{ident(), 0}; }
    template <class T, int DD>
    __device__ void pair(Acc& a, const QState&, const T (&pd)[DD], T d2, int j, const T*, int*) const {
        double val;
        if (fn == MR_ONE) val = 1.0;
        else if (fn == MR_INDEX) val = (double)((long long)j + base);
        else if (fn == MR_DISTANCE) val = (double)sqrt(d2);  // IEEE sqrt in the radius type
        else if (fn == MR_DISTANCE2) val = (double)d2;
        else {
            // the components pass through an empty asm first: a selection among plain loads of
            // pd was folded into one load at a selected offset, which put pd in scratch memory
            T p0 = pd[0], p1 = pd[DD > 1 ? 1 : 0], p2 = pd[DD > 2 ? 2 : 0];
            asm("" : "+v"(p0), "+v"(p1), "+v"(p2));
            const int k = fn - MR_POSDIFF;
            val = (double)(k == 2 ? p2 : (k == 1 ? p1 : p0));
        }
        a.v = apply(a.v, val);
        a.n += 1;
    }
    __device__ void combine(Acc& a, const Acc& b) const {
        a.v = apply(a.v, b.v);
        a.n += b.n;
    }
    __device__ void write(int64_t i, const QState&, const Acc& a) const {
        value[i] = a.v;
        count[i] = a.n;
    }
};

// ------------------------------------------------------------------------------------
// n-body over precomputed lists (src/nhs_precomputed.jl:210-247: stored neighbours,
// distance recomputed, no radius test).  Sixteen lanes per selected row; the 16 partial
// sums are combined in a fixed butterfly order.
// ------------------------------------------------------------------------------------
template <int D, class TC, class TR>
__global__ __launch_bounds__(256) void k_list_nbody(const int32_t* __restrict__ row_ptr,
                                                    const int32_t* __restrict__ col, int stored_base,
                                                    int64_t sel_first, const int32_t* __restrict__ sel_list,
                                                    int64_t n_sel, const TC* __restrict__ x,
                                                    const TC* __restrict__ y, const TR* __restrict__ mass, TR G,
                                                    TR* __restrict__ dv, int accumulate, int nblocks) {
    constexpr int LANES = 16, ROWS = 256 / LANES, UNROLL = 2;
    const int sub = threadIdx.x & (LANES - 1);
    const int64_t a = (int64_t)xcd_remap(blockIdx.x, nblocks) * ROWS + (threadIdx.x / LANES);
    if (a >= n_sel) return;  // whole 16-lane groups leave together
    const int64_t i = sel_point(a, sel_first, sel_list);
    const int b = row_ptr[i], e = row_ptr[i + 1];
    TC xi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = x[(int64_t)D * i + d];
    TR s[D];
#pragma unroll
    for (int d = 0; d < D; ++d) s[d] = (TR)0;
    for (int k0 = b + sub; k0 < e; k0 += LANES * UNROLL) {
        int64_t j[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int k = k0 + LANES * u;
            j[u] = k < e ? (int64_t)col[k] - stored_base : -1;
        }
        TC yv[UNROLL][D];
        TR mv[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int64_t jj = j[u] >= 0 ? j[u] : 0;
#pragma unroll
            for (int d = 0; d < D; ++d) yv[u][d] = y[(int64_t)D * jj + d];
            mv[u] = mass[jj];
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            TR pd[D];
#pragma unroll
            for (int d = 0; d < D; ++d) pd[d] = (TR)(xi[d] - yv[u][d]);
            TR d2 = pd[0] * pd[0];
#pragma unroll
            for (int d = 1; d < D; ++d) {
                const TR p2 = pd[d] * pd[d];
                d2 = d2 + p2;
            }
            const TR inv_d = prsqrt0(d2);
            const TR dist = d2 * inv_d;
            const TR c = (j[u] < 0 || dist < sqrt_eps<TR>()) ? (TR)0 : -G * mv[u] * (inv_d * inv_d * inv_d);
#pragma unroll
            for (int d = 0; d < D; ++d) s[d] = pfma(c, pd[d], s[d]);
        }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
#pragma unroll
        for (int o = LANES / 2; o > 0; o >>= 1) s[d] += __shfl_xor(s[d], o, LANES);
    }
    if (sub == 0) {
#pragma unroll
        for (int d = 0; d < D; ++d) dv[(int64_t)D * i + d] = accumulate ? dv[(int64_t)D * i + d] + s[d] : s[d];
    }
}

}  // namespace pn
