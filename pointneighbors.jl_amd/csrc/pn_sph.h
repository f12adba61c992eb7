// Interaction functors driven by the traversal (the closures of a WCSPH step and of a
// differentiable pair energy) and their adjoints.
//
// The physics is not in the reference (it lives in TrixiParticles.jl); the formulas are
// the textbook ones:
//   cubic spline (Monaghan 1992), support 2h:   W = s (1 - 3/2 q^2 + 3/4 q^3)  q < 1
//                                                   s (2 - q)^3 / 4            1 <= q < 2
//       s = 1/(pi h^3) in 3D, 10/(7 pi h^2) in 2D
//   Wendland C2 (Wendland 1995), support 2h:    W = s (1 - q/2)^4 (2q + 1)     q < 2
//       s = 21/(16 pi h^3) in 3D, 7/(4 pi h^2) in 2D
//   Tait equation of state:  p = B ((rho / rho0)^gamma - 1), B = rho0 c^2 / gamma
//   pressure acceleration:   a_i = -sum_{j != i} m_j (P_i + P_j) grad_i W_ij, P = p / rho^2
//   Wendland C2 pair energy: phi(r) = eps (1 - r/rc)^4 (1 + 4 r/rc),  E = 1/2 sum_{i != j} phi
// Adjoints (reverse mode, neighbour sets held fixed; the traversal is recomputed):
//   density:  rho_i = sum_j m_j W_ij, upstream g:
//             xbar_k = sum_{j != k} (g_k m_j + g_j m_k) W'(r) (x_k - x_j) / r
//             mbar_k = sum_j g_j W_kj
//   pressure: a_i = -sum_j m_j c_ij G_ij, c = P_i + P_j, G_ij = W'(r) v, v = (x_i - x_j)/r,
//             upstream abar; u = m_j abar_k - m_k abar_j:
//             Pbar_k = -sum_j u . G_kj
//             mbar_k =  sum_j c_kj abar_j . G_kj
//             xbar_k = -sum_j c_kj H_kj u,  H = W'' v v^T + (W'/r) (I - v v^T)
//   energy:   upstream w per point of e_i = 1/2 sum_j phi_ij:
//             xbar_k = sum_{j != k} (w_k + w_j)/2 phi'(r) (x_k - x_j) / r
#pragma once

#include "pn_traverse.h"

namespace pn {

enum { KERNEL_CUBIC_SPLINE = 0, KERNEL_WENDLAND_C2 = 1 };

// Arithmetic of the pair bodies.  The neighbour predicate is evaluated elsewhere in the
// reference's exact arithmetic; the physics only needs its stated tolerance, so Float32
// uses the hardware square root / reciprocal square root (about 1 ulp) and fused
// multiply-adds.  Float64 keeps IEEE square root and division.
__device__ __forceinline__ float psqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double psqrt(double x) { return sqrt(x); }
// 1/sqrt(x) for x > 0, 0 for x == 0
__device__ __forceinline__ float prsqrt0(float x) { return x > 0.0f ? __builtin_amdgcn_rsqf(x) : 0.0f; }
// Float64: the hardware estimate (v_rsq_f64) refined by two Newton steps
// y <- y + y (1 - x y^2) / 2 in FMAs, within an ulp or so of the rounded 1/sqrt(x) at a third
// of the cost of a correctly rounded square root and a division; x outside [2^-1000, 2^1000]
// (denormals included) takes the division.
__device__ __forceinline__ double prsqrt0(double x) {
    if (!(x > 0.0)) return 0.0;
    if (x < 0x1p-1000 || x > 0x1p+1000) return 1.0 / sqrt(x);
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double e = __builtin_fma(-(x * y), y, 1.0);
        y = __builtin_fma(0.5 * y, e, y);
    }
    return y;
}
// sqrt(eps(T)) in T, Julia's threshold below which a distance counts as zero
template <class T>
__device__ __forceinline__ T sqrt_eps();
template <>
__device__ __forceinline__ float sqrt_eps<float>() {
    return 3.4526698e-4f;  // sqrt(2^-23)
}
template <>
__device__ __forceinline__ double sqrt_eps<double>() {
    return 1.4901161193847656e-8;  // sqrt(2^-52) = 2^-26
}
// 1/x: the hardware reciprocal (about 1 ulp) in Float32, IEEE division in Float64
__device__ __forceinline__ float prcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double prcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float pfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double pfma(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <class TR, int D>
struct SmoothingKernel {
    int kind;
    TR inv_h;
    TR sigma;  // normalization

    __host__ static SmoothingKernel make(int kind, double h) {
        SmoothingKernel k;
        k.kind = kind;
        k.inv_h = (TR)(1.0 / h);
        const double pi = 3.14159265358979323846;
        double s;
        if (kind == KERNEL_CUBIC_SPLINE) s = D == 3 ? 1.0 / (pi * h * h * h) : 10.0 / (7.0 * pi * h * h);
        else s = D == 3 ? 21.0 / (16.0 * pi * h * h * h) : 7.0 / (4.0 * pi * h * h);
        k.sigma = (TR)s;
        return k;
    }

    __device__ __forceinline__ TR sqrt_(TR x) const { return psqrt(x); }

    // W(r), W'(r) = dW/dr, W''(r) without branches.  The cubic spline is written with the
    // truncated powers of its B-spline form, a = max(2 - q, 0), b = max(1 - q, 0):
    //     W = s (a^3 - 4 b^3) / 4,  dW/dq = s (3 b^2 - 3/4 a^2),  d2W/dq2 = s (3/2 a - 6 b),
    // which equal the piecewise polynomials on q < 1, 1 <= q < 2 and q >= 2.
    __device__ __forceinline__ TR w(TR r) const {
        const TR q = r * inv_h;
        if (kind == KERNEL_CUBIC_SPLINE) {
            const TR a = max((TR)2 - q, (TR)0), b = max((TR)1 - q, (TR)0);
            const TR b3 = b * b * b;
            return (TR)0.25 * sigma * pfma(a * a, a, (TR)-4 * b3);
        }
        const TR t = max((TR)1 - (TR)0.5 * q, (TR)0);
        const TR t2 = t * t;
        return sigma * t2 * t2 * pfma((TR)2, q, (TR)1);
    }

    __device__ __forceinline__ TR dw(TR r) const {
        const TR q = r * inv_h;
        const TR s1 = sigma * inv_h;
        if (kind == KERNEL_CUBIC_SPLINE) {
            const TR a = max((TR)2 - q, (TR)0), b = max((TR)1 - q, (TR)0);
            return s1 * pfma((TR)3 * b, b, (TR)-0.75 * a * a);
        }
        const TR t = max((TR)1 - (TR)0.5 * q, (TR)0);
        return s1 * ((TR)-5 * q * t * t * t);
    }

    __device__ __forceinline__ TR d2w(TR r) const {
        const TR q = r * inv_h;
        const TR s2 = sigma * inv_h * inv_h;
        if (kind == KERNEL_CUBIC_SPLINE) {
            const TR a = max((TR)2 - q, (TR)0), b = max((TR)1 - q, (TR)0);
            return s2 * pfma((TR)1.5, a, (TR)-6 * b);
        }
        const TR t = max((TR)1 - (TR)0.5 * q, (TR)0);
        return s2 * (TR)5 * t * t * pfma((TR)2, q, (TR)-1);
    }
};

// rho_i (+)= sum_j m_j W(|x_i - y_j|), self included.  Fields: m_j.
template <class TR, int D>
struct DensityF : FunctorBase {
    using Acc = TR;
    static constexpr bool kQueue = true;
    static constexpr bool kCache = true;
    static constexpr int kFields = 1;
    SmoothingKernel<TR, D> kern;
    TR* rho;
    int accumulate;
    __device__ Acc init() const { return (TR)0; }
    template <class T, int DD>
    __device__ void pair(Acc& a, const QState&, const T (&)[DD], T d2, int, const T* fj, int*) const {
        a = pfma(fj[0], kern.w(psqrt(d2)), a);
    }
    __device__ void combine(Acc& a, const Acc& b) const { a += b; }
    __device__ void write(int64_t i, const QState&, const Acc& a) const { rho[i] = accumulate ? rho[i] + a : a; }
};

// dv_i (+)= -sum_{j != i} m_j (P_i + P_j) W'(r) (x_i - x_j) / r.  Fields: m_j, P_j.
template <class TR, int D>
struct PressureAccelF : FunctorBase {
    using Acc = Vec<TR, D>;
    static constexpr bool kQueue = true;
    static constexpr bool kCache = true;
    struct QState {
        TR P;
    };
    static constexpr int kFields = 2;
    SmoothingKernel<TR, D> kern;
    const TR* P;
    TR* dv;
    int accumulate;
    __device__ QState query(int64_t i) const { return QState{P[i]}; }
    __device__ Acc init() const {
        Acc a;
#pragma unroll
        for (int d = 0; d < D; ++d) a.v[d] = (TR)0;
        return a;
    }
    template <class T, int DD>
    __device__ void pair(Acc& a, const QState& qs, const T (&pd)[DD], T d2, int, const T* fj, int*) const {
        // j == i (d2 == 0) contributes nothing: inv_r = 0 there
        const T inv_r = prsqrt0(d2);
        const T r = d2 * inv_r;
        const T c = -fj[0] * (qs.P + fj[1]) * kern.dw(r) * inv_r;
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

// e_i (+)= 1/2 sum_{j != i} phi(r).
template <class TR, int D>
struct EnergyF : FunctorBase {
    using Acc = TR;
    static constexpr bool kQueue = true;
    static constexpr bool kCache = true;
    TR eps, inv_rc;
    TR* e;
    int accumulate;
    __device__ Acc init() const { return (TR)0; }
    template <class T, int DD>
    __device__ void pair(Acc& a, const QState&, const T (&)[DD], T d2, int, const T*, int*) const {
        const T s = psqrt(d2) * inv_rc;
        const T t = max((T)1 - s, (T)0);
        const T t2 = t * t;
        const T phi = eps * t2 * t2 * pfma((T)4, s, (T)1);
        a += d2 > (T)0 ? phi : (T)0;
    }
    __device__ void combine(Acc& a, const Acc& b) const { a += b; }
    __device__ void write(int64_t i, const QState&, const Acc& a) const {
        e[i] = accumulate ? e[i] + (TR)0.5 * a : (TR)0.5 * a;
    }
};

// xbar_k (+)= sum_{j != k} (w_k + w_j)/2 phi'(r) (x_k - x_j) / r.  Fields: w_j.
template <class TR, int D>
struct EnergyVjpF : FunctorBase {
    using Acc = Vec<TR, D>;
    static constexpr bool kQueue = true;
    struct QState {
        TR w;
    };
    static constexpr int kFields = 1;
    TR eps, inv_rc;
    const TR* wq;
    TR* xbar;
    int accumulate;
    __device__ QState query(int64_t i) const { return QState{wq[i]}; }
    __device__ Acc init() const {
        Acc a;
#pragma unroll
        for (int d = 0; d < D; ++d) a.v[d] = (TR)0;
        return a;
    }
    template <class T, int DD>
    __device__ void pair(Acc& a, const QState& qs, const T (&pd)[DD], T d2, int, const T* fj, int*) const {
        const T inv_r = prsqrt0(d2);
        const T s = d2 * inv_r * inv_rc;
        const T t = max((T)1 - s, (T)0);
        const T dphi = (T)-20 * eps * s * t * t * t * inv_rc;
        const T c = (T)0.5 * (qs.w + fj[0]) * dphi * inv_r;
#pragma unroll
        for (int d = 0; d < D; ++d) a.v[d] = pfma(c, pd[d], a.v[d]);
    }
    __device__ void combine(Acc& a, const Acc& b) const {
#pragma unroll
        for (int d = 0; d < D; ++d) a.v[d] += b.v[d];
    }
    __device__ void write(int64_t i, const QState&, const Acc& a) const {
#pragma unroll
        for (int d = 0; d < D; ++d) xbar[D * i + d] = accumulate ? xbar[D * i + d] + a.v[d] : a.v[d];
    }
};

// Density adjoint.  Fields: m_j, g_j.
template <class TR, int D>
struct DensityVjpF : FunctorBase {
    static constexpr bool kQueue = true;
    struct Acc {
        TR x[D];
        TR m;
    };
    struct QState {
        TR m, g;
    };
    static constexpr int kFields = 2;
    SmoothingKernel<TR, D> kern;
    const TR* mass;
    const TR* g;
    TR* xbar;
    TR* mbar;
    int accumulate;
    __device__ QState query(int64_t i) const { return QState{mass[i], g[i]}; }
    __device__ Acc init() const {
        Acc a;
#pragma unroll
        for (int d = 0; d < D; ++d) a.x[d] = (TR)0;
        a.m = (TR)0;
        return a;
    }
    template <class T, int DD>
    __device__ void pair(Acc& a, const QState& qs, const T (&pd)[DD], T d2, int, const T* fj, int*) const {
        const T inv_r = prsqrt0(d2);
        const T r = d2 * inv_r;
        a.m = pfma(fj[1], kern.w(r), a.m);
        const T c = (qs.g * fj[0] + fj[1] * qs.m) * kern.dw(r) * inv_r;
#pragma unroll
        for (int d = 0; d < D; ++d) a.x[d] = pfma(c, pd[d], a.x[d]);
    }
    __device__ void combine(Acc& a, const Acc& b) const {
#pragma unroll
        for (int d = 0; d < D; ++d) a.x[d] += b.x[d];
        a.m += b.m;
    }
    __device__ void write(int64_t i, const QState&, const Acc& a) const {
#pragma unroll
        for (int d = 0; d < D; ++d) xbar[D * i + d] = accumulate ? xbar[D * i + d] + a.x[d] : a.x[d];
        if (mbar) mbar[i] = accumulate ? mbar[i] + a.m : a.m;
    }
};

// Pressure-acceleration adjoint.  Fields: m_j, P_j, abar_j (D components).
template <class TR, int D>
struct PressureVjpF : FunctorBase {
    static constexpr bool kQueue = true;
    struct Acc {
        TR x[D];
        TR P;
        TR m;
    };
    struct QState {
        TR m, P;
        TR ab[D];
    };
    static constexpr int kFields = 2 + D;
    SmoothingKernel<TR, D> kern;
    const TR* mass;
    const TR* P;
    const TR* abar;
    TR* xbar;
    TR* Pbar;
    TR* mbar;
    int accumulate;
    __device__ QState query(int64_t i) const {
        QState q;
        q.m = mass[i];
        q.P = P[i];
#pragma unroll
        for (int d = 0; d < D; ++d) q.ab[d] = abar[D * i + d];
        return q;
    }
    __device__ Acc init() const {
        Acc a;
#pragma unroll
        for (int d = 0; d < D; ++d) a.x[d] = (TR)0;
        a.P = (TR)0;
        a.m = (TR)0;
        return a;
    }
    template <class T, int DD>
    __device__ void pair(Acc& a, const QState& qs, const T (&pd)[DD], T d2, int, const T* fj, int*) const {
        if (d2 > (T)0) {
            const T inv_r = prsqrt0(d2);
            const T r = d2 * inv_r;
            const T w1 = kern.dw(r);
            const T w2 = kern.d2w(r);
            const T mj = fj[0], Pj = fj[1];
            T v[D], u[D];
            T vu = (T)0, abG = (T)0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                v[d] = pd[d] * inv_r;
                u[d] = mj * qs.ab[d] - qs.m * fj[2 + d];
                vu += v[d] * u[d];
                abG += fj[2 + d] * v[d];
            }
            const T c = qs.P + Pj;
            a.P -= w1 * vu;
            a.m += c * w1 * abG;
            const T wr = w1 * inv_r;
#pragma unroll
            for (int d = 0; d < D; ++d) a.x[d] -= c * (w2 * vu * v[d] + wr * (u[d] - vu * v[d]));
        }
    }
    __device__ void combine(Acc& a, const Acc& b) const {
#pragma unroll
        for (int d = 0; d < D; ++d) a.x[d] += b.x[d];
        a.P += b.P;
        a.m += b.m;
    }
    __device__ void write(int64_t i, const QState&, const Acc& a) const {
#pragma unroll
        for (int d = 0; d < D; ++d) xbar[D * i + d] = accumulate ? xbar[D * i + d] + a.x[d] : a.x[d];
        if (Pbar) Pbar[i] = accumulate ? Pbar[i] + a.P : a.P;
        if (mbar) mbar[i] = accumulate ? mbar[i] + a.m : a.m;
    }
};

// Tait equation of state, pointwise: p = B ((rho / rho0)^gamma - 1), B = rho0 c^2 / gamma.
// over_rho2 != 0 writes the pressure-acceleration term P = p / rho^2 instead of p.
template <class TR>
__global__ __launch_bounds__(256) void k_tait_eos(const TR* __restrict__ rho, TR* __restrict__ p, int64_t n, TR rho0,
                                                  TR c, TR gamma, int over_rho2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const TR B = rho0 * c * c / gamma;
    const TR r = rho[i];
    const TR pi = B * (pow(r / rho0, gamma) - (TR)1);
    p[i] = over_rho2 ? pi / (r * r) : pi;
}

}  // namespace pn
