// Oracle: a CPU restatement of PointNeighbors.jl v0.6.7 for the
// GridNeighborhoodSearch + FullGridCellList + PrecomputedNeighborhoodSearch path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
// pointneighbors.jl_amd/.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product path never does.
//
// Every function names the reference file:line it restates.  Arithmetic follows
// Julia semantics exactly: IEEE single/double, no fused multiply-add (build with
// -ffp-contract=off), left-to-right dot product (StaticArrays `dot` unrolled as
// ((p1*p1 + p2*p2) + p3*p3)), `r^2` evaluated as r*r, IEEE sqrt and division.
//
// Pinning (what checks each function, see tests/test_oracle_golden.py):
//   orc_floor_to_int*        test/nhs_grid.jl:62-86 saturation cases (fixture)
//   orc_fullgrid_geometry    test/cell_lists/full_grid.jl:25-29 padded corners (fixture)
//   orc_grid_build           test/nhs_grid.jl:197-225 FullGrid candidate list with
//                            eachindex_y (fixture); bounds/overflow errors
//                            test/cell_lists/full_grid.jl:31-64, test/vector_of_vectors.jl:53-58
//   orc_candidates_dict      test/nhs_grid.jl:88-195 DictionaryCellList candidate lists
//   orc_traverse_*           brute force (orc_trivial_*) on jittered clouds, as
//                            test/neighborhood_search.jl:186-469 does; closed-form
//                            lattice counts; test/gpu.jl:30-36 extrema (11, 29)
//   orc_trivial_*            numpy brute force on small clouds (independent code)
//   the d^2 evaluation order parity unpinned against the reference (StaticArrays is
//                            not vendored); pinned to the stated reading by a numpy
//                            step-rounded evaluation on adversarial pairs
//   orc_tlsph_force (stress)  parity unpinned against TrixiParticles (absent); pinned by the
//                            exact discrete energy identity m_a dv_a = -dE/dx_a with
//                            E = sum_c V_c Psi_SVK(F_c), F from orc_tlsph_grad (finite
//                            differences), two-point tension/compression signs, momentum
//   orc_tlsph_force (penalty) parity unpinned; pinned by momentum conservation, zero force
//                            under affine maps with F = A, a compressed pair repelling
//   orc_tlsph_pk1_rho2       pinned by the same energy identity (P = dPsi/dF)
//   orc_sph_* / orc_energy_* parity unpinned against the reference (the physics lives in
//                            TrixiParticles, absent); pinned by closed-form single-pair
//                            values, symmetry, and forward-mode duals vs finite differences

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr int OK = 0;
constexpr int ERR_ARG = 1;
constexpr int ERR_OUT_OF_BOUNDS = 2;   // src/cell_lists/full_grid.jl:187-213
constexpr int ERR_CELL_FULL = 3;       // src/vector_of_vectors.jl:114-121
constexpr int ERR_QUERY_BOUNDS = 4;    // BoundsError from LinearIndices in traversal

// src/util.jl:19-34.  Julia compares Float and Int exactly; 2^63 is exact in both
// float and double, so `rounded >= typemax(Int)` is `rounded >= 2^63`.
template <class T>
int64_t floor_to_int(T x) {
    T rounded = std::floor(x);
    if (std::isnan(rounded) || rounded >= T(9223372036854775808.0)) return INT64_MAX;
    if (rounded <= T(-9223372036854775808.0)) return INT64_MIN;
    return (int64_t)rounded;
}

// Int64 `+ 1` wraps in Julia (test/nhs_grid.jl:83-85).
inline int64_t wrap_add(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }

struct Geom {
    int ndims;
    int64_t n[3];
    double minc[3];   // padded min corner, exact in the cell type
    double cs;        // search radius, exact in the cell type
    double csd[3];    // cell size per dimension (the search radius unless periodic)
    bool periodic = false;
    int64_t per[3];   // periodic cells per dimension (src/nhs_grid.jl:116-117)
    double box[3];    // periodic box size, exact in T_r
    int64_t ncells() const {
        int64_t c = 1;
        for (int d = 0; d < ndims; ++d) c *= n[d];
        return c;
    }
};

// Julia `mod(a, n)` for Int64 with n > 0: floored, result in [0, n).
inline int64_t floored_mod(int64_t a, int64_t n) {
    int64_t r = a % n;
    return r < 0 ? r + n : r;
}

// src/nhs_grid.jl:609-620 periodic_cell_index: mod.(cell .- 2, n_cells) .+ 2 (Int64 wrap)
inline int64_t periodic_wrap(int64_t c, int64_t n) { return floored_mod(wrap_add(c, -2), n) + 2; }

// src/cell_lists/full_grid.jl:84-94: floor_to_int.((x .- min_corner) ./ cell_size) .+ 1
// The subtraction and division run in TCELL = promote(T_coords, T_min_corner).  With a
// periodic box the result is wrapped (cell_coords, src/nhs_grid.jl:622-628).
template <class TC, class TCELL>
void cell_coords(const Geom& g, const TC* xp, int64_t cell[3]) {
    for (int d = 0; d < g.ndims; ++d) {
        TCELL q = ((TCELL)xp[d] - (TCELL)g.minc[d]) / (TCELL)g.csd[d];
        cell[d] = wrap_add(floor_to_int<TCELL>(q), 1);
        if (g.periodic) cell[d] = periodic_wrap(cell[d], g.per[d]);
    }
}

// compute_periodic_distance, src/neighborhood_search.jl:428-436: if d2 > r^2, take the
// minimum image pos_diff -= size .* round.(pos_diff ./ size) (round half to even) and
// recompute d2 = dot(pos_diff, pos_diff).
template <class TR>
inline TR periodic_distance(const Geom& g, TR pd[3], TR d2, TR r2) {
    if (!g.periodic || !(d2 > r2)) return d2;
    for (int d = 0; d < g.ndims; ++d) {
        const TR L = (TR)g.box[d];
        const TR k = std::nearbyint(pd[d] / L);
        const TR s = L * k;
        pd[d] = pd[d] - s;
    }
    TR e = pd[0] * pd[0];
    for (int d = 1; d < g.ndims; ++d) {
        TR p = pd[d] * pd[d];
        e = e + p;
    }
    return e;
}

// src/cell_lists/full_grid.jl:187-213: every coordinate in 2:(n_d - 1)
inline bool cell_in_bounds(const Geom& g, const int64_t cell[3]) {
    for (int d = 0; d < g.ndims; ++d)
        if (cell[d] < 2 || cell[d] > g.n[d] - 1) return false;
    return true;
}

// LinearIndices (column-major, first dimension fastest), 0-based result.
inline int64_t linear0(const Geom& g, const int64_t cell[3]) {
    int64_t c = 0, stride = 1;
    for (int d = 0; d < g.ndims; ++d) {
        c += (cell[d] - 1) * stride;
        stride *= g.n[d];
    }
    return c;
}

// pos_diff = convert.(T_r, x_i - x_j); d2 = dot(pos_diff, pos_diff)  (src/nhs_grid.jl:547-549)
template <class TC, class TR>
inline TR dist2(int ndims, const TC* xi, const TC* xj, TR pd[3]) {
    for (int d = 0; d < ndims; ++d) pd[d] = (TR)(xi[d] - xj[d]);
    TR d2 = pd[0] * pd[0];
    for (int d = 1; d < ndims; ++d) {
        TR p = pd[d] * pd[d];
        d2 = d2 + p;
    }
    return d2;
}

// The DynamicVectorOfVectors column-per-cell storage (src/vector_of_vectors.jl:3-31).
struct CellStore {
    int64_t cap;                 // max_points_per_cell
    std::vector<int32_t> lengths;
    std::vector<int32_t> backend;  // cap x ncells, column per cell
};

struct Err {
    int64_t point = -1;
    int64_t cell[3] = {0, 0, 0};
};

// src/nhs_grid.jl:227-253 (serial build, insertion order) with the FullGrid push
// src/cell_lists/full_grid.jl:112-121 and pushat! src/vector_of_vectors.jl:70-85.
// The ParallelUpdate build (src/nhs_grid.jl:255-281) fills the same sets in a
// nondeterministic order; sorted neighbour sets are identical.
template <class TC, class TCELL>
int build_serial(const Geom& g, const TC* y, const int64_t* active, int64_t n_active,
                 CellStore& cs, Err& err) {
    const int D = g.ndims;
    cs.lengths.assign((size_t)g.ncells(), 0);
    cs.backend.assign((size_t)(cs.cap * g.ncells()), 0);
    for (int64_t a = 0; a < n_active; ++a) {
        int64_t p = active[a];
        int64_t cell[3];
        cell_coords<TC, TCELL>(g, y + D * p, cell);
        if (!cell_in_bounds(g, cell)) {
            err.point = p;
            for (int d = 0; d < D; ++d) err.cell[d] = cell[d];
            return ERR_OUT_OF_BOUNDS;
        }
        int64_t c = linear0(g, cell);
        int32_t len = ++cs.lengths[c];
        if (len > cs.cap) {
            --cs.lengths[c];
            err.point = p;
            return ERR_CELL_FULL;
        }
        cs.backend[(size_t)(c * cs.cap + (len - 1))] = (int32_t)p;
    }
    return OK;
}

// src/nhs_grid.jl:519-575 with neighboring_cells :577-583 (CartesianIndices,
// first dimension fastest) and points_in_cell :593-597.  Calls f(j, pd, d2) for
// every accepted neighbour in the reference's traversal order.
template <class TC, class TR, class TCELL, class F>
int traverse_point(const Geom& g, const CellStore& cs, const TC* y, const TC* xi, TR r, F&& f) {
    const int D = g.ndims;
    int64_t cell[3];
    cell_coords<TC, TCELL>(g, xi, cell);
    // LinearIndices getindex is bounds-checked in the safe variant: all 3^D neighbour
    // cells must exist, i.e. the query cell lies in 2:(n_d - 1).
    for (int d = 0; d < D; ++d)
        if (cell[d] < 2 || cell[d] > g.n[d] - 1) return ERR_QUERY_BOUNDS;
    const TR r2 = r * r;
    int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) {
        lo[d] = d < D ? cell[d] - 1 : 1;
        hi[d] = d < D ? cell[d] + 1 : 1;
    }
    for (int64_t c3 = lo[2]; c3 <= hi[2]; ++c3)
        for (int64_t c2 = lo[1]; c2 <= hi[1]; ++c2)
            for (int64_t c1 = lo[0]; c1 <= hi[0]; ++c1) {
                int64_t nc[3] = {c1, c2, c3};
                if (g.periodic)  // points_in_cell wraps the neighbour cell, src/nhs_grid.jl:593-597
                    for (int d = 0; d < D; ++d) nc[d] = periodic_wrap(nc[d], g.per[d]);
                int64_t c = linear0(g, nc);
                const int32_t* col = cs.backend.data() + c * cs.cap;
                for (int32_t k = 0; k < cs.lengths[c]; ++k) {
                    int32_t j = col[k];
                    TR pd[3] = {0, 0, 0};
                    TR d2 = dist2<TC, TR>(D, xi, y + D * (int64_t)j, pd);
                    d2 = periodic_distance<TR>(g, pd, d2, r2);
                    if (d2 <= r2) f(j, pd, d2);
                }
            }
    return OK;
}

Geom make_geom(int ndims, const double* minc, double cs, const int64_t* n) {
    Geom g;
    g.ndims = ndims;
    for (int d = 0; d < 3; ++d) {
        g.n[d] = d < ndims ? n[d] : 1;
        g.minc[d] = d < ndims ? minc[d] : 0.0;
        g.csd[d] = cs;
        g.per[d] = 1;
        g.box[d] = 0.0;
    }
    g.cs = cs;
    return g;
}

// One NHS instance kept alive across calls so that build and traversal are separate,
// like initialize!/update! and foreach_point_neighbor.
struct OracleGrid {
    Geom g;
    int tc, tr, tcell;   // 0 = Float32, 1 = Float64
    CellStore cs;
    double radius;
    bool built = false;
};

template <class TC, class TR, class TCELL>
struct Impl {
    static int build(OracleGrid& og, const void* y, const int64_t* active, int64_t n_active,
                     int64_t* err_point, int64_t* err_cell) {
        Err e;
        int st = build_serial<TC, TCELL>(og.g, (const TC*)y, active, n_active, og.cs, e);
        if (err_point) *err_point = e.point;
        if (err_cell)
            for (int d = 0; d < og.g.ndims; ++d) err_cell[d] = e.cell[d];
        og.built = (st == OK);
        return st;
    }

    // counts[i] (for i in points) and, if rowptr != null, neighbour lists in
    // traversal order.  The two-pass form mirrors count-then-fill.
    static int traverse(const OracleGrid& og, const void* xv, const void* yv,
                        const int64_t* points, int64_t n_points, int32_t* counts,
                        int64_t* index_sums, const int64_t* rowptr, int32_t* cols,
                        int64_t* err_point) {
        const TC* x = (const TC*)xv;
        const TC* y = (const TC*)yv;
        const int D = og.g.ndims;
        const TR r = (TR)og.radius;
        for (int64_t a = 0; a < n_points; ++a) {
            int64_t i = points[a];
            int32_t cnt = 0;
            int64_t isum = 0;
            int64_t pos = rowptr ? rowptr[a] : 0;
            int st = traverse_point<TC, TR, TCELL>(og.g, og.cs, y, x + D * i, r,
                                                   [&](int32_t j, const TR*, TR) {
                                                       ++cnt;
                                                       isum += j;
                                                       if (rowptr) cols[pos++] = j;
                                                   });
            if (st != OK) {
                if (err_point) *err_point = i;
                return st;
            }
            if (counts) counts[a] = cnt;
            if (index_sums) index_sums[a] = isum;
        }
        return OK;
    }
};

}  // namespace

extern "C" {

int64_t orc_floor_to_int_f64(double x) { return floor_to_int<double>(x); }
int64_t orc_floor_to_int_f32(float x) { return floor_to_int<float>(x); }

// FullGridCellList constructor geometry, src/cell_lists/full_grid.jl:48-82.
//   user_dtype: element type of the user's min/max corner (0 Float32, 1 Float64)
//   r_dtype:    type of search_radius
// The padding `1001 // 1000 * search_radius` evaluates in T_r
// (Rational{Int} * Float promotes to the float type, which converts 1001//1000 as
// Float(1001)/Float(1000)); `min_corner .- pad` evaluates in promote(T_user, T_r).
// n_cells = ceil.(Int, (max_corner .- min_corner) ./ search_radius).
// Returns the corner type (0/1) or -1 for a template (search_radius < eps()).
int orc_fullgrid_geometry(int ndims, int user_dtype, int r_dtype, const double* user_min,
                          const double* user_max, double radius, double* out_min,
                          double* out_max, int64_t* out_n) {
    int tm = (user_dtype == 1 || r_dtype == 1) ? 1 : 0;
    for (int d = 0; d < ndims; ++d) {
        if (r_dtype == 0) {
            float pad = (1001.0f / 1000.0f) * (float)radius;
            if (tm == 0) {
                out_min[d] = (double)((float)user_min[d] - pad);
                out_max[d] = (double)((float)user_max[d] + pad);
            } else {
                out_min[d] = user_min[d] - (double)pad;
                out_max[d] = user_max[d] + (double)pad;
            }
        } else {
            double pad = (1001.0 / 1000.0) * radius;
            out_min[d] = user_min[d] - pad;
            out_max[d] = user_max[d] + pad;
        }
    }
    if (radius < 2.220446049250313e-16) return -1;
    for (int d = 0; d < ndims; ++d) {
        if (tm == 0) {
            float q = ((float)out_max[d] - (float)out_min[d]) / (float)radius;
            out_n[d] = (int64_t)std::ceil(q);
        } else {
            double q = (out_max[d] - out_min[d]) / radius;
            out_n[d] = (int64_t)std::ceil(q);
        }
    }
    return tm;
}

// Cell coordinates of one point, src/cell_lists/full_grid.jl:84-94 (1-based, Int64 wrap).
void orc_cell_coords(int ndims, int tc, int tcell, const void* x, const double* minc,
                     double cs, int64_t* out) {
    const int64_t ones[3] = {1, 1, 1};
    Geom g = make_geom(ndims, minc, cs, ones);
    if (tc == 0 && tcell == 0) cell_coords<float, float>(g, (const float*)x, out);
    else if (tc == 0) cell_coords<float, double>(g, (const float*)x, out);
    else cell_coords<double, double>(g, (const double*)x, out);
}

// DictionaryCellList cell coordinates, src/nhs_grid.jl:636-638: floor_to_int.(x ./ cs).
// Used only to pin the floor-based cell indexing against test/nhs_grid.jl:88-195.
// Writes the sorted candidate list (points of the 3^D cells around `xq`) to out and
// returns its length.
int64_t orc_candidates_dict(int ndims, const double* y, int64_t n, const double* xq,
                            double cs, int64_t* out) {
    std::vector<std::vector<int64_t>> cellc((size_t)n, std::vector<int64_t>(3, 0));
    for (int64_t j = 0; j < n; ++j)
        for (int d = 0; d < ndims; ++d) cellc[j][d] = floor_to_int<double>(y[ndims * j + d] / cs);
    int64_t qc[3];
    for (int d = 0; d < ndims; ++d) qc[d] = floor_to_int<double>(xq[d] / cs);
    int64_t m = 0;
    for (int64_t j = 0; j < n; ++j) {
        bool ok = true;
        for (int d = 0; d < ndims; ++d)
            if (cellc[j][d] < qc[d] - 1 || cellc[j][d] > qc[d] + 1) ok = false;
        if (ok) out[m++] = j;
    }
    return m;
}

void* orc_grid_create(int ndims, int tc, int tr, int tcell, const double* minc,
                      const int64_t* n, double radius, int64_t max_points_per_cell) {
    OracleGrid* og = new OracleGrid();
    og->g = make_geom(ndims, minc, radius, n);
    og->tc = tc;
    og->tr = tr;
    og->tcell = tcell;
    og->radius = radius;
    og->cs.cap = max_points_per_cell;
    return og;
}

void orc_grid_destroy(void* h) { delete (OracleGrid*)h; }

// GridNeighborhoodSearch with a PeriodicBox (src/nhs_grid.jl:102-125): per-dimension
// cell size (size ./ n_cells, exact in T_r), periodic cell counts and the box size.
void orc_grid_set_periodic(void* h, const double* cell_size, const int64_t* n_periodic,
                           const double* box_size) {
    OracleGrid& og = *(OracleGrid*)h;
    og.g.periodic = true;
    for (int d = 0; d < og.g.ndims; ++d) {
        og.g.csd[d] = cell_size[d];
        og.g.per[d] = n_periodic[d];
        og.g.box[d] = box_size[d];
    }
}

// cell_coords of one point with the grid's geometry (periodic wrap included).
void orc_grid_cell_coords(void* h, const void* x, int64_t* out) {
    OracleGrid& og = *(OracleGrid*)h;
    if (og.tc == 0 && og.tcell == 0) cell_coords<float, float>(og.g, (const float*)x, out);
    else if (og.tc == 0) cell_coords<float, double>(og.g, (const float*)x, out);
    else cell_coords<double, double>(og.g, (const double*)x, out);
}

// Periodic grid sizing, src/nhs_grid.jl:109-117:
//   n_cells = floor.(Int, (size .+ 10eps()) / search_radius)   (Float64: 10eps() is Float64)
//   cell_size = size ./ n_cells                                  (in T_r)
void orc_periodic_sizing(int ndims, int r_dtype, const double* box_size, double radius,
                         int64_t* out_n, double* out_cs) {
    for (int d = 0; d < ndims; ++d) {
        const double q = (box_size[d] + 10.0 * 2.220446049250313e-16) / radius;
        out_n[d] = (int64_t)std::floor(q);
        if (r_dtype == 0) out_cs[d] = (double)((float)box_size[d] / (float)out_n[d]);
        else out_cs[d] = box_size[d] / (double)out_n[d];
    }
}

#define ORC_DISPATCH(og, CALL)                                                        \
    do {                                                                              \
        if (og.tc == 0 && og.tr == 0 && og.tcell == 0) return Impl<float, float, float>::CALL; \
        if (og.tc == 0 && og.tr == 0) return Impl<float, float, double>::CALL;       \
        if (og.tc == 0 && og.tr == 1) return Impl<float, double, double>::CALL;      \
        if (og.tc == 1 && og.tr == 0) return Impl<double, float, double>::CALL;      \
        return Impl<double, double, double>::CALL;                                    \
    } while (0)

// initialize!/update! with ParallelUpdate semantics (src/nhs_grid.jl:220-292, :470-477).
int orc_grid_build(void* h, const void* y, const int64_t* active, int64_t n_active,
                   int64_t* err_point, int64_t* err_cell) {
    OracleGrid& og = *(OracleGrid*)h;
    if (og.radius < 2.220446049250313e-16) {  // src/nhs_grid.jl:263-267
        og.cs.lengths.assign((size_t)og.g.ncells(), 0);
        og.built = true;
        return OK;
    }
    ORC_DISPATCH(og, build(og, y, active, n_active, err_point, err_cell));
}

// Lengths of each cell list after a build (DVoV `lengths`).
void orc_grid_lengths(void* h, int32_t* out) {
    OracleGrid& og = *(OracleGrid*)h;
    std::memcpy(out, og.cs.lengths.data(), og.cs.lengths.size() * sizeof(int32_t));
}

// Points of one cell (0-based linear index) in insertion order; returns the length.
int32_t orc_grid_cell(void* h, int64_t c, int32_t* out) {
    OracleGrid& og = *(OracleGrid*)h;
    int32_t len = og.cs.lengths[c];
    std::memcpy(out, og.cs.backend.data() + c * og.cs.cap, (size_t)len * sizeof(int32_t));
    return len;
}

// foreach_point_neighbor counting and sum-of-indices (the closures of
// benchmarks/count_neighbors.jl:24-27 and test/neighborhood_search.jl:389-399),
// optionally writing neighbour lists in traversal order.
int orc_traverse(void* h, const void* x, const void* y, const int64_t* points,
                 int64_t n_points, int32_t* counts, int64_t* index_sums, const int64_t* rowptr,
                 int32_t* cols, int64_t* err_point) {
    OracleGrid& og = *(OracleGrid*)h;
    ORC_DISPATCH(og, traverse(og, x, y, points, n_points, counts, index_sums, rowptr, cols,
                              err_point));
}

// TrivialNeighborhoodSearch (src/nhs_trivial.jl:55 with the generic loop
// src/neighborhood_search.jl:390-421): every y point is a candidate.
int orc_trivial_count(int ndims, int tc, int tr, const void* xv, const void* yv, int64_t nx,
                      int64_t ny, double radius, int32_t* counts, const int64_t* rowptr,
                      int32_t* cols) {
    auto run = [&](auto tcv, auto trv) {
        using TC = decltype(tcv);
        using TR = decltype(trv);
        const TC* x = (const TC*)xv;
        const TC* y = (const TC*)yv;
        const TR r = (TR)radius;
        const TR r2 = r * r;
        for (int64_t i = 0; i < nx; ++i) {
            int32_t cnt = 0;
            int64_t pos = rowptr ? rowptr[i] : 0;
            for (int64_t j = 0; j < ny; ++j) {
                TR pd[3] = {0, 0, 0};
                TR d2 = dist2<TC, TR>(ndims, x + ndims * i, y + ndims * j, pd);
                if (d2 <= r2) {
                    ++cnt;
                    if (rowptr) cols[pos++] = (int32_t)j;
                }
            }
            if (counts) counts[i] = cnt;
        }
    };
    if (tc == 0 && tr == 0) run(0.0f, 0.0f);
    else if (tc == 0) run(0.0f, 0.0);
    else if (tr == 0) run(0.0, 0.0f);
    else run(0.0, 0.0);
    return OK;
}

// TrivialNeighborhoodSearch with a PeriodicBox (src/nhs_trivial.jl:55, generic loop
// src/neighborhood_search.jl:390-421 with compute_periodic_distance :428-436).
int orc_trivial_count_periodic(int ndims, int tc, int tr, const void* xv, const void* yv, int64_t nx,
                               int64_t ny, double radius, const double* box_size, int32_t* counts,
                               const int64_t* rowptr, int32_t* cols) {
    const int64_t ones[3] = {1, 1, 1};
    const double zeros[3] = {0, 0, 0};
    Geom g = make_geom(ndims, zeros, radius, ones);
    g.periodic = true;
    for (int d = 0; d < ndims; ++d) g.box[d] = box_size[d];
    auto run = [&](auto tcv, auto trv) {
        using TC = decltype(tcv);
        using TR = decltype(trv);
        const TC* x = (const TC*)xv;
        const TC* y = (const TC*)yv;
        const TR r = (TR)radius;
        const TR r2 = r * r;
        for (int64_t i = 0; i < nx; ++i) {
            int32_t cnt = 0;
            int64_t pos = rowptr ? rowptr[i] : 0;
            for (int64_t j = 0; j < ny; ++j) {
                TR pd[3] = {0, 0, 0};
                TR d2 = dist2<TC, TR>(ndims, x + ndims * i, y + ndims * j, pd);
                d2 = periodic_distance<TR>(g, pd, d2, r2);
                if (d2 <= r2) {
                    ++cnt;
                    if (rowptr) cols[pos++] = (int32_t)j;
                }
            }
            if (counts) counts[i] = cnt;
        }
    };
    if (tc == 0 && tr == 0) run(0.0f, 0.0f);
    else if (tc == 0) run(0.0f, 0.0);
    else if (tr == 0) run(0.0, 0.0f);
    else run(0.0, 0.0);
    return OK;
}

// PrecomputedNeighborhoodSearch list build, src/nhs_precomputed.jl:187-206:
// lists filled in traversal order, overflow error from pushat!
// (src/vector_of_vectors.jl:79,114-121) when a list exceeds max_neighbors, then
// sorteach! (src/vector_of_vectors.jl:177-184).  rowptr must come from a count pass.
int orc_precomputed_lists(void* h, const void* x, const void* y, int64_t nx,
                          int64_t max_neighbors, const int64_t* rowptr, int32_t* cols,
                          int64_t* err_point) {
    OracleGrid& og = *(OracleGrid*)h;
    std::vector<int64_t> pts((size_t)nx);
    for (int64_t i = 0; i < nx; ++i) pts[i] = i;
    std::vector<int32_t> counts((size_t)nx);
    int st = orc_traverse(h, x, y, pts.data(), nx, counts.data(), nullptr, rowptr, cols, err_point);
    if (st != OK) return st;
    for (int64_t i = 0; i < nx; ++i) {
        if (counts[i] > max_neighbors) {
            if (err_point) *err_point = i;
            return ERR_CELL_FULL;
        }
        std::sort(cols + rowptr[i], cols + rowptr[i + 1]);
    }
    (void)og;
    return OK;
}

// ---------------------------------------------------------------------------
// Threaded CPU path with the reference's CPU semantics, used as the timed CPU
// baseline: ParallelUpdate build with atomic fetch-add pushes
// (src/nhs_grid.jl:255-281, src/vector_of_vectors.jl:93-112) and point-major
// traversal in original index order, statically chunked over points like
// Polyester.@batch (src/util.jl:133-137).  Float32 coordinates and radius, 3D.
// Returns seconds for build and traversal through the out pointers.
// ---------------------------------------------------------------------------
int orc_threaded_count_f32(void* h, const float* y, int64_t n, int nthreads, int32_t* counts,
                           double* t_build, double* t_trav) {
    OracleGrid& og = *(OracleGrid*)h;
    if (og.tc != 0 || og.tr != 0) return ERR_ARG;
    const Geom& g = og.g;
    const int D = g.ndims;
    const int64_t C = g.ncells();
    std::vector<std::atomic<int32_t>> lengths((size_t)C);
    og.cs.backend.assign((size_t)(og.cs.cap * C), 0);
    std::atomic<int> status{OK};
    auto now = [] {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
            .count();
    };
    auto chunked = [&](int64_t count, auto&& body) {
        std::vector<std::thread> th;
        for (int t = 0; t < nthreads; ++t) {
            int64_t b = count * t / nthreads, e = count * (t + 1) / nthreads;
            th.emplace_back([=, &body] {
                for (int64_t i = b; i < e; ++i) body(i);
            });
        }
        for (auto& x : th) x.join();
    };
    double t0 = now();
    chunked(C, [&](int64_t c) { lengths[c].store(0, std::memory_order_relaxed); });
    chunked(n, [&](int64_t p) {
        int64_t cell[3];
        if (og.tcell == 0) cell_coords<float, float>(g, y + D * p, cell);
        else cell_coords<float, double>(g, y + D * p, cell);
        if (!cell_in_bounds(g, cell)) {
            status.store(ERR_OUT_OF_BOUNDS);
            return;
        }
        int64_t c = linear0(g, cell);
        int32_t len = lengths[c].fetch_add(1, std::memory_order_relaxed) + 1;
        if (len > og.cs.cap) {
            lengths[c].fetch_sub(1, std::memory_order_relaxed);
            status.store(ERR_CELL_FULL);
            return;
        }
        og.cs.backend[(size_t)(c * og.cs.cap + len - 1)] = (int32_t)p;
    });
    double t1 = now();
    if (status.load() != OK) return status.load();
    og.cs.lengths.resize((size_t)C);
    for (int64_t c = 0; c < C; ++c) og.cs.lengths[c] = lengths[c].load();
    og.built = true;
    const float r = (float)og.radius;
    std::atomic<int> qstatus{OK};
    double t2 = now();
    chunked(n, [&](int64_t i) {
        int32_t cnt = 0;
        int st = og.tcell == 0
                     ? traverse_point<float, float, float>(g, og.cs, y, y + D * i, r,
                                                           [&](int32_t, const float*, float) { ++cnt; })
                     : traverse_point<float, float, double>(g, og.cs, y, y + D * i, r,
                                                            [&](int32_t, const float*, float) { ++cnt; });
        if (st != OK) qstatus.store(st);
        counts[i] = cnt;
    });
    double t3 = now();
    if (t_build) *t_build = t1 - t0;
    if (t_trav) *t_trav = t3 - t2;
    return qstatus.load();
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Interaction physics over fixed neighbour lists (CSR from orc_traverse, x == y).
// Not in the reference (TrixiParticles physics): textbook formulas, evaluated in double
// or in forward-mode dual numbers (value + one directional derivative), which stand in
// for ForwardDiff through the CPU path.
//   cubic spline (Monaghan 1992) and Wendland C2 (Wendland 1995) kernels, support 2h
//   Tait EOS p = B((rho/rho0)^gamma - 1), B = rho0 c^2/gamma
//   a_i = -sum_{j != i} m_j (P_i + P_j) W'(r) (x_i - x_j)/r,  P = p/rho^2
//   phi(r) = eps (1 - r/rc)^4 (1 + 4 r/rc)
// ---------------------------------------------------------------------------
namespace {

struct Dual {
    double v = 0.0, d = 0.0;
    Dual() = default;
    Dual(double a) : v(a), d(0.0) {}
    Dual(double a, double b) : v(a), d(b) {}
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
inline Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
inline Dual sqrt(Dual a) {
    double s = std::sqrt(a.v);
    return {s, s > 0 ? a.d / (2.0 * s) : 0.0};
}
inline Dual pow(Dual a, double e) { return {std::pow(a.v, e), e * std::pow(a.v, e - 1.0) * a.d}; }
inline double val(double a) { return a; }
inline double val(Dual a) { return a.v; }
inline double absval(double a) { return std::fabs(a); }
using std::pow;
using std::sqrt;

template <class T>
struct Kern {
    int kind;  // 0 cubic spline, 1 Wendland C2
    int nd;
    double h;
    double sigma() const {
        const double pi = 3.14159265358979323846;
        if (kind == 0) return nd == 3 ? 1.0 / (pi * h * h * h) : 10.0 / (7.0 * pi * h * h);
        return nd == 3 ? 21.0 / (16.0 * pi * h * h * h) : 7.0 / (4.0 * pi * h * h);
    }
    T w(T r) const {
        T q = r / T(h);
        double s = sigma();
        if (kind == 0) {
            if (q < T(1.0)) return T(s) * (T(1.0) - T(1.5) * q * q + T(0.75) * q * q * q);
            if (q < T(2.0)) {
                T t = T(2.0) - q;
                return T(s * 0.25) * t * t * t;
            }
            return T(0.0);
        }
        if (q < T(2.0)) {
            T t = T(1.0) - T(0.5) * q;
            return T(s) * t * t * t * t * (T(2.0) * q + T(1.0));
        }
        return T(0.0);
    }
    T dw(T r) const {
        T q = r / T(h);
        double s = sigma() / h;
        if (kind == 0) {
            if (q < T(1.0)) return T(s) * (T(-3.0) * q + T(2.25) * q * q);
            if (q < T(2.0)) {
                T t = T(2.0) - q;
                return T(-0.75 * s) * t * t;
            }
            return T(0.0);
        }
        if (q < T(2.0)) {
            T t = T(1.0) - T(0.5) * q;
            return T(-5.0 * s) * q * t * t * t;
        }
        return T(0.0);
    }
};

template <class T>
struct Phys {
    int nd;
    const T* x;  // N x nd
    int64_t n;
    const int64_t* rp;
    const int32_t* cols;

    T dist(int64_t i, int64_t j, T pd[3]) const {
        T d2 = T(0.0);
        for (int d = 0; d < nd; ++d) {
            pd[d] = x[nd * i + d] - x[nd * j + d];
            d2 += pd[d] * pd[d];
        }
        return d2;
    }

    void density(const Kern<T>& k, const T* mass, T* rho, double* scale) const {
        for (int64_t i = 0; i < n; ++i) {
            T s = T(0.0);
            double a = 0.0;
            for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {
                int32_t j = cols[e];
                T pd[3];
                T r = sqrt(dist(i, j, pd));
                T t = mass[j] * k.w(r);
                s += t;
                a += std::fabs(val(t));
            }
            rho[i] = s;
            if (scale) scale[i] = a;
        }
    }

    void pressure(const Kern<T>& k, const T* mass, const T* P, T* dv, double* scale) const {
        for (int64_t i = 0; i < n; ++i) {
            T acc[3] = {T(0.0), T(0.0), T(0.0)};
            double a = 0.0;
            for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {
                int32_t j = cols[e];
                T pd[3];
                T d2 = dist(i, j, pd);
                if (!(val(d2) > 0.0)) continue;
                T r = sqrt(d2);
                T c = -mass[j] * (P[i] + P[j]) * k.dw(r) / r;
                for (int d = 0; d < nd; ++d) {
                    acc[d] += c * pd[d];
                    a += std::fabs(val(c * pd[d]));
                }
            }
            for (int d = 0; d < nd; ++d) dv[nd * i + d] = acc[d];
            if (scale) scale[i] = a;
        }
    }

    void energy(double eps, double rc, T* e) const {
        for (int64_t i = 0; i < n; ++i) {
            T s = T(0.0);
            for (int64_t ed = rp[i]; ed < rp[i + 1]; ++ed) {
                int32_t j = cols[ed];
                T pd[3];
                T d2 = dist(i, j, pd);
                if (!(val(d2) > 0.0)) continue;
                T q = sqrt(d2) / T(rc);
                if (q < T(1.0)) {
                    T t = T(1.0) - q;
                    s += T(eps) * t * t * t * t * (T(1.0) + T(4.0) * q);
                }
            }
            e[i] = T(0.5) * s;
        }
    }
};

}  // namespace

extern "C" {

// Density, pressure acceleration and pair energy in double over CSR neighbour lists.
// scale (nullable) receives sum_j |term_j| per point: the summation-error scale.
void orc_sph_density(int nd, int kernel, double h, const double* x, int64_t n, const int64_t* rp,
                     const int32_t* cols, const double* mass, double* rho, double* scale) {
    Phys<double> p{nd, x, n, rp, cols};
    p.density(Kern<double>{kernel, nd, h}, mass, rho, scale);
}

void orc_sph_pressure_accel(int nd, int kernel, double h, const double* x, int64_t n, const int64_t* rp,
                            const int32_t* cols, const double* mass, const double* P, double* dv, double* scale) {
    Phys<double> p{nd, x, n, rp, cols};
    p.pressure(Kern<double>{kernel, nd, h}, mass, P, dv, scale);
}

// n-body acceleration of benchmarks/n_body.jl:34-51 over CSR lists, in double:
//     dv_i = sum_{j in list(i), d >= sqrt_eps} -G m_j (x_i - x_j) / d^3,
// scale_i = sum of |term| over the components (the summation-error scale).
void orc_nbody(int nd, double G, double sqrt_eps, const double* x, int64_t n, const int64_t* rp,
               const int32_t* cols, const double* mass, double* dv, double* scale) {
    for (int64_t i = 0; i < n; ++i) {
        double acc[3] = {0.0, 0.0, 0.0}, a = 0.0;
        for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {
            const int32_t j = cols[e];
            double pd[3], d2 = 0.0;
            for (int d = 0; d < nd; ++d) {
                pd[d] = x[nd * i + d] - x[nd * (int64_t)j + d];
                d2 += pd[d] * pd[d];
            }
            const double dist = std::sqrt(d2);
            if (dist < sqrt_eps) continue;  // `distance < sqrt(eps(ELTYPE)) && return`
            for (int d = 0; d < nd; ++d) {
                const double t = -G * mass[j] * pd[d] / (dist * dist * dist);
                acc[d] += t;
                a += std::fabs(t);
            }
        }
        for (int d = 0; d < nd; ++d) dv[nd * i + d] = acc[d];
        scale[i] = a;
    }
}

// The fluid-fluid interaction of TrixiParticles 0.5 timed by the reference's benchmark_wcsph
// (benchmarks/smoothed_particle_hydrodynamics.jl:44-102), restated term by term in double over
// CSR lists.  v and dv are (nd+1) values per point: velocity, then density (ContinuityDensity).
// scale[2 i] / scale[2 i + 1]: sums of |term| of the momentum / density equation of point i.
void orc_wcsph_interact(int nd, int kernel, double h, double c, double alpha, double beta, double eps_visc,
                        double delta, double sqrt_eps, const double* x, int64_t n, const int64_t* rp,
                        const int32_t* cols, const double* mass, const double* v, const double* p, double* dv,
                        double* scale) {
    const Kern<double> k{kernel, nd, h};
    const int nv = nd + 1;
    for (int64_t a = 0; a < n; ++a) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0}, sm = 0.0, sd = 0.0;
        const double rho_a = v[nv * a + nd];
        for (int64_t e = rp[a]; e < rp[a + 1]; ++e) {
            const int64_t b = cols[e];
            const double rho_b = v[nv * b + nd];
            double pd[3], d2 = 0.0;
            for (int d = 0; d < nd; ++d) {
                pd[d] = x[nd * a + d] - x[nd * b + d];
                d2 += pd[d] * pd[d];
            }
            const double dist = std::sqrt(d2);
            // smoothing_kernel_grad: W'(d) / d * pos_diff, zero below sqrt(eps)
            double grad[3] = {0.0, 0.0, 0.0};
            if (dist >= sqrt_eps)
                for (int d = 0; d < nd; ++d) grad[d] = k.dw(dist) / dist * pd[d];
            const double m_b = mass[b];
            // pressure acceleration (ContinuityDensity form)
            const double pa_coef = -m_b * (p[a] + p[b]) / (rho_a * rho_b);
            // Monaghan viscosity: approaching pairs only
            double vdiff[3], vr = 0.0;
            for (int d = 0; d < nd; ++d) {
                vdiff[d] = v[nv * a + d] - v[nv * b + d];
                vr += vdiff[d] * pd[d];
            }
            double pi_ab = 0.0;
            if (vr < 0.0) {
                const double mu = h * vr / (d2 + eps_visc * h * h);
                const double rho_mean = 0.5 * (rho_a + rho_b);
                pi_ab = (-alpha * c * mu + beta * mu * mu) / rho_mean;
            }
            for (int d = 0; d < nd; ++d) {
                const double tp = pa_coef * grad[d];
                const double tv = -m_b * pi_ab * grad[d];
                acc[d] += tp + tv;
                sm += std::fabs(tp) + std::fabs(tv);
            }
            // continuity equation
            double vg = 0.0;
            for (int d = 0; d < nd; ++d) vg += vdiff[d] * grad[d];
            const double tc = rho_a / rho_b * m_b * vg;
            // Molteni-Colagrossi density diffusion (zero at coincident points)
            double td = 0.0;
            if (dist >= sqrt_eps) {
                double psig = 0.0;
                for (int d = 0; d < nd; ++d) psig += 2.0 * (rho_a - rho_b) * pd[d] / d2 * grad[d];
                td = delta * h * c * psig * (m_b / rho_b);
            }
            acc[nd] += tc + td;
            sd += std::fabs(tc) + std::fabs(td);
        }
        for (int d = 0; d < nv; ++d) dv[nv * a + d] = acc[d];
        scale[2 * a] = sm;
        scale[2 * a + 1] = sd;
    }
}

void orc_pair_energy(int nd, double eps, double rc, const double* x, int64_t n, const int64_t* rp,
                     const int32_t* cols, double* e) {
    Phys<double> p{nd, x, n, rp, cols};
    p.energy(eps, rc, e);
}

// TLSPH deformation gradient over CSR lists of the initial coordinates X (TrixiParticles 0.5
// calc_deformation_grad!, run by benchmarks/smoothed_particle_hydrodynamics.jl:127-146):
//     G_a = sum_{b in list(a), |X_a - X_b| >= sqrt(eps(Float64))} V_b (x_b - x_a) (W'(r) X_ab / r)^T,
// out_a = G_a L_a (L != nullptr, row-major D x D per point here) or G_a.  TrixiParticles is not
// vendored; the sum is restated from its published algorithm and pinned by closed forms (an
// undeformed body gives the symmetric correction matrix M; with L = M^-1 an affine map x = A X + c
// gives A exactly).  Row-major D x D output per point.
void orc_tlsph_grad(int nd, int kernel, double h, const double* X, const double* x, int64_t n, const int64_t* rp,
                    const int32_t* cols, const double* vol, const double* L, double* out) {
    const Kern<double> k{kernel, nd, h};
    const double tiny = 1.4901161193847656e-8;
    for (int64_t a = 0; a < n; ++a) {
        double g[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int64_t q = rp[a]; q < rp[a + 1]; ++q) {
            const int64_t b = cols[q];
            double Xab[3] = {0, 0, 0}, d2 = 0.0;
            for (int d = 0; d < nd; ++d) {
                Xab[d] = X[a * nd + d] - X[b * nd + d];
                d2 += Xab[d] * Xab[d];
            }
            const double r = std::sqrt(d2);
            if (r < tiny) continue;
            const double s = vol[b] * k.dw(r) / r;
            for (int i = 0; i < nd; ++i)
                for (int j = 0; j < nd; ++j) g[i][j] += s * (x[b * nd + i] - x[a * nd + i]) * Xab[j];
        }
        double* o = out + a * nd * nd;
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j) {
                double v = 0.0;
                if (L) {
                    for (int m = 0; m < nd; ++m) v += g[i][m] * L[a * nd * nd + m * nd + j];
                } else {
                    v = g[i][j];
                }
                o[i * nd + j] = v;
            }
    }
}

// TLSPH interaction (TrixiParticles interact_structure_structure!, the loop of the reference's
// benchmark_tlsph, benchmarks/smoothed_particle_hydrodynamics.jl:115-121): stress term
// m_b (Pr_a + Pr_b) grad_a W(X_ab) plus, when F != nullptr, the Ganzenmueller penalty
// (alpha/2) V_a V_b W / |X_ab|^2 E delta x_ab / |x_ab| / m_a with
// delta = ((F_a + F_b) X_ab - 2 x_ab) . x_ab / |x_ab|.  hae = alpha E / 2.  Pr, F row-major
// D x D per point; X, x, dv: N x D.
void orc_tlsph_force(int nd, int kernel, double h, const double* X, const double* x, int64_t n, const int64_t* rp,
                     const int32_t* cols, const double* mass, const double* rho, const double* Pr, const double* F,
                     double hae, double* dv) {
    const Kern<double> k{kernel, nd, h};
    const double tiny = 1.4901161193847656e-8;
    for (int64_t a = 0; a < n; ++a) {
        double acc[3] = {0, 0, 0};
        for (int64_t q = rp[a]; q < rp[a + 1]; ++q) {
            const int64_t b = cols[q];
            double Xab[3] = {0, 0, 0}, d2 = 0.0;
            for (int d = 0; d < nd; ++d) {
                Xab[d] = X[a * nd + d] - X[b * nd + d];
                d2 += Xab[d] * Xab[d];
            }
            const double r = std::sqrt(d2);
            if (r < tiny) continue;
            double gw[3];
            for (int d = 0; d < nd; ++d) gw[d] = k.dw(r) * Xab[d] / r;
            for (int i = 0; i < nd; ++i) {
                double v = 0.0;
                for (int j = 0; j < nd; ++j) v += (Pr[a * nd * nd + i * nd + j] + Pr[b * nd * nd + i * nd + j]) * gw[j];
                acc[i] += mass[b] * v;
            }
            if (!F) continue;
            double xab[3] = {0, 0, 0}, c2 = 0.0;
            for (int d = 0; d < nd; ++d) {
                xab[d] = x[a * nd + d] - x[b * nd + d];
                c2 += xab[d] * xab[d];
            }
            const double cd = std::sqrt(c2);
            double delta = 0.0;
            for (int i = 0; i < nd; ++i) {
                double e = -2.0 * xab[i];
                for (int j = 0; j < nd; ++j) e += (F[a * nd * nd + i * nd + j] + F[b * nd * nd + i * nd + j]) * Xab[j];
                delta += e * xab[i];
            }
            delta /= cd;
            const double Va = mass[a] / rho[a], Vb = mass[b] / rho[b];
            const double f = hae * Va * Vb * k.w(r) / (r * r) * delta / cd;
            for (int d = 0; d < nd; ++d) acc[d] += f * xab[d] / mass[a];
        }
        for (int d = 0; d < nd; ++d) dv[a * nd + d] = acc[d];
    }
}

// Saint Venant-Kirchhoff corrected PK1 over rho^2 (TrixiParticles compute_pk1_corrected!):
// E = (F^T F - I)/2, S = lambda tr(E) I + 2 mu E, out = F S L / rho^2 (L == nullptr: identity).
// F, L, out row-major D x D per point.
void orc_tlsph_pk1_rho2(int nd, const double* F, const double* L, const double* rho, int64_t n, double lambda,
                        double mu, double* out) {
    for (int64_t a = 0; a < n; ++a) {
        const double* f = F + a * nd * nd;
        double E[3][3], S[3][3], P[3][3], tr = 0.0;
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j) {
                double v = 0.0;
                for (int m = 0; m < nd; ++m) v += f[m * nd + i] * f[m * nd + j];
                E[i][j] = 0.5 * (v - (i == j ? 1.0 : 0.0));
            }
        for (int i = 0; i < nd; ++i) tr += E[i][i];
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j) S[i][j] = 2.0 * mu * E[i][j] + (i == j ? lambda * tr : 0.0);
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j) {
                double v = 0.0;
                for (int m = 0; m < nd; ++m) v += f[i * nd + m] * S[m][j];
                P[i][j] = v;
            }
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j) {
                double v = P[i][j];
                if (L) {
                    v = 0.0;
                    for (int m = 0; m < nd; ++m) v += P[i][m] * L[a * nd * nd + m * nd + j];
                }
                out[a * nd * nd + i * nd + j] = v / (rho[a] * rho[a]);
            }
    }
}

// Scalar losses in dual numbers, derivative along coordinate (k, dim) (forward mode):
//   which 0: E = sum_i e_i (Wendland pair energy, eps, rc = cutoff)
//   which 1: L = sum_i (rho_i - rho0)^2, rho with `kernel`, mass m
//   which 2: L = sum_i |a_i|^2, a from rho -> Tait EOS (rho0, c, gamma) -> P = p/rho^2
// Writes the loss and dL/dx[k][dim].
void orc_loss_dual(int which, int nd, int kernel, double h, double eps, double rc, const double* x, int64_t n,
                   const int64_t* rp, const int32_t* cols, const double* mass, double rho0, double c, double gamma,
                   int64_t k, int dim, double* value, double* deriv) {
    std::vector<Dual> xd((size_t)(n * nd));
    for (int64_t i = 0; i < n * nd; ++i) xd[i] = Dual(x[i], 0.0);
    xd[(size_t)(k * nd + dim)].d = 1.0;
    std::vector<Dual> md((size_t)n);
    for (int64_t i = 0; i < n; ++i) md[i] = Dual(mass[i]);
    Phys<Dual> p{nd, xd.data(), n, rp, cols};
    Dual L(0.0);
    if (which == 0) {
        std::vector<Dual> e((size_t)n);
        p.energy(eps, rc, e.data());
        for (auto& v : e) L += v;
    } else {
        std::vector<Dual> rho((size_t)n);
        Kern<Dual> kern{kernel, nd, h};
        p.density(kern, md.data(), rho.data(), nullptr);
        if (which == 1) {
            for (auto& r : rho) L += (r - Dual(rho0)) * (r - Dual(rho0));
        } else {
            const double B = rho0 * c * c / gamma;
            std::vector<Dual> P((size_t)n), dv((size_t)(n * nd));
            for (int64_t i = 0; i < n; ++i) {
                Dual pr = Dual(B) * (pow(rho[i] / Dual(rho0), gamma) - Dual(1.0));
                P[i] = pr / (rho[i] * rho[i]);
            }
            p.pressure(kern, md.data(), P.data(), dv.data(), nullptr);
            for (auto& v : dv) L += v * v;
        }
    }
    *value = L.v;
    *deriv = L.d;
}

}  // extern "C"

extern "C" {

// Threaded CPU WCSPH step with the reference CPU path's structure (Float32, 3D):
// ParallelUpdate build with atomic pushes, then two point-major traversals statically
// chunked over points (density, then pressure acceleration), the Tait EOS in between.
// Timings (seconds) of the build and of the two traversals + EOS go to the out pointers.
int orc_threaded_wcsph_f32(void* h, const float* y, int64_t n, int nthreads, float mass, float hsml, float rho0,
                           float c, float gamma, float* rho, float* dv, double* t_build, double* t_phys) {
    OracleGrid& og = *(OracleGrid*)h;
    if (og.tc != 0 || og.tr != 0 || og.g.ndims != 3) return ERR_ARG;
    std::vector<int32_t> counts((size_t)n);
    double tb = 0, tt = 0;
    int st = orc_threaded_count_f32(h, y, n, nthreads, counts.data(), &tb, &tt);
    if (st != OK) return st;
    const Geom& g = og.g;
    const float r = (float)og.radius;
    const double pi = 3.14159265358979323846;
    const float sigma = (float)(1.0 / (pi * (double)hsml * hsml * hsml));
    const float inv_h = 1.0f / hsml;
    auto now = [] {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    };
    auto chunked = [&](int64_t count, auto&& body) {
        std::vector<std::thread> th;
        for (int t = 0; t < nthreads; ++t) {
            int64_t b = count * t / nthreads, e = count * (t + 1) / nthreads;
            th.emplace_back([=, &body] {
                for (int64_t i = b; i < e; ++i) body(i);
            });
        }
        for (auto& x : th) x.join();
    };
    auto w = [&](float d) {
        float q = d * inv_h;
        if (q < 1.0f) return sigma * (1.0f - 1.5f * q * q + 0.75f * q * q * q);
        if (q < 2.0f) {
            float t = 2.0f - q;
            return sigma * 0.25f * t * t * t;
        }
        return 0.0f;
    };
    auto dw = [&](float d) {
        float q = d * inv_h;
        if (q < 1.0f) return sigma * inv_h * (-3.0f * q + 2.25f * q * q);
        if (q < 2.0f) {
            float t = 2.0f - q;
            return sigma * inv_h * (-0.75f * t * t);
        }
        return 0.0f;
    };
    std::vector<float> P((size_t)n);
    const float B = rho0 * c * c / gamma;
    double t0 = now();
    chunked(n, [&](int64_t i) {
        float s = 0.0f;
        traverse_point<float, float, float>(g, og.cs, y, y + 3 * i, r,
                                            [&](int32_t, const float*, float d2) { s += mass * w(std::sqrt(d2)); });
        rho[i] = s;
        float p = B * (std::pow(s / rho0, gamma) - 1.0f);
        P[i] = p / (s * s);
    });
    chunked(n, [&](int64_t i) {
        float a[3] = {0, 0, 0};
        traverse_point<float, float, float>(g, og.cs, y, y + 3 * i, r, [&](int32_t j, const float* pd, float d2) {
            if (d2 > 0.0f) {
                float d = std::sqrt(d2);
                float cc = -mass * (P[i] + P[j]) * dw(d) / d;
                for (int k = 0; k < 3; ++k) a[k] += cc * pd[k];
            }
        });
        for (int k = 0; k < 3; ++k) dv[3 * i + k] = a[k];
    });
    double t1 = now();
    if (t_build) *t_build = tb;
    if (t_phys) *t_phys = t1 - t0;
    return OK;
}


// Density and pressure acceleration in double over the grid's neighbour sets (the exact
// Float32 predicate of traverse_point, src/nhs_grid.jl:519-575) for the query rows
// rows[0 .. nrows), threaded over rows.  The grid must have been built from y (for example
// by orc_threaded_count_f32).  Same sums as orc_sph_density / orc_sph_pressure_accel over
// CSR lists, without materialising the lists, so very large clouds can be checked.
// rho, rho_scale: nrows; dv: nrows x 3, dv_scale: nrows (sums of |term|).  P nullable.
int orc_threaded_sph_rows_f32(void* h, const float* y, int64_t n, const int64_t* rows, int64_t nrows,
                              int nthreads, int kernel, double hsml, const double* mass, const double* P,
                              double* rho, double* rho_scale, double* dv, double* dv_scale) {
    OracleGrid& og = *(OracleGrid*)h;
    if (og.tc != 0 || og.tr != 0 || og.g.ndims != 3 || !og.built) return ERR_ARG;
    const Geom& g = og.g;
    const float r = (float)og.radius;
    const Kern<double> k{kernel, 3, hsml};
    std::atomic<int> status{OK};
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) {
        const int64_t b = nrows * t / nthreads, e = nrows * (t + 1) / nthreads;
        th.emplace_back([&, b, e] {
            for (int64_t q = b; q < e; ++q) {
                const int64_t i = rows[q];
                if (i < 0 || i >= n) {
                    status.store(ERR_ARG);
                    return;
                }
                double s = 0.0, sa = 0.0, a[3] = {0.0, 0.0, 0.0}, aa = 0.0;
                auto body = [&](int32_t j, const float*, float) {
                    double pd[3], d2 = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        pd[d] = (double)y[3 * i + d] - (double)y[3 * (int64_t)j + d];
                        d2 += pd[d] * pd[d];
                    }
                    const double dist = std::sqrt(d2);
                    const double w = mass[j] * k.w(dist);
                    s += w;
                    sa += std::fabs(w);
                    if (P && d2 > 0.0) {
                        const double c = -mass[j] * (P[i] + P[j]) * k.dw(dist) / dist;
                        for (int d = 0; d < 3; ++d) {
                            a[d] += c * pd[d];
                            aa += std::fabs(c * pd[d]);
                        }
                    }
                };
                const int st = og.tcell == 0 ? traverse_point<float, float, float>(g, og.cs, y, y + 3 * i, r, body)
                                             : traverse_point<float, float, double>(g, og.cs, y, y + 3 * i, r, body);
                if (st != OK) status.store(st);
                rho[q] = s;
                if (rho_scale) rho_scale[q] = sa;
                if (P) {
                    for (int d = 0; d < 3; ++d) dv[3 * q + d] = a[d];
                    if (dv_scale) dv_scale[q] = aa;
                }
            }
        });
    }
    for (auto& x : th) x.join();
    return status.load();
}

}  // extern "C"
