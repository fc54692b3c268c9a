// Horizontal metrics of an icosahedral grid (host code, set-up time).
//
// nicam_gmtr_setup restates the reference's metrics kernel
// (kernels/NICAM/dyn_metrics/src/mod_gmtr.f90: GMTR_p_setup :46-233,
// GMTR_t_setup :237-400, GMTR_a_setup :404-741, GMTR_TNvec :744-812, with the
// vector helpers of mod_vector.f90:142-208 and VECTR_triangle :345-461) for a
// grid on the sphere with spherical polygons, the reference's configuration
// (problem_size.inc:73-78).  The steps follow the reference's order, so the
// metrics match its compiled kernel to rounding.
//
// nicam_grid_vertices forms what that kernel reads besides GRD_x: the triangle
// vertices GRD_xt, the latitude/longitude GRD_s, and the cell-wall midpoints
// GRD_xr that the flux kernel uses.  The reference reads these from files; the
// definitions here are the usual ones (normalised centroids and midpoints).
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

namespace nicam {
namespace {

struct V {
    double x[3];
};

// mod_vector.f90:142-157, 161-174, 178-189, 193-208 (same operand order)
V vcross(const V &a, const V &b, const V &c, const V &d)
{
    V n;
    n.x[0] = (b.x[1] - a.x[1]) * (d.x[2] - c.x[2]) - (b.x[2] - a.x[2]) * (d.x[1] - c.x[1]);
    n.x[1] = (b.x[2] - a.x[2]) * (d.x[0] - c.x[0]) - (b.x[0] - a.x[0]) * (d.x[2] - c.x[2]);
    n.x[2] = (b.x[0] - a.x[0]) * (d.x[1] - c.x[1]) - (b.x[1] - a.x[1]) * (d.x[0] - c.x[0]);
    return n;
}
double vdot(const V &a, const V &b, const V &c, const V &d)
{
    return (b.x[0] - a.x[0]) * (d.x[0] - c.x[0]) + (b.x[1] - a.x[1]) * (d.x[1] - c.x[1]) +
           (b.x[2] - a.x[2]) * (d.x[2] - c.x[2]);
}
double vabs(const V &a) { return std::sqrt(a.x[0] * a.x[0] + a.x[1] * a.x[1] + a.x[2] * a.x[2]); }
double vangle(const V &a, const V &b, const V &c)
{
    const double nvlenC = vdot(b, a, b, c);
    const double nvlenS = vabs(vcross(b, a, b, c));
    return std::atan2(nvlenS, nvlenC);
}
const V O = {{0.0, 0.0, 0.0}};

// VECTR_triangle, ON_SPHERE branch (mod_vector.f90:389-404): l'Huilier
double triangle(const V &a, const V &b, const V &c, double radius)
{
    double l1 = vangle(a, O, b), l2 = vangle(b, O, c), l3 = vangle(c, O, a);
    l1 = 0.5 * l1;
    l2 = 0.5 * l2;
    l3 = 0.5 * l3;
    const double s = 0.5 * (l1 + l2 + l3);
    double x = std::tan(s) * std::tan(s - l1) * std::tan(s - l2) * std::tan(s - l3);
    x = (0.0 > x) ? 0.0 : x;  // max(x, 0)
    return 4.0 * std::atan(std::sqrt(x)) * radius * radius;
}

// GMTR_TNvec, on the sphere with ON_SPHERE polygons (mod_gmtr.f90:772-802)
void tnvec(V &vT, V &vN, const V &from, const V &to, double radius)
{
    for (int d = 0; d < 3; ++d) vT.x[d] = to.x[d] - from.x[d];
    const double distance = vangle(from, O, to) * radius;
    double length = vabs(vT);
    for (int d = 0; d < 3; ++d) vT.x[d] = vT.x[d] * distance / length;
    vN = vcross(O, from, O, to);
    length = vabs(vN);
    for (int d = 0; d < 3; ++d) vN.x[d] = vN.x[d] * distance / length;
}

// sin and cos of one argument through glibc's sincos, as the reference's
// compiled kernel evaluates sin(GRD_s(..LON)) and cos(GRD_s(..LON)) (the
// Fortran compiler merges the pair); the two separate routines can differ from
// it in the last bit.
void sincos_lon(double lon, double &s, double &c) { ::sincos(lon, &s, &c); }

V scale(V a, double r)
{
    for (double &c : a.x) c = c / r;
    return a;
}

struct Grid {
    int iall, gall, lall, npl;
    int gmin, gmax;  // 1-based, as the reference
    int suf(int i, int j) const { return (j - 1) * iall + i; }
};

}  // namespace
}  // namespace nicam

using namespace nicam;

extern "C" int nicam_gmtr_setup(int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp,
                                double radius, const double *GRD_x, const double *GRD_x_pl, const double *GRD_xt,
                                const double *GRD_xt_pl, const double *GRD_s, const double *GRD_s_pl,
                                double *GMTR_p, double *GMTR_p_pl, double *GMTR_t, double *GMTR_t_pl,
                                double *GMTR_a, double *GMTR_a_pl)
{
    if (gall_1d < 4 || lall < 1 || gall_pl < 3 || gall_pl > 15) return fail(NICAM_ERR_ARG, "sizes out of range");
    if (!have_sgp || !GRD_x || !GRD_xt || !GRD_s || !GMTR_p || !GMTR_t || !GMTR_a)
        return fail(NICAM_ERR_ARG, "null argument");
    if (have_pl && (!GRD_x_pl || !GRD_xt_pl || !GRD_s_pl || !GMTR_p_pl || !GMTR_t_pl || !GMTR_a_pl))
        return fail(NICAM_ERR_ARG, "null pole argument");
    if (!(radius > 0.0)) return fail(NICAM_ERR_ARG, "radius must be positive");
    const Grid G{gall_1d, gall_1d * gall_1d, lall, gall_pl, 2, gall_1d - 1};
    const size_t gall = G.gall, npl = gall_pl;
    // 1-based accessors in the reference's shapes
    auto X = [&](int g, int l) {
        V v;
        for (int d = 0; d < 3; ++d) v.x[d] = GRD_x[((size_t)d * lall + (l - 1)) * gall + (g - 1)];
        return v;
    };
    auto XT = [&](int g, int l, int t) {  // t = 1 (TI) or 2 (TJ)
        V v;
        for (int d = 0; d < 3; ++d) v.x[d] = GRD_xt[(((size_t)d * 2 + (t - 1)) * lall + (l - 1)) * gall + (g - 1)];
        return v;
    };
    auto XPL = [&](int v_, int l) {
        V v;
        for (int d = 0; d < 3; ++d) v.x[d] = GRD_x_pl[((size_t)d * 2 + (l - 1)) * npl + (v_ - 1)];
        return v;
    };
    auto XTPL = [&](int v_, int l) {
        V v;
        for (int d = 0; d < 3; ++d) v.x[d] = GRD_xt_pl[((size_t)d * 2 + (l - 1)) * npl + (v_ - 1)];
        return v;
    };
    auto P = [&](int g, int l, int c) -> double & { return GMTR_p[((size_t)(c - 1) * lall + (l - 1)) * gall + (g - 1)]; };
    auto T = [&](int g, int l, int t, int c) -> double & {
        return GMTR_t[(((size_t)(c - 1) * 2 + (t - 1)) * lall + (l - 1)) * gall + (g - 1)];
    };
    auto A = [&](int g, int l, int a, int c) -> double & {
        return GMTR_a[(((size_t)(c - 1) * 3 + (a - 1)) * lall + (l - 1)) * gall + (g - 1)];
    };
    auto PPL = [&](int v, int l, int c) -> double & { return GMTR_p_pl[((size_t)(c - 1) * 2 + (l - 1)) * npl + (v - 1)]; };
    auto TPL = [&](int v, int l, int c) -> double & { return GMTR_t_pl[((size_t)(c - 1) * 2 + (l - 1)) * npl + (v - 1)]; };
    auto APL = [&](int v, int l, int c) -> double & { return GMTR_a_pl[((size_t)(c - 1) * 2 + (l - 1)) * npl + (v - 1)]; };
    const int gmin = G.gmin, gmax = G.gmax;
    auto suf = [&](int i, int j) { return G.suf(i, j); };
    std::memset(GMTR_p, 0, sizeof(double) * gall * lall * 8);
    std::memset(GMTR_t, 0, sizeof(double) * gall * lall * 10);
    std::memset(GMTR_a, 0, sizeof(double) * gall * lall * 36);
    if (have_pl) {
        std::memset(GMTR_p_pl, 0, sizeof(double) * npl * 2 * 8);
        std::memset(GMTR_t_pl, 0, sizeof(double) * npl * 2 * 5);
        std::memset(GMTR_a_pl, 0, sizeof(double) * npl * 2 * 18);
    }
    enum { AI = 1, AIJ = 2, AJ = 3, TI = 1, TJ = 2 };

    std::vector<V> wk((size_t)gall * 8);
    for (int l = 1; l <= lall; ++l) {
        const bool sgp = have_sgp[l - 1] != 0;
        // ---- GMTR_p_setup (mod_gmtr.f90:84-170) ----
        auto W8 = [&](int g, int v) -> V & { return wk[(size_t)(g - 1) * 8 + v]; };
        for (int j = gmin; j <= gmax; ++j)
            for (int i = gmin; i <= gmax; ++i) {
                const int ij = suf(i, j), im1j = suf(i - 1, j), im1jm1 = suf(i - 1, j - 1), ijm1 = suf(i, j - 1);
                W8(ij, 0) = X(ij, l);
                W8(ij, 1) = XT(ijm1, l, TJ);
                W8(ij, 2) = XT(ij, l, TI);
                W8(ij, 3) = XT(ij, l, TJ);
                W8(ij, 4) = XT(im1j, l, TI);
                W8(ij, 5) = XT(im1jm1, l, TJ);
                W8(ij, 6) = XT(im1jm1, l, TI);
                W8(ij, 7) = W8(ij, 1);
            }
        if (sgp) {
            W8(suf(gmin, gmin), 6) = W8(suf(gmin, gmin), 1);
            W8(suf(gmin, gmin), 7) = W8(suf(gmin, gmin), 1);
        }
        for (int j = gmin; j <= gmax; ++j)
            for (int i = gmin; i <= gmax; ++i) {
                const int ij = suf(i, j);
                for (int v = 0; v < 8; ++v) W8(ij, v) = scale(W8(ij, v), radius);
                double area = 0.0;
                for (int v = 1; v <= 6; ++v) area = area + triangle(W8(ij, 0), W8(ij, v), W8(ij, v + 1), radius);
                P(ij, l, 1) = area;
                P(ij, l, 2) = 1.0 / P(ij, l, 1);
            }
        for (int j = gmin; j <= gmax; ++j)
            for (int i = gmin; i <= gmax; ++i) {
                const int ij = suf(i, j);
                const double lon = GRD_s[((size_t)1 * lall + (l - 1)) * gall + (ij - 1)];
                double sl, cl;
                sincos_lon(lon, sl, cl);
                const V x = X(ij, l);
                P(ij, l, 3) = -sl;
                P(ij, l, 4) = cl;
                P(ij, l, 5) = 0.0;
                P(ij, l, 6) = -(x.x[2] * cl) / radius;
                P(ij, l, 7) = -(x.x[2] * sl) / radius;
                P(ij, l, 8) = (x.x[0] * cl + x.x[1] * sl) / radius;
            }

        // ---- GMTR_t_setup (mod_gmtr.f90:268-352) ----
        auto W4 = [&](int g, int t, int v) -> V & { return wk[((size_t)(g - 1) * 2 + (t - 1)) * 4 + v]; };
        for (int j = gmin - 1; j <= gmax; ++j)
            for (int i = gmin - 1; i <= gmax; ++i) {
                const int ij = suf(i, j), ip1j = suf(i + 1, j), ip1jp1 = suf(i + 1, j + 1), ijp1 = suf(i, j + 1);
                W4(ij, TI, 0) = XT(ij, l, TI);
                W4(ij, TI, 1) = X(ij, l);
                W4(ij, TI, 2) = X(ip1j, l);
                W4(ij, TI, 3) = X(ip1jp1, l);
                W4(ij, TJ, 0) = XT(ij, l, TJ);
                W4(ij, TJ, 1) = X(ij, l);
                W4(ij, TJ, 2) = X(ip1jp1, l);
                W4(ij, TJ, 3) = X(ijp1, l);
            }
        for (int v = 0; v < 4; ++v) {
            W4(suf(gmax, gmin - 1), TI, v) = W4(suf(gmax, gmin - 1), TJ, v);
            W4(suf(gmin - 1, gmax), TJ, v) = W4(suf(gmin - 1, gmax), TI, v);
        }
        if (sgp)
            for (int v = 0; v < 4; ++v) W4(suf(gmin - 1, gmin - 1), TI, v) = W4(suf(gmin, gmin - 1), TJ, v);
        for (int t = TI; t <= TJ; ++t)
            for (int j = gmin - 1; j <= gmax; ++j)
                for (int i = gmin - 1; i <= gmax; ++i) {
                    const int ij = suf(i, j);
                    for (int v = 0; v < 4; ++v) W4(ij, t, v) = scale(W4(ij, t, v), radius);
                    const double a1 = triangle(W4(ij, t, 0), W4(ij, t, 2), W4(ij, t, 3), radius);
                    const double a2 = triangle(W4(ij, t, 0), W4(ij, t, 3), W4(ij, t, 1), radius);
                    const double a3 = triangle(W4(ij, t, 0), W4(ij, t, 1), W4(ij, t, 2), radius);
                    const double area = a1 + a2 + a3;
                    T(ij, l, t, 1) = area;
                    T(ij, l, t, 2) = 1.0 / area;
                    T(ij, l, t, 3) = a1 / area;
                    T(ij, l, t, 4) = a2 / area;
                    T(ij, l, t, 5) = a3 / area;
                }

        // ---- GMTR_a_setup (mod_gmtr.f90:435-650) ----
        auto W2 = [&](int g, int v) -> V & { return wk[(size_t)(g - 1) * 2 + v]; };
        auto put_t = [&](int ij, int a, int base) {  // base 7: TN/TT, 1: HN/HT
            V vt, vn;
            tnvec(vt, vn, W2(ij, 0), W2(ij, 1), radius);
            for (int d = 0; d < 3; ++d) {
                A(ij, l, a, base + d) = vn.x[d];
                A(ij, l, a, base + 3 + d) = vt.x[d];
            }
        };
        // AI, triangle edge from x(ij) to x(ip1j)
        for (int j = gmin - 1; j <= gmax + 1; ++j)
            for (int i = gmin - 1; i <= gmax; ++i) {
                const int ij = suf(i, j);
                W2(ij, 0) = X(ij, l);
                W2(ij, 1) = X(suf(i + 1, j), l);
            }
        W2(suf(gmax, gmin - 1), 0) = X(suf(gmax, gmin - 1), l);
        W2(suf(gmax, gmin - 1), 1) = X(suf(gmax, gmin), l);
        W2(suf(gmin - 1, gmax + 1), 0) = X(suf(gmin, gmax + 1), l);
        W2(suf(gmin - 1, gmax + 1), 1) = X(suf(gmin, gmax), l);
        if (sgp) {
            W2(suf(gmin - 1, gmin - 1), 0) = X(suf(gmin, gmin - 1), l);
            W2(suf(gmin - 1, gmin - 1), 1) = X(suf(gmin + 1, gmin), l);
        }
        for (int j = gmin - 1; j <= gmax + 1; ++j)
            for (int i = gmin - 1; i <= gmax; ++i) put_t(suf(i, j), AI, 7);
        // AIJ, from x(ij) to x(ip1jp1)
        for (int j = gmin - 1; j <= gmax; ++j)
            for (int i = gmin - 1; i <= gmax; ++i) {
                const int ij = suf(i, j);
                W2(ij, 0) = X(ij, l);
                W2(ij, 1) = X(suf(i + 1, j + 1), l);
            }
        for (int j = gmin - 1; j <= gmax; ++j)
            for (int i = gmin - 1; i <= gmax; ++i) put_t(suf(i, j), AIJ, 7);
        // AJ, from x(ij) to x(ijp1)
        for (int j = gmin - 1; j <= gmax; ++j)
            for (int i = gmin - 1; i <= gmax + 1; ++i) {
                const int ij = suf(i, j);
                W2(ij, 0) = X(ij, l);
                W2(ij, 1) = X(suf(i, j + 1), l);
            }
        W2(suf(gmax + 1, gmin - 1), 0) = X(suf(gmax + 1, gmin), l);
        W2(suf(gmax + 1, gmin - 1), 1) = X(suf(gmax, gmin), l);
        W2(suf(gmin - 1, gmax), 0) = X(suf(gmin - 1, gmax), l);
        W2(suf(gmin - 1, gmax), 1) = X(suf(gmin, gmax), l);
        for (int j = gmin - 1; j <= gmax; ++j)
            for (int i = gmin - 1; i <= gmax + 1; ++i) put_t(suf(i, j), AJ, 7);
    }
    for (int l = 1; l <= lall; ++l) {
        const bool sgp = have_sgp[l - 1] != 0;
        auto W2 = [&](int g, int v) -> V & { return wk[(size_t)(g - 1) * 2 + v]; };
        auto put_h = [&](int ij, int a) {
            V vt, vn;
            tnvec(vt, vn, W2(ij, 0), W2(ij, 1), radius);
            for (int d = 0; d < 3; ++d) {
                A(ij, l, a, 1 + d) = vn.x[d];
                A(ij, l, a, 4 + d) = vt.x[d];
            }
        };
        // AI, hexagon edge from xt(ij,TI) to xt(ijm1,TJ)
        for (int j = gmin; j <= gmax; ++j)
            for (int i = gmin - 1; i <= gmax; ++i) {
                const int ij = suf(i, j);
                W2(ij, 0) = XT(ij, l, TI);
                W2(ij, 1) = XT(suf(i, j - 1), l, TJ);
            }
        for (int j = gmin; j <= gmax; ++j)
            for (int i = gmin - 1; i <= gmax; ++i) put_h(suf(i, j), AI);
        // AIJ, from xt(ij,TJ) to xt(ij,TI)
        for (int j = gmin - 1; j <= gmax; ++j)
            for (int i = gmin - 1; i <= gmax; ++i) {
                const int ij = suf(i, j);
                W2(ij, 0) = XT(ij, l, TJ);
                W2(ij, 1) = XT(ij, l, TI);
            }
        W2(suf(gmax, gmin - 1), 0) = XT(suf(gmax, gmin - 1), l, TJ);
        W2(suf(gmax, gmin - 1), 1) = XT(suf(gmax, gmin), l, TI);
        W2(suf(gmin - 1, gmax), 0) = XT(suf(gmin, gmax), l, TJ);
        W2(suf(gmin - 1, gmax), 1) = XT(suf(gmin - 1, gmax), l, TI);
        for (int j = gmin - 1; j <= gmax; ++j)
            for (int i = gmin - 1; i <= gmax; ++i) put_h(suf(i, j), AIJ);
        // AJ, from xt(im1j,TI) to xt(ij,TJ)
        for (int j = gmin - 1; j <= gmax; ++j)
            for (int i = gmin; i <= gmax; ++i) {
                const int ij = suf(i, j);
                W2(ij, 0) = XT(suf(i - 1, j), l, TI);
                W2(ij, 1) = XT(ij, l, TJ);
            }
        if (sgp) {
            W2(suf(gmin, gmin - 1), 0) = XT(suf(gmin, gmin), l, TI);
            W2(suf(gmin, gmin - 1), 1) = XT(suf(gmin, gmin - 1), l, TJ);
        }
        for (int j = gmin - 1; j <= gmax; ++j)
            for (int i = gmin; i <= gmax; ++i) put_h(suf(i, j), AJ);
    }

    if (have_pl) {
        const int n = 1, gmin_pl = 2, gmax_pl = gall_pl;
        for (int l = 1; l <= 2; ++l) {
            // GMTR_p_setup poles (:172-229)
            V w[16];
            w[0] = XPL(n, l);
            for (int v = 1; v <= gall_pl - 1; ++v) w[v] = XTPL(v + 1, l);
            w[gall_pl] = w[1];
            for (int v = 0; v <= gall_pl; ++v) w[v] = scale(w[v], radius);
            double area = 0.0;
            for (int v = 1; v <= gall_pl - 1; ++v) area = area + triangle(w[0], w[v], w[v + 1], radius);
            PPL(n, l, 1) = area;
            PPL(n, l, 2) = 1.0 / PPL(n, l, 1);
            const double lon = GRD_s_pl[((size_t)1 * 2 + (l - 1)) * npl + (n - 1)];
            double sl, cl;
                sincos_lon(lon, sl, cl);
            const V x = XPL(n, l);
            PPL(n, l, 3) = -sl;
            PPL(n, l, 4) = cl;
            PPL(n, l, 5) = 0.0;
            PPL(n, l, 6) = -(x.x[2] * cl) / radius;
            PPL(n, l, 7) = -(x.x[2] * sl) / radius;
            PPL(n, l, 8) = (x.x[0] * cl + x.x[1] * sl) / radius;
            // GMTR_t_setup poles (:355-395)
            for (int v = gmin_pl; v <= gmax_pl; ++v) {
                const int ij = v;
                int ijp1 = v + 1;
                if (ijp1 == gmax_pl + 1) ijp1 = gmin_pl;
                V t4[4] = {XTPL(ij, l), XPL(n, l), XPL(ij, l), XPL(ijp1, l)};
                for (V &q : t4) q = scale(q, radius);
                const double a1 = triangle(t4[0], t4[2], t4[3], radius);
                const double a2 = triangle(t4[0], t4[3], t4[1], radius);
                const double a3 = triangle(t4[0], t4[1], t4[2], radius);
                const double ar = a1 + a2 + a3;
                TPL(ij, l, 1) = ar;
                TPL(ij, l, 2) = 1.0 / ar;
                TPL(ij, l, 3) = a1 / ar;
                TPL(ij, l, 4) = a2 / ar;
                TPL(ij, l, 5) = a3 / ar;
            }
            // GMTR_a_setup poles (:653-738)
            for (int v = gmin_pl; v <= gmax_pl; ++v) {
                V vt, vn;
                tnvec(vt, vn, XPL(n, l), XPL(v, l), radius);
                for (int d = 0; d < 3; ++d) {
                    APL(v, l, 7 + d) = vn.x[d];
                    APL(v, l, 10 + d) = vt.x[d];
                }
            }
            for (int v = gmin_pl; v <= gmax_pl; ++v) {
                int ijp1 = v + 1;
                if (ijp1 == gmax_pl + 1) ijp1 = gmin_pl;
                V vt, vn;
                tnvec(vt, vn, XPL(v, l), XPL(ijp1, l), radius);
                for (int d = 0; d < 3; ++d) {
                    APL(v, l, 13 + d) = vn.x[d];
                    APL(v, l, 16 + d) = vt.x[d];
                }
            }
            for (int v = gmin_pl; v <= gmax_pl; ++v) {
                int ijm1 = v - 1;
                if (ijm1 == gmin_pl - 1) ijm1 = gmax_pl;
                V vt, vn;
                tnvec(vt, vn, XTPL(ijm1, l), XTPL(v, l), radius);
                for (int d = 0; d < 3; ++d) {
                    APL(v, l, 1 + d) = vn.x[d];
                    APL(v, l, 4 + d) = vt.x[d];
                }
            }
        }
    }
    return NICAM_OK;
}

// Triangle vertices, latitude/longitude and cell-wall midpoints of a grid
// whose GRD_x halo is filled.  GRD_xt(gall,1,lall,2,3): TI of cell ij is the
// normalised centroid of x(ij), x(ip1j), x(ip1jp1), TJ of x(ij), x(ip1jp1),
// x(ijp1), for i, j in [gmin-1, gmax]; GRD_xt_pl ring v is the centroid of the
// pole, x(v) and x(v+1).  GRD_s(gall,1,lall,2) = (lat, lon) as
// VECTR_xyz2latlon (mod_vector.f90:56-113) gives them.  GRD_xr(gall,1,lall,3,3)
// arc a of cell ij is the normalised midpoint of the hexagon edge that GMTR_a
// gives its HN/HT (the two vertices of mod_gmtr.f90:561-650); GRD_xr_pl ring v
// is the midpoint of xt(v-1) and xt(v).  All points scaled to the radius.
extern "C" int nicam_grid_vertices(int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp,
                                   double radius, const double *GRD_x, const double *GRD_x_pl, double *GRD_xt,
                                   double *GRD_xt_pl, double *GRD_s, double *GRD_s_pl, double *GRD_xr,
                                   double *GRD_xr_pl)
{
    if (gall_1d < 4 || lall < 1 || gall_pl < 3 || gall_pl > 15) return fail(NICAM_ERR_ARG, "sizes out of range");
    if (!have_sgp || !GRD_x || !GRD_xt || !GRD_s || !GRD_xr) return fail(NICAM_ERR_ARG, "null argument");
    if (have_pl && (!GRD_x_pl || !GRD_xt_pl || !GRD_s_pl || !GRD_xr_pl)) return fail(NICAM_ERR_ARG, "null pole argument");
    const int iall = gall_1d;
    const size_t gall = (size_t)iall * iall, npl = gall_pl;
    const int gmin = 2, gmax = iall - 1;
    auto suf = [&](int i, int j) { return (j - 1) * iall + i; };
    auto X = [&](int g, int l) {
        V v;
        for (int d = 0; d < 3; ++d) v.x[d] = GRD_x[((size_t)d * lall + (l - 1)) * gall + (g - 1)];
        return v;
    };
    auto unit3 = [&](const V &a, const V &b, const V &c) {
        V s;
        for (int d = 0; d < 3; ++d) s.x[d] = a.x[d] + b.x[d] + c.x[d];
        const double r = vabs(s);
        for (int d = 0; d < 3; ++d) s.x[d] = s.x[d] / r * radius;
        return s;
    };
    auto latlon = [&](const V &p, double &lat, double &lon) {  // mod_vector.f90:56-113
        const double x = p.x[0], y = p.x[1], z = p.x[2];
        const double length = std::sqrt(x * x + y * y + z * z);
        lat = lon = 0.0;
        if (length < DBL_EPSILON) return;
        if (z / length >= 1.0) { lat = std::asin(1.0); return; }
        if (z / length <= -1.0) { lat = std::asin(-1.0); return; }
        lat = std::asin(z / length);
        const double lh = std::sqrt(x * x + y * y);
        if (lh < DBL_EPSILON) return;
        if (x / lh >= 1.0) lon = std::acos(1.0);
        else if (x / lh <= -1.0) lon = std::acos(-1.0);
        else lon = std::acos(x / lh);
        if (y < 0.0) lon = -lon;
    };
    std::memset(GRD_xt, 0, sizeof(double) * gall * lall * 6);
    std::memset(GRD_s, 0, sizeof(double) * gall * lall * 2);
    std::memset(GRD_xr, 0, sizeof(double) * gall * lall * 9);
    for (int l = 1; l <= lall; ++l) {
        auto put_xt = [&](int g, int t, const V &v) {
            for (int d = 0; d < 3; ++d) GRD_xt[(((size_t)d * 2 + t) * lall + (l - 1)) * gall + (g - 1)] = v.x[d];
        };
        auto XT = [&](int g, int t) {
            V v;
            for (int d = 0; d < 3; ++d) v.x[d] = GRD_xt[(((size_t)d * 2 + t) * lall + (l - 1)) * gall + (g - 1)];
            return v;
        };
        for (int j = gmin - 1; j <= gmax; ++j)
            for (int i = gmin - 1; i <= gmax; ++i) {
                const int ij = suf(i, j);
                put_xt(ij, 0, unit3(X(ij, l), X(suf(i + 1, j), l), X(suf(i + 1, j + 1), l)));
                put_xt(ij, 1, unit3(X(ij, l), X(suf(i + 1, j + 1), l), X(suf(i, j + 1), l)));
            }
        // At a pentagon the triangle TI(gmin-1,gmin-1) is TJ(gmin,gmin-1), as
        // GMTR_t_setup treats it (mod_gmtr.f90:312-314): the fifth wall of the
        // pentagon then runs from xt(gmin-1,gmin-1,TJ) to xt(gmin,gmin-1,TJ).
        if (have_sgp[l - 1]) put_xt(suf(gmin - 1, gmin - 1), 0, XT(suf(gmin, gmin - 1), 1));
        for (int g = 1; g <= (int)gall; ++g) {
            double lat, lon;
            latlon(X(g, l), lat, lon);
            GRD_s[((size_t)0 * lall + (l - 1)) * gall + (g - 1)] = lat;
            GRD_s[((size_t)1 * lall + (l - 1)) * gall + (g - 1)] = lon;
        }
        auto put_xr = [&](int g, int a, const V &p, const V &q) {
            V s;
            for (int d = 0; d < 3; ++d) s.x[d] = p.x[d] + q.x[d];
            const double r = vabs(s);
            for (int d = 0; d < 3; ++d) GRD_xr[(((size_t)d * 3 + a) * lall + (l - 1)) * gall + (g - 1)] = s.x[d] / r * radius;
        };
        const bool sgp = have_sgp[l - 1] != 0;
        for (int j = gmin; j <= gmax; ++j)
            for (int i = gmin - 1; i <= gmax; ++i) put_xr(suf(i, j), 0, XT(suf(i, j), 0), XT(suf(i, j - 1), 1));
        for (int j = gmin - 1; j <= gmax; ++j)
            for (int i = gmin - 1; i <= gmax; ++i) put_xr(suf(i, j), 1, XT(suf(i, j), 1), XT(suf(i, j), 0));
        put_xr(suf(gmax, gmin - 1), 1, XT(suf(gmax, gmin - 1), 1), XT(suf(gmax, gmin), 0));
        put_xr(suf(gmin - 1, gmax), 1, XT(suf(gmin, gmax), 1), XT(suf(gmin - 1, gmax), 0));
        for (int j = gmin - 1; j <= gmax; ++j)
            for (int i = gmin; i <= gmax; ++i) put_xr(suf(i, j), 2, XT(suf(i - 1, j), 0), XT(suf(i, j), 1));
        if (sgp) put_xr(suf(gmin, gmin - 1), 2, XT(suf(gmin, gmin), 0), XT(suf(gmin, gmin - 1), 1));
    }
    if (have_pl) {
        std::memset(GRD_xt_pl, 0, sizeof(double) * npl * 2 * 3);
        std::memset(GRD_s_pl, 0, sizeof(double) * npl * 2 * 2);
        std::memset(GRD_xr_pl, 0, sizeof(double) * npl * 2 * 3);
        for (int l = 1; l <= 2; ++l) {
            auto XP = [&](int v) {
                V p;
                for (int d = 0; d < 3; ++d) p.x[d] = GRD_x_pl[((size_t)d * 2 + (l - 1)) * npl + (v - 1)];
                return p;
            };
            V xt[16];
            for (int v = 2; v <= gall_pl; ++v) {
                const int vp1 = v + 1 == gall_pl + 1 ? 2 : v + 1;
                xt[v] = unit3(XP(1), XP(v), XP(vp1));
                for (int d = 0; d < 3; ++d) GRD_xt_pl[((size_t)d * 2 + (l - 1)) * npl + (v - 1)] = xt[v].x[d];
            }
            for (int v = 1; v <= gall_pl; ++v) {
                double lat, lon;
                latlon(XP(v), lat, lon);
                GRD_s_pl[((size_t)0 * 2 + (l - 1)) * npl + (v - 1)] = lat;
                GRD_s_pl[((size_t)1 * 2 + (l - 1)) * npl + (v - 1)] = lon;
            }
            for (int v = 2; v <= gall_pl; ++v) {
                const int vm1 = v - 1 == 1 ? gall_pl : v - 1;
                V s;
                for (int d = 0; d < 3; ++d) s.x[d] = xt[vm1].x[d] + xt[v].x[d];
                const double r = vabs(s);
                for (int d = 0; d < 3; ++d) GRD_xr_pl[((size_t)d * 2 + (l - 1)) * npl + (v - 1)] = s.x[d] / r * radius;
            }
        }
    }
    return NICAM_OK;
}
