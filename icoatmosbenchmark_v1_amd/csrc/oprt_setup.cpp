// Stencil coefficients of the horizontal operators (host code, set-up time).
//
// nicam_oprt_diffusion_setup restates OPRT_diffusion_setup
// (kernels/NICAM/dyn_metrics/src/mod_oprt.f90:1235-1390): the coefficients
// coef_intp (vector at the triangle centres from the arc normals) and coef_diff
// (the flux through the six hexagon walls) that OPRT_diffusion and
// OPRT3D_divdamp read.  The flat index loops and the operand order are the
// reference's, so the coefficients match its compiled kernel bit for bit,
// halo entries included.
#include <cstring>

#include "internal.h"

using namespace nicam;

namespace {

// index positions of the metric arrays (dyn_metrics/src/problem_size.inc:80-125)
constexpr int P_RAREA = 2, T_RAREA = 2, HNX = 1, TNX = 7, TN2X = 13;
constexpr int TI = 1, TJ = 2, AI = 1, AIJ = 2, AJ = 3;

struct Metrics {
    const double *p, *p_pl, *t, *t_pl, *a, *a_pl;
    size_t gall, lall, npl;
    // 1-based accessors in the reference's shapes (k0 = 1 dropped)
    double P(size_t g, size_t l, int c) const { return p[((size_t)(c - 1) * lall + (l - 1)) * gall + (g - 1)]; }
    double T(size_t g, size_t l, int t_, int c) const
    {
        return t[(((size_t)(c - 1) * 2 + (t_ - 1)) * lall + (l - 1)) * gall + (g - 1)];
    }
    double A(size_t g, size_t l, int a_, int c) const
    {
        return a[(((size_t)(c - 1) * 3 + (a_ - 1)) * lall + (l - 1)) * gall + (g - 1)];
    }
    double PPL(size_t v, size_t l, int c) const { return p_pl[((size_t)(c - 1) * 2 + (l - 1)) * npl + (v - 1)]; }
    double TPL(size_t v, size_t l, int c) const { return t_pl[((size_t)(c - 1) * 2 + (l - 1)) * npl + (v - 1)]; }
    double APL(size_t v, size_t l, int c) const { return a_pl[((size_t)(c - 1) * 2 + (l - 1)) * npl + (v - 1)]; }
};

}  // namespace

extern "C" int nicam_oprt_diffusion_setup(int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp,
                                          const double *GMTR_p, const double *GMTR_p_pl, const double *GMTR_t,
                                          const double *GMTR_t_pl, const double *GMTR_a, const double *GMTR_a_pl,
                                          double *coef_intp, double *coef_intp_pl, double *coef_diff,
                                          double *coef_diff_pl)
{
    if (gall_1d < 4 || lall < 1 || gall_pl < 3 || gall_pl > 15) return fail(NICAM_ERR_ARG, "sizes out of range");
    if (!have_sgp || !GMTR_p || !GMTR_t || !GMTR_a || !coef_intp || !coef_diff)
        return fail(NICAM_ERR_ARG, "null argument");
    if (have_pl && (!GMTR_p_pl || !GMTR_t_pl || !GMTR_a_pl || !coef_intp_pl || !coef_diff_pl))
        return fail(NICAM_ERR_ARG, "null pole argument");
    const size_t iall = gall_1d, gall = iall * iall, npl = gall_pl, vlink = npl - 1;
    const Metrics M{GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, gall, (size_t)lall, npl};
    // coef_intp(gall,3,3,TI:TJ,lall), coef_diff(gall,6,3,lall), 1-based
    auto CI = [&](size_t g, int m, int d, int t, size_t l) -> double & {
        return coef_intp[((((l - 1) * 2 + (t - 1)) * 3 + (d - 1)) * 3 + (m - 1)) * gall + (g - 1)];
    };
    auto CD = [&](size_t g, int m, int d, size_t l) -> double & {
        return coef_diff[(((l - 1) * 3 + (d - 1)) * 6 + (m - 1)) * gall + (g - 1)];
    };
    std::memset(coef_intp, 0, sizeof(double) * gall * 18 * lall);
    std::memset(coef_diff, 0, sizeof(double) * gall * 18 * lall);
    if (have_pl) {
        std::memset(coef_intp_pl, 0, sizeof(double) * npl * 9 * 2);
        std::memset(coef_diff_pl, 0, sizeof(double) * vlink * 3 * 2);
    }
    const size_t gmin = iall + 2, gmax = (iall - 2) * iall + (iall - 1), gminm1 = 1;  // flat, 1-based

    for (size_t l = 1; l <= (size_t)lall; ++l)
        for (int d = 1; d <= 3; ++d) {
            const int tn = d + TNX - 1;
            for (size_t g = gminm1; g <= gmax; ++g) {
                const size_t ij = g, ip1j = g + 1, ijp1 = g + iall;
                const double rti = M.T(ij, l, TI, T_RAREA), rtj = M.T(ij, l, TJ, T_RAREA);
                CI(ij, 1, d, TI, l) = (+M.A(ij, l, AIJ, tn) - M.A(ij, l, AI, tn)) * 0.5 * rti;
                CI(ij, 2, d, TI, l) = (-M.A(ij, l, AI, tn) - M.A(ip1j, l, AJ, tn)) * 0.5 * rti;
                CI(ij, 3, d, TI, l) = (-M.A(ip1j, l, AJ, tn) + M.A(ij, l, AIJ, tn)) * 0.5 * rti;
                CI(ij, 1, d, TJ, l) = (+M.A(ij, l, AJ, tn) - M.A(ij, l, AIJ, tn)) * 0.5 * rtj;
                CI(ij, 2, d, TJ, l) = (-M.A(ij, l, AIJ, tn) + M.A(ijp1, l, AI, tn)) * 0.5 * rtj;
                CI(ij, 3, d, TJ, l) = (+M.A(ijp1, l, AI, tn) + M.A(ij, l, AJ, tn)) * 0.5 * rtj;
            }
        }
    for (size_t l = 1; l <= (size_t)lall; ++l)
        for (int d = 1; d <= 3; ++d) {
            const int hn = d + HNX - 1;
            for (size_t g = gmin; g <= gmax; ++g) {
                const size_t ij = g, im1j = g - 1, im1jm1 = g - iall - 1, ijm1 = g - iall;
                const double rp = M.P(ij, l, P_RAREA);
                CD(ij, 1, d, l) = +M.A(ij, l, AIJ, hn) * 0.5 * rp;
                CD(ij, 2, d, l) = +M.A(ij, l, AJ, hn) * 0.5 * rp;
                CD(ij, 3, d, l) = -(M.A(im1j, l, AI, hn) * 0.5 * rp);
                CD(ij, 4, d, l) = -(M.A(im1jm1, l, AIJ, hn) * 0.5 * rp);
                CD(ij, 5, d, l) = -(M.A(ijm1, l, AJ, hn) * 0.5 * rp);
                CD(ij, 6, d, l) = +M.A(ij, l, AI, hn) * 0.5 * rp;
            }
            if (have_sgp[l - 1]) CD(gmin, 5, d, l) = 0.0;
        }
    if (have_pl) {
        const size_t n = 1, gmin_pl = 2, gmax_pl = npl;
        // coef_intp_pl(gall_pl,3,3,2), coef_diff_pl(vlink,3,2)
        auto CIP = [&](size_t v, int m, int d, size_t l) -> double & {
            return coef_intp_pl[(((l - 1) * 3 + (d - 1)) * 3 + (m - 1)) * npl + (v - 1)];
        };
        for (size_t l = 1; l <= 2; ++l)
            for (int d = 1; d <= 3; ++d) {
                const int hn = d + HNX - 1, tn = d + TNX - 1, tn2 = d + TN2X - 1;
                for (size_t v = gmin_pl; v <= gmax_pl; ++v) {
                    const size_t ij = v, ijp1 = v + 1 == gmax_pl + 1 ? gmin_pl : v + 1;
                    double c[3];
                    c[0] = -M.APL(ijp1, l, tn) + M.APL(ij, l, tn);
                    c[1] = +M.APL(ij, l, tn) + M.APL(ij, l, tn2);
                    c[2] = +M.APL(ij, l, tn2) - M.APL(ijp1, l, tn);
                    const double rt = M.TPL(v, l, T_RAREA);
                    for (int m = 0; m < 3; ++m) CIP(v, m + 1, d, l) = c[m] * 0.5 * rt;
                    coef_diff_pl[((l - 1) * 3 + (d - 1)) * vlink + (v - 2)] = M.APL(v, l, hn) * 0.5 * M.PPL(n, l, P_RAREA);
                }
            }
    }
    return NICAM_OK;
}
