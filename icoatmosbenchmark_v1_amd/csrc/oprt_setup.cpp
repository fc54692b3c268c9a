// Stencil coefficients of the horizontal operators (host code, set-up time).
//
// nicam_oprt_diffusion_setup restates OPRT_diffusion_setup
// (kernels/NICAM/dyn_metrics/src/mod_oprt.f90:1235-1390): the coefficients
// coef_intp (vector at the triangle centres from the arc normals) and coef_diff
// (the flux through the six hexagon walls) that OPRT_diffusion and
// OPRT3D_divdamp read.  The flat index loops and the operand order are the
// reference's, so the coefficients match its compiled kernel bit for bit,
// halo entries included.
#include <cstdint>
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

// ---------------------------------------------------------------------------
// OPRT_divergence_setup, OPRT_rotation_setup, OPRT_gradient_setup and
// OPRT_laplacian_setup (dyn_metrics/src/mod_oprt.f90:45-269, 272-496,
// 499-734, 737-1232).  Each coefficient of the reference is a signed sum of
// metric products evaluated left to right; the tables below list those terms
// in the reference's order (the order fixes the rounding), and one evaluator
// runs them.  Cells are named by their offset from ij on the flat index.
// ---------------------------------------------------------------------------
namespace {

enum Cell : int8_t { IJ, IM1J, IM1JM1, IJM1, IP1J, IJP1 };
constexpr int W1 = 3, W2 = 4, W3 = 5;

long cell_offset(int c, long iall)
{
    switch (c) {
    case IM1J: return -1;
    case IM1JM1: return -iall - 1;
    case IJM1: return -iall;
    case IP1J: return 1;
    case IJP1: return iall;
    default: return 0;
    }
}

// One term of a divergence/rotation/gradient coefficient: neg ? - : +
// GMTR_t(tc,tt,w) * GMTR_a(ac,arc,comp), or (tt == 0) 2 * GMTR_a(ac,arc,comp).
struct VTerm {
    int8_t neg, tc, tt, w, ac, arc;
};
#define TP(tc, tt, w, ac, arc) {0, tc, tt, w, ac, arc}
#define TM(tc, tt, w, ac, arc) {1, tc, tt, w, ac, arc}
#define P2(ac, arc) {0, IJ, 0, 0, ac, arc}
#define M2(ac, arc) {1, IJ, 0, 0, ac, arc}

struct VCoef {
    int n;
    VTerm t[18];
};

// mod_oprt.f90:111-171 (divergence; rotation :338-398 has the same terms)
const VCoef kDivHex[7] = {
    {12, {TP(IJ, TI, W1, IJ, AI), TP(IJ, TI, W1, IJ, AIJ), TP(IJ, TJ, W1, IJ, AIJ), TP(IJ, TJ, W1, IJ, AJ),
          TP(IM1J, TI, W2, IJ, AJ), TM(IM1J, TI, W2, IM1J, AI), TM(IM1JM1, TJ, W2, IM1J, AI),
          TM(IM1JM1, TJ, W2, IM1JM1, AIJ), TM(IM1JM1, TI, W3, IM1JM1, AIJ), TM(IM1JM1, TI, W3, IJM1, AJ),
          TM(IJM1, TJ, W3, IJM1, AJ), TP(IJM1, TJ, W3, IJ, AI)}},
    {4, {TM(IJM1, TJ, W2, IJM1, AJ), TP(IJM1, TJ, W2, IJ, AI), TP(IJ, TI, W2, IJ, AI), TP(IJ, TI, W2, IJ, AIJ)}},
    {4, {TP(IJ, TI, W3, IJ, AI), TP(IJ, TI, W3, IJ, AIJ), TP(IJ, TJ, W2, IJ, AIJ), TP(IJ, TJ, W2, IJ, AJ)}},
    {4, {TP(IJ, TJ, W3, IJ, AIJ), TP(IJ, TJ, W3, IJ, AJ), TP(IM1J, TI, W3, IJ, AJ), TM(IM1J, TI, W3, IM1J, AI)}},
    {4, {TP(IM1J, TI, W1, IJ, AJ), TM(IM1J, TI, W1, IM1J, AI), TM(IM1JM1, TJ, W3, IM1J, AI),
         TM(IM1JM1, TJ, W3, IM1JM1, AIJ)}},
    {4, {TM(IM1JM1, TJ, W1, IM1J, AI), TM(IM1JM1, TJ, W1, IM1JM1, AIJ), TM(IM1JM1, TI, W1, IM1JM1, AIJ),
         TM(IM1JM1, TI, W1, IJM1, AJ)}},
    {4, {TM(IM1JM1, TI, W2, IM1JM1, AIJ), TM(IM1JM1, TI, W2, IJM1, AJ), TM(IJM1, TJ, W1, IJM1, AJ),
         TP(IJM1, TJ, W1, IJ, AI)}},
};
// mod_oprt.f90:183-236: the pentagon's five walls
const VCoef kDivPent[7] = {
    {10, {TP(IJ, TI, W1, IJ, AI), TP(IJ, TI, W1, IJ, AIJ), TP(IJ, TJ, W1, IJ, AIJ), TP(IJ, TJ, W1, IJ, AJ),
          TP(IM1J, TI, W2, IJ, AJ), TM(IM1J, TI, W2, IM1J, AI), TM(IM1JM1, TJ, W2, IM1J, AI),
          TM(IM1JM1, TJ, W2, IM1JM1, AIJ), TM(IJM1, TJ, W3, IM1JM1, AIJ), TP(IJM1, TJ, W3, IJ, AI)}},
    {4, {TM(IJM1, TJ, W2, IM1JM1, AIJ), TP(IJM1, TJ, W2, IJ, AI), TP(IJ, TI, W2, IJ, AI), TP(IJ, TI, W2, IJ, AIJ)}},
    kDivHex[2],
    kDivHex[3],
    kDivHex[4],
    {2, {TM(IM1JM1, TJ, W1, IM1J, AI), TM(IM1JM1, TJ, W1, IM1JM1, AIJ)}},
    {2, {TM(IJM1, TJ, W1, IM1JM1, AIJ), TP(IJM1, TJ, W1, IJ, AI)}},
};
// mod_oprt.f90:565-631 (gradient): the Q4/Q5 terms of coefficient 0 come in
// another order than in the divergence, and six P0 terms follow
const VCoef kGradHex0 = {
    18, {TP(IJ, TI, W1, IJ, AI), TP(IJ, TI, W1, IJ, AIJ), TP(IJ, TJ, W1, IJ, AIJ), TP(IJ, TJ, W1, IJ, AJ),
         TP(IM1J, TI, W2, IJ, AJ), TM(IM1J, TI, W2, IM1J, AI), TM(IM1JM1, TJ, W2, IM1J, AI),
         TM(IM1JM1, TI, W3, IM1JM1, AIJ), TM(IM1JM1, TJ, W2, IM1JM1, AIJ), TM(IM1JM1, TI, W3, IJM1, AJ),
         TM(IJM1, TJ, W3, IJM1, AJ), TP(IJM1, TJ, W3, IJ, AI), M2(IJ, AIJ), M2(IJ, AJ), P2(IM1J, AI),
         P2(IM1JM1, AIJ), P2(IJM1, AJ), M2(IJ, AI)}};
// mod_oprt.f90:643-660
const VCoef kGradPent0 = {
    15, {TP(IJ, TI, W1, IJ, AI), TP(IJ, TI, W1, IJ, AIJ), TP(IJ, TJ, W1, IJ, AIJ), TP(IJ, TJ, W1, IJ, AJ),
         TP(IM1J, TI, W2, IJ, AJ), TM(IM1J, TI, W2, IM1J, AI), TM(IM1JM1, TJ, W2, IM1J, AI),
         TM(IM1JM1, TJ, W2, IM1JM1, AIJ), TM(IJM1, TJ, W3, IM1JM1, AIJ), TP(IJM1, TJ, W3, IJ, AI), M2(IJ, AIJ),
         M2(IJ, AJ), P2(IM1J, AI), P2(IM1JM1, AIJ), M2(IJ, AI)}};
#undef TP
#undef TM
#undef P2
#undef M2

// One block of a laplacian coefficient: coef(m) += GMTR_t(tc,tt,RAREA) *
// (sum of six terms neg ? - : +  c * GMTR_a(c1,a1,TN) * GMTR_a(c2,a2,HN)).
struct LTerm {
    int8_t neg, two, c1, a1, c2, a2;
};
struct LBlock {
    int8_t m, tc, tt;
    LTerm t[6];
};
#define L(s, c, c1, a1, c2, a2) {(s) < 0, (c) == 2, c1, a1, c2, a2}

// mod_oprt.f90:800-981, in the reference's order
const LBlock kLapHex[18] = {
    {0, IJ, TI, {L(-1, 1, IJ, AI, IJ, AI), L(1, 2, IP1J, AJ, IJ, AI), L(1, 1, IJ, AIJ, IJ, AI),
                 L(-1, 1, IJ, AI, IJ, AIJ), L(1, 2, IP1J, AJ, IJ, AIJ), L(1, 1, IJ, AIJ, IJ, AIJ)}},
    {0, IJ, TJ, {L(-1, 1, IJ, AIJ, IJ, AIJ), L(-1, 2, IJP1, AI, IJ, AIJ), L(1, 1, IJ, AJ, IJ, AIJ),
                 L(-1, 1, IJ, AIJ, IJ, AJ), L(-1, 2, IJP1, AI, IJ, AJ), L(1, 1, IJ, AJ, IJ, AJ)}},
    {0, IM1J, TI, {L(-1, 1, IJ, AJ, IJ, AJ), L(-1, 2, IM1J, AIJ, IJ, AJ), L(-1, 1, IM1J, AI, IJ, AJ),
                   L(1, 1, IJ, AJ, IM1J, AI), L(1, 2, IM1J, AIJ, IM1J, AI), L(1, 1, IM1J, AI, IM1J, AI)}},
    {0, IM1JM1, TJ, {L(-1, 1, IM1J, AI, IM1J, AI), L(1, 2, IM1JM1, AJ, IM1J, AI), L(1, 1, IM1JM1, AIJ, IM1J, AI),
                     L(-1, 1, IM1J, AI, IM1JM1, AIJ), L(1, 2, IM1JM1, AJ, IM1JM1, AIJ),
                     L(1, 1, IM1JM1, AIJ, IM1JM1, AIJ)}},
    {0, IM1JM1, TI, {L(-1, 1, IM1JM1, AIJ, IM1JM1, AIJ), L(-1, 2, IM1JM1, AI, IM1JM1, AIJ),
                     L(1, 1, IJM1, AJ, IM1JM1, AIJ), L(-1, 1, IM1JM1, AIJ, IJM1, AJ), L(-1, 2, IM1JM1, AI, IJM1, AJ),
                     L(1, 1, IJM1, AJ, IJM1, AJ)}},
    {0, IJM1, TJ, {L(-1, 1, IJM1, AJ, IJM1, AJ), L(-1, 2, IJM1, AIJ, IJM1, AJ), L(-1, 1, IJ, AI, IJM1, AJ),
                   L(1, 1, IJM1, AJ, IJ, AI), L(1, 2, IJM1, AIJ, IJ, AI), L(1, 1, IJ, AI, IJ, AI)}},
    {1, IJM1, TJ, {L(-1, 1, IJ, AI, IJM1, AJ), L(1, 2, IJM1, AJ, IJM1, AJ), L(1, 1, IJM1, AIJ, IJM1, AJ),
                   L(1, 1, IJ, AI, IJ, AI), L(-1, 2, IJM1, AJ, IJ, AI), L(-1, 1, IJM1, AIJ, IJ, AI)}},
    {1, IJ, TI, {L(-1, 1, IP1J, AJ, IJ, AI), L(-1, 2, IJ, AIJ, IJ, AI), L(-1, 1, IJ, AI, IJ, AI),
                 L(-1, 1, IP1J, AJ, IJ, AIJ), L(-1, 2, IJ, AIJ, IJ, AIJ), L(-1, 1, IJ, AI, IJ, AIJ)}},
    {2, IJ, TI, {L(1, 1, IJ, AIJ, IJ, AI), L(1, 2, IJ, AI, IJ, AI), L(-1, 1, IP1J, AJ, IJ, AI),
                 L(1, 1, IJ, AIJ, IJ, AIJ), L(1, 2, IJ, AI, IJ, AIJ), L(-1, 1, IP1J, AJ, IJ, AIJ)}},
    {2, IJ, TJ, {L(1, 1, IJP1, AI, IJ, AIJ), L(-1, 2, IJ, AJ, IJ, AIJ), L(-1, 1, IJ, AIJ, IJ, AIJ),
                 L(1, 1, IJP1, AI, IJ, AJ), L(-1, 1, IJ, AIJ, IJ, AJ), L(-1, 2, IJ, AJ, IJ, AJ)}},
    {3, IJ, TJ, {L(1, 1, IJ, AJ, IJ, AIJ), L(1, 2, IJ, AIJ, IJ, AIJ), L(1, 1, IJP1, AI, IJ, AIJ),
                 L(1, 1, IJ, AJ, IJ, AJ), L(1, 2, IJ, AIJ, IJ, AJ), L(1, 1, IJP1, AI, IJ, AJ)}},
    {3, IM1J, TI, {L(1, 1, IM1J, AIJ, IJ, AJ), L(1, 2, IM1J, AI, IJ, AJ), L(-1, 1, IJ, AJ, IJ, AJ),
                   L(-1, 1, IM1J, AIJ, IM1J, AI), L(-1, 2, IM1J, AI, IM1J, AI), L(1, 1, IJ, AJ, IM1J, AI)}},
    {4, IM1J, TI, {L(-1, 1, IM1J, AI, IJ, AJ), L(1, 2, IJ, AJ, IJ, AJ), L(1, 1, IM1J, AIJ, IJ, AJ),
                   L(1, 1, IM1J, AI, IM1J, AI), L(-1, 2, IJ, AJ, IM1J, AI), L(-1, 1, IM1J, AIJ, IM1J, AI)}},
    {4, IM1JM1, TJ, {L(-1, 1, IM1JM1, AJ, IM1J, AI), L(-1, 2, IM1JM1, AIJ, IM1J, AI), L(-1, 1, IM1J, AI, IM1J, AI),
                     L(-1, 1, IM1JM1, AJ, IM1JM1, AIJ), L(-1, 2, IM1JM1, AIJ, IM1JM1, AIJ),
                     L(-1, 1, IM1J, AI, IM1JM1, AIJ)}},
    {5, IM1JM1, TJ, {L(1, 1, IM1JM1, AIJ, IM1J, AI), L(1, 2, IM1J, AI, IM1J, AI), L(-1, 1, IM1JM1, AJ, IM1J, AI),
                     L(1, 1, IM1JM1, AIJ, IM1JM1, AIJ), L(1, 2, IM1J, AI, IM1JM1, AIJ),
                     L(-1, 1, IM1JM1, AJ, IM1JM1, AIJ)}},
    {5, IM1JM1, TI, {L(1, 1, IM1JM1, AI, IM1JM1, AIJ), L(-1, 2, IJM1, AJ, IM1JM1, AIJ),
                     L(-1, 1, IM1JM1, AIJ, IM1JM1, AIJ), L(1, 1, IM1JM1, AI, IJM1, AJ), L(-1, 2, IJM1, AJ, IJM1, AJ),
                     L(-1, 1, IM1JM1, AIJ, IJM1, AJ)}},
    {6, IM1JM1, TI, {L(1, 1, IJM1, AJ, IM1JM1, AIJ), L(1, 2, IM1JM1, AIJ, IM1JM1, AIJ),
                     L(1, 1, IM1JM1, AI, IM1JM1, AIJ), L(1, 1, IJM1, AJ, IJM1, AJ), L(1, 2, IM1JM1, AIJ, IJM1, AJ),
                     L(1, 1, IM1JM1, AI, IJM1, AJ)}},
    {6, IJM1, TJ, {L(1, 1, IJM1, AIJ, IJM1, AJ), L(1, 2, IJ, AI, IJM1, AJ), L(-1, 1, IJM1, AJ, IJM1, AJ),
                   L(-1, 1, IJM1, AIJ, IJ, AI), L(-1, 2, IJ, AI, IJ, AI), L(1, 1, IJM1, AJ, IJ, AI)}},
};
// mod_oprt.f90:1006-1161: the pentagon (coefficients reset to zero first)
const LBlock kLapPent[15] = {
    kLapHex[0],
    kLapHex[1],
    kLapHex[2],
    kLapHex[3],
    {0, IJM1, TJ, {L(-1, 1, IJM1, AJ, IM1JM1, AIJ), L(-1, 2, IJM1, AIJ, IM1JM1, AIJ), L(-1, 1, IJ, AI, IM1JM1, AIJ),
                   L(1, 1, IJM1, AJ, IJ, AI), L(1, 2, IJM1, AIJ, IJ, AI), L(1, 1, IJ, AI, IJ, AI)}},
    {1, IJM1, TJ, {L(1, 1, IJM1, AIJ, IM1JM1, AIJ), L(1, 2, IJM1, AJ, IM1JM1, AIJ), L(-1, 1, IJ, AI, IM1JM1, AIJ),
                   L(-1, 1, IJM1, AIJ, IJ, AI), L(-1, 2, IJM1, AJ, IJ, AI), L(1, 1, IJ, AI, IJ, AI)}},
    kLapHex[7],
    kLapHex[8],
    {2, IJ, TJ, {L(1, 1, IJP1, AI, IJ, AIJ), L(-1, 2, IJ, AJ, IJ, AIJ), L(-1, 1, IJ, AIJ, IJ, AIJ),
                 L(1, 1, IJP1, AI, IJ, AJ), L(-1, 2, IJ, AJ, IJ, AJ), L(-1, 1, IJ, AIJ, IJ, AJ)}},
    {3, IJ, TJ, {L(1, 1, IJP1, AI, IJ, AIJ), L(1, 2, IJ, AIJ, IJ, AIJ), L(1, 1, IJ, AJ, IJ, AIJ),
                 L(1, 1, IJP1, AI, IJ, AJ), L(1, 2, IJ, AIJ, IJ, AJ), L(1, 1, IJ, AJ, IJ, AJ)}},
    kLapHex[11],
    {4, IM1J, TI, {L(1, 1, IM1J, AIJ, IJ, AJ), L(1, 2, IJ, AJ, IJ, AJ), L(-1, 1, IM1J, AI, IJ, AJ),
                   L(-1, 1, IM1J, AIJ, IM1J, AI), L(-1, 2, IJ, AJ, IM1J, AI), L(1, 1, IM1J, AI, IM1J, AI)}},
    kLapHex[13],
    {5, IM1JM1, TJ, {L(-1, 1, IM1JM1, AJ, IM1J, AI), L(1, 2, IM1J, AI, IM1J, AI), L(1, 1, IM1JM1, AIJ, IM1J, AI),
                     L(-1, 1, IM1JM1, AJ, IM1JM1, AIJ), L(1, 2, IM1J, AI, IM1JM1, AIJ),
                     L(1, 1, IM1JM1, AIJ, IM1JM1, AIJ)}},
    {6, IJM1, TJ, {L(1, 1, IJM1, AIJ, IM1JM1, AIJ), L(1, 2, IJ, AI, IM1JM1, AIJ), L(-1, 1, IJM1, AJ, IM1JM1, AIJ),
                   L(-1, 1, IJM1, AIJ, IJ, AI), L(-1, 2, IJ, AI, IJ, AI), L(1, 1, IJM1, AJ, IJ, AI)}},
};
#undef L

// The signed left-to-right sum of a coefficient's terms; the first term's
// sign is a negation, as in Fortran's "( - a*b + ...)".
double eval_vcoef(const VCoef &c, const Metrics &M, size_t g, size_t l, long iall, int comp)
{
    double s = 0.0;
    for (int k = 0; k < c.n; ++k) {
        const VTerm &t = c.t[k];
        const size_t ga = g + cell_offset(t.ac, iall);
        const double a = M.A(ga, l, t.arc, comp);
        const double v = t.tt == 0 ? 2.0 * a : M.T(g + cell_offset(t.tc, iall), l, t.tt, t.w) * a;
        if (k == 0) s = t.neg ? -v : v;
        else s = t.neg ? s - v : s + v;
    }
    return s;
}

double eval_lblock(const LBlock &b, const Metrics &M, size_t g, size_t l, long iall, int tn, int hn)
{
    double s = 0.0;
    for (int k = 0; k < 6; ++k) {
        const LTerm &t = b.t[k];
        const double a1 = M.A(g + cell_offset(t.c1, iall), l, t.a1, tn);
        const double a2 = M.A(g + cell_offset(t.c2, iall), l, t.a2, hn);
        const double v = (t.two ? 2.0 : 1.0) * a1 * a2;
        if (k == 0) s = t.neg ? -v : v;
        else s = t.neg ? s - v : s + v;
    }
    return s;
}

int check_setup_args(int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp, const void *p,
                     const void *t, const void *a, const void *out, const void *p_pl, const void *t_pl,
                     const void *a_pl, const void *out_pl)
{
    if (gall_1d < 4 || lall < 1 || gall_pl < 3 || gall_pl > 15) return fail(NICAM_ERR_ARG, "sizes out of range");
    if (!have_sgp || !p || !t || !a || !out) return fail(NICAM_ERR_ARG, "null argument");
    if (have_pl && (!p_pl || !t_pl || !a_pl || !out_pl)) return fail(NICAM_ERR_ARG, "null pole argument");
    return NICAM_OK;
}

enum class VKind { div, rot, grad };

// coef(gall,0:6,3,lall) and coef_pl(0:vlink,3,2)
int vector_setup(VKind kind, int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp,
                 const double *GMTR_p, const double *GMTR_p_pl, const double *GMTR_t, const double *GMTR_t_pl,
                 const double *GMTR_a, const double *GMTR_a_pl, double *coef, double *coef_pl)
{
    if (int e = check_setup_args(gall_1d, lall, gall_pl, have_pl, have_sgp, GMTR_p, GMTR_t, GMTR_a, coef, GMTR_p_pl,
                                 GMTR_t_pl, GMTR_a_pl, coef_pl))
        return e;
    const long iall = gall_1d;
    const size_t gall = (size_t)iall * iall, npl = gall_pl, vlink = npl - 1;
    const Metrics M{GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, gall, (size_t)lall, npl};
    const int base = kind == VKind::rot ? 4 : HNX;  // HTX = 4 (problem_size.inc:105)
    auto C = [&](size_t g, int m, int d, size_t l) -> double & {
        return coef[(((l - 1) * 3 + (d - 1)) * 7 + m) * gall + (g - 1)];
    };
    std::memset(coef, 0, sizeof(double) * gall * 21 * lall);
    if (have_pl) std::memset(coef_pl, 0, sizeof(double) * (vlink + 1) * 3 * 2);
    const size_t gmin = iall + 2, gmax = (iall - 2) * iall + (iall - 1);
    for (size_t l = 1; l <= (size_t)lall; ++l)
        for (int d = 1; d <= 3; ++d) {
            const int comp = d + base - 1;
            for (size_t g = gmin; g <= gmax; ++g) {
                const double rp = M.P(g, l, P_RAREA);
                for (int m = 0; m < 7; ++m) {
                    const VCoef &c = (kind == VKind::grad && m == 0) ? kGradHex0 : kDivHex[m];
                    C(g, m, d, l) = eval_vcoef(c, M, g, l, iall, comp) * 0.5 * rp;
                }
            }
            if (have_sgp[l - 1]) {
                const double rp = M.P(gmin, l, P_RAREA);
                for (int m = 0; m < 7; ++m) {
                    const VCoef &c = (kind == VKind::grad && m == 0) ? kGradPent0 : kDivPent[m];
                    C(gmin, m, d, l) = eval_vcoef(c, M, gmin, l, iall, comp) * 0.5 * rp;
                }
            }
        }
    if (have_pl) {
        const size_t n = 1, gmin_pl = 2, gmax_pl = npl;
        auto CP = [&](size_t v, int d, size_t l) -> double & { return coef_pl[((l - 1) * 3 + (d - 1)) * (vlink + 1) + v]; };
        for (size_t l = 1; l <= 2; ++l)
            for (int d = 1; d <= 3; ++d) {
                const int comp = d + base - 1;
                const double rp = M.PPL(n, l, P_RAREA);
                double s = 0.0;
                for (size_t v = gmin_pl; v <= gmax_pl; ++v) {
                    const size_t ij = v, ijp1 = v + 1 == gmax_pl + 1 ? gmin_pl : v + 1;
                    if (kind == VKind::grad)
                        s = s + 2.0 * (M.TPL(ij, l, W1) - 1.0) * M.APL(ijp1, l, comp);
                    else
                        s = s + (M.TPL(ij, l, W1) * M.APL(ij, l, comp) + M.TPL(ij, l, W1) * M.APL(ijp1, l, comp));
                }
                CP(0, d, l) = s * 0.5 * rp;
                for (size_t v = gmin_pl; v <= gmax_pl; ++v) {
                    const size_t ij = v, ijp1 = v + 1 == gmax_pl + 1 ? gmin_pl : v + 1;
                    const size_t ijm1 = v - 1 == gmin_pl - 1 ? gmax_pl : v - 1;
                    CP(v - 1, d, l) = (M.TPL(ijm1, l, W3) * M.APL(ijm1, l, comp) + M.TPL(ijm1, l, W3) * M.APL(ij, l, comp) +
                                       M.TPL(ij, l, W2) * M.APL(ij, l, comp) + M.TPL(ij, l, W2) * M.APL(ijp1, l, comp)) *
                                      0.5 * rp;
                }
            }
    }
    return NICAM_OK;
}

}  // namespace

extern "C" int nicam_oprt_divergence_setup(int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp,
                                           const double *GMTR_p, const double *GMTR_p_pl, const double *GMTR_t,
                                           const double *GMTR_t_pl, const double *GMTR_a, const double *GMTR_a_pl,
                                           double *coef_div, double *coef_div_pl)
{
    return vector_setup(VKind::div, gall_1d, lall, gall_pl, have_pl, have_sgp, GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl,
                        GMTR_a, GMTR_a_pl, coef_div, coef_div_pl);
}

extern "C" int nicam_oprt_rotation_setup(int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp,
                                         const double *GMTR_p, const double *GMTR_p_pl, const double *GMTR_t,
                                         const double *GMTR_t_pl, const double *GMTR_a, const double *GMTR_a_pl,
                                         double *coef_rot, double *coef_rot_pl)
{
    return vector_setup(VKind::rot, gall_1d, lall, gall_pl, have_pl, have_sgp, GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl,
                        GMTR_a, GMTR_a_pl, coef_rot, coef_rot_pl);
}

extern "C" int nicam_oprt_gradient_setup(int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp,
                                         const double *GMTR_p, const double *GMTR_p_pl, const double *GMTR_t,
                                         const double *GMTR_t_pl, const double *GMTR_a, const double *GMTR_a_pl,
                                         double *coef_grad, double *coef_grad_pl)
{
    return vector_setup(VKind::grad, gall_1d, lall, gall_pl, have_pl, have_sgp, GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl,
                        GMTR_a, GMTR_a_pl, coef_grad, coef_grad_pl);
}

// coef_lap(gall,0:6,lall), coef_lap_pl(0:vlink,2)
extern "C" int nicam_oprt_laplacian_setup(int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp,
                                          const double *GMTR_p, const double *GMTR_p_pl, const double *GMTR_t,
                                          const double *GMTR_t_pl, const double *GMTR_a, const double *GMTR_a_pl,
                                          double *coef_lap, double *coef_lap_pl)
{
    if (int e = check_setup_args(gall_1d, lall, gall_pl, have_pl, have_sgp, GMTR_p, GMTR_t, GMTR_a, coef_lap,
                                 GMTR_p_pl, GMTR_t_pl, GMTR_a_pl, coef_lap_pl))
        return e;
    const long iall = gall_1d;
    const size_t gall = (size_t)iall * iall, npl = gall_pl, vlink = npl - 1;
    const Metrics M{GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, gall, (size_t)lall, npl};
    auto C = [&](size_t g, int m, size_t l) -> double & { return coef_lap[((l - 1) * 7 + m) * gall + (g - 1)]; };
    std::memset(coef_lap, 0, sizeof(double) * gall * 7 * lall);
    const size_t gmin = iall + 2, gmax = (iall - 2) * iall + (iall - 1);
    for (size_t l = 1; l <= (size_t)lall; ++l) {
        for (int d = 1; d <= 3; ++d) {
            const int hn = d + HNX - 1, tn = d + TNX - 1;
            for (size_t g = gmin; g <= gmax; ++g)
                for (const LBlock &b : kLapHex)
                    C(g, b.m, l) = C(g, b.m, l) +
                                   M.T(g + cell_offset(b.tc, iall), l, b.tt, T_RAREA) * eval_lblock(b, M, g, l, iall, tn, hn);
        }
        if (have_sgp[l - 1]) {
            for (int m = 0; m < 7; ++m) C(gmin, m, l) = 0.0;
            for (int d = 1; d <= 3; ++d) {
                const int hn = d + HNX - 1, tn = d + TNX - 1;
                for (const LBlock &b : kLapPent)
                    C(gmin, b.m, l) = C(gmin, b.m, l) + M.T(gmin + cell_offset(b.tc, iall), l, b.tt, T_RAREA) *
                                                            eval_lblock(b, M, gmin, l, iall, tn, hn);
            }
        }
        for (size_t g = 1; g <= gall; ++g)
            for (int m = 0; m < 7; ++m) C(g, m, l) = C(g, m, l) * M.P(g, l, P_RAREA) / 12.0;
    }
    if (have_pl) {
        std::memset(coef_lap_pl, 0, sizeof(double) * (vlink + 1) * 2);
        const size_t n = 1, gmin_pl = 2, gmax_pl = npl;
        for (size_t l = 1; l <= 2; ++l) {
            double *cp = coef_lap_pl + (l - 1) * (vlink + 1);
            for (int d = 1; d <= 3; ++d) {
                const int hn = d + HNX - 1, tn = d + TNX - 1, tn2 = d + TN2X - 1;
                auto A = [&](size_t v, int c) { return M.APL(v, l, c); };
                for (size_t v = gmin_pl; v <= gmax_pl; ++v) {
                    const size_t ij = v, ijp1 = v + 1 == gmax_pl + 1 ? gmin_pl : v + 1;
                    const size_t ijm1 = v - 1 == gmin_pl - 1 ? gmax_pl : v - 1;
                    cp[0] = cp[0] + M.TPL(ijm1, l, T_RAREA) * (1.0 * A(ijm1, tn) * A(ij, hn) - 2.0 * A(ijm1, tn2) * A(ij, hn) -
                                                               1.0 * A(ij, tn) * A(ij, hn));
                    cp[0] = cp[0] + M.TPL(ij, l, T_RAREA) * (1.0 * A(ij, tn) * A(ij, hn) - 2.0 * A(ij, tn2) * A(ij, hn) -
                                                             1.0 * A(ijp1, tn) * A(ij, hn));
                }
                for (size_t v = gmin_pl; v <= gmax_pl; ++v) {
                    const size_t ij = v, ijp1 = v + 1 == gmax_pl + 1 ? gmin_pl : v + 1;
                    const size_t ijm1 = v - 1 == gmin_pl - 1 ? gmax_pl : v - 1;
                    cp[v - 1] = cp[v - 1] + M.TPL(ijm1, l, T_RAREA) *
                                                (-(2.0 * A(ijm1, tn) * A(ijm1, hn)) + 1.0 * A(ijm1, tn2) * A(ijm1, hn) -
                                                 1.0 * A(ij, tn) * A(ijm1, hn) - 2.0 * A(ijm1, tn) * A(ij, hn) +
                                                 1.0 * A(ijm1, tn2) * A(ij, hn) - 1.0 * A(ij, tn) * A(ij, hn));
                    cp[v - 1] = cp[v - 1] + M.TPL(ij, l, T_RAREA) *
                                                (1.0 * A(ij, tn) * A(ij, hn) + 1.0 * A(ij, tn2) * A(ij, hn) +
                                                 2.0 * A(ijp1, tn) * A(ij, hn) + 1.0 * A(ij, tn) * A(ijp1, hn) +
                                                 1.0 * A(ij, tn2) * A(ijp1, hn) + 2.0 * A(ijp1, tn) * A(ijp1, hn));
                }
            }
            for (size_t v = n; v <= gmax_pl; ++v) cp[v - 1] = cp[v - 1] * M.PPL(n, l, P_RAREA) / 12.0;
        }
    }
    return NICAM_OK;
}
