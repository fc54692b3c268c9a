// Adjoints of the horizontal step glue (glue.hip) and of the halo fill
// (comm.hip) on gfx950, for the differentiable tracer step.
//
// The reference has no glue (its kernels take ch, cmask and q_a as inputs),
// so these have no reference counterpart; each is the transpose of its
// forward's derivative, written as a gather so that every output is summed by
// one work-item in a fixed order (the oracle's, nicam_oracle.c
// orc_*_adjoint), with no atomics:
//   gradient  scl_b(g) = sum over d, m of coef_grad(g', m, d) grad_b(g', d),
//             g' = g - offset(m) on the flat range [gmin, gmax];
//   courant   flx_b = ch_b / rhog, rhog_b = -(sum of ch_b (flx/rhog)) / rhog;
//   remap     each arc's two face cotangents give its two nodes' (the upwind
//             selection with cmask fixed); a cell gathers them from the arcs
//             of origin c-iall-1, c-iall, c-1 (as their neighbour) and c (as
//             their origin), in that order, the order of the forward's sweep;
//   update    flx_b = -(rhogq_b q_a), q_a_b = -(rhogq_b flx) on the inner
//             cells; rhogq_b passes through;
//   halo      each source adds its destinations' cotangents in ascending
//             order of their codes and clears them.
#include <algorithm>
#include <vector>

#include "internal.h"

namespace nicam {
namespace {

constexpr int BLK = 256;

struct GradAdjArgs {
    double *sb, *sb_pl;          // scl_b (gall,kall,lall), (gall_pl,kall,2)
    const double *gb, *gb_pl;    // grad_b (gall,kall,lall,3), (gall_pl,kall,2,3)
    const double *cg, *cg_pl;    // coef_grad(gall,0:6,3,lall), coef_grad_pl(0:gall_pl-1,3,2)
    int iall, gall, kall, lall, gall_pl, have_pl;
    long nreg, npole;
};

__global__ void __launch_bounds__(BLK) gradient_adjoint_kernel(GradAdjArgs A)
{
    // consecutive work-groups on one XCD: the rows above and below and the region's coefficients stay in its L2
    const long t = xcd_contiguous(blockIdx.x, gridDim.x) * BLK + threadIdx.x;
    const int iall = A.iall, gall = A.gall, kall = A.kall;
    if (t < A.nreg) {
        const int g = (int)(t % gall);
        const long lk = t / gall;
        const int l = (int)(lk / kall);
        const int gmin = iall + 1, gmax = gall - iall - 2;  // 0-based flat range of the forward
        const int off[7] = {0, 1, iall + 1, iall, -1, -iall - 1, -iall};
        // every term is loaded (addresses clamped into the range) and the ones outside it add +0.0: the
        // sum starts at +0.0 and so never holds -0.0, and adding +0.0 leaves it as it is
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int m = 0; m < 7; ++m) {
                const int gp = g - off[m];
                const bool ok = gp >= gmin && gp <= gmax;
                const int gc = gp < gmin ? gmin : (gp > gmax ? gmax : gp);
                const double term = A.cg[(((size_t)l * 3 + d) * 7 + m) * gall + gc] *
                                    A.gb[(size_t)d * A.nreg + (size_t)lk * gall + gc];
                s = s + (ok ? term : 0.0);
            }
        A.sb[t] = s;
        return;
    }
    const long p = t - A.nreg;  // one work-item per pole point and level
    if (p >= A.npole) return;
    const int npl = A.gall_pl;
    const int v = (int)(p % npl);
    const long lk = p / npl;
    const int l = (int)(lk / kall);
    double s = 0.0;
    if (A.have_pl)
        for (int d = 0; d < 3; ++d) s = s + A.cg_pl[((size_t)l * 3 + d) * npl + v] * A.gb_pl[(size_t)d * A.npole + lk * npl];
    A.sb_pl[p] = s;
}

struct CourantAdjArgs {
    double *fb, *fb_pl, *rb, *rb_pl;  // flx_b (gall,kall,lall,6), rhog_b
    const double *chb, *chb_pl, *flx, *flx_pl, *rhog, *rhog_pl;
    long nreg;
    int gall_pl, kall, have_pl;
};

__global__ void __launch_bounds__(BLK) courant_adjoint_kernel(CourantAdjArgs A)
{
    const long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t < A.nreg) {
        const double r = A.rhog[t];
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            const double b = A.chb[m * A.nreg + t];
            A.fb[m * A.nreg + t] = b / r;
            s = s + b * (A.flx[m * A.nreg + t] / r);
        }
        A.rb[t] = -s / r;
        return;
    }
    const long p = t - A.nreg;  // one work-item per pole level
    if (p >= (long)A.kall * kLallPl) return;
    const int npl = A.gall_pl;
    const size_t fb = (size_t)p * npl;
    for (int v = 0; v < npl; ++v) {
        A.fb_pl[fb + v] = 0.0;
        A.rb_pl[fb + v] = 0.0;
    }
    if (!A.have_pl) return;
    const double r = A.rhog_pl[fb];
    double s = 0.0;
    for (int v = 1; v < npl; ++v) {
        const double b = A.chb_pl[fb + v];
        A.fb_pl[fb + v] = b / r;
        s = s + b * (A.flx_pl[fb + v] / r);
    }
    A.rb_pl[fb] = -s / r;
}

struct RemapAdjArgs {
    double *qb, *qb_pl, *gb, *gb_pl, *xb, *xb_pl;  // q_b, grad_b (.., 3), GRD_xc_b (.., 3, 3) / (.., 3)
    const double *qab, *qab_pl, *grad, *grad_pl, *cm, *cm_pl, *xc, *xc_pl, *x, *x_pl;
    int iall, gall, kall, lall, gall_pl, have_pl;
    long nreg;
};

__global__ void __launch_bounds__(BLK) remap_adjoint_kernel(RemapAdjArgs A)
{
    // consecutive work-groups on one XCD: the row below, read by the neighbour arcs, stays in its L2
    const long t = xcd_contiguous(blockIdx.x, gridDim.x) * BLK + threadIdx.x;
    const int iall = A.iall, gall = A.gall, kall = A.kall, lall = A.lall;
    const long n = A.nreg;
    if (t < n) {
        const int g = (int)(t % gall);
        const long lk = t / gall;
        const int l = (int)(lk / kall);
        const int i = g % iall, j = g / iall;
        const size_t base = (size_t)lk * gall;  // + g: this level's plane entry
        const size_t xl = (size_t)l * gall;      // GRD_x(g, l, d) = x[d][l][g]
        // node cotangents of arc a (0 = AI, 1 = AIJ, 2 = AJ) with origin o
        auto nodes = [&](int o, int a, double &qs_b, double &qn_b) {
            const int nb = o + (a == 0 ? 1 : (a == 1 ? iall + 1 : iall));
            const double fa = A.qab[a * n + base + o], fb = A.qab[(a + 3) * n + base + nb];
            const double cs = A.cm[a * n + base + o], cn = A.cm[(a + 3) * n + base + nb];
            qs_b = (1.0 - cs) * fa + cn * fb;
            qn_b = cs * fa + (1.0 - cn) * fb;
        };
        double qb = 0.0, gb[3] = {0.0, 0.0, 0.0};
        // c as the neighbour of the arcs of origin c-iall-1 (AIJ), c-iall (AJ), c-1 (AI)
        const int no[3] = {g - iall - 1, g - iall, g - 1};
        const int na[3] = {1, 2, 0};
        const bool has[3] = {i >= 1 && j >= 1, j >= 1 && i <= iall - 2, i >= 1 && j <= iall - 2};
#pragma unroll
        for (int h = 0; h < 3; ++h) {
            if (!has[h]) continue;
            double qs_b, qn_b;
            nodes(no[h], na[h], qs_b, qn_b);
            qb = qb + qn_b;
#pragma unroll
            for (int d = 0; d < 3; ++d)
                gb[d] = gb[d] + qn_b * (A.xc[(d * 3 + na[h]) * n + base + no[h]] - A.x[(size_t)d * lall * gall + xl + g]);
        }
        // c as the origin of its three arcs
        double xb[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        if (i <= iall - 2 && j <= iall - 2) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int nb = g + (a == 0 ? 1 : (a == 1 ? iall + 1 : iall));
                double qs_b, qn_b;
                nodes(g, a, qs_b, qn_b);
                qb = qb + qs_b;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    gb[d] = gb[d] + qs_b * (A.xc[(d * 3 + a) * n + base + g] - A.x[(size_t)d * lall * gall + xl + g]);
                    xb[a][d] = qs_b * A.grad[d * n + base + g] + qn_b * A.grad[d * n + base + nb];
                }
            }
        }
        stream_store(&A.qb[t], qb);
#pragma unroll
        for (int d = 0; d < 3; ++d) stream_store(&A.gb[d * n + t], gb[d]);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) stream_store(&A.xb[(d * 3 + a) * n + t], xb[a][d]);
        return;
    }
    const long p = t - n;  // one work-item per pole level, in the oracle's order
    if (p >= (long)kall * kLallPl) return;
    const int npl = A.gall_pl, l = (int)(p / kall);
    const size_t fb = (size_t)p * npl, pp = (size_t)kLallPl * kall * npl;
    for (int v = 0; v < npl; ++v) {
        A.qb_pl[fb + v] = 0.0;
        for (int d = 0; d < 3; ++d) {
            A.gb_pl[d * pp + fb + v] = 0.0;
            A.xb_pl[d * pp + fb + v] = 0.0;
        }
    }
    if (!A.have_pl) return;
    auto X = [&](int v, int d) { return A.x_pl[((size_t)d * kLallPl + l) * npl + v]; };
    for (int v = 1; v < npl; ++v) {
        const double f = A.qab_pl[fb + v], c = A.cm_pl[fb + v];
        const double qs_b = (1.0 - c) * f, qn_b = c * f;
        A.qb_pl[fb] = A.qb_pl[fb] + qs_b;
        A.qb_pl[fb + v] = A.qb_pl[fb + v] + qn_b;
        for (int d = 0; d < 3; ++d) {
            const double x = A.xc_pl[d * pp + fb + v];
            A.gb_pl[d * pp + fb] = A.gb_pl[d * pp + fb] + qs_b * (x - X(0, d));
            A.gb_pl[d * pp + fb + v] = A.gb_pl[d * pp + fb + v] + qn_b * (x - X(v, d));
            A.xb_pl[d * pp + fb + v] = qs_b * A.grad_pl[d * pp + fb] + qn_b * A.grad_pl[d * pp + fb + v];
        }
    }
}

struct UpdateAdjArgs {
    double *fb, *fb_pl, *qab, *qab_pl;
    const double *rqb, *rqb_pl, *flx, *flx_pl, *qa, *qa_pl;
    int iall, gall, kall, gall_pl, have_pl;
    long nreg;
};

__global__ void __launch_bounds__(BLK) update_adjoint_kernel(UpdateAdjArgs A)
{
    const long t = (long)blockIdx.x * BLK + threadIdx.x;
    const long n = A.nreg;
    if (t < n) {
        const int g = (int)(t % A.gall), iall = A.iall;
        const int i = g % iall, j = g / iall;
        const bool inner = i >= 1 && i <= iall - 2 && j >= 1 && j <= iall - 2;
        const double b = A.rqb[t];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            A.fb[m * n + t] = inner ? -(b * A.qa[m * n + t]) : 0.0;
            A.qab[m * n + t] = inner ? -(b * A.flx[m * n + t]) : 0.0;
        }
        return;
    }
    const long p = t - n;  // one work-item per pole point and level
    const int npl = A.gall_pl;
    if (p >= (long)npl * A.kall * kLallPl) return;
    const int v = (int)(p % npl);
    const long cen = p - v;
    const double b = A.have_pl ? A.rqb_pl[cen] : 0.0;
    const bool ring = A.have_pl && v >= 1;
    A.fb_pl[p] = ring ? -(b * A.qa_pl[p]) : 0.0;
    A.qab_pl[p] = ring ? -(b * A.flx_pl[p]) : 0.0;
}

struct HaloAdjArgs {
    double *var, *var_pl;
    const int *src, *off, *dst;  // sources, and each one's destinations (CSR, ascending codes)
    int nsrc, gall, kmax, lall, gall_pl, nvar;
};

__device__ __forceinline__ double *halo_locate(const HaloAdjArgs &A, int code, int k, int v)
{
    const int nreg = A.gall * A.lall;
    if (code < nreg) {
        const int l = code / A.gall, g = code - l * A.gall;
        return A.var + (((size_t)v * A.lall + l) * A.kmax + k) * A.gall + g;
    }
    const int p = code - nreg;
    const int l = p / A.gall_pl, g = p - l * A.gall_pl;
    return A.var_pl + (((size_t)v * kLallPl + l) * A.kmax + k) * A.gall_pl + g;
}

// One work-item per (source, level): it owns its destinations (each has one
// source), so it reads and clears them itself.
__global__ void __launch_bounds__(BLK) halo_adjoint_kernel(HaloAdjArgs A)
{
    const long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)A.nsrc * A.kmax) return;
    const int si = (int)(t % A.nsrc), k = (int)(t / A.nsrc);
    const int s = A.src[si], e0 = A.off[si], e1 = A.off[si + 1];
    for (int v = 0; v < A.nvar; ++v) {
        double *ps = halo_locate(A, s, k, v);
        double acc = *ps;
        for (int e = e0; e < e1; ++e) {
            double *pd = halo_locate(A, A.dst[e], k, v);
            acc = acc + *pd;
            *pd = 0.0;
        }
        *ps = acc;
    }
}

int grid1(long n, long &nb)
{
    nb = (n + BLK - 1) / BLK;
    if (nb > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    return NICAM_OK;
}

}  // namespace

int launch_gradient_adjoint(nicam_ctx c, double *scl_b, double *scl_b_pl, const double *grad_b,
                            const double *grad_b_pl, const double *coef_grad, const double *coef_grad_pl)
{
    const Dims &D = c->d;
    GradAdjArgs A{scl_b, scl_b_pl, grad_b, grad_b_pl, coef_grad, coef_grad_pl, D.gall_1d, D.gall, D.kall, D.lall,
                  D.gall_pl, D.have_pl, (long)D.gall * D.kall * D.lall,
                  scl_b_pl ? (long)D.gall_pl * D.kall * kLallPl : 0};
    long nb;
    if (grid1(A.nreg + A.npole, nb) != NICAM_OK) return NICAM_ERR_ARG;
    hipLaunchKernelGGL(gradient_adjoint_kernel, dim3((unsigned)nb), dim3(BLK), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

int launch_courant_adjoint(nicam_ctx c, double *flx_b, double *flx_b_pl, double *rhog_b, double *rhog_b_pl,
                           const double *ch_b, const double *ch_b_pl, const double *flx_h, const double *flx_h_pl,
                           const double *rhog, const double *rhog_pl)
{
    const Dims &D = c->d;
    CourantAdjArgs A{flx_b, flx_b_pl, rhog_b, rhog_b_pl, ch_b, ch_b_pl, flx_h, flx_h_pl, rhog, rhog_pl,
                     (long)D.gall * D.kall * D.lall, D.gall_pl, D.kall, D.have_pl};
    long nb;
    if (grid1(A.nreg + (flx_b_pl ? (long)D.kall * kLallPl : 0), nb) != NICAM_OK) return NICAM_ERR_ARG;
    hipLaunchKernelGGL(courant_adjoint_kernel, dim3((unsigned)nb), dim3(BLK), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

int launch_remap_adjoint(nicam_ctx c, double *q_b, double *q_b_pl, double *grad_b, double *grad_b_pl, double *xc_b,
                         double *xc_b_pl, const double *q_a_b, const double *q_a_b_pl, const double *grad,
                         const double *grad_pl, const double *cmask, const double *cmask_pl, const double *xc,
                         const double *xc_pl, const double *x, const double *x_pl)
{
    const Dims &D = c->d;
    RemapAdjArgs A{q_b, q_b_pl, grad_b, grad_b_pl, xc_b, xc_b_pl, q_a_b, q_a_b_pl, grad, grad_pl, cmask, cmask_pl,
                   xc, xc_pl, x, x_pl, D.gall_1d, D.gall, D.kall, D.lall, D.gall_pl, D.have_pl,
                   (long)D.gall * D.kall * D.lall};
    long nb;
    if (grid1(A.nreg + (q_b_pl ? (long)D.kall * kLallPl : 0), nb) != NICAM_OK) return NICAM_ERR_ARG;
    hipLaunchKernelGGL(remap_adjoint_kernel, dim3((unsigned)nb), dim3(BLK), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

int launch_update_adjoint(nicam_ctx c, double *flx_b, double *flx_b_pl, double *q_a_b, double *q_a_b_pl,
                          const double *rhogq_b, const double *rhogq_b_pl, const double *flx_h,
                          const double *flx_h_pl, const double *q_a, const double *q_a_pl)
{
    const Dims &D = c->d;
    UpdateAdjArgs A{flx_b, flx_b_pl, q_a_b, q_a_b_pl, rhogq_b, rhogq_b_pl, flx_h, flx_h_pl, q_a, q_a_pl,
                    D.gall_1d, D.gall, D.kall, D.gall_pl, D.have_pl, (long)D.gall * D.kall * D.lall};
    long nb;
    if (grid1(A.nreg + (flx_b_pl ? (long)D.gall_pl * D.kall * kLallPl : 0), nb) != NICAM_OK) return NICAM_ERR_ARG;
    hipLaunchKernelGGL(update_adjoint_kernel, dim3((unsigned)nb), dim3(BLK), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

int halo_adjoint_setup(nicam_ctx c, const std::vector<int> &dst, const std::vector<int> &src)
{
    // sources in ascending order, each with its destinations in ascending order
    std::vector<std::pair<int, int>> e(dst.size());
    for (size_t i = 0; i < dst.size(); ++i) e[i] = {src[i], dst[i]};
    std::sort(e.begin(), e.end());
    std::vector<int> s, off, d;
    for (size_t i = 0; i < e.size(); ++i) {
        if (i == 0 || e[i].first != e[i - 1].first) {
            s.push_back(e[i].first);
            off.push_back((int)i);
        }
        d.push_back(e[i].second);
    }
    off.push_back((int)e.size());
    int **arr[] = {&c->hadj_src, &c->hadj_off, &c->hadj_dst};
    const std::vector<int> *vec[] = {&s, &off, &d};
    for (int a = 0; a < 3; ++a) {
        if (*arr[a]) (void)hipFree(*arr[a]);
        *arr[a] = nullptr;
        NICAM_HIP(hipMalloc(arr[a], sizeof(int) * std::max<size_t>(1, vec[a]->size())));
        NICAM_HIP(hipMemcpy(*arr[a], vec[a]->data(), sizeof(int) * vec[a]->size(), hipMemcpyHostToDevice));
    }
    c->hadj_nsrc = (int)s.size();
    return NICAM_OK;
}

int launch_halo_adjoint(nicam_ctx c, double *var, double *var_pl, int kmax, int nvar)
{
    if (!c->hadj_nsrc) return fail(NICAM_ERR_STATE, "halo map not set (nicam_comm_setup)");
    if (kmax < 1 || nvar < 1) return fail(NICAM_ERR_ARG, "kmax and vmax must be positive");
    const Dims &D = c->d;
    HaloAdjArgs A{var, var_pl, c->hadj_src, c->hadj_off, c->hadj_dst, c->hadj_nsrc, D.gall, kmax, D.lall,
                  D.gall_pl, nvar};
    long nb;
    if (grid1((long)A.nsrc * kmax, nb) != NICAM_OK) return NICAM_ERR_ARG;
    hipLaunchKernelGGL(halo_adjoint_kernel, dim3((unsigned)nb), dim3(BLK), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

}  // namespace nicam
