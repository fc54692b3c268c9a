// The step glue around the reference's kernels, on gfx950.
//
// The reference's kernels take the Courant numbers, upwind masks and face
// values as inputs, computed "in advance by other subroutine"; the reference
// has none of that code.  These kernels are the NICAM forms of the Miura
// upwind scheme (docs/NICAM/src/20_intro_nicam.tex:368-438):
//   gradient  grad = the 7-point stencil of coef_grad (OPRT_gradient_setup's
//             coefficients, dyn_metrics/src/mod_oprt.f90:499-734) on q;
//   courant   ch = flx_h / rhog, cmask = 0.5 - sign(0.5, ch - EPS);
//   remap     the face value of an arc is the upwind node's q plus its
//             gradient dotted with (GRD_xc - GRD_x(node)), GRD_xc being the
//             mass centroid of the swept area that horizontal_flux returns;
//   update    rhogq -= sum over the six faces of flx_h * q_a.
// Each sum is evaluated left to right in the order the oracle states.
#include "internal.h"

namespace nicam {
namespace {

constexpr int BLK = 256;

struct GradArgs {
    double *grad, *grad_pl;  // (gall,kall,lall,3), (gall_pl,kall,2,3)
    const double *scl, *scl_pl;
    const double *cg, *cg_pl;  // coef_grad(gall,0:6,3,lall), coef_grad_pl(0:gall_pl-1,3,2)
    int iall, gall, kall, lall, gall_pl, have_pl;
    int nchunk, nkr, kper, nmain;
};

__device__ void gradient_pole(const GradArgs &A, int t)
{
    const int kall = A.kall, npl = A.gall_pl;
    if (t >= kall * kLallPl) return;
    const int k = t % kall, l = t / kall;
    const size_t fb = ((size_t)l * kall + k) * npl, plane = (size_t)kLallPl * kall * npl;
    for (int d = 0; d < 3; ++d) {
        double s = 0.0;
        if (A.have_pl)
            for (int v = 0; v < npl; ++v) s = s + A.cg_pl[((size_t)l * 3 + d) * npl + v] * A.scl_pl[fb + v];
        A.grad_pl[d * plane + fb] = s;
        for (int v = 1; v < npl; ++v) A.grad_pl[d * plane + fb + v] = 0.0;
    }
}

// A work-group owns BLK consecutive cells of a region and a range of levels;
// each work-item keeps its cell's 21 coefficients in registers.
__global__ void __launch_bounds__(BLK) gradient_kernel(GradArgs A)
{
    int b = blockIdx.x;
    if (b >= A.nmain) {
        gradient_pole(A, (b - A.nmain) * BLK + threadIdx.x);
        return;
    }
    b = (int)xcd_contiguous(b, A.nmain);  // a region's chunks share halo rows: one XCD
    const int iall = A.iall, gall = A.gall, kall = A.kall;
    const int chunk = b % A.nchunk, kr = (b / A.nchunk) % A.nkr, l = b / (A.nchunk * A.nkr);
    const int k0 = kr * A.kper, k1 = min(kall, k0 + A.kper);
    if (l >= A.lall || k0 >= k1) return;
    const int g = chunk * BLK + threadIdx.x;
    if (g >= gall) return;
    const int gmin = iall + 1, gmax = gall - iall - 2;  // 0-based flat range of the reference's loops
    const bool in = g >= gmin && g <= gmax;
    const size_t plane = (size_t)A.lall * kall * gall;
    double c[3][7];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int m = 0; m < 7; ++m) c[d][m] = in ? A.cg[(((size_t)l * 3 + d) * 7 + m) * gall + g] : 0.0;
    const int off[7] = {0, 1, iall + 1, iall, -1, -iall - 1, -iall};  // ij ip1j ip1jp1 ijp1 im1j im1jm1 ijm1
    for (int k = k0; k < k1; ++k) {
        const size_t fb = ((size_t)l * kall + k) * gall;
        if (!in) {
#pragma unroll
            for (int d = 0; d < 3; ++d) stream_store(&A.grad[d * plane + fb + g], 0.0);
            continue;
        }
        double s[7];
#pragma unroll
        for (int m = 0; m < 7; ++m) s[m] = A.scl[fb + g + off[m]];
#pragma unroll
        for (int d = 0; d < 3; ++d)
            stream_store(&A.grad[d * plane + fb + g], c[d][0] * s[0] + c[d][1] * s[1] + c[d][2] * s[2] +
                                                          c[d][3] * s[3] + c[d][4] * s[4] + c[d][5] * s[5] +
                                                          c[d][6] * s[6]);
    }
}

struct CourantArgs {
    double *ch, *ch_pl, *cmask, *cmask_pl;
    const double *flx, *flx_pl, *rhog, *rhog_pl;
    long nreg, npole;  // cell-levels of the regular and pole regions
    int gall_pl, have_pl;
    double eps;
};

__global__ void __launch_bounds__(BLK) courant_kernel(CourantArgs A)
{
    const long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t < A.nreg) {
        const double r = A.rhog[t];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            const double c = A.flx[m * A.nreg + t] / r;
            stream_store(&A.ch[m * A.nreg + t], c);
            stream_store(&A.cmask[m * A.nreg + t], 0.5 - copysign(0.5, c - A.eps));
        }
        return;
    }
    const long p = t - A.nreg;  // one work-item per pole point and level
    if (p >= A.npole) return;
    const int npl = A.gall_pl;
    const long v = p % npl, cen = p - v;
    double c = 0.0, m = 0.0;
    if (v > 0 && A.have_pl) {
        c = A.flx_pl[p] / A.rhog_pl[cen];
        m = 0.5 - copysign(0.5, c - A.eps);
    }
    A.ch_pl[p] = c;
    A.cmask_pl[p] = m;
}

constexpr int RCH = 256;   // output cells per remap work-group
constexpr int RMAXI = 160; // largest gall_1d the remap tiles hold
constexpr int RNO = 2;     // arc origins per work-item: RCH*RNO >= RCH + RMAXI + 1

struct RemapArgs {
    double *q_a, *q_a_pl;
    const double *q, *q_pl, *grad, *grad_pl, *cmask, *cmask_pl, *xc, *xc_pl, *x, *x_pl;
    int iall, gall, kall, lall, gall_pl, have_pl;
    int nchunk, nkr, kper, nmain;  // chunks per slab, level ranges, levels per range, regular work-groups
};

// Pole regions: one work-item per (k, pole region); each ring face takes the
// upwind one of the centre's and the ring point's reconstructions.
__device__ void remap_pole(const RemapArgs &A, long p)
{
    const int kall = A.kall;
    if (!A.have_pl || p >= (long)kall * kLallPl) return;
    const int npl = A.gall_pl, k = (int)(p % kall), l = (int)(p / kall);
    const size_t fb = ((size_t)l * kall + k) * npl, plane = (size_t)kLallPl * kall * npl;
    const size_t xb = (size_t)l * npl, xplane = (size_t)kLallPl * npl;
    for (int v = 1; v < npl; ++v) {
        const double x0 = A.xc_pl[0 * plane + fb + v], x1 = A.xc_pl[1 * plane + fb + v],
                     x2 = A.xc_pl[2 * plane + fb + v];
        const double qs = A.q_pl[fb] + (A.grad_pl[0 * plane + fb] * (x0 - A.x_pl[0 * xplane + xb]) +
                                        A.grad_pl[1 * plane + fb] * (x1 - A.x_pl[1 * xplane + xb]) +
                                        A.grad_pl[2 * plane + fb] * (x2 - A.x_pl[2 * xplane + xb]));
        const double qn = A.q_pl[fb + v] + (A.grad_pl[0 * plane + fb + v] * (x0 - A.x_pl[0 * xplane + xb + v]) +
                                            A.grad_pl[1 * plane + fb + v] * (x1 - A.x_pl[1 * xplane + xb + v]) +
                                            A.grad_pl[2 * plane + fb + v] * (x2 - A.x_pl[2 * xplane + xb + v]));
        const double c = A.cmask_pl[fb + v];
        A.q_a_pl[fb + v] = c * qn + (1.0 - c) * qs;
    }
}

// A work-group owns RCH consecutive cells of one region and marches up a
// range of levels.  Positions of the cells [g0-iall-1, g0+RCH+iall+1) sit in
// LDS for the whole range; at each level q and its gradient over the same
// window are staged, then (phase 1) every arc with origin in
// [g0-iall-1, g0+RCH) forms its two reconstructions once, qs from the origin
// and qn from the far end, into LDS, and (phase 2) every cell picks each of
// its six faces by that face's own mask: faces 1-3 from its own arcs, faces
// 4-6 from the arcs of im1j (AI), im1jm1 (AIJ) and ijm1 (AJ).  A face no arc
// reaches keeps its value.
__global__ void __launch_bounds__(RCH) remap_kernel(RemapArgs A)
{
    extern __shared__ double lds[];
    int b = blockIdx.x;
    if (b >= A.nmain) {
        remap_pole(A, (long)(b - A.nmain) * RCH + threadIdx.x);
        return;
    }
    b = (int)xcd_contiguous(b, A.nmain);  // a region's chunks share halo rows: one XCD
    const int iall = A.iall, gall = A.gall, kall = A.kall, lall = A.lall;
    const int chunk = b % A.nchunk, kr = (b / A.nchunk) % A.nkr, l = b / (A.nchunk * A.nkr);
    const int k0 = kr * A.kper, k1 = min(kall, k0 + A.kper);
    if (l >= lall || k0 >= k1) return;
    const int g0 = chunk * RCH, lo = g0 - iall - 1;
    const int n_in = RCH + 2 * iall + 2;  // staged cells [lo, lo + n_in)
    const int n_o = RCH + iall + 1;       // arc origins [lo, lo + n_o)
    double *s_q = lds;                    // [n_in]
    double *s_g = lds + n_in;             // [3][n_in]
    double *s_x = lds + 4 * n_in;         // [3][n_in]
    double *s_a = lds + 7 * n_in;         // [6][n_o]: qs of AI, AIJ, AJ, then qn of AI, AIJ, AJ
    const int tid = threadIdx.x;
    const size_t mb = (size_t)l * gall, xplane = (size_t)lall * gall;
    const size_t plane = (size_t)lall * kall * gall;

    for (int s = tid; s < n_in; s += RCH) {  // GRD_x(g,1,l,d), K-invariant
        const int p = lo + s;
        const bool ok = p >= 0 && p < gall;
#pragma unroll
        for (int d = 0; d < 3; ++d) s_x[d * n_in + s] = ok ? A.x[d * xplane + mb + p] : 0.0;
    }
    int org[RNO];  // my arc origins (cell index), -1 if none
#pragma unroll
    for (int r = 0; r < RNO; ++r) {
        const int s = tid + r * RCH, p = lo + s;
        org[r] = (s < n_o && p >= 0 && p < gall && p % iall <= iall - 2 && p / iall <= iall - 2) ? p : -1;
    }
    const int g = g0 + tid;
    const bool active = g < gall;
    const int i = active ? g % iall : 0, j = active ? g / iall : 0;
    const int last = iall - 2;
    const bool own = active && i <= last && j <= last;
    const bool f4 = active && i >= 1 && j <= last, f5 = active && i >= 1 && j >= 1,
               f6 = active && i <= last && j >= 1;

    for (int k = k0; k < k1; ++k) {
        const size_t fb = ((size_t)l * kall + k) * gall;
        // loads that do not wait for the tile: my origins' arc centroids and my faces' masks
        double xcv[RNO][9];
#pragma unroll
        for (int r = 0; r < RNO; ++r)
#pragma unroll
            for (int m = 0; m < 9; ++m) xcv[r][m] = org[r] >= 0 ? A.xc[(size_t)m * plane + fb + org[r]] : 0.0;
        double cm[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) cm[f] = active ? A.cmask[(size_t)f * plane + fb + g] : 0.0;
        for (int s = tid; s < n_in; s += RCH) {
            const int p = lo + s;
            const bool ok = p >= 0 && p < gall;
            s_q[s] = ok ? A.q[fb + p] : 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) s_g[d * n_in + s] = ok ? A.grad[d * plane + fb + p] : 0.0;
        }
        __syncthreads();
        // phase 1: the two reconstructions of each arc (GRD_xc(g,k,l,a,d) at (d*3+a)*plane)
#pragma unroll
        for (int r = 0; r < RNO; ++r) {
            if (org[r] < 0) continue;
            const int s = tid + r * RCH;
            const int far[3] = {s + 1, s + iall + 1, s + iall};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double x0 = xcv[r][0 * 3 + a], x1 = xcv[r][1 * 3 + a], x2 = xcv[r][2 * 3 + a];
                const int n = far[a];
                const double qs = s_q[s] + (s_g[0 * n_in + s] * (x0 - s_x[0 * n_in + s]) +
                                            s_g[1 * n_in + s] * (x1 - s_x[1 * n_in + s]) +
                                            s_g[2 * n_in + s] * (x2 - s_x[2 * n_in + s]));
                const double qn = s_q[n] + (s_g[0 * n_in + n] * (x0 - s_x[0 * n_in + n]) +
                                            s_g[1 * n_in + n] * (x1 - s_x[1 * n_in + n]) +
                                            s_g[2 * n_in + n] * (x2 - s_x[2 * n_in + n]));
                s_a[a * n_o + s] = qs;
                s_a[(3 + a) * n_o + s] = qn;
            }
        }
        __syncthreads();
        // phase 2: my faces, each by its own mask
        if (active) {
            const int s = g - lo;
            if (own) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double c = cm[a];
                    stream_store(&A.q_a[(size_t)a * plane + fb + g],
                                 c * s_a[(3 + a) * n_o + s] + (1.0 - c) * s_a[a * n_o + s]);
                }
            }
            const int o4 = s - 1, o5 = s - iall - 1, o6 = s - iall;  // origins of the arcs of faces 4-6
            if (f4) stream_store(&A.q_a[3 * plane + fb + g], cm[3] * s_a[0 * n_o + o4] + (1.0 - cm[3]) * s_a[3 * n_o + o4]);
            if (f5) stream_store(&A.q_a[4 * plane + fb + g], cm[4] * s_a[1 * n_o + o5] + (1.0 - cm[4]) * s_a[4 * n_o + o5]);
            if (f6) stream_store(&A.q_a[5 * plane + fb + g], cm[5] * s_a[2 * n_o + o6] + (1.0 - cm[5]) * s_a[5 * n_o + o6]);
        }
        __syncthreads();  // the tiles are restaged for the next level
    }
}

struct UpdateArgs {
    double *rq, *rq_pl;
    const double *flx, *flx_pl, *qa, *qa_pl;
    int iall, gall, kall, lall, gall_pl, have_pl;
    long nreg;
};

__global__ void __launch_bounds__(BLK) update_kernel(UpdateArgs A)
{
    const long t = (long)blockIdx.x * BLK + threadIdx.x;
    const int iall = A.iall, gall = A.gall, kall = A.kall;
    if (t >= A.nreg) {
        const long p = t - A.nreg;
        if (!A.have_pl || p >= (long)kall * kLallPl) return;
        const int npl = A.gall_pl;
        const size_t fb = (size_t)p * npl;  // (k, l) planes are consecutive
        double s = A.flx_pl[fb + 1] * A.qa_pl[fb + 1];
        for (int v = 2; v < npl; ++v) s = s + A.flx_pl[fb + v] * A.qa_pl[fb + v];
        A.rq_pl[fb] = A.rq_pl[fb] - s;
        return;
    }
    const int g = (int)(t % gall);
    const int i = g % iall, j = g / iall;
    if (i < 1 || i > iall - 2 || j < 1 || j > iall - 2) return;
    const size_t plane = (size_t)A.lall * kall * gall;
    const double *f = A.flx + t, *q = A.qa + t;
    A.rq[t] = A.rq[t] - (f[0] * q[0] + f[plane] * q[plane] + f[2 * plane] * q[2 * plane] +
                         f[3 * plane] * q[3 * plane] + f[4 * plane] * q[4 * plane] + f[5 * plane] * q[5 * plane]);
}

}  // namespace

int launch_gradient(nicam_ctx c, double *grad, double *grad_pl, const double *scl, const double *scl_pl,
                    const double *coef_grad, const double *coef_grad_pl)
{
    const Dims &D = c->d;
    GradArgs A;
    A.grad = grad;
    A.grad_pl = grad_pl;
    A.scl = scl;
    A.scl_pl = scl_pl;
    A.cg = coef_grad;
    A.cg_pl = coef_grad_pl;
    A.iall = D.gall_1d;
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.gall_pl = D.gall_pl;
    A.have_pl = D.have_pl;
    A.nchunk = (D.gall + BLK - 1) / BLK;
    const long base = (long)A.nchunk * D.lall;
    int nkr = (int)((4096 + base - 1) / base);
    nkr = std::max(1, std::min(nkr, D.kall));
    A.kper = (D.kall + nkr - 1) / nkr;
    A.nkr = (D.kall + A.kper - 1) / A.kper;
    const long nmain = base * A.nkr;
    const long npole = grad_pl ? (D.kall * kLallPl + BLK - 1) / BLK : 0;
    if (nmain + npole > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    A.nmain = (int)nmain;
    hipLaunchKernelGGL(gradient_kernel, dim3((unsigned)(nmain + npole)), dim3(BLK), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

int launch_courant(nicam_ctx c, double *ch, double *ch_pl, double *cmask, double *cmask_pl, const double *flx_h,
                   const double *flx_h_pl, const double *rhog, const double *rhog_pl)
{
    const Dims &D = c->d;
    CourantArgs A;
    A.ch = ch;
    A.ch_pl = ch_pl;
    A.cmask = cmask;
    A.cmask_pl = cmask_pl;
    A.flx = flx_h;
    A.flx_pl = flx_h_pl;
    A.rhog = rhog;
    A.rhog_pl = rhog_pl;
    A.nreg = (long)D.gall * D.kall * D.lall;
    A.npole = ch_pl ? (long)D.gall_pl * D.kall * kLallPl : 0;
    A.gall_pl = D.gall_pl;
    A.have_pl = D.have_pl;
    A.eps = c->eps;
    const long n = A.nreg + A.npole, nb = (n + BLK - 1) / BLK;
    if (nb > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    hipLaunchKernelGGL(courant_kernel, dim3((unsigned)nb), dim3(BLK), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

int launch_remap(nicam_ctx c, double *q_a, double *q_a_pl, const double *q, const double *q_pl, const double *grad,
                 const double *grad_pl, const double *cmask, const double *cmask_pl, const double *xc,
                 const double *xc_pl, const double *x, const double *x_pl)
{
    const Dims &D = c->d;
    if (D.gall_1d > RMAXI) return fail(NICAM_ERR_ARG, "gall_1d larger than the remap tiles (160)");
    RemapArgs A;
    A.q_a = q_a;
    A.q_a_pl = q_a_pl;
    A.q = q;
    A.q_pl = q_pl;
    A.grad = grad;
    A.grad_pl = grad_pl;
    A.cmask = cmask;
    A.cmask_pl = cmask_pl;
    A.xc = xc;
    A.xc_pl = xc_pl;
    A.x = x;
    A.x_pl = x_pl;
    A.iall = D.gall_1d;
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.gall_pl = D.gall_pl;
    A.have_pl = D.have_pl && q_a_pl;
    A.nchunk = (D.gall + RCH - 1) / RCH;
    const long base = (long)A.nchunk * D.lall;
    int nkr = (int)((4096 + base - 1) / base);  // enough work-groups to fill the chip
    nkr = std::max(1, std::min(nkr, D.kall));
    A.kper = (D.kall + nkr - 1) / nkr;
    A.nkr = (D.kall + A.kper - 1) / A.kper;
    const long nmain = base * A.nkr;
    const long npole = A.have_pl ? ((long)D.kall * kLallPl + RCH - 1) / RCH : 0;
    if (nmain + npole > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    A.nmain = (int)nmain;
    const int n_in = RCH + 2 * D.gall_1d + 2, n_o = RCH + D.gall_1d + 1;
    const size_t lds = sizeof(double) * (7 * (size_t)n_in + 6 * (size_t)n_o);
    hipLaunchKernelGGL(remap_kernel, dim3((unsigned)(nmain + npole)), dim3(RCH), lds, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

int launch_update(nicam_ctx c, double *rhogq, double *rhogq_pl, const double *flx_h, const double *flx_h_pl,
                  const double *q_a, const double *q_a_pl)
{
    const Dims &D = c->d;
    UpdateArgs A;
    A.rq = rhogq;
    A.rq_pl = rhogq_pl;
    A.flx = flx_h;
    A.flx_pl = flx_h_pl;
    A.qa = q_a;
    A.qa_pl = q_a_pl;
    A.iall = D.gall_1d;
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.gall_pl = D.gall_pl;
    A.have_pl = D.have_pl && rhogq_pl;
    A.nreg = (long)D.gall * D.kall * D.lall;
    const long n = A.nreg + (A.have_pl ? (long)D.kall * kLallPl : 0), nb = (n + BLK - 1) / BLK;
    if (nb > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    hipLaunchKernelGGL(update_kernel, dim3((unsigned)nb), dim3(BLK), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

}  // namespace nicam
