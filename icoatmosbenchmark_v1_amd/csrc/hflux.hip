// horizontal_flux on gfx950.
//
// Reference: kernels/NICAM/dyn_horiz_adv_flux/src/mod_src_tracer.f90:42-379.
//
// The reference builds the triangle-vertex values rho/rhov at TI and TJ for a
// whole (k, l) slab, then sweeps three arcs, each arc writing its own face
// (1..3) of cell ij and the mirrored face (4..6) of the neighbour.  Here each
// work-item owns one output cell and produces all 15 of its outputs (6 faces,
// 3x3 mass-centroid components): the mirrored faces are recomputed from the
// neighbour's arc instead of being scattered, so every store has exactly one
// producer and the stores of a wavefront are contiguous in g.
//
// Each work-group covers CHUNK consecutive cells g of one (k, l) slab.  The
// triangle-vertex values TI/TJ needed by the chunk (cells g0-gall_1d-1 ..
// g0+CHUNK-1) are computed once into LDS from the rho/rhov values, which are
// themselves staged in LDS; the 3 arcs of every cell then read only LDS.
// Floating-point operations follow the reference's order term by term (the
// library is compiled with -ffp-contract=off), so the FP64 results are the
// reference's to the last bit.
#include "internal.h"

namespace nicam {
namespace {

constexpr int CHUNK = 256;          // output cells per work-group
constexpr int MAXI = 160;           // largest gall_1d the LDS tiles are sized for

struct HfluxArgs {
    double *flx;       // (gall,kall,lall,6)
    double *xc;        // (gall,kall,lall,3,3)
    const double *f[4];  // rho, rhovx, rhovy, rhovz (gall,kall,lall)
    const double *xr, *gp, *gt, *ga;
    const int *have_sgp;
    int iall, gall, kall, lall;
    double dt, eps;
};

__device__ __forceinline__ double dmax(double a, double b) { return (b > a) ? b : a; }

// One arc: flux through it and the centroid shift, in the reference's order
// (FLX:241-255 for AI; :266-280 AIJ; :292-306 AJ).
struct Arc {
    double flux, x2, y2, z2, rr2;
};

__device__ __forceinline__ Arc make_arc(const double (&ta)[4], const double (&tb)[4], double hnx,
                                        double hny, double hnz, double eps)
{
    Arc a;
    a.rr2 = 1.0 / dmax(ta[0] + tb[0], eps);
    a.x2 = ta[1] + tb[1];
    a.y2 = ta[2] + tb[2];
    a.z2 = ta[3] + tb[3];
    a.flux = 0.5 * (a.x2 * hnx + a.y2 * hny + a.z2 * hnz);
    return a;
}

__global__ void __launch_bounds__(CHUNK) hflux_kernel(HfluxArgs A)
{
    // LDS: rho/rhov for cells [g0 - iall - 1, g0 + CHUNK + iall + 1)
    //      TI/TJ (4 fields each) for cells [g0 - iall - 1, g0 + CHUNK)
    constexpr int SPAN_IN = CHUNK + 2 * (MAXI + 1);
    constexpr int SPAN_T = CHUNK + MAXI + 1;
    __shared__ double s_in[4][SPAN_IN];
    __shared__ double s_ti[4][SPAN_T];
    __shared__ double s_tj[4][SPAN_T];

    const int iall = A.iall, gall = A.gall, kall = A.kall, lall = A.lall;
    const int nchunk = (gall + CHUNK - 1) / CHUNK;
    // blocks of one region are consecutive, and consecutive blocks land on
    // different XCDs; group a region's (k, chunk) blocks onto one XCD so its
    // metrics stay in that XCD's L2.
    const int nb = gridDim.x;
    int b = blockIdx.x;
    {
        const int nxcd = 8;
        const int per = nb / nxcd;
        if (per > 0 && b < per * nxcd) b = (b % nxcd) * per + b / nxcd;
    }
    const int chunk = b % nchunk;
    const int kl = b / nchunk;
    const int k = kl % kall;
    const int l = kl / kall;
    if (l >= lall) return;
    const int g0 = chunk * CHUNK;
    const int lo_in = g0 - iall - 1;  // first cell staged
    const int n_in = CHUNK + 2 * (iall + 1);
    const int n_t = CHUNK + iall + 1;
    const size_t fb = ((size_t)l * kall + k) * gall;
    const size_t mb = (size_t)l * gall;
    const int tid = threadIdx.x;

    // stage inputs
    for (int s = tid; s < n_in; s += CHUNK) {
        const int g = lo_in + s;
        const bool ok = (g >= 0 && g < gall);
#pragma unroll
        for (int f = 0; f < 4; ++f) s_in[f][s] = ok ? A.f[f][fb + g] : 0.0;
    }
    __syncthreads();

    // triangle-vertex values at points p = lo_in + s (valid where i,j <= iall-2)
    const size_t tstride = (size_t)lall * gall;  // between GMTR_t(…,t,w) planes
    for (int s = tid; s < n_t; s += CHUNK) {
        const int p = lo_in + s;
        if (p < 0 || p >= gall) continue;
        const int i = p % iall, j = p / iall;
        if (i > iall - 2 || j > iall - 2) continue;
        // GMTR_t(g,1,l,t,w): index ((w*2 + t)*lall + l)*gall + g, w = 2,3,4 for W1..W3
        const double wi1 = A.gt[(2 * 2 + 0) * tstride + mb + p];
        const double wi2 = A.gt[(3 * 2 + 0) * tstride + mb + p];
        const double wi3 = A.gt[(4 * 2 + 0) * tstride + mb + p];
        const double wj1 = A.gt[(2 * 2 + 1) * tstride + mb + p];
        const double wj2 = A.gt[(3 * 2 + 1) * tstride + mb + p];
        const double wj3 = A.gt[(4 * 2 + 1) * tstride + mb + p];
        const int s00 = s, s10 = s + 1, s11 = s + iall + 1, s01 = s + iall;
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            // FLX:146-153 then :171-174
            double ti = s_in[f][s00] * wi1 + s_in[f][s10] * wi2;
            ti = ti + s_in[f][s11] * wi3;
            // FLX:155-158 then :176-183
            double tj = s_in[f][s00] * wj1;
            tj = tj + s_in[f][s11] * wj2 + s_in[f][s01] * wj3;
            s_ti[f][s] = ti;
            s_tj[f][s] = tj;
        }
    }
    __syncthreads();
    // singular point: TI(gmin-1,gmin-1) := TJ(gmin,gmin-1)  (FLX:188-201)
    if (A.have_sgp[l] && lo_in <= 0 && 1 - lo_in < n_t && tid < 4) {
        s_ti[tid][0 - lo_in] = s_tj[tid][1 - lo_in];
    }
    __syncthreads();

    const int g = g0 + tid;
    if (g >= gall) return;
    const int i = g % iall, j = g / iall;
    const int gmx = iall - 2;  // last index of the TI/TJ range (0-based gmax-1... = iall-2)
    const int s = g - lo_in;   // LDS slot of this cell
    const double dt = A.dt, eps = A.eps;

    double face[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double xcv[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};

    auto T = [&](double (&dst)[4], const double (*src)[SPAN_T], int slot) {
#pragma unroll
        for (int f = 0; f < 4; ++f) dst[f] = src[f][slot];
    };
    const size_t astride = (size_t)lall * gall;
    auto HN = [&](int arc, int c, int p) { return A.ga[((size_t)c * 3 + arc) * astride + mb + p]; };
    auto XR = [&](int arc, int d) { return A.xr[((size_t)d * 3 + arc) * astride + mb + g]; };
    const double rarea = A.gp[1 * astride + mb + g];  // GMTR_p(:,:,:,P_RAREA)

    double ta[4], tb[4];
    // own arcs (origin = this cell)
    if (i <= gmx && j >= 1 && j <= gmx) {  // AI: TJ(ijm1) + TI(ij)
        T(ta, s_tj, s - iall);
        T(tb, s_ti, s);
        const Arc a = make_arc(ta, tb, HN(0, 0, g), HN(0, 1, g), HN(0, 2, g), eps);
        face[0] = a.flux * rarea * dt;
        xcv[0][0] = XR(0, 0) - a.x2 * a.rr2 * dt * 0.5;
        xcv[0][1] = XR(0, 1) - a.y2 * a.rr2 * dt * 0.5;
        xcv[0][2] = XR(0, 2) - a.z2 * a.rr2 * dt * 0.5;
    }
    if (i <= gmx && j <= gmx) {  // AIJ: TI(ij) + TJ(ij)
        T(ta, s_ti, s);
        T(tb, s_tj, s);
        const Arc a = make_arc(ta, tb, HN(1, 0, g), HN(1, 1, g), HN(1, 2, g), eps);
        face[1] = a.flux * rarea * dt;
        xcv[1][0] = XR(1, 0) - a.x2 * a.rr2 * dt * 0.5;
        xcv[1][1] = XR(1, 1) - a.y2 * a.rr2 * dt * 0.5;
        xcv[1][2] = XR(1, 2) - a.z2 * a.rr2 * dt * 0.5;
    }
    if (i >= 1 && i <= gmx && j <= gmx) {  // AJ: TJ(ij) + TI(im1j)
        T(ta, s_tj, s);
        T(tb, s_ti, s - 1);
        const Arc a = make_arc(ta, tb, HN(2, 0, g), HN(2, 1, g), HN(2, 2, g), eps);
        face[2] = a.flux * rarea * dt;
        xcv[2][0] = XR(2, 0) - a.x2 * a.rr2 * dt * 0.5;
        xcv[2][1] = XR(2, 1) - a.y2 * a.rr2 * dt * 0.5;
        xcv[2][2] = XR(2, 2) - a.z2 * a.rr2 * dt * 0.5;
    }
    // mirrored faces (origin = a neighbour)
    if (i >= 1 && i - 1 <= gmx && j >= 1 && j <= gmx) {  // face 4 <- AI at (i-1, j)
        const int o = g - 1;
        T(ta, s_tj, s - 1 - iall);
        T(tb, s_ti, s - 1);
        const Arc a = make_arc(ta, tb, HN(0, 0, o), HN(0, 1, o), HN(0, 2, o), eps);
        face[3] = -a.flux * rarea * dt;
    }
    if (i >= 1 && j >= 1 && i - 1 <= gmx && j - 1 <= gmx) {  // face 5 <- AIJ at (i-1, j-1)
        const int o = g - 1 - iall;
        T(ta, s_ti, s - 1 - iall);
        T(tb, s_tj, s - 1 - iall);
        const Arc a = make_arc(ta, tb, HN(1, 0, o), HN(1, 1, o), HN(1, 2, o), eps);
        face[4] = -a.flux * rarea * dt;
    }
    if (i >= 1 && i <= gmx && j >= 1 && j - 1 <= gmx) {  // face 6 <- AJ at (i, j-1)
        const int o = g - iall;
        T(ta, s_tj, s - iall);
        T(tb, s_ti, s - iall - 1);
        const Arc a = make_arc(ta, tb, HN(2, 0, o), HN(2, 1, o), HN(2, 2, o), eps);
        face[5] = -a.flux * rarea * dt;
    }
    if (A.have_sgp[l] && i == 1 && j == 1) face[5] = 0.0;  // FLX:311-320

    const size_t plane = (size_t)lall * kall * gall;
#pragma unroll
    for (int m = 0; m < 6; ++m) A.flx[m * plane + fb + g] = face[m];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int a = 0; a < 3; ++a) A.xc[((size_t)d * 3 + a) * plane + fb + g] = xcv[a][d];
}

// Pole regions: one work-item per (k, l) walks the ring (FLX:326-374).
struct HfluxPoleArgs {
    double *flx_pl, *xc_pl;
    const double *f[4];
    const double *xr, *gp, *gt, *ga;
    int gall_pl, kall;
    double dt, eps;
};

__global__ void hflux_pole_kernel(HfluxPoleArgs A)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int kall = A.kall, npl = A.gall_pl;
    if (t >= kall * kLallPl) return;
    const int k = t % kall, l = t / kall;
    const size_t fb = ((size_t)l * kall + k) * npl;
    // GMTR_*_pl(v,1,l,c): index (c*2 + l)*gall_pl + v
    auto M = [&](const double *m, int c, int v) { return m[((size_t)c * kLallPl + l) * npl + v]; };
    double tv[16][4];
    for (int v = 1; v < npl; ++v) {
        const int vp1 = (v + 1 == npl) ? 1 : v + 1;
        const double w1 = M(A.gt, 2, v), w2 = M(A.gt, 3, v), w3 = M(A.gt, 4, v);
        for (int f = 0; f < 4; ++f)
            tv[v][f] = A.f[f][fb + 0] * w1 + A.f[f][fb + v] * w2 + A.f[f][fb + vp1] * w3;
    }
    const double rarea = M(A.gp, 1, 0);
    const size_t plane = (size_t)kLallPl * kall * npl;
    for (int v = 1; v < npl; ++v) {
        const int vm1 = (v - 1 == 0) ? npl - 1 : v - 1;
        const double rr2 = 1.0 / dmax(tv[vm1][0] + tv[v][0], A.eps);
        const double x2 = tv[vm1][1] + tv[v][1];
        const double y2 = tv[vm1][2] + tv[v][2];
        const double z2 = tv[vm1][3] + tv[v][3];
        const double flux = 0.5 * (x2 * M(A.ga, 0, v) + y2 * M(A.ga, 1, v) + z2 * M(A.ga, 2, v));
        A.flx_pl[fb + v] = flux * rarea * A.dt;
        A.xc_pl[0 * plane + fb + v] = M(A.xr, 0, v) - x2 * rr2 * A.dt * 0.5;
        A.xc_pl[1 * plane + fb + v] = M(A.xr, 1, v) - y2 * rr2 * A.dt * 0.5;
        A.xc_pl[2 * plane + fb + v] = M(A.xr, 2, v) - z2 * rr2 * A.dt * 0.5;
    }
}

}  // namespace

int launch_hflux(nicam_ctx c, double *flx_h, double *flx_h_pl, double *xc, double *xc_pl,
                 const double *rho, const double *rho_pl, const double *rvx, const double *rvx_pl,
                 const double *rvy, const double *rvy_pl, const double *rvz, const double *rvz_pl,
                 double dt)
{
    const Dims &D = c->d;
    if (D.gall_1d > MAXI) return fail(NICAM_ERR_ARG, "gall_1d larger than the LDS tiles (160)");
    HfluxArgs A;
    A.flx = flx_h;
    A.xc = xc;
    A.f[0] = rho;
    A.f[1] = rvx;
    A.f[2] = rvy;
    A.f[3] = rvz;
    A.xr = c->GRD_xr;
    A.gp = c->GMTR_p;
    A.gt = c->GMTR_t;
    A.ga = c->GMTR_a;
    A.have_sgp = c->have_sgp;
    A.iall = D.gall_1d;
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.dt = dt;
    A.eps = c->eps;
    const long nchunk = (D.gall + CHUNK - 1) / CHUNK;
    const long nblk = nchunk * D.kall * D.lall;
    if (nblk > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    hipLaunchKernelGGL(hflux_kernel, dim3((unsigned)nblk), dim3(CHUNK), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    if (D.have_pl) {
        HfluxPoleArgs P;
        P.flx_pl = flx_h_pl;
        P.xc_pl = xc_pl;
        P.f[0] = rho_pl;
        P.f[1] = rvx_pl;
        P.f[2] = rvy_pl;
        P.f[3] = rvz_pl;
        P.xr = c->GRD_xr_pl;
        P.gp = c->GMTR_p_pl;
        P.gt = c->GMTR_t_pl;
        P.ga = c->GMTR_a_pl;
        P.gall_pl = D.gall_pl;
        P.kall = D.kall;
        P.dt = dt;
        P.eps = c->eps;
        const int n = D.kall * kLallPl;
        hipLaunchKernelGGL(hflux_pole_kernel, dim3((n + 63) / 64), dim3(64), 0, c->stream, P);
        NICAM_HIP(hipGetLastError());
    }
    return NICAM_OK;
}

}  // namespace nicam
