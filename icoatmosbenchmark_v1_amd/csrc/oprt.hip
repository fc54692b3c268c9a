// OPRT_diffusion on gfx950.
//
// Reference: kernels/NICAM/dyn_diffusion/src/mod_oprt.f90:41-296.
//
// The reference first forms, for a whole (k, l) slab, the gradient of the
// scalar interpolated to the two triangle vertices TI and TJ of every cell
// (vt, three components each, :97-132), then sums over the six edges of each
// hexagon in three passes, one per component (:138-216).  Here a work-group
// owns CHUNK consecutive cells of one region and marches up a range of
// levels.  Its coefficients (coef_intp of the vt points it forms, coef_diff
// of its cells) are K-invariant and stay in registers.  At each level the
// group stages scl and kh of the cells its stencils touch in LDS, forms vt
// of cells [g0-gall_1d-1, g0+CHUNK) in LDS, and each work-item then produces
// its cell's dscl.  The expressions keep the reference's operand order
// (including the X component that the Y and Z passes read as the first point
// of each pair), so with -ffp-contract=off the FP64 results are the
// reference's to the last bit.  The pole regions (:220-291) are extra
// work-groups of the same launch.
#include "internal.h"

namespace nicam {
namespace {

constexpr int CHUNK = 256;  // output cells per work-group
constexpr int MAXI = 255;   // largest gall_1d: CHUNK*NTP >= CHUNK + MAXI + 1
constexpr int NTP = 2;      // vt points per work-item
constexpr int NST = 3;      // staged cells per work-item: CHUNK*NST >= CHUNK + 2*MAXI + 2

struct DiffArgs {
    double *dscl, *dscl_pl;
    const double *scl, *scl_pl, *kh, *kh_pl;
    const double *ci, *ci_pl;  // coef_intp(gall,3,3,2,lall), coef_intp_pl(gall_pl,3,3,2)
    const double *cd, *cd_pl;  // coef_diff(gall,6,3,lall), coef_diff_pl(5,3,2)
    const int *have_sgp;
    int iall, gall, kall, lall, gall_pl, have_pl;
    int nchunk, nkr, kper, nmain;
};

// vt of one triangle vertex: the reference's expression, term by term
__device__ __forceinline__ double vt_point(double c1, double c2, double c3, double s0, double s1, double s2)
{
    return ((+2.0 * c1 - 1.0 * c2 - 1.0 * c3) * s0 + (-1.0 * c1 + 2.0 * c2 - 1.0 * c3) * s1 +
            (-1.0 * c1 - 1.0 * c2 + 2.0 * c3) * s2) /
           3.0;
}

__device__ void diffusion_pole(const DiffArgs &A, int t)
{
    // one work-item per (k, pole region): :220-287; zeros without pole regions (:288-290)
    const int kall = A.kall, npl = A.gall_pl;
    if (t >= kall * kLallPl) return;
    const int k = t % kall, l = t / kall;
    const size_t fb = ((size_t)l * kall + k) * npl;
    if (!A.have_pl) {
        for (int v = 0; v < npl; ++v) A.dscl_pl[fb + v] = 0.0;
        return;
    }
    const int gmin_pl = 1, gmax_pl = npl - 1;  // 0-based ring 1..npl-1, centre 0
    double vt[16][3];
    for (int d = 0; d < 3; ++d)
        for (int v = gmin_pl; v <= gmax_pl; ++v) {
            const int vp1 = (v + 1 == gmax_pl + 1) ? gmin_pl : v + 1;
            const double *c = A.ci_pl + ((size_t)(l * 3 + d) * 3) * npl;
            vt[v][d] = vt_point(c[0 * npl + v], c[1 * npl + v], c[2 * npl + v], A.scl_pl[fb + 0], A.scl_pl[fb + v],
                                A.scl_pl[fb + vp1]);
        }
    double acc = 0.0;
    for (int v = gmin_pl; v <= gmax_pl; ++v) {
        const int vm1 = (v - 1 == gmin_pl - 1) ? gmax_pl : v - 1;
        const double *cd = A.cd_pl + (size_t)l * 3 * 5;
        acc = acc + (cd[0 * 5 + (v - 1)] * (vt[vm1][0] + vt[v][0]) + cd[1 * 5 + (v - 1)] * (vt[vm1][1] + vt[v][1]) +
                     cd[2 * 5 + (v - 1)] * (vt[vm1][2] + vt[v][2])) *
                        0.5 * (A.kh_pl[fb + 0] + A.kh_pl[fb + v]);
    }
    A.dscl_pl[fb + 0] = acc;
    for (int v = 1; v < npl; ++v) A.dscl_pl[fb + v] = 0.0;
}

__global__ void __launch_bounds__(CHUNK) diffusion_kernel(DiffArgs A)
{
    extern __shared__ double lds[];
    const int iall = A.iall, gall = A.gall, kall = A.kall, lall = A.lall;
    const int n_in = CHUNK + 2 * iall + 2;  // staged scl/kh window
    const int n_t = CHUNK + iall + 1;       // vt window
    double *s_s = lds, *s_k = lds + n_in, *s_vt = lds + 2 * n_in;  // s_vt: [t*3+d][n_t]

    int b = blockIdx.x;
    if (b >= A.nmain) {  // pole work-groups
        diffusion_pole(A, (b - A.nmain) * CHUNK + threadIdx.x);
        return;
    }
    {  // keep a region's chunks and level ranges on one XCD (block b runs on XCD b % 8)
        const int nxcd = 8, per = A.nmain / nxcd;
        if (per > 0 && b < per * nxcd) b = (b % nxcd) * per + b / nxcd;
    }
    const int chunk = b % A.nchunk;
    const int kr = (b / A.nchunk) % A.nkr;
    const int l = b / (A.nchunk * A.nkr);
    const int k0 = kr * A.kper, k1 = min(kall, k0 + A.kper);
    if (l >= lall || k0 >= k1) return;
    const int tid = threadIdx.x;
    const int g0 = chunk * CHUNK, lo = g0 - iall - 1;
    const int gmin = iall + 1, gmax = gall - iall - 2;  // 0-based (:85-86); vt runs over [0, gmax]
    const bool sgp = A.have_sgp[l] != 0;

    // ---- K-invariant coefficients ----
    // vt points of this work-item; with a singular point, TI of point 0 is TJ of point 1 (:128-132)
    double cv[NTP][2][3][3];
    int vp[NTP], vsh[NTP];
#pragma unroll
    for (int r = 0; r < NTP; ++r) {
        const int s = tid + r * CHUNK, p = lo + s;
        vp[r] = (s < n_t && p >= 0 && p <= gmax) ? p : -1;
        vsh[r] = (vp[r] == 0 && sgp) ? 1 : 0;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int tt = vsh[r] ? 1 : t, pp = vsh[r] && t == 0 ? 1 : p;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    cv[r][t][d][c] =
                        vp[r] >= 0 ? A.ci[((((size_t)l * 2 + tt) * 3 + d) * 3 + c) * gall + pp] : 0.0;
        }
    }
    const int g = g0 + tid;
    const bool active = g < gall;
    const bool inner = g >= gmin && g <= gmax;
    double cdv[3][6];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int m = 0; m < 6; ++m) cdv[d][m] = inner ? A.cd[(((size_t)l * 3 + d) * 6 + m) * gall + g] : 0.0;

    // ---- level loop, the next level's window prefetched into registers ----
    double nx[NST][2];
    auto load_level = [&](int k) {
        const size_t fb = ((size_t)l * kall + k) * gall;
#pragma unroll
        for (int r = 0; r < NST; ++r) {
            const int s = tid + r * CHUNK, p = lo + s;
            const bool ok = s < n_in && p >= 0 && p < gall;
            nx[r][0] = ok ? A.scl[fb + p] : 0.0;
            nx[r][1] = ok ? A.kh[fb + p] : 0.0;
        }
    };
    auto store_level = [&]() {
#pragma unroll
        for (int r = 0; r < NST; ++r) {
            const int s = tid + r * CHUNK;
            if (s < n_in) {
                s_s[s] = nx[r][0];
                s_k[s] = nx[r][1];
            }
        }
    };
    load_level(k0);
    store_level();
    __syncthreads();
    if (k0 + 1 < k1) load_level(k0 + 1);

    for (int k = k0; k < k1; ++k) {
        // vt of my points (:97-125)
#pragma unroll
        for (int r = 0; r < NTP; ++r) {
            if (vp[r] < 0) continue;
            const int s = tid + r * CHUNK;
            const int sj = s + vsh[r];  // TJ stencil origin of the TI slot at a singular point
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double ti;
                if (vsh[r])
                    ti = vt_point(cv[r][0][d][0], cv[r][0][d][1], cv[r][0][d][2], s_s[sj], s_s[sj + iall + 1],
                                  s_s[sj + iall]);
                else
                    ti = vt_point(cv[r][0][d][0], cv[r][0][d][1], cv[r][0][d][2], s_s[s], s_s[s + 1],
                                  s_s[s + iall + 1]);
                const double tj = vt_point(cv[r][1][d][0], cv[r][1][d][1], cv[r][1][d][2], s_s[s],
                                           s_s[s + iall + 1], s_s[s + iall]);
                s_vt[(0 * 3 + d) * n_t + s] = ti;
                s_vt[(1 * 3 + d) * n_t + s] = tj;
            }
        }
        __syncthreads();
        const double kc = s_k[g - lo];
        double kf[6];
        if (inner) {  // read kh before the window is overwritten
            const int s = g - lo;
            kf[0] = 0.5 * (kc + s_k[s + iall + 1]);
            kf[1] = 0.5 * (kc + s_k[s + iall]);
            kf[2] = 0.5 * (s_k[s - 1] + kc);
            kf[3] = 0.5 * (s_k[s - iall - 1] + kc);
            kf[4] = 0.5 * (s_k[s - iall] + kc);
            kf[5] = 0.5 * (kc + s_k[s + 1]);
        }
        __syncthreads();
        if (k + 1 < k1) {  // the staging tiles are free
            store_level();
            if (k + 2 < k1) load_level(k + 2);
        }
        if (active) {
            double out = 0.0;
            if (inner) {
                // vt window slots of ij, im1j, im1jm1, ijm1
                const int s0 = g - lo, s1 = s0 - 1, s2 = s0 - iall - 1, s3 = s0 - iall;
                auto V = [&](int t, int d, int s) { return s_vt[(t * 3 + d) * n_t + s]; };
#pragma unroll
                for (int d = 0; d < 3; ++d) {  // :138-216, X then Y then Z
                    const double sum = kf[0] * cdv[d][0] * (V(0, 0, s0) + V(1, d, s0)) +
                                       kf[1] * cdv[d][1] * (V(1, 0, s0) + V(0, d, s1)) +
                                       kf[2] * cdv[d][2] * (V(0, 0, s1) + V(1, d, s2)) +
                                       kf[3] * cdv[d][3] * (V(1, 0, s2) + V(0, d, s2)) +
                                       kf[4] * cdv[d][4] * (V(0, 0, s2) + V(1, d, s3)) +
                                       kf[5] * cdv[d][5] * (V(1, 0, s3) + V(0, d, s0));
                    out = d == 0 ? sum : out + sum;
                }
            }
            stream_store(&A.dscl[((size_t)l * kall + k) * gall + g], out);
        }
        __syncthreads();
    }
}


// ---------------------------------------------------------------------------
// OPRT3D_divdamp (kernels/NICAM/dyn_divdamp/src/mod_oprt3d.f90:41-402).
//
// The reference first forms rhogw_vm, the vertical mass flux divided by the
// vertical metrics, for the whole region (:143-162), then per level the
// scalar divergence sclt at the TI/TJ vertices (horizontal part through
// coef_intp plus the vertical difference of rhogw_vm, :166-229) and its
// gradient ddivd{x,y,z} through coef_diff (:237-282).  Here a work-group owns
// CHUNK cells of one region and marches up a level range: at level k it
// forms rhogw_vm(k+1) of its staged window (rhogw_vm(k) and the raw rhogv of
// level k are carried from the previous level), rhogv*RGAM of level k, then
// sclt, then the three outputs.  Each input is read once per level.
struct DdmpArgs {
    double *out[3], *out_pl[3];  // ddivdx, ddivdy, ddivdz
    const double *v[3], *v_pl[3], *w, *w_pl;
    const double *ci, *ci_pl, *cd, *cd_pl;
    const double *rgsqrth, *rgsqrth_pl, *rgam, *rgam_pl, *rgamh, *rgamh_pl;
    const double *c2w, *c2w_pl;  // VMTR_C2WfactGz(gall,kall,6,lall), _pl(gall_pl,kall,6,2)
    const double *rdgz;          // GRD_rdgz(kall), device copy
    const int *have_sgp;
    int iall, gall, kall, lall, gall_pl, have_pl;
    int nchunk, nkr, kper, nmain;
};

__device__ void ddmp_pole(const DdmpArgs &A, int t)
{
    // one work-item per (k, pole region): :313-392; zeros without pole regions (:393-397)
    const int kall = A.kall, npl = A.gall_pl;
    if (t >= kall * kLallPl) return;
    const int k = t % kall, l = t / kall;
    const size_t fb = ((size_t)l * kall + k) * npl;
    const bool edge = k == 0 || k == kall - 1;
    if (!A.have_pl || edge) {
        for (int d = 0; d < 3; ++d)
            for (int v = 0; v < npl; ++v) A.out_pl[d][fb + v] = 0.0;
        return;
    }
    const int kmin = 1, kmax = kall - 2;
    const size_t cstr = (size_t)kall * npl;           // between C2WfactGz_pl planes
    const double *c2w = A.c2w_pl + (size_t)l * 6 * cstr;
    auto W = [&](int p, int kk) -> double {
        if (kk == kmin || kk == kmax + 1) return 0.0;
        const size_t f = ((size_t)l * kall + kk) * npl + p, fm = ((size_t)l * kall + kk - 1) * npl + p;
        return (c2w[0 * cstr + (size_t)kk * npl + p] * A.v_pl[0][f] + c2w[1 * cstr + (size_t)kk * npl + p] * A.v_pl[0][fm] +
                c2w[2 * cstr + (size_t)kk * npl + p] * A.v_pl[1][f] + c2w[3 * cstr + (size_t)kk * npl + p] * A.v_pl[1][fm] +
                c2w[4 * cstr + (size_t)kk * npl + p] * A.v_pl[2][f] + c2w[5 * cstr + (size_t)kk * npl + p] * A.v_pl[2][fm]) *
                   A.rgamh_pl[f] +
               A.w_pl[f] * A.rgsqrth_pl[f];
    };
    double vm[3][16], sp[16];
    for (int v = 0; v < npl; ++v)
        for (int d = 0; d < 3; ++d) vm[d][v] = A.v_pl[d][fb + v] * A.rgam_pl[fb + v];
    const double rdgz = A.rdgz[k];
    for (int v = 1; v < npl; ++v) {
        const int vp1 = (v + 1 == npl) ? 1 : v + 1;
        const double sw = ((W(0, k + 1) + W(v, k + 1) + W(vp1, k + 1)) - (W(0, k) + W(v, k) + W(vp1, k))) / 3.0 * rdgz;
        const double *c = A.ci_pl + (size_t)l * 9 * npl;  // coef_intp_pl(v,c,d,l): ((l*3+d)*3+c)*npl+v
        sp[v] = c[(0 * 3 + 0) * npl + v] * vm[0][0] + c[(0 * 3 + 1) * npl + v] * vm[0][v] +
                c[(0 * 3 + 2) * npl + v] * vm[0][vp1] + c[(1 * 3 + 0) * npl + v] * vm[1][0] +
                c[(1 * 3 + 1) * npl + v] * vm[1][v] + c[(1 * 3 + 2) * npl + v] * vm[1][vp1] +
                c[(2 * 3 + 0) * npl + v] * vm[2][0] + c[(2 * 3 + 1) * npl + v] * vm[2][v] +
                c[(2 * 3 + 2) * npl + v] * vm[2][vp1] + sw;
    }
    (void)kmin;
    for (int d = 0; d < 3; ++d) {
        const double *cd = A.cd_pl + (size_t)l * 15 + d * 5;
        double acc = 0.0;
        for (int v = 1; v < npl; ++v) {
            const int vm1 = (v == 1) ? npl - 1 : v - 1;
            acc = acc + cd[v - 1] * (sp[vm1] + sp[v]);
        }
        A.out_pl[d][fb + 0] = acc;
        for (int v = 1; v < npl; ++v) A.out_pl[d][fb + v] = 0.0;
    }
}

template <int NSTT>
__global__ void __launch_bounds__(CHUNK) ddmp_kernel(DdmpArgs A)
{
    extern __shared__ double lds[];
    const int iall = A.iall, gall = A.gall, kall = A.kall, lall = A.lall;
    const int n_in = CHUNK + 2 * iall + 2, n_t = CHUNK + iall + 1;
    double *s_vm = lds;                  // [3][n_in] rhogv*RGAM of level k
    double *s_w = lds + 3 * n_in;        // [2][n_in] rhogw_vm of levels k and k+1 (alternating)
    double *s_st = lds + 5 * n_in;       // [2][n_t]  sclt at TI, TJ

    int b = blockIdx.x;
    if (b >= A.nmain) {
        ddmp_pole(A, (b - A.nmain) * CHUNK + threadIdx.x);
        return;
    }
    {
        const int nxcd = 8, per = A.nmain / nxcd;
        if (per > 0 && b < per * nxcd) b = (b % nxcd) * per + b / nxcd;
    }
    const int chunk = b % A.nchunk;
    const int kr = (b / A.nchunk) % A.nkr;
    const int l = b / (A.nchunk * A.nkr);
    const int kmin = 1, kmax = kall - 2;
    const int k0 = kmin + kr * A.kper, k1 = min(kmax + 1, k0 + A.kper);  // levels [k0, k1) of [kmin, kmax]
    if (l >= lall || k0 >= k1) return;
    const int tid = threadIdx.x;
    const int g0 = chunk * CHUNK, lo = g0 - iall - 1;
    const int gmin = iall + 1, gmax = gall - iall - 2;
    const bool sgp = A.have_sgp[l] != 0;
    const int g = g0 + tid;
    const bool active = g < gall;
    const bool inner = g >= gmin && g <= gmax;
    const size_t slab = (size_t)gall, lk = (size_t)l * kall;
    const size_t cstr = (size_t)kall * gall;  // between C2WfactGz planes
    const double *c2w = A.c2w + (size_t)l * 6 * cstr;

    // the level planes outside [kmin, kmax] are zero (:303-310)
    if (active && k0 == kmin)
        for (int d = 0; d < 3; ++d) A.out[d][(lk + 0) * slab + g] = 0.0;
    if (active && k1 == kmax + 1)
        for (int d = 0; d < 3; ++d) A.out[d][(lk + kall - 1) * slab + g] = 0.0;

    // ---- K-invariant coefficients (TI of point 0 is TJ of point 1 at a singular point, :227-229) ----
    double cv[NTP][2][3][3];
    int vp[NTP], vsh[NTP];
#pragma unroll
    for (int r = 0; r < NTP; ++r) {
        const int s = tid + r * CHUNK, p = lo + s;
        vp[r] = (s < n_t && p >= 0 && p <= gmax) ? p : -1;
        vsh[r] = (vp[r] == 0 && sgp) ? 1 : 0;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int tt = vsh[r] ? 1 : t, pp = vsh[r] && t == 0 ? 1 : p;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    cv[r][t][d][c] = vp[r] >= 0 ? A.ci[((((size_t)l * 2 + tt) * 3 + d) * 3 + c) * gall + pp] : 0.0;
        }
    }
    double cdv[3][6];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int m = 0; m < 6; ++m) cdv[d][m] = inner ? A.cd[(((size_t)l * 3 + d) * 6 + m) * gall + g] : 0.0;

    // ---- window state: raw rhogv of the current level in registers ----
    double vr[NSTT][3];
    auto in_win = [&](int r, int &p) {
        const int s = tid + r * CHUNK;
        p = lo + s;
        return s < n_in && p >= 0 && p < gall;
    };
    // prologue: rhogv(k0) and rhogw_vm(k0)
    int cur = 0;
#pragma unroll
    for (int r = 0; r < NSTT; ++r) {
        int p;
        const bool ok = in_win(r, p);
        const size_t f = (lk + k0) * slab + p;
#pragma unroll
        for (int d = 0; d < 3; ++d) vr[r][d] = ok ? A.v[d][f] : 0.0;
        const int s = tid + r * CHUNK;
        if (s < n_in) {
            double wv = 0.0;
            if (ok && k0 > kmin) {  // :148-156 at level k0
                const size_t fk = (size_t)k0 * slab + p, fm = f - slab;  // fk: within the region's C2W planes
                wv = (c2w[0 * cstr + fk] * A.v[0][f] + c2w[1 * cstr + fk] * A.v[0][fm] +
                      c2w[2 * cstr + fk] * A.v[1][f] + c2w[3 * cstr + fk] * A.v[1][fm] +
                      c2w[4 * cstr + fk] * A.v[2][f] + c2w[5 * cstr + fk] * A.v[2][fm]) *
                         A.rgamh[f] +
                     A.w[f] * A.rgsqrth[f];
            }
            s_w[cur * n_in + s] = wv;
        }
    }

    for (int k = k0; k < k1; ++k) {
        const int nxt = cur ^ 1;
        // rhogw_vm(k+1) and rhogv*RGAM(k) of the window
#pragma unroll
        for (int r = 0; r < NSTT; ++r) {
            int p;
            const bool ok = in_win(r, p);
            const int s = tid + r * CHUNK;
            if (s >= n_in) continue;
            const size_t f = (lk + k) * slab + p, f1 = f + slab;
            double v1[3] = {0.0, 0.0, 0.0};
            double wv = 0.0;
            if (ok && k + 1 <= kmax) {
#pragma unroll
                for (int d = 0; d < 3; ++d) v1[d] = A.v[d][f1];
                const size_t fk = (size_t)(k + 1) * slab + p;  // within the region's C2W planes
                wv = (c2w[0 * cstr + fk] * v1[0] + c2w[1 * cstr + fk] * vr[r][0] + c2w[2 * cstr + fk] * v1[1] +
                      c2w[3 * cstr + fk] * vr[r][1] + c2w[4 * cstr + fk] * v1[2] + c2w[5 * cstr + fk] * vr[r][2]) *
                         A.rgamh[f1] +
                     A.w[f1] * A.rgsqrth[f1];
            }
            s_w[nxt * n_in + s] = wv;
            const double rg = ok ? A.rgam[f] : 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                s_vm[d * n_in + s] = vr[r][d] * rg;
                vr[r][d] = v1[d];
            }
        }
        __syncthreads();
        // sclt of my vertices (:183-225)
        const double rdgz = A.rdgz[k];
        const double *w0 = s_w + cur * n_in, *w1 = s_w + nxt * n_in;
#pragma unroll
        for (int r = 0; r < NTP; ++r) {
            if (vp[r] < 0) continue;
            const int s = tid + r * CHUNK;
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                // TI: (ij, ip1j, ip1jp1); TJ: (ij, ip1jp1, ijp1); the singular TI slot is TJ of point 1
                const bool tj = t == 1 || vsh[r];
                const int a = tj && t == 0 ? s + 1 : s;
                const int b2 = tj ? a + iall + 1 : a + 1;
                const int c3 = tj ? a + iall : a + iall + 1;
                const double sw = ((w1[a] + w1[b2] + w1[c3]) - (w0[a] + w0[b2] + w0[c3])) / 3.0 * rdgz;
                const double *vx = s_vm, *vy = s_vm + n_in, *vz = s_vm + 2 * n_in;
                s_st[t * n_t + s] = cv[r][t][0][0] * vx[a] + cv[r][t][0][1] * vx[b2] + cv[r][t][0][2] * vx[c3] +
                                    cv[r][t][1][0] * vy[a] + cv[r][t][1][1] * vy[b2] + cv[r][t][1][2] * vy[c3] +
                                    cv[r][t][2][0] * vz[a] + cv[r][t][2][1] * vz[b2] + cv[r][t][2][2] * vz[c3] + sw;
            }
        }
        __syncthreads();
        if (active) {
            const size_t fo = (lk + k) * slab + g;
            if (inner) {
                const int s0 = g - lo, s1 = s0 - 1, s2 = s0 - iall - 1, s3 = s0 - iall;
                const double *ti = s_st, *tj = s_st + n_t;
                const double e1 = ti[s0] + tj[s0], e2 = tj[s0] + ti[s1], e3 = ti[s1] + tj[s2];
                const double e4 = tj[s2] + ti[s2], e5 = ti[s2] + tj[s3], e6 = tj[s3] + ti[s0];
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    stream_store(&A.out[d][fo], cdv[d][0] * e1 + cdv[d][1] * e2 + cdv[d][2] * e3 + cdv[d][3] * e4 +
                                                    cdv[d][4] * e5 + cdv[d][5] * e6);
            } else {
#pragma unroll
                for (int d = 0; d < 3; ++d) stream_store(&A.out[d][fo], 0.0);
            }
        }
        __syncthreads();
        cur = nxt;
    }
}

}  // namespace

int launch_diffusion(nicam_ctx c, double *dscl, double *dscl_pl, const double *scl, const double *scl_pl,
                     const double *kh, const double *kh_pl, const double *coef_intp, const double *coef_intp_pl,
                     const double *coef_diff, const double *coef_diff_pl)
{
    const Dims &D = c->d;
    if (D.gall_1d > MAXI) return fail(NICAM_ERR_ARG, "gall_1d larger than the supported 255");
    DiffArgs A;
    A.dscl = dscl;
    A.dscl_pl = dscl_pl;
    A.scl = scl;
    A.scl_pl = scl_pl;
    A.kh = kh;
    A.kh_pl = kh_pl;
    A.ci = coef_intp;
    A.ci_pl = coef_intp_pl;
    A.cd = coef_diff;
    A.cd_pl = coef_diff_pl;
    A.have_sgp = c->have_sgp;
    A.iall = D.gall_1d;
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.gall_pl = D.gall_pl;
    A.have_pl = D.have_pl;
    A.nchunk = (D.gall + CHUNK - 1) / CHUNK;
    const long base = (long)A.nchunk * D.lall;
    int nkr = (int)((2048 + base - 1) / base);
    nkr = std::max(1, std::min(nkr, D.kall));
    if (c->tune.diff_kper > 0) nkr = (D.kall + c->tune.diff_kper - 1) / c->tune.diff_kper;
    A.kper = (D.kall + nkr - 1) / nkr;
    A.nkr = (D.kall + A.kper - 1) / A.kper;
    const long nmain = base * A.nkr;
    const long npole = dscl_pl ? (D.kall * kLallPl + CHUNK - 1) / CHUNK : 0;
    if (nmain + npole > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    A.nmain = (int)nmain;
    const size_t lds = sizeof(double) * (2 * (CHUNK + 2 * D.gall_1d + 2) + 6 * (CHUNK + D.gall_1d + 1));
    if (lds > 64 * 1024)
        NICAM_HIP(hipFuncSetAttribute((const void *)diffusion_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    hipLaunchKernelGGL(diffusion_kernel, dim3((unsigned)(nmain + npole)), dim3(CHUNK), lds, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

int launch_divdamp(nicam_ctx c, double *const out[3], double *const out_pl[3], const double *const v[3],
                   const double *const v_pl[3], const double *w, const double *w_pl, const double *coef_intp,
                   const double *coef_intp_pl, const double *coef_diff, const double *coef_diff_pl,
                   const double *rgsqrth, const double *rgsqrth_pl, const double *rgam, const double *rgam_pl,
                   const double *rgamh, const double *rgamh_pl, const double *c2w, const double *c2w_pl,
                   const double *rdgz_dev)
{
    const Dims &D = c->d;
    if (D.gall_1d > MAXI) return fail(NICAM_ERR_ARG, "gall_1d larger than the supported 255");
    if (D.kall < 3) return fail(NICAM_ERR_ARG, "OPRT3D_divdamp needs kall >= 3");
    DdmpArgs A;
    for (int d = 0; d < 3; ++d) {
        A.out[d] = out[d];
        A.out_pl[d] = out_pl[d];
        A.v[d] = v[d];
        A.v_pl[d] = v_pl[d];
    }
    A.w = w;
    A.w_pl = w_pl;
    A.ci = coef_intp;
    A.ci_pl = coef_intp_pl;
    A.cd = coef_diff;
    A.cd_pl = coef_diff_pl;
    A.rgsqrth = rgsqrth;
    A.rgsqrth_pl = rgsqrth_pl;
    A.rgam = rgam;
    A.rgam_pl = rgam_pl;
    A.rgamh = rgamh;
    A.rgamh_pl = rgamh_pl;
    A.c2w = c2w;
    A.c2w_pl = c2w_pl;
    A.rdgz = rdgz_dev;
    A.have_sgp = c->have_sgp;
    A.iall = D.gall_1d;
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.gall_pl = D.gall_pl;
    A.have_pl = D.have_pl;
    A.nchunk = (D.gall + CHUNK - 1) / CHUNK;
    const int nlev = D.kall - 2;  // kmin..kmax
    const long base = (long)A.nchunk * D.lall;
    int nkr = (int)((2048 + base - 1) / base);
    nkr = std::max(1, std::min(nkr, nlev));
    if (c->tune.ddmp_kper > 0) nkr = (nlev + c->tune.ddmp_kper - 1) / c->tune.ddmp_kper;
    A.kper = (nlev + nkr - 1) / nkr;
    A.nkr = (nlev + A.kper - 1) / A.kper;
    const long nmain = base * A.nkr;
    const long npole = out_pl[0] ? (D.kall * kLallPl + CHUNK - 1) / CHUNK : 0;
    if (nmain + npole > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    A.nmain = (int)nmain;
    const size_t lds = sizeof(double) * (5 * (CHUNK + 2 * D.gall_1d + 2) + 2 * (CHUNK + D.gall_1d + 1));
    const bool small = CHUNK + 2 * D.gall_1d + 2 <= 2 * CHUNK;
    const void *fn = small ? (const void *)ddmp_kernel<2> : (const void *)ddmp_kernel<3>;
    if (lds > 64 * 1024) NICAM_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (small)
        hipLaunchKernelGGL(ddmp_kernel<2>, dim3((unsigned)(nmain + npole)), dim3(CHUNK), lds, c->stream, A);
    else
        hipLaunchKernelGGL(ddmp_kernel<3>, dim3((unsigned)(nmain + npole)), dim3(CHUNK), lds, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

}  // namespace nicam
