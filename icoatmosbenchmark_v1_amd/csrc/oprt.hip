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
// of its cells) are K-invariant and are read once.  At each level the group
// stages scl and kh of the cells its stencils touch in LDS, forms vt of cells
// [g0-gall_1d-1, g0+CHUNK) in LDS, and each work-item then produces its
// cell's dscl.  The expressions keep the reference's operand order
// (including the X component that the Y and Z passes read as the first point
// of each pair), so with -ffp-contract=off the FP64 results are the
// reference's to the last bit.  The pole regions (:220-291) are extra
// work-groups of the same launch.
#include "internal.h"

namespace nicam {
namespace {

constexpr int CHUNK = 256;  // output cells per work-group
constexpr int MAXI = 255;   // largest gall_1d: the gall_1d+1 vt points below a chunk fit one slot
constexpr int NPL = 6;      // pole ring: the centre and ADM_vlink = 5 points (problem_size.inc:41)
#ifndef NICAM_DIFF_DEPTH
#define NICAM_DIFF_DEPTH 4
#endif
constexpr int DEPTH = NICAM_DIFF_DEPTH;  // levels of OPRT_diffusion input in flight per work-item

struct DiffArgs {
    double *dscl, *dscl_pl;
    const double *scl, *scl_pl, *kh, *kh_pl;
    const double *ci, *ci_pl;  // coef_intp(gall,3,3,2,lall), coef_intp_pl(gall_pl,3,3,2)
    const double *cd, *cd_pl;  // coef_diff(gall,6,3,lall), coef_diff_pl(5,3,2)
    const int *have_sgp;
    int iall, gall, kall, lall, gall_pl, have_pl;
    int nchunk, nkr, kper, nmain;
};

// The three factors of a triangle vertex's vt (:97-125), from its coef_intp
// c1..c3: the reference's expressions, evaluated once per column (they do not
// depend on the level, and each evaluation gives the same bits).
__device__ __forceinline__ void vt_factors(double c1, double c2, double c3, double &e0, double &e1, double &e2)
{
    e0 = +2.0 * c1 - 1.0 * c2 - 1.0 * c3;
    e1 = -1.0 * c1 + 2.0 * c2 - 1.0 * c3;
    e2 = -1.0 * c1 - 1.0 * c2 + 2.0 * c3;
}
// vt = (e0*s0 + e1*s1 + e2*s2) / 3
__device__ __forceinline__ double vt_value(double e0, double e1, double e2, double s0, double s1, double s2)
{
    return div3(e0 * s0 + e1 * s1 + e2 * s2);
}
__device__ __forceinline__ double vt_point(double c1, double c2, double c3, double s0, double s1, double s2)
{
    double e0, e1, e2;
    vt_factors(c1, c2, c3, e0, e1, e2);
    return vt_value(e0, e1, e2, s0, s1, s2);
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
    // ring points 1..5 around the centre 0 (the launcher checks gall_pl == NPL)
    double vt[NPL][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int v = 1; v < NPL; ++v) {
            const int vp1 = (v + 1 == NPL) ? 1 : v + 1;
            const double *c = A.ci_pl + ((size_t)(l * 3 + d) * 3) * NPL;
            vt[v][d] = vt_point(c[0 * NPL + v], c[1 * NPL + v], c[2 * NPL + v], A.scl_pl[fb + 0], A.scl_pl[fb + v],
                                A.scl_pl[fb + vp1]);
        }
    double acc = 0.0;
    const double *cd = A.cd_pl + (size_t)l * 3 * (NPL - 1);
#pragma unroll
    for (int v = 1; v < NPL; ++v) {
        const int vm1 = (v == 1) ? NPL - 1 : v - 1;
        acc = acc + (cd[0 * 5 + (v - 1)] * (vt[vm1][0] + vt[v][0]) + cd[1 * 5 + (v - 1)] * (vt[vm1][1] + vt[v][1]) +
                     cd[2 * 5 + (v - 1)] * (vt[vm1][2] + vt[v][2])) *
                        0.5 * (A.kh_pl[fb + 0] + A.kh_pl[fb + v]);
    }
    A.dscl_pl[fb + 0] = acc;
#pragma unroll
    for (int v = 1; v < NPL; ++v) A.dscl_pl[fb + v] = 0.0;
}

// A work-group owns CHUNK consecutive cells [g0, g0+CHUNK) of one region and
// a range of levels.  Window index s names point lo + s, lo = g0 - gall_1d - 1:
// the inputs are staged for s in [0, n_in), vt is formed for s in [0, n_t),
// and work-item tid owns cell g0 + tid at s = n_x + tid.  The vt points below
// the chunk (s < n_x) are formed by the first n_x work-items with their
// coefficients read from LDS; the own point's coefficients and coef_diff stay
// in registers.  The inputs of the next DEPTH levels are in flight in a
// register ring while a level is computed, and the staging tiles alternate
// between two LDS buffers, so a level needs two barriers.
template <int NST>
__global__ void __launch_bounds__(CHUNK) diffusion_kernel(DiffArgs A)
{
    extern __shared__ double lds[];
    const int iall = A.iall, gall = A.gall, kall = A.kall, lall = A.lall;
    const int n_x = iall + 1;               // vt points below the chunk
    const int n_in = CHUNK + 2 * n_x;       // staged scl/kh window
    const int n_t = CHUNK + n_x;            // vt window
    double *s_in = lds;                     // [2 buffers][scl, kh][n_in]
    double *s_vt = lds + 4 * n_in;          // [t*3+d][n_t]
    double *s_cx = s_vt + 6 * n_t;          // [(t*3+d)*3+c][n_x]: coef_intp of the points below

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
    const int g0 = chunk * CHUNK, lo = g0 - n_x;
    const int gmin = iall + 1, gmax = gall - iall - 2;  // 0-based (:85-86); vt runs over [0, gmax]
    const bool sgp = A.have_sgp[l] != 0;

    // ---- K-invariant coefficients ----
    const int g = g0 + tid, so = n_x + tid;  // own cell and its window index
    const bool own_vt = g <= gmax;
    // with a singular point, TI of point 0 is TJ of point 1 (:128-132)
    const bool vsh = g == 0 && sgp;
    double cv[2][3][3];  // vt factors e0..e2 of my TI, TJ per component
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int tt = vsh ? 1 : t, pp = vsh && t == 0 ? 1 : g;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double c[3];
#pragma unroll
            for (int q = 0; q < 3; ++q)
                c[q] = own_vt ? A.ci[((((size_t)l * 2 + tt) * 3 + d) * 3 + q) * gall + pp] : 0.0;
            vt_factors(c[0], c[1], c[2], cv[t][d][0], cv[t][d][1], cv[t][d][2]);
        }
    }
    const int px = lo + tid;  // my point below the chunk, if tid < n_x
    const bool x_vt = tid < n_x && px >= 0 && px <= gmax;
    if (tid < n_x) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double c[3], e[3];
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    c[q] = x_vt ? A.ci[((((size_t)l * 2 + t) * 3 + d) * 3 + q) * gall + px] : 0.0;
                vt_factors(c[0], c[1], c[2], e[0], e[1], e[2]);
#pragma unroll
                for (int q = 0; q < 3; ++q) s_cx[((t * 3 + d) * 3 + q) * n_x + tid] = e[q];
            }
    }
    const bool active = g < gall;
    const bool inner = g >= gmin && g <= gmax;
    double cdv[3][6];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int m = 0; m < 6; ++m) cdv[d][m] = inner ? A.cd[(((size_t)l * 3 + d) * 6 + m) * gall + g] : 0.0;

    // ---- staging: slot 0 is the own cell, slots 1.. the 2*n_x cells below and above ----
    auto slot_s = [&](int r) { return r == 0 ? so : ((r - 1) * CHUNK + tid < n_x ? (r - 1) * CHUNK + tid : (r - 1) * CHUNK + tid + CHUNK); };
    auto slot_ok = [&](int r) { return r == 0 || (r - 1) * CHUNK + tid < 2 * n_x; };
    double ring[DEPTH][NST][2];
    auto load_level = [&](double (&dst)[NST][2], int k) {
        const size_t fb = ((size_t)l * kall + k) * gall;
#pragma unroll
        for (int r = 0; r < NST; ++r) {
            const int p = lo + slot_s(r);
            const bool ok = slot_ok(r) && p >= 0 && p < gall;
            dst[r][0] = ok ? A.scl[fb + p] : 0.0;
            dst[r][1] = ok ? A.kh[fb + p] : 0.0;
        }
    };
#pragma unroll
    for (int u = 0; u < DEPTH; ++u)
        if (k0 + u < k1) load_level(ring[u], k0 + u);

    for (int kb = k0; kb < k1; kb += DEPTH) {
#pragma unroll
        for (int u = 0; u < DEPTH; ++u) {
            const int k = kb + u;
            if (k >= k1) break;
            double *s_s = s_in + (k & 1) * 2 * n_in, *s_k = s_s + n_in;
#pragma unroll
            for (int r = 0; r < NST; ++r)
                if (slot_ok(r)) {
                    s_s[slot_s(r)] = ring[u][r][0];
                    s_k[slot_s(r)] = ring[u][r][1];
                }
            if (k + DEPTH < k1) load_level(ring[u], k + DEPTH);
            __syncthreads();

            // vt (:97-125): the own point, then the point below the chunk
            if (own_vt) {
                const int sj = so + (vsh ? 1 : 0);  // TJ stencil origin of the TI slot at a singular point
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double ti =
                        vsh ? vt_value(cv[0][d][0], cv[0][d][1], cv[0][d][2], s_s[sj], s_s[sj + iall + 1], s_s[sj + iall])
                            : vt_value(cv[0][d][0], cv[0][d][1], cv[0][d][2], s_s[so], s_s[so + 1], s_s[so + iall + 1]);
                    const double tj = vt_value(cv[1][d][0], cv[1][d][1], cv[1][d][2], s_s[so], s_s[so + iall + 1],
                                               s_s[so + iall]);
                    s_vt[(0 * 3 + d) * n_t + so] = ti;
                    s_vt[(1 * 3 + d) * n_t + so] = tj;
                }
            }
            if (x_vt) {
                const int s = tid;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double *c0 = s_cx + ((0 * 3 + d) * 3) * n_x + tid, *c1 = s_cx + ((1 * 3 + d) * 3) * n_x + tid;
                    s_vt[(0 * 3 + d) * n_t + s] =
                        vt_value(c0[0], c0[n_x], c0[2 * n_x], s_s[s], s_s[s + 1], s_s[s + iall + 1]);
                    s_vt[(1 * 3 + d) * n_t + s] =
                        vt_value(c1[0], c1[n_x], c1[2 * n_x], s_s[s], s_s[s + iall + 1], s_s[s + iall]);
                }
            }
            __syncthreads();

            if (active) {
                double out = 0.0;
                if (inner) {
                    const int s0 = so, s1 = s0 - 1, s2 = s0 - iall - 1, s3 = s0 - iall;
                    const double kc = s_k[s0];
                    const double kf0 = 0.5 * (kc + s_k[s0 + iall + 1]), kf1 = 0.5 * (kc + s_k[s0 + iall]),
                                 kf2 = 0.5 * (s_k[s0 - 1] + kc), kf3 = 0.5 * (s_k[s0 - iall - 1] + kc),
                                 kf4 = 0.5 * (s_k[s0 - iall] + kc), kf5 = 0.5 * (kc + s_k[s0 + 1]);
                    auto V = [&](int t, int d, int s) { return s_vt[(t * 3 + d) * n_t + s]; };
#pragma unroll
                    for (int d = 0; d < 3; ++d) {  // :138-216, X then Y then Z
                        const double sum = kf0 * cdv[d][0] * (V(0, 0, s0) + V(1, d, s0)) +
                                           kf1 * cdv[d][1] * (V(1, 0, s0) + V(0, d, s1)) +
                                           kf2 * cdv[d][2] * (V(0, 0, s1) + V(1, d, s2)) +
                                           kf3 * cdv[d][3] * (V(1, 0, s2) + V(0, d, s2)) +
                                           kf4 * cdv[d][4] * (V(0, 0, s2) + V(1, d, s3)) +
                                           kf5 * cdv[d][5] * (V(1, 0, s3) + V(0, d, s0));
                        out = d == 0 ? sum : out + sum;
                    }
                }
                stream_store(&A.dscl[((size_t)l * kall + k) * gall + g], out);
            }
        }
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
    // ring points 1..5 around the centre 0 (the launcher checks gall_pl == NPL)
    const int kmin = 1, kmax = kall - 2;
    const size_t cstr = (size_t)kall * NPL;  // between C2WfactGz_pl planes
    const double *c2w = A.c2w_pl + (size_t)l * 6 * cstr;
    auto W = [&](int p, int kk) -> double {
        if (kk == kmin || kk == kmax + 1) return 0.0;
        const size_t f = ((size_t)l * kall + kk) * NPL + p, fm = ((size_t)l * kall + kk - 1) * NPL + p;
        return (c2w[0 * cstr + (size_t)kk * NPL + p] * A.v_pl[0][f] + c2w[1 * cstr + (size_t)kk * NPL + p] * A.v_pl[0][fm] +
                c2w[2 * cstr + (size_t)kk * NPL + p] * A.v_pl[1][f] + c2w[3 * cstr + (size_t)kk * NPL + p] * A.v_pl[1][fm] +
                c2w[4 * cstr + (size_t)kk * NPL + p] * A.v_pl[2][f] + c2w[5 * cstr + (size_t)kk * NPL + p] * A.v_pl[2][fm]) *
                   A.rgamh_pl[f] +
               A.w_pl[f] * A.rgsqrth_pl[f];
    };
    double vm[3][NPL], sp[NPL];
#pragma unroll
    for (int v = 0; v < NPL; ++v)
#pragma unroll
        for (int d = 0; d < 3; ++d) vm[d][v] = A.v_pl[d][fb + v] * A.rgam_pl[fb + v];
    const double rdgz = A.rdgz[k];
    const double w0 = W(0, k + 1), w0m = W(0, k);
    sp[0] = 0.0;
#pragma unroll
    for (int v = 1; v < NPL; ++v) {
        const int vp1 = (v + 1 == NPL) ? 1 : v + 1;
        const double sw = div3((w0 + W(v, k + 1) + W(vp1, k + 1)) - (w0m + W(v, k) + W(vp1, k))) * rdgz;
        const double *c = A.ci_pl + (size_t)l * 9 * NPL;  // coef_intp_pl(v,c,d,l): ((l*3+d)*3+c)*npl+v
        sp[v] = c[(0 * 3 + 0) * NPL + v] * vm[0][0] + c[(0 * 3 + 1) * NPL + v] * vm[0][v] +
                c[(0 * 3 + 2) * NPL + v] * vm[0][vp1] + c[(1 * 3 + 0) * NPL + v] * vm[1][0] +
                c[(1 * 3 + 1) * NPL + v] * vm[1][v] + c[(1 * 3 + 2) * NPL + v] * vm[1][vp1] +
                c[(2 * 3 + 0) * NPL + v] * vm[2][0] + c[(2 * 3 + 1) * NPL + v] * vm[2][v] +
                c[(2 * 3 + 2) * NPL + v] * vm[2][vp1] + sw;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double *cd = A.cd_pl + (size_t)l * 15 + d * 5;
        double acc = 0.0;
#pragma unroll
        for (int v = 1; v < NPL; ++v) {
            const int vm1 = (v == 1) ? NPL - 1 : v - 1;
            acc = acc + cd[v - 1] * (sp[vm1] + sp[v]);
        }
        A.out_pl[d][fb + 0] = acc;
#pragma unroll
        for (int v = 1; v < NPL; ++v) A.out_pl[d][fb + v] = 0.0;
    }
}

// A work-group owns CHUNK consecutive cells [g0, g0+CHUNK) of one region and
// marches up a range of levels; the window indexing is OPRT_diffusion's.  The
// per-level record of a window point (rhogv x/y/z, RGAM, C2WfactGz(1:6),
// RGAMH, rhogw, RGSQRTH) of the next level is in flight in registers while a
// level is computed; rhogv of the level below and rhogv*RGAM of the current
// level are carried in registers, rhogw_vm of two levels alternates in LDS.
template <int NSTT>
__global__ void __launch_bounds__(CHUNK) ddmp_kernel(DdmpArgs A)
{
    extern __shared__ double lds[];
    const int iall = A.iall, gall = A.gall, kall = A.kall, lall = A.lall;
    const int n_x = iall + 1, n_in = CHUNK + 2 * n_x, n_t = CHUNK + n_x;
    double *s_vm = lds;                  // [3][n_in] rhogv*RGAM of level k
    double *s_w = lds + 3 * n_in;        // [2][n_in] rhogw_vm of levels k and k+1 (alternating)
    double *s_st = lds + 5 * n_in;       // [2][n_t]  sclt at TI, TJ
    double *s_cx = s_st + 2 * n_t;       // [(t*3+d)*3+c][n_x]: coef_intp of the points below

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
    const int g0 = chunk * CHUNK, lo = g0 - n_x;
    const int gmin = iall + 1, gmax = gall - iall - 2;
    const bool sgp = A.have_sgp[l] != 0;
    const int g = g0 + tid, so = n_x + tid;
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
    const bool own_vt = g <= gmax;
    const bool vsh = g == 0 && sgp;
    double cv[2][3][3];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int tt = vsh ? 1 : t, pp = vsh && t == 0 ? 1 : g;
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                cv[t][d][c] = own_vt ? A.ci[((((size_t)l * 2 + tt) * 3 + d) * 3 + c) * gall + pp] : 0.0;
    }
    const int px = lo + tid;
    const bool x_vt = tid < n_x && px >= 0 && px <= gmax;
    if (tid < n_x) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    s_cx[((t * 3 + d) * 3 + c) * n_x + tid] =
                        x_vt ? A.ci[((((size_t)l * 2 + t) * 3 + d) * 3 + c) * gall + px] : 0.0;
    }
    double cdv[3][6];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int m = 0; m < 6; ++m) cdv[d][m] = inner ? A.cd[(((size_t)l * 3 + d) * 6 + m) * gall + g] : 0.0;

    // ---- window slots: 0 is the own cell, 1.. the 2*n_x cells below and above ----
    auto slot_s = [&](int r) {
        return r == 0 ? so : ((r - 1) * CHUNK + tid < n_x ? (r - 1) * CHUNK + tid : (r - 1) * CHUNK + tid + CHUNK);
    };
    auto slot_ok = [&](int r) { return r == 0 || (r - 1) * CHUNK + tid < 2 * n_x; };
    auto slot_p = [&](int r, int &p) {
        p = lo + slot_s(r);
        return slot_ok(r) && p >= 0 && p < gall;
    };
    // record of level kk: v[3], rgam, c2w[6], rgamh, w, rgsqrth
    double nx[NSTT][13];
    auto load_rec = [&](int kk) {
#pragma unroll
        for (int r = 0; r < NSTT; ++r) {
            int p;
            const bool ok = slot_p(r, p) && kk <= kall - 1;
            const size_t f = (lk + kk) * slab + p, fk = (size_t)kk * slab + p;
#pragma unroll
            for (int d = 0; d < 3; ++d) nx[r][d] = ok ? A.v[d][f] : 0.0;
            nx[r][3] = ok ? A.rgam[f] : 0.0;
#pragma unroll
            for (int j = 0; j < 6; ++j) nx[r][4 + j] = ok ? c2w[j * cstr + fk] : 0.0;
            nx[r][10] = ok ? A.rgamh[f] : 0.0;
            nx[r][11] = ok ? A.w[f] : 0.0;
            nx[r][12] = ok ? A.rgsqrth[f] : 0.0;
        }
    };
    // rhogw_vm of a level from its record and rhogv of the level below (:148-156)
    auto wvm = [&](const double *rec, const double *vb) {
        return (rec[4] * rec[0] + rec[5] * vb[0] + rec[6] * rec[1] + rec[7] * vb[1] + rec[8] * rec[2] +
                rec[9] * vb[2]) *
                   rec[10] +
               rec[11] * rec[12];
    };
    // prologue: rhogv(k0-1) for rhogw_vm(k0); record of k0 gives rhogw_vm(k0) and rhogv*RGAM(k0)
    double vb[NSTT][3], vmc[NSTT][3];
    int cur = 0;
#pragma unroll
    for (int r = 0; r < NSTT; ++r) {
        int p;
        const bool ok = slot_p(r, p);
        const size_t f = (lk + k0 - 1) * slab + p;
#pragma unroll
        for (int d = 0; d < 3; ++d) vb[r][d] = ok ? A.v[d][f] : 0.0;
    }
    load_rec(k0);
#pragma unroll
    for (int r = 0; r < NSTT; ++r) {
        if (!slot_ok(r)) continue;
        int p;
        const bool ok = slot_p(r, p);
        s_w[cur * n_in + slot_s(r)] = (ok && k0 > kmin) ? wvm(nx[r], vb[r]) : 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            vmc[r][d] = nx[r][d] * nx[r][3];
            vb[r][d] = nx[r][d];
        }
    }
    load_rec(k0 + 1);

    for (int k = k0; k < k1; ++k) {
        const int nxt = cur ^ 1;
        // rhogw_vm(k+1) from the record of k+1, and rhogv*RGAM of level k
#pragma unroll
        for (int r = 0; r < NSTT; ++r) {
            if (!slot_ok(r)) continue;
            int p;
            const bool ok = slot_p(r, p);
            const int s = slot_s(r);
            s_w[nxt * n_in + s] = (ok && k + 1 <= kmax) ? wvm(nx[r], vb[r]) : 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                s_vm[d * n_in + s] = vmc[r][d];
                vmc[r][d] = nx[r][d] * nx[r][3];
                vb[r][d] = nx[r][d];
            }
        }
        if (k + 1 < k1) load_rec(k + 2);
        __syncthreads();
        // sclt of my vertices (:183-225)
        const double rdgz = A.rdgz[k];
        const double *w0 = s_w + cur * n_in, *w1 = s_w + nxt * n_in;
        const double *vx = s_vm, *vy = s_vm + n_in, *vz = s_vm + 2 * n_in;
        if (own_vt) {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                // TI: (ij, ip1j, ip1jp1); TJ: (ij, ip1jp1, ijp1); the singular TI slot is TJ of point 1
                const bool tj = t == 1 || vsh;
                const int a = tj && t == 0 ? so + 1 : so;
                const int b2 = tj ? a + iall + 1 : a + 1;
                const int c3 = tj ? a + iall : a + iall + 1;
                const double sw = div3((w1[a] + w1[b2] + w1[c3]) - (w0[a] + w0[b2] + w0[c3])) * rdgz;
                s_st[t * n_t + so] = cv[t][0][0] * vx[a] + cv[t][0][1] * vx[b2] + cv[t][0][2] * vx[c3] +
                                     cv[t][1][0] * vy[a] + cv[t][1][1] * vy[b2] + cv[t][1][2] * vy[c3] +
                                     cv[t][2][0] * vz[a] + cv[t][2][1] * vz[b2] + cv[t][2][2] * vz[c3] + sw;
            }
        }
        if (x_vt) {
            const int a0 = tid;
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int a = a0, b2 = t == 1 ? a + iall + 1 : a + 1, c3 = t == 1 ? a + iall : a + iall + 1;
                const double sw = div3((w1[a] + w1[b2] + w1[c3]) - (w0[a] + w0[b2] + w0[c3])) * rdgz;
                const double *c = s_cx + (t * 9) * n_x + tid;  // [(t*3+d)*3+c][n_x]
                s_st[t * n_t + a0] = c[0 * n_x] * vx[a] + c[1 * n_x] * vx[b2] + c[2 * n_x] * vx[c3] +
                                     c[3 * n_x] * vy[a] + c[4 * n_x] * vy[b2] + c[5 * n_x] * vy[c3] +
                                     c[6 * n_x] * vz[a] + c[7 * n_x] * vz[b2] + c[8 * n_x] * vz[c3] + sw;
            }
        }
        __syncthreads();
        if (active) {
            const size_t fo = (lk + k) * slab + g;
            if (inner) {
                const int s0 = so, s1 = s0 - 1, s2 = s0 - iall - 1, s3 = s0 - iall;
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
    if (dscl_pl && D.have_pl && D.gall_pl != NPL)
        return fail(NICAM_ERR_ARG, "OPRT_diffusion's pole regions need gall_pl = 6 (ADM_vlink = 5)");
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
    // about 1024 work-groups: longer level marches read the coefficients fewer times (gl7: two ranges
    // of 49 levels 290 us, three of 33 298 us, six of 17 337 us)
    int nkr = (int)((1024 + base - 1) / base);
    nkr = std::max(1, std::min(nkr, D.kall));
    if (c->tune.diff_kper > 0) nkr = (D.kall + c->tune.diff_kper - 1) / c->tune.diff_kper;
    A.kper = (D.kall + nkr - 1) / nkr;
    A.nkr = (D.kall + A.kper - 1) / A.kper;
    const long nmain = base * A.nkr;
    const long npole = dscl_pl ? (D.kall * kLallPl + CHUNK - 1) / CHUNK : 0;
    if (nmain + npole > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    A.nmain = (int)nmain;
    const int n_x = D.gall_1d + 1;
    const size_t lds = sizeof(double) * (4 * (CHUNK + 2 * n_x) + 6 * (CHUNK + n_x) + 18 * n_x);
    const bool two = 2 * n_x <= CHUNK;  // the cells below and above the chunk fit one extra slot
    const void *kern = two ? (const void *)diffusion_kernel<2> : (const void *)diffusion_kernel<3>;
    if (lds > 64 * 1024)
        NICAM_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (two)
        hipLaunchKernelGGL(diffusion_kernel<2>, dim3((unsigned)(nmain + npole)), dim3(CHUNK), lds, c->stream, A);
    else
        hipLaunchKernelGGL(diffusion_kernel<3>, dim3((unsigned)(nmain + npole)), dim3(CHUNK), lds, c->stream, A);
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
    if (out_pl[0] && D.have_pl && D.gall_pl != NPL)
        return fail(NICAM_ERR_ARG, "OPRT3D_divdamp's pole regions need gall_pl = 6 (ADM_vlink = 5)");
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
    // about 1024 work-groups (gl7: two ranges of 48 levels 566 us, three of 32 603 us)
    int nkr = (int)((1024 + base - 1) / base);
    nkr = std::max(1, std::min(nkr, nlev));
    if (c->tune.ddmp_kper > 0) nkr = (nlev + c->tune.ddmp_kper - 1) / c->tune.ddmp_kper;
    A.kper = (nlev + nkr - 1) / nkr;
    A.nkr = (nlev + A.kper - 1) / A.kper;
    const long nmain = base * A.nkr;
    const long npole = out_pl[0] ? (D.kall * kLallPl + CHUNK - 1) / CHUNK : 0;
    if (nmain + npole > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    A.nmain = (int)nmain;
    const int n_x = D.gall_1d + 1;
    const size_t lds = sizeof(double) * (5 * (CHUNK + 2 * n_x) + 2 * (CHUNK + n_x) + 18 * n_x);
    const bool small = 2 * n_x <= CHUNK;  // the cells below and above the chunk fit one extra slot
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
