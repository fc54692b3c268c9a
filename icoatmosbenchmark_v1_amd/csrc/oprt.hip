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

}  // namespace nicam
