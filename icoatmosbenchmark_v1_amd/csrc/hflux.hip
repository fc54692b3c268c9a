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
// A work-group owns CHUNK consecutive cells g of one region and marches up a
// range of levels.  The K-invariant metrics of its cells (arc normals, arc
// midpoints, cell areas, triangle weights) are loaded once into registers.
// At each level the group stages rho/rhov of cells [g0-gall_1d-1,
// g0+CHUNK+gall_1d+1) in LDS, builds TI/TJ of cells [g0-gall_1d-1, g0+CHUNK)
// in LDS, and every work-item then forms its 6 arcs from LDS.  The next
// level's rho/rhov are loaded into registers while the current level is
// computed.  Floating-point operations follow the reference's order term by
// term (the library is compiled with -ffp-contract=off), so the FP64 results
// are the reference's to the last bit.
#include "internal.h"

namespace nicam {
namespace {

constexpr int CHUNK = 256;  // output cells per work-group
constexpr int MAXI = 514;   // largest gall_1d supported (gl8/rl1 regions are 130, gl9/rl1 258, gl9/rl0 514)
// Per work-item: NTP TI/TJ points, CHUNK*NTP >= CHUNK + gall_1d + 1, and NST
// staged cells, CHUNK*NST >= CHUNK + 2*(gall_1d+1).  The launcher picks the
// smallest pair that covers the region width, to spare registers: (2, 2) up to
// gall_1d = 127, (3, 2) to 255, (4, 3) to 383, (6, 4) to 514.
#ifndef NICAM_HFLUX_WT_LDS
#define NICAM_HFLUX_WT_LDS 0  // 1: triangle weights in LDS (154 VGPRs, 3 waves per SIMD) measured no faster at gl7
#endif
#ifndef NICAM_HFLUX_XR_GLOBAL
#define NICAM_HFLUX_XR_GLOBAL 0  // 1: arc midpoints read per level from L2 rather than held (no fewer registers)
#endif

template <class T>
struct HfluxArgs {
    T *flx;         // (gall,kall,lall,6)
    T *xc;          // (gall,kall,lall,3,3)
    const T *f[4];  // rho, rhovx, rhovy, rhovz (gall,kall,lall)
    const T *xr, *gp, *gt, *ga;
    const int *have_sgp;
    int iall, gall, kall, lall;
    int nchunk, nkr, kper;  // chunks per slab, level ranges, levels per range
    int nmain;              // work-groups of the regular regions; the rest do the poles
    T dt, eps;
    // pole regions (FLX:326-374)
    T *flx_pl, *xc_pl;
    const T *f_pl[4];
    const T *xr_pl, *gp_pl, *gt_pl, *ga_pl;
    int gall_pl;
};

template <class T>
__device__ __forceinline__ T dmax(T a, T b) { return (b > a) ? b : a; }

// Pole regions: one work-item per (k, l) walks the ring (FLX:326-374).  Ring
// point v takes the TI value of v-1 (of the last point for v = 1), which is
// carried from the previous step instead of stored: no per-lane array.
template <class T>
__device__ void hflux_pole(const HfluxArgs<T> &A, int t)
{
    const int kall = A.kall, npl = A.gall_pl;
    if (t >= kall * kLallPl) return;
    const int k = t % kall, l = t / kall;
    const size_t fb = ((size_t)l * kall + k) * npl;
    // GMTR_*_pl(v,1,l,c): index (c*2 + l)*gall_pl + v
    auto M = [&](const T *m, int c, int v) { return m[((size_t)c * kLallPl + l) * npl + v]; };
    auto tvp = [&](int v, T (&tv)[4]) {
        const int vp1 = (v + 1 == npl) ? 1 : v + 1;
        const T w1 = M(A.gt_pl, 2, v), w2 = M(A.gt_pl, 3, v), w3 = M(A.gt_pl, 4, v);
#pragma unroll
        for (int f = 0; f < 4; ++f)
            tv[f] = A.f_pl[f][fb + 0] * w1 + A.f_pl[f][fb + v] * w2 + A.f_pl[f][fb + vp1] * w3;
    };
    T prev[4], cur[4];
    tvp(npl - 1, prev);
    const T rarea = M(A.gp_pl, 1, 0);
    const size_t plane = (size_t)kLallPl * kall * npl;
    for (int v = 1; v < npl; ++v) {
        tvp(v, cur);
        const T rr2 = T(1) / dmax(prev[0] + cur[0], A.eps);
        const T x2 = prev[1] + cur[1];
        const T y2 = prev[2] + cur[2];
        const T z2 = prev[3] + cur[3];
        const T flux = T(0.5) * (x2 * M(A.ga_pl, 0, v) + y2 * M(A.ga_pl, 1, v) + z2 * M(A.ga_pl, 2, v));
        A.flx_pl[fb + v] = flux * rarea * A.dt;
        A.xc_pl[0 * plane + fb + v] = M(A.xr_pl, 0, v) - x2 * rr2 * A.dt * T(0.5);
        A.xc_pl[1 * plane + fb + v] = M(A.xr_pl, 1, v) - y2 * rr2 * A.dt * T(0.5);
        A.xc_pl[2 * plane + fb + v] = M(A.xr_pl, 2, v) - z2 * rr2 * A.dt * T(0.5);
#pragma unroll
        for (int f = 0; f < 4; ++f) prev[f] = cur[f];
    }
}

// Work-group b of one grid's launch (the grouped kernel of several grids
// passes the index within the grid).
template <int NST, int NTP, class T>
__device__ __forceinline__ void hflux_block(const HfluxArgs<T> &A, int b)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T *lds = reinterpret_cast<T *>(lds_raw);
    const int iall = A.iall, gall = A.gall, kall = A.kall, lall = A.lall;
    const int n_in = CHUNK + 2 * (iall + 1);
    const int n_t = CHUNK + iall + 1;
    T *s_in = lds;             // [4][n_in]
    T *s_t = lds + 4 * n_in;   // [8][n_t]: TI fields 0..3, TJ fields 4..7
#if NICAM_HFLUX_WT_LDS
    T *s_w = s_t + 8 * n_t;    // [6][n_t]: GMTR_t W1..W3 of TI, then of TJ
#endif

    // Logical block -> (region, level range, chunk); the chunks of one region
    // and level range are consecutive and are kept on one XCD (hardware
    // dispatch sends block b to XCD b % 8) so their shared halo rows hit L2.
    if (b >= A.nmain) {  // pole work-groups
        hflux_pole(A, (b - A.nmain) * CHUNK + threadIdx.x);
        return;
    }
    {
        const int nxcd = 8, per = A.nmain / nxcd;
        if (per > 0 && b < per * nxcd) b = (b % nxcd) * per + b / nxcd;
    }
    const int chunk = b % A.nchunk;
    const int kr = (b / A.nchunk) % A.nkr;
    const int l = b / (A.nchunk * A.nkr);
    if (l >= lall) return;
    const int k0 = kr * A.kper;
    const int k1 = min(kall, k0 + A.kper);
    if (k0 >= k1) return;

    const int g0 = chunk * CHUNK;
    const int lo = g0 - iall - 1;  // first staged / TI-TJ cell
    const int tid = threadIdx.x;
    const size_t slab = (size_t)gall;
    const size_t mb = (size_t)l * gall;
    const size_t mstr = (size_t)lall * gall;  // between metric planes
    const bool sgp = A.have_sgp[l] != 0;
    const T dt = A.dt, eps = A.eps;

    // ---- K-invariant metrics into registers ----
    // triangle weights of my TI/TJ points (GMTR_t(g,1,l,t,w): ((w*2+t)*lall+l)*gall+g)
#if !NICAM_HFLUX_WT_LDS
    T wt[NTP][6];
#endif
    int tp[NTP];
#pragma unroll
    for (int r = 0; r < NTP; ++r) {
        const int s = tid + r * CHUNK;
        const int p = lo + s;
        tp[r] = -1;
        if (s < n_t && p >= 0 && p < gall && (p % iall) <= iall - 2 && (p / iall) <= iall - 2) tp[r] = p;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const int w = 2 + (q % 3), t = q / 3;
            const T v = tp[r] >= 0 ? A.gt[(size_t)(w * 2 + t) * mstr + mb + p] : T(0);
#if NICAM_HFLUX_WT_LDS
            if (s < n_t) s_w[q * n_t + s] = v;  // read only by this work-item
#else
            wt[r][q] = v;
#endif
        }
    }
    const int g = g0 + tid;
    const bool active = g < gall;
    const int i = active ? g % iall : 0, j = active ? g / iall : 0;
    const int gmx = iall - 2;
    // which of the 6 arcs touching this cell exist (FLX:234-309 loop bounds)
    const bool a_ai = active && i <= gmx && j >= 1 && j <= gmx;
    const bool a_aij = active && i <= gmx && j <= gmx;
    const bool a_aj = active && i >= 1 && i <= gmx && j <= gmx;
    const bool a_f4 = active && i >= 1 && i - 1 <= gmx && j >= 1 && j <= gmx;
    const bool a_f5 = active && i >= 1 && j >= 1 && i - 1 <= gmx && j - 1 <= gmx;
    const bool a_f6 = active && i >= 1 && i <= gmx && j >= 1 && j - 1 <= gmx;
    // GMTR_a(g,1,l,a,c): ((c*3+a)*lall+l)*gall+g; GRD_xr(g,1,l,a,d): ((d*3+a)*lall+l)*gall+g
    T hn[6][3], rarea = T(0);
#if NICAM_HFLUX_XR_GLOBAL
#define XR(a, d) A.xr[(size_t)((d) * 3 + (a)) * mstr + mb + g]
#else
    T xr[3][3];
#define XR(a, d) xr[a][d]
#endif
    {
        const bool have[6] = {a_ai, a_aij, a_aj, a_f4, a_f5, a_f6};
        const int org[6] = {g, g, g, g - 1, g - 1 - iall, g - iall};
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                hn[a][c] = have[a] ? A.ga[(size_t)(c * 3 + (a % 3)) * mstr + mb + org[a]] : T(0);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
#if !NICAM_HFLUX_XR_GLOBAL
                xr[a][d] = active ? A.xr[(size_t)(d * 3 + a) * mstr + mb + g] : T(0);
#endif
            }
        if (active) rarea = A.gp[1 * mstr + mb + g];
    }

    // ---- level loop ----
    T nxt[NST][4];
    auto load_level = [&](int k) {
        const size_t fb = ((size_t)l * kall + k) * slab;
#pragma unroll
        for (int r = 0; r < NST; ++r) {
            const int s = tid + r * CHUNK;
            const int p = lo + s;
            const bool ok = s < n_in && p >= 0 && p < gall;
#pragma unroll
            for (int f = 0; f < 4; ++f) nxt[r][f] = ok ? A.f[f][fb + p] : T(0);
        }
    };
    auto store_level = [&]() {
#pragma unroll
        for (int r = 0; r < NST; ++r) {
            const int s = tid + r * CHUNK;
            if (s < n_in) {
#pragma unroll
                for (int f = 0; f < 4; ++f) s_in[f * n_in + s] = nxt[r][f];
            }
        }
    };

    load_level(k0);
    store_level();
    __syncthreads();
    if (k0 + 1 < k1) load_level(k0 + 1);

    const size_t plane = (size_t)lall * kall * gall;
    for (int k = k0; k < k1; ++k) {
        // TI/TJ of my points (FLX:146-158 then :171-183)
#pragma unroll
        for (int r = 0; r < NTP; ++r) {
            if (tp[r] < 0) continue;
            const int s = tid + r * CHUNK;
            const int s10 = s + 1, s11 = s + iall + 1, s01 = s + iall;
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const T *in = s_in + f * n_in;
#if NICAM_HFLUX_WT_LDS
                const T *w = s_w + s;
                T ti = in[s] * w[0 * n_t] + in[s10] * w[1 * n_t];
                ti = ti + in[s11] * w[2 * n_t];
                T tj = in[s] * w[3 * n_t];
                tj = tj + in[s11] * w[4 * n_t] + in[s01] * w[5 * n_t];
#else
                T ti = in[s] * wt[r][0] + in[s10] * wt[r][1];
                ti = ti + in[s11] * wt[r][2];
                T tj = in[s] * wt[r][3];
                tj = tj + in[s11] * wt[r][4] + in[s01] * wt[r][5];
#endif
                s_t[f * n_t + s] = ti;
                s_t[(4 + f) * n_t + s] = tj;
            }
            if (sgp && tp[r] == 0) {
                // FLX:188-201: TI(gmin-1,gmin-1) := TJ(gmin,gmin-1), i.e. TJ of point 1
                const T w1 = A.gt[(size_t)(2 * 2 + 1) * mstr + mb + 1];
                const T w2 = A.gt[(size_t)(3 * 2 + 1) * mstr + mb + 1];
                const T w3 = A.gt[(size_t)(4 * 2 + 1) * mstr + mb + 1];
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    const T *in = s_in + f * n_in;
                    T tj = in[s + 1] * w1;
                    tj = tj + in[s + 1 + iall + 1] * w2 + in[s + 1 + iall] * w3;
                    s_t[f * n_t + s] = tj;
                }
            }
        }
        __syncthreads();
        // the input tile is free: put the next level in it and fetch the one after
        if (k + 1 < k1) {
            store_level();
            if (k + 2 < k1) load_level(k + 2);
        }

        if (active) {
            const int s = g - lo;
            const size_t fb = ((size_t)l * kall + k) * slab + g;
            // arc a between TI/TJ points (sa, ta) and (sb, tb); ta/tb = 0 for TI, 4 for TJ
            auto arc = [&](int sa, int ta, int sb, int tb, const T (&h)[3], T &x2, T &y2, T &z2,
                           T &rr2) {
                rr2 = T(1) / dmax(s_t[ta * n_t + sa] + s_t[tb * n_t + sb], eps);
                x2 = s_t[(ta + 1) * n_t + sa] + s_t[(tb + 1) * n_t + sb];
                y2 = s_t[(ta + 2) * n_t + sa] + s_t[(tb + 2) * n_t + sb];
                z2 = s_t[(ta + 3) * n_t + sa] + s_t[(tb + 3) * n_t + sb];
                return T(0.5) * (x2 * h[0] + y2 * h[1] + z2 * h[2]);
            };
            // each own arc's face and mass centroid are stored as soon as they are formed
            auto own_arc = [&](int m, bool have, int sa, int ta, int sb, int tb) {
                T f = T(0), c0 = T(0), c1 = T(0), c2 = T(0);
                if (have) {
                    T x2, y2, z2, rr2;
                    const T flux = arc(sa, ta, sb, tb, hn[m], x2, y2, z2, rr2);
                    f = flux * rarea * dt;
                    c0 = XR(m, 0) - x2 * rr2 * dt * T(0.5);
                    c1 = XR(m, 1) - y2 * rr2 * dt * T(0.5);
                    c2 = XR(m, 2) - z2 * rr2 * dt * T(0.5);
                }
                A.flx[m * plane + fb] = f;
                A.xc[(size_t)(0 * 3 + m) * plane + fb] = c0;
                A.xc[(size_t)(1 * 3 + m) * plane + fb] = c1;
                A.xc[(size_t)(2 * 3 + m) * plane + fb] = c2;
            };
            own_arc(0, a_ai, s - iall, 4, s, 0);   // AI: TJ(ijm1) + TI(ij)   FLX:241-255
            own_arc(1, a_aij, s, 0, s, 4);         // AIJ: TI(ij) + TJ(ij)    FLX:266-280
            own_arc(2, a_aj, s, 4, s - 1, 0);      // AJ: TJ(ij) + TI(im1j)   FLX:292-306
            T x2, y2, z2, rr2;
            T f4 = T(0), f5 = T(0), f6 = T(0);
            if (a_f4) f4 = -arc(s - 1 - iall, 4, s - 1, 0, hn[3], x2, y2, z2, rr2) * rarea * dt;
            if (a_f5) f5 = -arc(s - 1 - iall, 0, s - 1 - iall, 4, hn[4], x2, y2, z2, rr2) * rarea * dt;
            if (a_f6) f6 = -arc(s - iall, 4, s - iall - 1, 0, hn[5], x2, y2, z2, rr2) * rarea * dt;
            if (sgp && i == 1 && j == 1) f6 = T(0);  // FLX:311-320
            A.flx[3 * plane + fb] = f4;
            A.flx[4 * plane + fb] = f5;
            A.flx[5 * plane + fb] = f6;
        }
        __syncthreads();
    }
#undef XR
}

template <int NST, int NTP, class T>
__global__ void __launch_bounds__(CHUNK) hflux_kernel(HfluxArgs<T> A)
{
    hflux_block<NST, NTP, T>(A, blockIdx.x);
}

// several grids in one launch: work-groups [off[n], off[n+1]) belong to grid n
template <class T>
struct HfluxGroup {
    HfluxArgs<T> a[kMaxGroup];
    int off[kMaxGroup + 1];
    int n;
};

template <int NST, int NTP, class T>
__global__ void __launch_bounds__(CHUNK) hflux_group_kernel(HfluxGroup<T> G)
{
    const int b = blockIdx.x;
    int n = 0;
    while (n + 1 < G.n && b >= G.off[n + 1]) ++n;
    hflux_block<NST, NTP, T>(G.a[n], b - G.off[n]);
}

// The (NST, NTP) pair for a region width: 0 (2, 2), 1 (3, 2), 2 (4, 3), 3 (6, 4).
inline int hflux_variant(int iall) { return iall <= 127 ? 0 : (iall <= 255 ? 1 : (iall <= 383 ? 2 : 3)); }

// One launch of variant v of the four instances k[0..3].
template <class Arg>
int hflux_launch(void (*const k[4])(Arg), int v, long nblk, size_t lds, hipStream_t stream, const Arg &A)
{
    void (*fn)(Arg) = k[v];
    if (lds > 64 * 1024)
        NICAM_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(fn, dim3((unsigned)nblk), dim3(CHUNK), lds, stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

// The launch arguments of one grid: nblk work-groups, lds bytes, nst the
// variant (hflux_variant) its width needs.
template <class T>
int hflux_make_args(nicam_ctx c, HfluxArgs<T> &A, long &nblk, size_t &lds, int &nst, T *flx_h, T *flx_h_pl, T *xc,
                    T *xc_pl, const T *rho, const T *rho_pl, const T *rvx, const T *rvx_pl, const T *rvy,
                    const T *rvy_pl, const T *rvz, const T *rvz_pl, T dt, T eps, const HMetrics<T> &M)
{
    const Dims &D = c->d;
    if (D.gall_1d > MAXI) return fail(NICAM_ERR_ARG, "gall_1d larger than the supported 514");
    A.flx = flx_h;
    A.xc = xc;
    A.f[0] = rho;
    A.f[1] = rvx;
    A.f[2] = rvy;
    A.f[3] = rvz;
    A.xr = M.xr;
    A.gp = M.gp;
    A.gt = M.gt;
    A.ga = M.ga;
    A.have_sgp = c->have_sgp;
    A.iall = D.gall_1d;
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.dt = dt;
    A.eps = eps;
    A.nchunk = (D.gall + CHUNK - 1) / CHUNK;
    // enough work-groups to fill 256 CUs several times over
    const long base = (long)A.nchunk * D.lall;
    int nkr = (int)((4096 + base - 1) / base);
    nkr = std::max(1, std::min(nkr, D.kall));
    if (c->tune.hflux_kper > 0) nkr = (D.kall + c->tune.hflux_kper - 1) / c->tune.hflux_kper;
    A.kper = (D.kall + nkr - 1) / nkr;
    A.nkr = (D.kall + A.kper - 1) / A.kper;
    const long nmain = base * A.nkr;
    const long npole = D.have_pl ? (D.kall * kLallPl + CHUNK - 1) / CHUNK : 0;
    if (nmain + npole > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    A.nmain = (int)nmain;
    A.flx_pl = flx_h_pl;
    A.xc_pl = xc_pl;
    A.f_pl[0] = rho_pl;
    A.f_pl[1] = rvx_pl;
    A.f_pl[2] = rvy_pl;
    A.f_pl[3] = rvz_pl;
    A.xr_pl = M.xr_pl;
    A.gp_pl = M.gp_pl;
    A.gt_pl = M.gt_pl;
    A.ga_pl = M.ga_pl;
    A.gall_pl = D.gall_pl;
    const size_t n_t = CHUNK + D.gall_1d + 1;
    lds = sizeof(T) * (4 * (CHUNK + 2 * (D.gall_1d + 1)) + 8 * n_t + (NICAM_HFLUX_WT_LDS ? 6 * n_t : 0));
    nst = hflux_variant(D.gall_1d);
    nblk = nmain + npole;
    return NICAM_OK;
}

template <class T>
int launch_hflux_t(nicam_ctx c, T *flx_h, T *flx_h_pl, T *xc, T *xc_pl, const T *rho, const T *rho_pl,
                   const T *rvx, const T *rvx_pl, const T *rvy, const T *rvy_pl, const T *rvz, const T *rvz_pl,
                   T dt, T eps, const HMetrics<T> &M)
{
    HfluxArgs<T> A;
    long nblk = 0;
    size_t lds = 0;
    int nst = 2;
    const int rc = hflux_make_args(c, A, nblk, lds, nst, flx_h, flx_h_pl, xc, xc_pl, rho, rho_pl, rvx, rvx_pl, rvy,
                                   rvy_pl, rvz, rvz_pl, dt, eps, M);
    if (rc != NICAM_OK) return rc;
    void (*const k[4])(HfluxArgs<T>) = {hflux_kernel<2, 2, T>, hflux_kernel<3, 2, T>, hflux_kernel<4, 3, T>,
                                         hflux_kernel<6, 4, T>};
    return hflux_launch(k, nst, nblk, lds, c->stream, A);
}

}  // namespace

int launch_hflux(nicam_ctx c, double *flx_h, double *flx_h_pl, double *xc, double *xc_pl,
                 const double *rho, const double *rho_pl, const double *rvx, const double *rvx_pl,
                 const double *rvy, const double *rvy_pl, const double *rvz, const double *rvz_pl,
                 double dt)
{
    const HMetrics<double> M{c->GRD_xr, c->GMTR_p, c->GMTR_t, c->GMTR_a,
                             c->GRD_xr_pl, c->GMTR_p_pl, c->GMTR_t_pl, c->GMTR_a_pl};
    return launch_hflux_t<double>(c, flx_h, flx_h_pl, xc, xc_pl, rho, rho_pl, rvx, rvx_pl, rvy, rvy_pl, rvz, rvz_pl,
                                  dt, c->eps, M);
}

int launch_hflux_f32(nicam_ctx c, float *flx_h, float *flx_h_pl, float *xc, float *xc_pl, const float *rho,
                     const float *rho_pl, const float *rvx, const float *rvx_pl, const float *rvy,
                     const float *rvy_pl, const float *rvz, const float *rvz_pl, float dt)
{
    // the FP32 CONST_EPS (nicam_set_const_f32; by default epsilon(0._SP), the -DSINGLE build's)
    return launch_hflux_t<float>(c, flx_h, flx_h_pl, xc, xc_pl, rho, rho_pl, rvx, rvx_pl, rvy, rvy_pl, rvz, rvz_pl,
                                 dt, c->eps32, c->hm32);
}

}  // namespace nicam

namespace nicam {

int launch_hflux_group_f32(const nicam_ctx *cs, int n, const nicam_tracer_step_f32 *g, float dt, hipStream_t stream)
{
    if (n < 1 || n > kMaxGroup) return fail(NICAM_ERR_ARG, "a group holds 1..8 grids");
    HfluxGroup<float> G;
    G.n = n;
    G.off[0] = 0;
    size_t lds_max = 0;
    int nst_max = 0;
    for (int i = 0; i < n; ++i) {
        const nicam_tracer_step_f32 &p = g[i];
        long nb = 0;
        size_t lds = 0;
        int nst = 0;
        const int rc = hflux_make_args<float>(cs[i], G.a[i], nb, lds, nst, p.flx_h, p.flx_h_pl, p.GRD_xc, p.GRD_xc_pl,
                                              p.rho, p.rho_pl, p.rhovx, p.rhovx_pl, p.rhovy, p.rhovy_pl, p.rhovz,
                                              p.rhovz_pl, dt, cs[i]->eps32, cs[i]->hm32);
        if (rc != NICAM_OK) return rc;
        if ((long)G.off[i] + nb > 0x7fffffffL) return fail(NICAM_ERR_ARG, "group too large for one launch");
        G.off[i + 1] = G.off[i] + (int)nb;
        lds_max = std::max(lds_max, lds);
        nst_max = std::max(nst_max, nst);
    }
    // the staging ring and TI/TJ points of the widest grid serve them all
    void (*const k[4])(HfluxGroup<float>) = {hflux_group_kernel<2, 2, float>, hflux_group_kernel<3, 2, float>,
                                             hflux_group_kernel<4, 3, float>, hflux_group_kernel<6, 4, float>};
    return hflux_launch(k, nst_max, G.off[n], lds_max, stream, G);
}

}  // namespace nicam
