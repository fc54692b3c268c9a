// horizontal_limiter_thuburn on gfx950.
//
// Reference: kernels/NICAM/dyn_horiz_adv_limiter/src/mod_src_tracer.f90:42-430.
//
// The reference stores Qin(gall,kall,lall,2,6) (the inflow bounds of the six
// faces of every cell) and Qout(gall,kall,lall,2) in automatic arrays.  Qin of
// a face is a function of the 4 tracer values around its arc and of that
// arc's cmask, so it is recomputed here where it is used instead of being
// stored: one work-item per cell computes the min/max of its 3 own arcs and of
// the 3 arcs it shares with its lower-left neighbours from a tracer tile held
// in LDS, derives Qout of the cell (phase 1, written to Qout_prev) and clamps
// the face values of its own 3 arcs into the inflow bounds and the
// post-exchange range Qout_post (phase 2), writing both the arc's own face and
// the neighbour's mirrored face.  With Qout_post given, the two phases are
// independent, so one launch does both.
//
// Many tracers share d, ch and cmask (NICAM keeps its tracers as the last
// dimension, q(gall,kall,lall,ntr)).  A work-item loads its cell's share of
// those fields once, with the tracer-independent sums Cin, Cout and the
// zerosw switch, and then limits every tracer in turn, restaging the tracer
// tile in LDS for each.  One tracer is the same kernel with ntr = 1, and
// every tracer's arithmetic is the single-tracer kernel's, in its order.
#include "internal.h"

namespace nicam {
namespace {

#ifndef NICAM_HLIM_CHUNK
#define NICAM_HLIM_CHUNK 512  // 256: 678 us, 512: 659 us, 1024: 684 us at gl7
#endif
#ifndef NICAM_HLIM_NTLOAD
#define NICAM_HLIM_NTLOAD 0
#endif
#ifndef NICAM_HLIM_STORE
#define NICAM_HLIM_STORE 0  // 0: nontemporal stores (668 us at gl7), 1: plain mirrored-face stores (697), 2: plain throughout (710)
#endif
constexpr int CHUNK = NICAM_HLIM_CHUNK;  // cells (work-items) per work-group
constexpr int MAXI_TILE = 160;  // widest region of the common instances' q tile (gl8/rl1 regions are 130)
constexpr int MAXI = 514;       // widest region of all (gl9/rl0, gl10/rl1), with the wide instances' tile

// A load of a streamed input read once (optionally with the nontemporal hint).
template <class T>
__device__ __forceinline__ T ld_once(const T *p)
{
#if NICAM_HLIM_NTLOAD
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}

template <class T>
__device__ __forceinline__ T dmin(T a, T b) { return (b < a) ? b : a; }
template <class T>
__device__ __forceinline__ T dmax(T a, T b) { return (b > a) ? b : a; }
// Fortran SIGN(0.5, x): 0.5 with the sign bit of x
__device__ __forceinline__ double sign_half(double x) { return copysign(0.5, x); }
__device__ __forceinline__ float sign_half(float x) { return copysignf(0.5f, x); }
__device__ __forceinline__ double tabs(double x) { return fabs(x); }
__device__ __forceinline__ float tabs(float x) { return fabsf(x); }

template <class T>
struct HlimArgs {
    T *q_a;          // (gall,kall,lall,6) inout
    const T *q, *d;  // (gall,kall,lall)
    const T *ch, *cmask;  // (gall,kall,lall,6)
    T *Qprev;        // (gall,kall,lall,2)
    const T *Qpost;  // (gall,kall,lall,2)
    const int *have_sgp;
    int iall, gall, kall, lall;
    int phases;  // 1: allowable range only, 2: clamp only, 3: both
    int ntr;     // tracers; tracer t's fields start t*ts (regular) and t*tps (pole) elements on,
    long ts, tps;  // times 6 for q_a and 2 for Qout
    int nmain;   // work-groups of the regular regions; the rest do the poles
    T eps, big;
    // pole regions (HLIM:307-363, 367-373, 412-425)
    T *q_a_pl, *Qprev_pl;
    const T *q_pl, *d_pl, *ch_pl, *cmask_pl, *Qpost_pl;
    int gall_pl, have_pl;
    // branch mask (masked variant): uint8 per cell-level, bit m set where face
    // m+1 of an own arc changed; uint8 per pole point, bit 0 for ring faces
    uint8_t *mask, *mask_pl;
};

template <class T>
__device__ void hlim_pole(const HlimArgs<T> &A, int t, int tr);

// MULTI: the tracer loop with the shared fields loaded once before it; the
// single-tracer instance loads them where the tracer's own work needs them
// (fewer registers live across the tile's barrier), unless PRE, which issues
// every streamed load (ch, cmask, d, Qout_post, q_a) before the barrier so
// that their latency overlaps the tile's.
// Work-group b of one grid's launch (the single-grid kernel passes its block
// index; the grouped kernel of several grids passes the index within the grid).
// TW: the widest region the q tile holds.
template <class T, bool MULTI, bool PRE = false, int TW = MAXI_TILE>
__device__ __forceinline__ void hlim_block(const HlimArgs<T> &A, int b)
{
    // q for cells [g0 - iall - 1, g0 + CHUNK + iall + 2)
    // MULTI keeps two tiles: tracer t+1's tile is loaded into registers while
    // tracer t is limited, and written to the other tile after it
    constexpr int SPAN = CHUNK + 2 * TW + 3;
    constexpr int NLD = (SPAN + CHUNK - 1) / CHUNK;
    __shared__ T s_q[MULTI ? 2 : 1][SPAN];

    const int iall = A.iall, gall = A.gall, kall = A.kall, lall = A.lall;
    const int nchunk = (gall + CHUNK - 1) / CHUNK;
    if (b >= A.nmain) {  // pole work-groups
        for (int tr = 0; tr < (MULTI ? A.ntr : 1); ++tr) hlim_pole(A, (b - A.nmain) * CHUNK + threadIdx.x, tr);
        return;
    }
    {
        const int nxcd = 8, per = A.nmain / nxcd;
        if (per > 0 && b < per * nxcd) b = (b % nxcd) * per + b / nxcd;
    }
    const int chunk = b % nchunk;
    const int kl = b / nchunk;
    const int k = kl % kall, l = kl / kall;
    if (l >= lall) return;
    const int g0 = chunk * CHUNK;
    const int lo = g0 - iall - 1;
    const int n_in = CHUNK + 2 * iall + 3;
    const size_t fb = ((size_t)l * kall + k) * gall;
    const int tid = threadIdx.x;
    const int g = g0 + tid;
    const bool active = g < gall;
    const int i = g % iall, j = g / iall;
    const bool sgp = A.have_sgp[l] != 0;
    const size_t plane = (size_t)lall * kall * gall;
    const T BIG = A.big;
    const bool own = active && i <= iall - 2 && j <= iall - 2;  // own arcs, origin g
    const bool inner = active && i >= 1 && i <= iall - 2 && j >= 1 && j <= iall - 2;
    const int org[3] = {g - 1, g - 1 - iall, g - iall};  // arcs at (i-1,j), (i-1,j-1), (i,j-1)

    // ---- what every tracer shares: masks, Courant sums, zerosw, d ----
    T ocm[3] = {T(0), T(0), T(0)}, ncm[3] = {T(0), T(0), T(0)}, chm[6];
    T Cin_sum = T(0), Cout_sum = T(0), zerosw = T(0), dc = T(0);
    if (own) {
#pragma unroll
        for (int m = 0; m < 3; ++m) ocm[m] = A.cmask[(size_t)m * plane + fb + g];  // re-read by the neighbours
    }
    const bool p1 = (A.phases & 1) && inner;
    auto load_shared = [&]() {
        if (!p1) return;
#pragma unroll
        for (int m = 0; m < 3; ++m) ncm[m] = A.cmask[(size_t)m * plane + fb + org[m]];
        T chv[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            chv[f] = ld_once(&A.ch[(size_t)f * plane + fb + g]);
            chm[f] = dmin(chv[f], T(0));
        }
        Cin_sum = chm[0] + chm[1] + chm[2] + chm[3] + chm[4] + chm[5];
        Cout_sum = chv[0] - chm[0] + chv[1] - chm[1] + chv[2] - chm[2] + chv[3] - chm[3] + chv[4] - chm[4] +
                   chv[5] - chm[5];
        zerosw = T(0.5) - sign_half(tabs(Cout_sum) - A.eps);
        dc = ld_once(&A.d[fb + g]);
    };
    if constexpr (MULTI || PRE) load_shared();
    // PRE: the clamp's inputs too (one tracer); MULTI: tracer tr+1's clamp inputs are loaded with its
    // tile, while tracer tr is limited
    constexpr bool PF = MULTI || PRE;
    T pre_qa[3] = {T(0), T(0), T(0)}, pre_qpn[3][2] = {}, pre_qpc[2] = {T(0), T(0)};
    auto load_clamp = [&](int tr, T (&qa)[3], T (&qpn)[3][2], T (&qpc)[2]) {
        if (own && (A.phases & 2)) {
            const int nb[3] = {g + 1, g + iall + 1, g + iall};
            const T *Qp = A.Qpost + tr * 2 * A.ts;
            const T *qa_t = A.q_a + tr * 6 * A.ts;
            qpc[0] = Qp[fb + g];
            qpc[1] = Qp[plane + fb + g];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                qpn[m][0] = Qp[fb + nb[m]];
                qpn[m][1] = Qp[plane + fb + nb[m]];
                qa[m] = ld_once(&qa_t[(size_t)m * plane + fb + g]);
            }
        }
    };
    if constexpr (PF) load_clamp(0, pre_qa, pre_qpn, pre_qpc);
    T nx_qa[3] = {T(0), T(0), T(0)}, nx_qpn[3][2] = {}, nx_qpc[2] = {T(0), T(0)};

    T nxt[NLD];  // the next tracer's share of its tile
    for (int tr = 0; tr < (MULTI ? A.ntr : 1); ++tr) {
        if (MULTI && tr > 0) {  // this tracer's clamp inputs, loaded during the previous one
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                pre_qa[m] = nx_qa[m];
                pre_qpn[m][0] = nx_qpn[m][0];
                pre_qpn[m][1] = nx_qpn[m][1];
            }
            pre_qpc[0] = nx_qpc[0];
            pre_qpc[1] = nx_qpc[1];
        }
        T *q_a = A.q_a + tr * 6 * A.ts;
        T *Qprev = A.Qprev + tr * 2 * A.ts;
        const T *Qpost = A.Qpost + tr * 2 * A.ts;
        T *sq = s_q[MULTI ? (tr & 1) : 0];
        // tile tr-2, the last one read from sq, was read before the previous barrier
        if (tr == 0) {
            const T *q = A.q;
            for (int s = tid; s < n_in; s += CHUNK) {
                const int gg = lo + s;
                sq[s] = (gg >= 0 && gg < gall) ? q[fb + gg] : T(0);
            }
        } else {
#pragma unroll
            for (int u = 0; u < NLD; ++u)
                if (tid + u * CHUNK < n_in) sq[tid + u * CHUNK] = nxt[u];
        }
        __syncthreads();
        if (MULTI && tr + 1 < A.ntr) {
            const T *qn = A.q + (tr + 1) * A.ts;
#pragma unroll
            for (int u = 0; u < NLD; ++u) {
                const int gg = lo + tid + u * CHUNK;
                nxt[u] = (tid + u * CHUNK < n_in && gg >= 0 && gg < gall) ? qn[fb + gg] : T(0);
            }
            load_clamp(tr + 1, nx_qa, nx_qpn, nx_qpc);
        }
        if (!active) continue;

        // tracer at absolute cell index p (must lie in the staged window)
        auto Q = [&](int p) { return sq[p - lo]; };

        // min/max of the 4 tracer values around arc m (0=AI,1=AIJ,2=AJ) with
        // origin p; HLIM:145-160 with the index clamps of :152-153 and the
        // singular-point stencil of :197-203.
        auto arc_minmax = [&](int m, int p, T &qmn, T &qmx) {
            T a, b2, c2, d2;
            if (m == 0) {
                const int ijm1 = max(p - iall, 0);
                a = Q(p); b2 = Q(ijm1); c2 = Q(p + 1); d2 = Q(p + iall + 1);
            } else if (m == 1) {
                if (sgp && p == 0) {
                    a = Q(p); b2 = Q(p + iall + 1); c2 = Q(p + iall + 2); d2 = Q(p + iall);
                } else {
                    a = Q(p); b2 = Q(p + 1); c2 = Q(p + iall + 1); d2 = Q(p + iall);
                }
            } else {
                const int im1j = max(p - 1, 0);
                a = Q(p); b2 = Q(p + iall + 1); c2 = Q(p + iall); d2 = Q(im1j);
            }
            qmn = dmin(dmin(dmin(a, b2), c2), d2);
            qmx = dmax(dmax(dmax(a, b2), c2), d2);
        };

        T omn[3], omx[3];
        if (own) {
#pragma unroll
            for (int m = 0; m < 3; ++m) arc_minmax(m, g, omn[m], omx[m]);
        }

        // ---- phase 1: allowable range of the cell (HLIM:216-301) ----
        T qout_min, qout_max;
        const T qc = Q(g);
        if constexpr (!MULTI && !PRE) load_shared();
        if (p1) {
            T qinmn[6], qinmx[6];
#pragma unroll
            for (int m = 0; m < 3; ++m) {  // faces 1..3: own side of own arcs (:162-187)
                const T c = ocm[m];
                qinmn[m] = (c)*omn[m] + (T(1) - c) * BIG;
                qinmx[m] = (c)*omx[m] + (T(1) - c) * (-BIG);
            }
            // faces 4..6: neighbour side of the arcs at (i-1,j), (i-1,j-1), (i,j-1)
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                T mn, mx;
                arc_minmax(m, org[m], mn, mx);
                const T c = ncm[m];
                qinmn[m + 3] = (c)*BIG + (T(1) - c) * mn;
                qinmx[m + 3] = (c) * (-BIG) + (T(1) - c) * mx;
            }
            T qnext_min = qc, qnext_max = qc;
#pragma unroll
            for (int f = 0; f < 6; ++f) qnext_min = dmin(qnext_min, qinmn[f]);
#pragma unroll
            for (int f = 0; f < 6; ++f) qnext_max = dmax(qnext_max, qinmx[f]);
            const T CQin_min_sum = chm[0] * qinmn[0] + chm[1] * qinmn[1] + chm[2] * qinmn[2] +
                                   chm[3] * qinmn[3] + chm[4] * qinmn[4] + chm[5] * qinmn[5];
            const T CQin_max_sum = chm[0] * qinmx[0] + chm[1] * qinmx[1] + chm[2] * qinmx[2] +
                                   chm[3] * qinmx[3] + chm[4] * qinmx[4] + chm[5] * qinmx[5];
            qout_min = (qc - CQin_max_sum - qnext_max * (T(1) - Cin_sum - Cout_sum + dc)) / (Cout_sum + zerosw) *
                           (T(1) - zerosw) + qc * zerosw;
            qout_max = (qc - CQin_min_sum - qnext_min * (T(1) - Cin_sum - Cout_sum + dc)) / (Cout_sum + zerosw) *
                           (T(1) - zerosw) + qc * zerosw;
        } else {  // halo (:286-301)
            qout_min = qc;
            qout_max = qc;
        }
        if (A.phases & 1) {
#if NICAM_HLIM_STORE == 2
            Qprev[fb + g] = qout_min;
            Qprev[plane + fb + g] = qout_max;
#else
            stream_store(&Qprev[fb + g], qout_min);
            stream_store(&Qprev[plane + fb + g], qout_max);
#endif
        }

        // ---- phase 2: clamp the face values of the own arcs (HLIM:377-410) ----
        if (A.mask && (A.phases & 2) && !own) A.mask[fb + g] = 0;
        if (own && (A.phases & 2)) {
            uint32_t bits = 0;
            const int nb[3] = {g + 1, g + iall + 1, g + iall};
            const T qo_min_c = PF ? pre_qpc[0] : Qpost[fb + g];  // re-read by neighbours
            const T qo_max_c = PF ? pre_qpc[1] : Qpost[plane + fb + g];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const T c = ocm[m];
                const T in_mn_s = (c)*omn[m] + (T(1) - c) * BIG;
                const T in_mx_s = (c)*omx[m] + (T(1) - c) * (-BIG);
                const T in_mn_n = (c)*BIG + (T(1) - c) * omn[m];
                const T in_mx_n = (c) * (-BIG) + (T(1) - c) * omx[m];
                const T qo_min_n = PF ? pre_qpn[m][0] : Qpost[fb + nb[m]];
                const T qo_max_n = PF ? pre_qpn[m][1] : Qpost[plane + fb + nb[m]];
                T qa = PF ? pre_qa[m] : ld_once(&q_a[(size_t)m * plane + fb + g]);
                const T qa_in = qa;
                qa = (c)*dmin(dmax(qa, in_mn_s), in_mx_s) + (T(1) - c) * dmin(dmax(qa, in_mn_n), in_mx_n);
                qa = (c)*dmax(dmin(qa, qo_max_n), qo_min_n) + (T(1) - c) * dmax(dmin(qa, qo_max_c), qo_min_c);
                bits |= (qa != qa_in ? 1u : 0u) << m;
#if NICAM_HLIM_STORE == 2
                q_a[(size_t)m * plane + fb + g] = qa;
#else
                stream_store(&q_a[(size_t)m * plane + fb + g], qa);
#endif
#if NICAM_HLIM_STORE >= 1
                q_a[(size_t)(m + 3) * plane + fb + nb[m]] = qa;  // one element off the wave's alignment
#else
                stream_store(&q_a[(size_t)(m + 3) * plane + fb + nb[m]], qa);
#endif
            }
            if (A.mask) A.mask[fb + g] = (uint8_t)bits;
        }
    }
}

template <class T, bool MULTI, bool PRE = false, int TW = MAXI_TILE>
__global__ void __launch_bounds__(CHUNK) hlim_kernel(HlimArgs<T> A)
{
    hlim_block<T, MULTI, PRE, TW>(A, blockIdx.x);
}

// Several grids in one launch: work-groups [off[n], off[n+1]) belong to grid n.
template <class T>
struct HlimGroup {
    HlimArgs<T> a[kMaxGroup];
    int off[kMaxGroup + 1];
    int n;
};

template <class T>
__global__ void __launch_bounds__(CHUNK) hlim_group_kernel(HlimGroup<T> G)
{
    const int b = blockIdx.x;
    int n = 0;
    while (n + 1 < G.n && b >= G.off[n + 1]) ++n;
    hlim_block<T, true>(G.a[n], b - G.off[n]);
}

// Pole regions: one work-item per (k, l) (HLIM:307-363, 367-373, 412-425).
// Qin of ring point v depends only on q around it and its cmask, so it is
// recomputed in each of the three passes over the ring instead of stored.
template <class T>
__device__ void hlim_pole(const HlimArgs<T> &A0, int t, int tr)
{
    const int kall = A0.kall, npl = A0.gall_pl;
    if (t >= kall * kLallPl) return;
    // tracer tr's pole fields
    HlimArgs<T> A = A0;
    A.q_pl = A0.q_pl + tr * A0.tps;
    A.q_a_pl = A0.q_a_pl ? A0.q_a_pl + tr * A0.tps : nullptr;
    A.Qprev_pl = A0.Qprev_pl + tr * 2 * A0.tps;
    A.Qpost_pl = A0.Qpost_pl ? A0.Qpost_pl + tr * 2 * A0.tps : nullptr;
    const int k = t % kall, l = t / kall;
    const size_t fb = ((size_t)l * kall + k) * npl;
    const size_t plane = (size_t)kLallPl * kall * npl;
    if (!A.have_pl) {  // only the ring zeroing of :367-370 happens
        if (!(A.phases & 1)) return;
        for (int v = 1; v < npl; ++v) {
            A.Qprev_pl[fb + v] = T(0);
            A.Qprev_pl[plane + fb + v] = T(0);
        }
        return;
    }
    const T BIG = A.big;
    // [min,max][side 1,2] of ring point v
    auto qin = [&](int v, T &mn1, T &mn2, T &mx1, T &mx2) {
        int vp1 = v + 1, vm1 = v - 1;
        if (vp1 == npl) vp1 = 1;
        if (vm1 == 0) vm1 = npl - 1;
        const T a = A.q_pl[fb + 0], b = A.q_pl[fb + v], c2 = A.q_pl[fb + vm1], d2 = A.q_pl[fb + vp1];
        const T mn = dmin(dmin(dmin(a, b), c2), d2);
        const T mx = dmax(dmax(dmax(a, b), c2), d2);
        const T c = A.cmask_pl[fb + v];
        mn1 = (c)*mn + (T(1) - c) * BIG;
        mn2 = (c)*BIG + (T(1) - c) * mn;
        mx1 = (c)*mx + (T(1) - c) * (-BIG);
        mx2 = (c) * (-BIG) + (T(1) - c) * mx;
    };
    const T qn = A.q_pl[fb + 0];
    T qnext_min = qn, qnext_max = qn;
    for (int v = 1; v < npl; ++v) {
        T mn1, mn2, mx1, mx2;
        qin(v, mn1, mn2, mx1, mx2);
        qnext_min = dmin(qnext_min, mn1);
        qnext_max = dmax(qnext_max, mx1);
    }
    T Cin = T(0), Cout = T(0), CQmax = T(0), CQmin = T(0);
    for (int v = 1; v < npl; ++v) {
        T mn1, mn2, mx1, mx2;
        qin(v, mn1, mn2, mx1, mx2);
        const T chm = A.cmask_pl[fb + v] * A.ch_pl[fb + v];
        Cin = Cin + chm;
        Cout = Cout - chm + A.ch_pl[fb + v];
        CQmin = CQmin + chm * mn1;
        CQmax = CQmax + chm * mx1;
    }
    const T zerosw = T(0.5) - sign_half(tabs(Cout) - A.eps);
    const T dn = A.d_pl[fb + 0];
    const T qo_min = (qn - CQmax - qnext_max * (T(1) - Cin - Cout + dn)) / (Cout + zerosw) * (T(1) - zerosw) + qn * zerosw;
    const T qo_max = (qn - CQmin - qnext_min * (T(1) - Cin - Cout + dn)) / (Cout + zerosw) * (T(1) - zerosw) + qn * zerosw;
    if (A.phases & 1) {
        A.Qprev_pl[fb + 0] = qo_min;
        A.Qprev_pl[plane + fb + 0] = qo_max;
        for (int v = 1; v < npl; ++v) {
            A.Qprev_pl[fb + v] = T(0);
            A.Qprev_pl[plane + fb + v] = T(0);
        }
    }
    if (!(A.phases & 2)) return;
    const T pn_min = A.Qpost_pl[fb + 0], pn_max = A.Qpost_pl[plane + fb + 0];
    if (A.mask_pl) A.mask_pl[fb + 0] = 0;
    for (int v = 1; v < npl; ++v) {
        T mn1, mn2, mx1, mx2;
        qin(v, mn1, mn2, mx1, mx2);
        const T c = A.cmask_pl[fb + v];
        T qa = A.q_a_pl[fb + v];
        const T qa_in = qa;
        qa = (c)*dmin(dmax(qa, mn1), mx1) + (T(1) - c) * dmin(dmax(qa, mn2), mx2);
        qa = (c)*dmax(dmin(qa, A.Qpost_pl[plane + fb + v]), A.Qpost_pl[fb + v]) + (T(1) - c) * dmax(dmin(qa, pn_max), pn_min);
        A.q_a_pl[fb + v] = qa;
        if (A.mask_pl) A.mask_pl[fb + v] = (qa != qa_in) ? 1 : 0;
    }
}

}  // namespace

namespace {

int halo_fill(nicam_ctx c, double *v, double *v_pl, int kmax, int nvar)
{
    return launch_halo_fill(c, v, v_pl, kmax, nvar, false);
}
int halo_fill(nicam_ctx c, float *v, float *v_pl, int kmax, int nvar)
{
    return launch_halo_fill_f32(c, v, v_pl, kmax, nvar, false);
}

// The launch arguments of one grid (nblocks: work-groups of one launch).
template <class T>
int hlim_make_args(nicam_ctx c, HlimArgs<T> &A, long &nblocks, T *q_a, T *q_a_pl, const T *q, const T *q_pl,
                   const T *d, const T *d_pl, const T *ch, const T *ch_pl, const T *cmask, const T *cmask_pl,
                   T *Qprev, T *Qprev_pl, const T *Qpost, const T *Qpost_pl, T eps, T big, uint8_t *mask,
                   uint8_t *mask_pl, int ntr)
{
    const Dims &D = c->d;
    if (D.gall_1d > MAXI) return fail(NICAM_ERR_ARG, "gall_1d larger than the LDS tiles (514)");
    if (ntr < 1 || (mask && ntr != 1)) return fail(NICAM_ERR_ARG, "tracer count must be >= 1 (1 with a mask)");
    A.ntr = ntr;
    A.ts = (long)D.nreg();
    A.tps = (long)D.npl();
    A.mask = mask;
    A.mask_pl = mask_pl;
    A.q_a = q_a;
    A.q = q;
    A.d = d;
    A.ch = ch;
    A.cmask = cmask;
    A.Qprev = Qprev;
    A.Qpost = Qpost;
    A.have_sgp = c->have_sgp;
    A.iall = D.gall_1d;
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.eps = eps;
    A.big = big;
    const long nmain = (long)((D.gall + CHUNK - 1) / CHUNK) * D.kall * D.lall;
    const long npole = Qprev_pl ? (D.kall * kLallPl + CHUNK - 1) / CHUNK : 0;
    if (nmain + npole > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    A.nmain = (int)nmain;
    A.q_a_pl = q_a_pl;
    A.q_pl = q_pl;
    A.d_pl = d_pl;
    A.ch_pl = ch_pl;
    A.cmask_pl = cmask_pl;
    A.Qprev_pl = Qprev_pl;
    A.Qpost_pl = Qpost_pl;
    A.gall_pl = D.gall_pl;
    A.have_pl = D.have_pl;
    A.phases = 3;
    nblocks = nmain + npole;
    return NICAM_OK;
}

template <class T>
int launch_hlim_t(nicam_ctx c, T *q_a, T *q_a_pl, const T *q, const T *q_pl, const T *d, const T *d_pl,
                  const T *ch, const T *ch_pl, const T *cmask, const T *cmask_pl, T *Qprev, T *Qprev_pl,
                  const T *Qpost, const T *Qpost_pl, T eps, T big, uint8_t *mask = nullptr,
                  uint8_t *mask_pl = nullptr, int ntr = 1)
{
    const Dims &D = c->d;
    HlimArgs<T> A;
    long nblk = 0;
    int rc0 = hlim_make_args(c, A, nblk, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl, Qprev, Qprev_pl,
                             Qpost, Qpost_pl, eps, big, mask, mask_pl, ntr);
    if (rc0 != NICAM_OK) return rc0;
    // regions wider than the common tile take the wide instances (a larger static tile)
    const bool wide = D.gall_1d > MAXI_TILE;
    auto run = [&](int phases) -> int {
        A.phases = phases;
        const dim3 grid((unsigned)nblk), blk(CHUNK);
        if (ntr > 1) {
            if (wide) hipLaunchKernelGGL((hlim_kernel<T, true, false, MAXI>), grid, blk, 0, c->stream, A);
            else hipLaunchKernelGGL((hlim_kernel<T, true>), grid, blk, 0, c->stream, A);
        }
        // the loads before the barrier (gl7 FP32: 344 against 359-369 us; FP64, in the step after the flux:
        // 710 against 719 us); variant 1 keeps them after it (fewer registers live across the barrier)
        else if (c->tune.hlim_variant != 1) {
            if (wide) hipLaunchKernelGGL((hlim_kernel<T, false, true, MAXI>), grid, blk, 0, c->stream, A);
            else hipLaunchKernelGGL((hlim_kernel<T, false, true>), grid, blk, 0, c->stream, A);
        } else {
            if (wide) hipLaunchKernelGGL((hlim_kernel<T, false, false, MAXI>), grid, blk, 0, c->stream, A);
            else hipLaunchKernelGGL((hlim_kernel<T, false>), grid, blk, 0, c->stream, A);
        }
        NICAM_HIP(hipGetLastError());
        return NICAM_OK;
    };
    if (Qpost) return run(3);  // the reference's kernel: Qout_post supplied

    // No Qout_post: do the halo exchange the reference cut out (HLIM:365-375,
    // the commented COMM_data_transfer of Qout) on the device, in place on
    // Qout_prev as NICAM itself exchanges Qout: sources are inner points and
    // pole centres, destinations halo points and pole rings, so the gather
    // reads nothing it writes.  The clamp then reads the exchanged range.
    if (!c->halo_n) return fail(NICAM_ERR_STATE, "Qout_post is NULL but no halo map is set (nicam_comm_setup)");
    if (!Qprev_pl) return fail(NICAM_ERR_ARG, "Qout_prev_pl is needed to exchange Qout with the poles");
    int rc = run(1);
    if (rc != NICAM_OK) return rc;
    rc = halo_fill(c, Qprev, Qprev_pl, D.kall, 2 * ntr);
    if (rc != NICAM_OK) return rc;
    A.Qpost = Qprev;
    A.Qpost_pl = Qprev_pl;
    return run(2);
}

}  // namespace

int launch_hlim(nicam_ctx c, double *q_a, double *q_a_pl, const double *q, const double *q_pl,
                const double *d, const double *d_pl, const double *ch, const double *ch_pl,
                const double *cmask, const double *cmask_pl, double *Qprev, double *Qprev_pl,
                const double *Qpost, const double *Qpost_pl)
{
    return launch_hlim_t<double>(c, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl, Qprev, Qprev_pl, Qpost,
                                 Qpost_pl, c->eps, c->big);
}

int launch_hlim_mask(nicam_ctx c, double *q_a, double *q_a_pl, const double *q, const double *q_pl,
                     const double *d, const double *d_pl, const double *ch, const double *ch_pl,
                     const double *cmask, const double *cmask_pl, double *Qprev, double *Qprev_pl,
                     const double *Qpost, const double *Qpost_pl, uint8_t *mask, uint8_t *mask_pl)
{
    return launch_hlim_t<double>(c, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl, Qprev, Qprev_pl, Qpost,
                                 Qpost_pl, c->eps, c->big, mask, mask_pl);
}

int launch_hlim_f32(nicam_ctx c, float *q_a, float *q_a_pl, const float *q, const float *q_pl, const float *d,
                    const float *d_pl, const float *ch, const float *ch_pl, const float *cmask,
                    const float *cmask_pl, float *Qprev, float *Qprev_pl, const float *Qpost,
                    const float *Qpost_pl)
{
    // the FP32 CONST_EPS and CONST_HUGE (nicam_set_const_f32; by default the -DSINGLE build's)
    return launch_hlim_t<float>(c, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl, Qprev, Qprev_pl, Qpost,
                                Qpost_pl, c->eps32, c->big32);
}

}  // namespace nicam

namespace nicam {

int launch_hlim_batch(nicam_ctx c, int ntr, double *q_a, double *q_a_pl, const double *q, const double *q_pl,
                      const double *d, const double *d_pl, const double *ch, const double *ch_pl,
                      const double *cmask, const double *cmask_pl, double *Qprev, double *Qprev_pl,
                      const double *Qpost, const double *Qpost_pl)
{
    return launch_hlim_t<double>(c, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl, Qprev, Qprev_pl, Qpost,
                                 Qpost_pl, c->eps, c->big, nullptr, nullptr, ntr);
}

int launch_hlim_batch_f32(nicam_ctx c, int ntr, float *q_a, float *q_a_pl, const float *q, const float *q_pl,
                          const float *d, const float *d_pl, const float *ch, const float *ch_pl,
                          const float *cmask, const float *cmask_pl, float *Qprev, float *Qprev_pl,
                          const float *Qpost, const float *Qpost_pl)
{
    return launch_hlim_t<float>(c, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl, Qprev, Qprev_pl, Qpost,
                                Qpost_pl, c->eps32, c->big32, nullptr, nullptr, ntr);
}

}  // namespace nicam

namespace nicam {

// FP32 limiter of ntr tracers on n grids in one launch (Qout_post supplied)
int launch_hlim_group_f32(const nicam_ctx *cs, int n, int ntr, const nicam_tracer_step_f32 *g, hipStream_t stream)
{
    if (n < 1 || n > kMaxGroup) return fail(NICAM_ERR_ARG, "a group holds 1..8 grids");
    HlimGroup<float> G;
    G.n = n;
    G.off[0] = 0;
    for (int i = 0; i < n; ++i) {
        const nicam_tracer_step_f32 &p = g[i];
        if (!p.Qout_post) return fail(NICAM_ERR_ARG, "the grouped limiter needs Qout_post");
        if (cs[i]->d.gall_1d > MAXI_TILE) return fail(NICAM_ERR_ARG, "gall_1d larger than the grouped limiter's tiles (160)");
        long nb = 0;
        const int rc = hlim_make_args<float>(cs[i], G.a[i], nb, p.q_a, p.q_a_pl, p.q, p.q_pl, p.d, p.d_pl, p.ch,
                                             p.ch_pl, p.cmask, p.cmask_pl, p.Qout_prev, p.Qout_prev_pl, p.Qout_post,
                                             p.Qout_post_pl, cs[i]->eps32, cs[i]->big32, nullptr, nullptr, ntr);
        if (rc != NICAM_OK) return rc;
        if ((long)G.off[i] + nb > 0x7fffffffL) return fail(NICAM_ERR_ARG, "group too large for one launch");
        G.off[i + 1] = G.off[i] + (int)nb;
    }
    hipLaunchKernelGGL(hlim_group_kernel<float>, dim3((unsigned)G.off[n]), dim3(CHUNK), 0, stream, G);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

}  // namespace nicam
