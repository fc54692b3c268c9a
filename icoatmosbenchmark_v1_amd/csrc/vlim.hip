// vertical_limiter_thuburn on gfx950.
//
// Reference: kernels/NICAM/dyn_vert_adv_limiter/src/mod_src_tracer.f90:41-243.
//
// The limited half-level value q_h(k) needs the allowable ranges Qout of
// levels k-1 and k; the reference carries Qout(k-1) from one level to the
// next (:140-141, :182-183), but every (g, k) is independent.  One work-item
// handles a run of KC consecutive levels of one column: it recomputes Qout of
// the level just below its run (the reference's peel at kmin plays that role
// for the first run), then marches up carrying Qout, q(k-1), q(k) and
// ck(k,1) in registers.  A wavefront covers 64 consecutive g at one level, so
// every load and store is contiguous.  The pole columns are appended as the
// last work-items.
//
// Many tracers (q(gall,kall,lall,ntr), NICAM's tracer-last layout) share ck
// and d.  The batched kernel gives a work-item up to TB tracers of one
// column run: per level it loads ck and d once and forms the
// tracer-independent flags, Courant sums and zerosw once, then limits each
// tracer with the single-tracer arithmetic.
#include "internal.h"

namespace nicam {
namespace {

constexpr int BLOCK = 256;
#ifndef NICAM_VLIM_KC
#define NICAM_VLIM_KC 4  // levels per run of the single-tracer kernel (0: runtime-length runs of 4); gl7: 4 164 us, 8 168, 16 168, runtime-length 169 (FP64)
#endif

template <class T>
__device__ __forceinline__ T dmin(T a, T b) { return (b < a) ? b : a; }
template <class T>
__device__ __forceinline__ T dmax(T a, T b) { return (b > a) ? b : a; }
__device__ __forceinline__ double sign_half(double x) { return copysign(0.5, x); }
__device__ __forceinline__ float sign_half(float x) { return copysignf(0.5f, x); }
__device__ __forceinline__ double tabs(double x) { return fabs(x); }
__device__ __forceinline__ float tabs(float x) { return fabsf(x); }

template <class T>
struct VlimArgs {
    T *q_h;
    const T *q, *d, *ck;
    T *q_h_pl;
    const T *q_pl, *d_pl, *ck_pl;
    int gall, kall, lall, gall_pl, have_pl;
    int kc, nrun;  // levels per work-item, runs per column
    T eps, big;
    int ntr;          // tracers (batched kernel)
    long ts, tps;     // tracer strides of the regular and pole fields
    uint32_t *mask, *mask_pl;  // clip bits (masked variant only; kc = 32)
};

// Levels [kb, ke) of one column are limited (kb >= kmin+1, 0-based).
// base = offset of level 0, s = stride between levels, ck2 = offset of the
// upper-face plane of ck.  POLE selects the pole loop's product grouping of
// CQin (VLIM:212-215 versus :126-129).
// MASK: also record one bit per limited face, set where the limited value
// differs from q_h's input, in the word at mask_word (the run is the 32
// faces of that word).
template <bool POLE, class T, bool MASK = false>
__device__ __forceinline__ void vlim_run(T *__restrict__ q_h, const T *__restrict__ q, const T *__restrict__ d,
                                         const T *__restrict__ ck, size_t base, size_t s, size_t ck2, int kb,
                                         int ke, T EPS, T BIG, uint32_t *mask_word = nullptr)
{
    uint32_t bits = 0;
    int k = kb - 1;  // first level whose Qout is needed
    T qkm1 = q[base + (size_t)(k - 1) * s];
    T qk = q[base + (size_t)k * s];
    T ck1k = ck[base + (size_t)k * s];
    T Qmin_km1 = T(0), Qmax_km1 = T(0);
    for (; k < ke; ++k) {
        const size_t o = base + (size_t)k * s;
        const T qkp1 = q[o + s];
        const T ck1kp1 = ck[o + s];
        const T ck2k = ck[ck2 + o];
        const T dk = d[o];
        const T inflagL = T(0.5) - sign_half(ck1k);
        const T inflagU = T(0.5) + sign_half(ck1kp1);
        const T Qin_minL = dmin(qk, qkm1) + (T(1) - inflagL) * BIG;
        const T Qin_minU = dmin(qk, qkp1) + (T(1) - inflagU) * BIG;
        const T Qin_maxL = dmax(qk, qkm1) - (T(1) - inflagL) * BIG;
        const T Qin_maxU = dmax(qk, qkp1) - (T(1) - inflagU) * BIG;
        const T qnext_min = dmin(dmin(Qin_minL, Qin_minU), qk);
        const T qnext_max = dmax(dmax(Qin_maxL, Qin_maxU), qk);
        const T Cin = (inflagL)*ck1k + (inflagU)*ck2k;
        const T Cout = (T(1) - inflagL) * ck1k + (T(1) - inflagU) * ck2k;
        T CQin_min, CQin_max;
        if constexpr (POLE) {
            CQin_max = (inflagL) * (ck1k * Qin_maxL) + (inflagU) * (ck2k * Qin_maxU);
            CQin_min = (inflagL) * (ck1k * Qin_minL) + (inflagU) * (ck2k * Qin_minU);
        } else {
            CQin_min = (inflagL)*ck1k * Qin_minL + (inflagU)*ck2k * Qin_minU;
            CQin_max = (inflagL)*ck1k * Qin_maxL + (inflagU)*ck2k * Qin_maxU;
        }
        const T zerosw = T(0.5) - sign_half(tabs(Cout) - EPS);
        const T Qmin_k = ((qk - qnext_max) + qnext_max * (Cin + Cout - dk) - CQin_max) / (Cout + zerosw) *
                                  (T(1) - zerosw) + qk * zerosw;
        const T Qmax_k = ((qk - qnext_min) + qnext_min * (Cin + Cout - dk) - CQin_min) / (Cout + zerosw) *
                                  (T(1) - zerosw) + qk * zerosw;
        if (k >= kb) {  // VLIM:179-180
            const T qh = q_h[o];
            const T out = (inflagL)*dmax(dmin(qh, Qmax_km1), Qmin_km1) + (T(1) - inflagL) * dmax(dmin(qh, Qmax_k), Qmin_k);
            stream_store(&q_h[o], out);
            if constexpr (MASK) bits |= (out != qh ? 1u : 0u) << (k - kb);
        }
        Qmin_km1 = Qmin_k;
        Qmax_km1 = Qmax_k;
        qkm1 = qk;
        qk = qkp1;
        ck1k = ck1kp1;
    }
    if constexpr (MASK) *mask_word = bits;
}

// vlim_run with the run's length fixed at KC levels at compile time: every
// input of the run (q on kb-2..ke, ck on kb-1..ke, d on kb-1..ke-1, q_h on
// kb..ke-1) is loaded before the first level is formed, so a work-item waits
// for memory once per run rather than once per level, and runs can be long
// (fewer levels re-read at run edges) without losing loads in flight.  A
// shorter last run (ke - kb < KC) reads level ke in place of the levels past
// it and stores nothing there.  Same operations, in the same order, as
// vlim_run.
template <bool POLE, class T, int KC>
__device__ __forceinline__ void vlim_run_kc(T *__restrict__ q_h, const T *__restrict__ q, const T *__restrict__ d,
                                            const T *__restrict__ ck, size_t base, size_t s, size_t ck2, int kb,
                                            int ke, T EPS, T BIG)
{
    auto at = [&](int k) { return base + (size_t)(k < ke ? k : ke) * s; };
    T qv[KC + 3], c1[KC + 2], c2[KC + 1], dv[KC + 1], qh[KC];
#pragma unroll
    for (int t = 0; t < KC + 3; ++t) qv[t] = q[at(kb - 2 + t)];
#pragma unroll
    for (int t = 0; t < KC + 2; ++t) c1[t] = ck[at(kb - 1 + t)];
#pragma unroll
    for (int t = 0; t < KC + 1; ++t) {
        c2[t] = ck[ck2 + at(kb - 1 + t)];
        dv[t] = d[at(kb - 1 + t)];
    }
#pragma unroll
    for (int t = 0; t < KC; ++t) qh[t] = q_h[at(kb + t)];
    T Qmin_km1 = T(0), Qmax_km1 = T(0);
#pragma unroll
    for (int t = 0; t <= KC; ++t) {  // level k = kb - 1 + t
        const int k = kb - 1 + t;
        const T qkm1 = qv[t], qk = qv[t + 1], qkp1 = qv[t + 2];
        const T ck1k = c1[t], ck1kp1 = c1[t + 1], ck2k = c2[t], dk = dv[t];
        const T inflagL = T(0.5) - sign_half(ck1k);
        const T inflagU = T(0.5) + sign_half(ck1kp1);
        const T Qin_minL = dmin(qk, qkm1) + (T(1) - inflagL) * BIG;
        const T Qin_minU = dmin(qk, qkp1) + (T(1) - inflagU) * BIG;
        const T Qin_maxL = dmax(qk, qkm1) - (T(1) - inflagL) * BIG;
        const T Qin_maxU = dmax(qk, qkp1) - (T(1) - inflagU) * BIG;
        const T qnext_min = dmin(dmin(Qin_minL, Qin_minU), qk);
        const T qnext_max = dmax(dmax(Qin_maxL, Qin_maxU), qk);
        const T Cin = (inflagL)*ck1k + (inflagU)*ck2k;
        const T Cout = (T(1) - inflagL) * ck1k + (T(1) - inflagU) * ck2k;
        T CQin_min, CQin_max;
        if constexpr (POLE) {
            CQin_max = (inflagL) * (ck1k * Qin_maxL) + (inflagU) * (ck2k * Qin_maxU);
            CQin_min = (inflagL) * (ck1k * Qin_minL) + (inflagU) * (ck2k * Qin_minU);
        } else {
            CQin_min = (inflagL)*ck1k * Qin_minL + (inflagU)*ck2k * Qin_minU;
            CQin_max = (inflagL)*ck1k * Qin_maxL + (inflagU)*ck2k * Qin_maxU;
        }
        const T zerosw = T(0.5) - sign_half(tabs(Cout) - EPS);
        const T Qmin_k = ((qk - qnext_max) + qnext_max * (Cin + Cout - dk) - CQin_max) / (Cout + zerosw) *
                                  (T(1) - zerosw) + qk * zerosw;
        const T Qmax_k = ((qk - qnext_min) + qnext_min * (Cin + Cout - dk) - CQin_min) / (Cout + zerosw) *
                                  (T(1) - zerosw) + qk * zerosw;
        if (t >= 1 && k < ke) {  // VLIM:179-180
            const T qhv = qh[t - 1];
            const T out = (inflagL)*dmax(dmin(qhv, Qmax_km1), Qmin_km1) + (T(1) - inflagL) * dmax(dmin(qhv, Qmax_k), Qmin_k);
            stream_store(&q_h[base + (size_t)k * s], out);
        }
        Qmin_km1 = Qmin_k;
        Qmax_km1 = Qmax_k;
    }
}

// vlim_run for up to TB tracers (nb of them) of one column run; tracer b's
// fields start b*ts elements after q / q_h.
template <bool POLE, int TB, class T>
__device__ __forceinline__ void vlim_run_multi(T *__restrict__ q_h, const T *__restrict__ q,
                                               const T *__restrict__ d, const T *__restrict__ ck, size_t base,
                                               size_t s, size_t ck2, int kb, int ke, T EPS, T BIG, int nb, long ts)
{
    int k = kb - 1;
    T qkm1[TB], qk[TB], Qmin_km1[TB], Qmax_km1[TB];
#pragma unroll
    for (int b = 0; b < TB; ++b) {
        const int bb = b < nb ? b : 0;  // idle slots repeat tracer 0 and store nothing
        qkm1[b] = q[bb * ts + base + (size_t)(k - 1) * s];
        qk[b] = q[bb * ts + base + (size_t)k * s];
        Qmin_km1[b] = T(0);
        Qmax_km1[b] = T(0);
    }
    T ck1k = ck[base + (size_t)k * s];
    for (; k < ke; ++k) {
        const size_t o = base + (size_t)k * s;
        const T ck1kp1 = ck[o + s];
        const T ck2k = ck[ck2 + o];
        const T dk = d[o];
        const T inflagL = T(0.5) - sign_half(ck1k);
        const T inflagU = T(0.5) + sign_half(ck1kp1);
        const T Cin = (inflagL)*ck1k + (inflagU)*ck2k;
        const T Cout = (T(1) - inflagL) * ck1k + (T(1) - inflagU) * ck2k;
        const T zerosw = T(0.5) - sign_half(tabs(Cout) - EPS);
#pragma unroll
        for (int b = 0; b < TB; ++b) {
            const int bb = b < nb ? b : 0;
            const T qkp1 = q[bb * ts + o + s];
            const T Qin_minL = dmin(qk[b], qkm1[b]) + (T(1) - inflagL) * BIG;
            const T Qin_minU = dmin(qk[b], qkp1) + (T(1) - inflagU) * BIG;
            const T Qin_maxL = dmax(qk[b], qkm1[b]) - (T(1) - inflagL) * BIG;
            const T Qin_maxU = dmax(qk[b], qkp1) - (T(1) - inflagU) * BIG;
            const T qnext_min = dmin(dmin(Qin_minL, Qin_minU), qk[b]);
            const T qnext_max = dmax(dmax(Qin_maxL, Qin_maxU), qk[b]);
            T CQin_min, CQin_max;
            if constexpr (POLE) {
                CQin_max = (inflagL) * (ck1k * Qin_maxL) + (inflagU) * (ck2k * Qin_maxU);
                CQin_min = (inflagL) * (ck1k * Qin_minL) + (inflagU) * (ck2k * Qin_minU);
            } else {
                CQin_min = (inflagL)*ck1k * Qin_minL + (inflagU)*ck2k * Qin_minU;
                CQin_max = (inflagL)*ck1k * Qin_maxL + (inflagU)*ck2k * Qin_maxU;
            }
            const T Qmin_k = ((qk[b] - qnext_max) + qnext_max * (Cin + Cout - dk) - CQin_max) / (Cout + zerosw) *
                                 (T(1) - zerosw) + qk[b] * zerosw;
            const T Qmax_k = ((qk[b] - qnext_min) + qnext_min * (Cin + Cout - dk) - CQin_min) / (Cout + zerosw) *
                                 (T(1) - zerosw) + qk[b] * zerosw;
            if (k >= kb && b < nb) {  // VLIM:179-180
                const T qh = q_h[b * ts + o];
                const T out = (inflagL)*dmax(dmin(qh, Qmax_km1[b]), Qmin_km1[b]) +
                              (T(1) - inflagL) * dmax(dmin(qh, Qmax_k), Qmin_k);
                stream_store(&q_h[b * ts + o], out);
            }
            Qmin_km1[b] = Qmin_k;
            Qmax_km1[b] = Qmax_k;
            qkm1[b] = qk[b];
            qk[b] = qkp1;
        }
        ck1k = ck1kp1;
    }
}

constexpr int VTB = 4;  // tracers per work-item in the batched kernel

// the batched kernel: work-group = BLOCK columns at one run for one group of
// VTB tracers; the tracer groups of one column block and run are neighbours
template <class T>
__device__ __forceinline__ void vlim_multi_block(const VlimArgs<T> &A, long blk)
{
    const long ncol_reg = (long)A.gall * A.lall;
    const long ncol = ncol_reg + (A.have_pl ? (long)A.gall_pl * kLallPl : 0);
    const long ncb = (ncol + BLOCK - 1) / BLOCK;
    const int ngrp = (A.ntr + VTB - 1) / VTB;
    const long L = xcd_contiguous(blk, ncb * A.nrun * ngrp);
    const int grp = (int)(L % ngrp);
    const int run = (int)((L / ngrp) % A.nrun);
    const long col = (L / ngrp / A.nrun) * BLOCK + threadIdx.x;
    if (col >= ncol) return;
    const int kmin = 1, kmax = A.kall - 2;
    const int kb = kmin + 1 + run * A.kc;
    const int ke = min(kmax + 1, kb + A.kc);
    if (kb >= ke) return;
    const int t0 = grp * VTB, nb = min(VTB, A.ntr - t0);
    if (col < ncol_reg) {
        const int g = (int)(col % A.gall), l = (int)(col / A.gall);
        const size_t base = (size_t)l * A.kall * A.gall + g;
        const size_t ck2 = (size_t)A.gall * A.kall * A.lall;
        vlim_run_multi<false, VTB, T>(A.q_h + t0 * A.ts, A.q + t0 * A.ts, A.d, A.ck, base, A.gall, ck2, kb, ke,
                                      A.eps, A.big, nb, A.ts);
    } else {
        const int p = (int)(col - ncol_reg);
        const int g = p % A.gall_pl, l = p / A.gall_pl;
        const size_t base = (size_t)l * A.kall * A.gall_pl + g;
        const size_t ck2 = (size_t)A.gall_pl * A.kall * kLallPl;
        vlim_run_multi<true, VTB, T>(A.q_h_pl + t0 * A.tps, A.q_pl + t0 * A.tps, A.d_pl, A.ck_pl, base, A.gall_pl,
                                     ck2, kb, ke, A.eps, A.big, nb, A.tps);
    }
}

template <class T>
__global__ void __launch_bounds__(BLOCK) vlim_multi_kernel(VlimArgs<T> A)
{
    vlim_multi_block<T>(A, blockIdx.x);
}

// several grids in one launch: work-groups [off[n], off[n+1]) belong to grid n
template <class T>
struct VlimGroup {
    VlimArgs<T> a[kMaxGroup];
    int off[kMaxGroup + 1];
    int n;
};

template <class T>
__global__ void __launch_bounds__(BLOCK) vlim_group_kernel(VlimGroup<T> G)
{
    const int b = blockIdx.x;
    int n = 0;
    while (n + 1 < G.n && b >= G.off[n + 1]) ++n;
    vlim_multi_block<T>(G.a[n], b - G.off[n]);
}

// A work-group is BLOCK consecutive columns at one run of levels.  The runs
// of one column block are consecutive work-groups on one XCD: each run also
// reads the two levels below it, which the run below has just brought into L2.
template <class T, bool MASK = false>
__global__ void __launch_bounds__(BLOCK) vlim_kernel(VlimArgs<T> A)
{
    const long ncol_reg = (long)A.gall * A.lall;
    const long ncol = ncol_reg + (A.have_pl ? (long)A.gall_pl * kLallPl : 0);
    const long ncb = (ncol + BLOCK - 1) / BLOCK;  // column blocks
    const long L = xcd_contiguous(blockIdx.x, ncb * A.nrun);
    const int run = (int)(L % A.nrun);
    const long col = (L / A.nrun) * BLOCK + threadIdx.x;
    if (col >= ncol) return;
    const int kmin = 1, kmax = A.kall - 2;  // 0-based
    const int kb = kmin + 1 + run * A.kc;
    const int ke = min(kmax + 1, kb + A.kc);
    if (kb >= ke) return;
    if (col < ncol_reg) {
        const int g = (int)(col % A.gall), l = (int)(col / A.gall);
        const size_t base = (size_t)l * A.kall * A.gall + g;
        const size_t ck2 = (size_t)A.gall * A.kall * A.lall;
        uint32_t *mw = MASK ? A.mask + ((size_t)l * A.nrun + run) * A.gall + g : nullptr;
        vlim_run<false, T, MASK>(A.q_h, A.q, A.d, A.ck, base, A.gall, ck2, kb, ke, A.eps, A.big, mw);
    } else {
        const int p = (int)(col - ncol_reg);
        const int g = p % A.gall_pl, l = p / A.gall_pl;
        const size_t base = (size_t)l * A.kall * A.gall_pl + g;
        const size_t ck2 = (size_t)A.gall_pl * A.kall * kLallPl;
        uint32_t *mw = MASK ? A.mask_pl + ((size_t)l * A.nrun + run) * A.gall_pl + g : nullptr;
        vlim_run<true, T, MASK>(A.q_h_pl, A.q_pl, A.d_pl, A.ck_pl, base, A.gall_pl, ck2, kb, ke, A.eps, A.big, mw);
    }
}

// vlim_kernel with KC-level runs (vlim_run_kc); no mask.
template <class T, int KC>
__global__ void __launch_bounds__(BLOCK) vlim_kernel_kc(VlimArgs<T> A)
{
    const long ncol_reg = (long)A.gall * A.lall;
    const long ncol = ncol_reg + (A.have_pl ? (long)A.gall_pl * kLallPl : 0);
    const long ncb = (ncol + BLOCK - 1) / BLOCK;  // column blocks
    const long L = xcd_contiguous(blockIdx.x, ncb * A.nrun);
    const int run = (int)(L % A.nrun);
    const long col = (L / A.nrun) * BLOCK + threadIdx.x;
    if (col >= ncol) return;
    const int kmin = 1, kmax = A.kall - 2;  // 0-based
    const int kb = kmin + 1 + run * KC;
    const int ke = min(kmax + 1, kb + KC);
    if (kb >= ke) return;
    if (col < ncol_reg) {
        const int g = (int)(col % A.gall), l = (int)(col / A.gall);
        const size_t base = (size_t)l * A.kall * A.gall + g;
        const size_t ck2 = (size_t)A.gall * A.kall * A.lall;
        vlim_run_kc<false, T, KC>(A.q_h, A.q, A.d, A.ck, base, A.gall, ck2, kb, ke, A.eps, A.big);
    } else {
        const int p = (int)(col - ncol_reg);
        const int g = p % A.gall_pl, l = p / A.gall_pl;
        const size_t base = (size_t)l * A.kall * A.gall_pl + g;
        const size_t ck2 = (size_t)A.gall_pl * A.kall * kLallPl;
        vlim_run_kc<true, T, KC>(A.q_h_pl, A.q_pl, A.d_pl, A.ck_pl, base, A.gall_pl, ck2, kb, ke, A.eps, A.big);
    }
}

}  // namespace

namespace {

template <class T>
int vlim_make_args(nicam_ctx c, VlimArgs<T> &A, long &nblk, T *q_h, T *q_h_pl, const T *q, const T *q_pl,
                   const T *d, const T *d_pl, const T *ck, const T *ck_pl, T eps, T big, uint32_t *mask,
                   uint32_t *mask_pl, int ntr, bool multi)
{
    const Dims &D = c->d;
    if (ntr < 1 || (mask && ntr != 1)) return fail(NICAM_ERR_ARG, "tracer count must be >= 1 (1 with a mask)");
    A.ntr = ntr;
    A.ts = (long)D.nreg();
    A.tps = (long)D.npl();
    A.q_h = q_h;
    A.q = q;
    A.d = d;
    A.ck = ck;
    A.q_h_pl = q_h_pl;
    A.q_pl = q_pl;
    A.d_pl = d_pl;
    A.ck_pl = ck_pl;
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.gall_pl = D.gall_pl;
    A.have_pl = D.have_pl;
    A.eps = eps;
    A.big = big;
    A.mask = mask;
    A.mask_pl = mask_pl;
    nblk = 0;
    const int nlev = D.kall - 3;  // limited levels kmin+1..kmax
    if (nlev <= 0) return NICAM_OK;
    // runtime-length runs: short runs give more independent work-items (measured best at gl7: 4); the
    // single-tracer kernel without a mask takes compile-time runs of NICAM_VLIM_KC levels instead
    int kc = c->tune.vlim_kchunk > 0 ? c->tune.vlim_kchunk : (!multi && !mask && NICAM_VLIM_KC > 0 ? NICAM_VLIM_KC : 4);
    if (mask) kc = 32;  // one mask word per run
    kc = std::min(kc, nlev);
    A.kc = kc;
    A.nrun = (nlev + kc - 1) / kc;
    const long ncol = (long)D.gall * D.lall + (D.have_pl ? (long)D.gall_pl * kLallPl : 0);
    nblk = (ncol + BLOCK - 1) / BLOCK * A.nrun * (multi ? (ntr + VTB - 1) / VTB : 1);
    if (nblk > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    return NICAM_OK;
}

template <class T>
int launch_vlim_t(nicam_ctx c, T *q_h, T *q_h_pl, const T *q, const T *q_pl, const T *d, const T *d_pl,
                  const T *ck, const T *ck_pl, T eps, T big, uint32_t *mask = nullptr, uint32_t *mask_pl = nullptr,
                  int ntr = 1)
{
    VlimArgs<T> A;
    long nblk = 0;
    const int rc = vlim_make_args(c, A, nblk, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl, eps, big, mask, mask_pl, ntr,
                                  ntr > 1);
    if (rc != NICAM_OK || nblk == 0) return rc;
    if (ntr > 1)
        hipLaunchKernelGGL(vlim_multi_kernel<T>, dim3((unsigned)nblk), dim3(BLOCK), 0, c->stream, A);
    else if (mask) hipLaunchKernelGGL((vlim_kernel<T, true>), dim3((unsigned)nblk), dim3(BLOCK), 0, c->stream, A);
    else if (NICAM_VLIM_KC > 0 && A.kc == NICAM_VLIM_KC)
        hipLaunchKernelGGL((vlim_kernel_kc<T, (NICAM_VLIM_KC > 0 ? NICAM_VLIM_KC : 1)>), dim3((unsigned)nblk),
                           dim3(BLOCK), 0, c->stream, A);
    else hipLaunchKernelGGL((vlim_kernel<T, false>), dim3((unsigned)nblk), dim3(BLOCK), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

}  // namespace

int launch_vlim(nicam_ctx c, double *q_h, double *q_h_pl, const double *q, const double *q_pl,
                const double *d, const double *d_pl, const double *ck, const double *ck_pl)
{
    return launch_vlim_t<double>(c, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl, c->eps, c->big);
}

int launch_vlim_mask(nicam_ctx c, double *q_h, double *q_h_pl, const double *q, const double *q_pl,
                     const double *d, const double *d_pl, const double *ck, const double *ck_pl, uint32_t *mask,
                     uint32_t *mask_pl)
{
    return launch_vlim_t<double>(c, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl, c->eps, c->big, mask, mask_pl);
}

int launch_vlim_f32(nicam_ctx c, float *q_h, float *q_h_pl, const float *q, const float *q_pl, const float *d,
                    const float *d_pl, const float *ck, const float *ck_pl)
{
    // the FP32 CONST_EPS and CONST_HUGE (nicam_set_const_f32; by default the -DSINGLE build's)
    return launch_vlim_t<float>(c, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl, c->eps32, c->big32);
}

}  // namespace nicam

namespace nicam {

int launch_vlim_batch(nicam_ctx c, int ntr, double *q_h, double *q_h_pl, const double *q, const double *q_pl,
                      const double *d, const double *d_pl, const double *ck, const double *ck_pl)
{
    return launch_vlim_t<double>(c, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl, c->eps, c->big, nullptr, nullptr, ntr);
}

int launch_vlim_batch_f32(nicam_ctx c, int ntr, float *q_h, float *q_h_pl, const float *q, const float *q_pl,
                          const float *d, const float *d_pl, const float *ck, const float *ck_pl)
{
    return launch_vlim_t<float>(c, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl, c->eps32, c->big32, nullptr, nullptr,
                                ntr);
}

}  // namespace nicam

namespace nicam {

int launch_vlim_group_f32(const nicam_ctx *cs, int n, int ntr, const nicam_tracer_step_f32 *g, hipStream_t stream)
{
    if (n < 1 || n > kMaxGroup) return fail(NICAM_ERR_ARG, "a group holds 1..8 grids");
    VlimGroup<float> G;
    G.n = n;
    G.off[0] = 0;
    for (int i = 0; i < n; ++i) {
        const nicam_tracer_step_f32 &p = g[i];
        long nb = 0;
        const int rc = vlim_make_args<float>(cs[i], G.a[i], nb, p.q_h, p.q_h_pl, p.qv, p.qv_pl, p.dv, p.dv_pl, p.ck,
                                             p.ck_pl, cs[i]->eps32, cs[i]->big32, nullptr, nullptr, ntr, true);
        if (rc != NICAM_OK) return rc;
        if ((long)G.off[i] + nb > 0x7fffffffL) return fail(NICAM_ERR_ARG, "group too large for one launch");
        G.off[i + 1] = G.off[i] + (int)nb;
    }
    if (G.off[n] == 0) return NICAM_OK;
    hipLaunchKernelGGL(vlim_group_kernel<float>, dim3((unsigned)G.off[n]), dim3(BLOCK), 0, stream, G);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

}  // namespace nicam
