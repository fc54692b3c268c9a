"""A differentiable tracer step: the adjoint test harness around the
hand-written adjoint kernels.

One step advances ntr tracers q and the density rhog of an icosahedral grid
(all regions and both poles on one GPU) over fixed winds:

    horizontal_flux(rhog, rho*v)              -> flx_h, GRD_xc    [HIP; adjoint HIP]
    ch = flx_h / rhog, cmask = upwind mask                          [glue]
    grad = gradient(q), halo exchange; q_a = upwind remap            [glue]
    Qout = phase 1 of the limiter, halo exchange, q_a limited        [HIP; adjoint HIP]
    rhog*q and rhog updated by the face fluxes, q = rhog*q / rhog    [glue]
    rho*w = vi_rhow_solver(..., rhog0 = rhog, ...)                   [HIP; adjoint HIP]
    flx_v = alpha * rho*w * dt; ck; q_h = half-level interpolation   [glue]
    q_h limited by vertical_limiter_thuburn                          [HIP; adjoint HIP]
    rhog*q and rhog updated by the vertical fluxes; halo exchange    [glue]

The four reference kernels run as torch.autograd.Functions over the
library's forward kernels (the masked forwards for the limiters) and their
nicam_*_adjoint entry points.  The glue between them, which the reference
does not have (its kernels take ch, cmask, q_a, ck, q_h as inputs), is
written here as torch operations on the same arrays in the oracle's layout
and operation order, so its derivatives come from torch's autograd.  The
limiters' adjoints return the cotangents of the Courant numbers ch and ck,
through which their bounds depend on rhog; what stays fixed is piecewise
constant (cmask and the inflow flags, the signs of the Courant numbers) or
zero here (d), so the gradient in q and rhog is the step's own wherever no
switch sits exactly at its threshold.

Arrays are C arrays with the Fortran dimensions reversed: q[t][l][k][g],
q_pl[t][p][k][v]; rhog[l][k][g], rhog_pl[p][k][v].
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import numpy as np
import torch
import torch.nn.functional as F

from .autodiff import RhowSolve, VerticalLimiter
from .drivers import VI_IN
from .runtime import Context


class HorizontalFlux(torch.autograd.Function):
    """(flx_h, GRD_xc) = horizontal_flux(rho, rho*v) (FLX:42-379)."""

    @staticmethod
    def forward(fctx, nctx: Context, dt: float, *X):
        rho, rho_pl = X[0], X[1]
        flx, flx_pl = rho.new_zeros((6,) + rho.shape), rho_pl.new_zeros(rho_pl.shape)
        xc, xc_pl = rho.new_zeros((3, 3) + rho.shape), rho_pl.new_zeros((3,) + rho_pl.shape)
        nctx.horizontal_flux(flx, flx_pl, xc, xc_pl, *X, dt)
        fctx.nctx, fctx.dt = nctx, dt
        fctx.save_for_backward(*X)
        return flx, flx_pl, xc, xc_pl

    @staticmethod
    def backward(fctx, gf, gf_pl, gx, gx_pl):
        X = fctx.saved_tensors
        rho, rho_pl = X[0], X[1]
        c = lambda g, shape, ref: ref.new_zeros(shape) if g is None else g.contiguous()  # noqa: E731
        gf = c(gf, (6,) + rho.shape, rho)
        gf_pl = c(gf_pl, rho_pl.shape, rho_pl)
        gx = c(gx, (3, 3) + rho.shape, rho)
        gx_pl = c(gx_pl, (3,) + rho_pl.shape, rho_pl)
        b = [torch.empty_like(x) for x in X]
        fctx.nctx.horizontal_flux_adjoint(gf, gf_pl, gx, gx_pl, *b, *X, fctx.dt)
        return (None, None, *b)


class HorizontalLimiter(torch.autograd.Function):
    """(q_a', Qout_prev) = horizontal_limiter_thuburn(q_a, q, Qout_post, ch; d,
    cmask) (HLIM:42-430) with Qout_post supplied; d and cmask held fixed."""

    @staticmethod
    def forward(fctx, nctx: Context, q_a, q_a_pl, q, q_pl, post, post_pl, d, d_pl, ch, ch_pl, cm, cm_pl):
        out, out_pl = q_a.clone(), q_a_pl.clone()
        prev, prev_pl = post.new_zeros(post.shape), post_pl.new_zeros(post_pl.shape)
        mask = torch.empty(q.shape, dtype=torch.uint8, device=q.device)
        mask_pl = torch.empty(q_pl.shape, dtype=torch.uint8, device=q.device)
        nctx.horizontal_limiter_thuburn_mask(out, out_pl, q, q_pl, d, d_pl, ch, ch_pl, cm, cm_pl, prev, prev_pl,
                                             post, post_pl, mask, mask_pl)
        fctx.nctx = nctx
        fctx.save_for_backward(out, out_pl, q, q_pl, post, post_pl, d, d_pl, ch, ch_pl, cm, cm_pl, mask, mask_pl)
        return out, out_pl, prev, prev_pl

    @staticmethod
    def backward(fctx, g_qa, g_qa_pl, g_prev, g_prev_pl):
        out, out_pl, q, q_pl, post, post_pl, d, d_pl, ch, ch_pl, cm, cm_pl, mask, mask_pl = fctx.saved_tensors
        c = lambda g, ref: torch.zeros_like(ref) if g is None else g.contiguous().clone()  # noqa: E731
        qab, qab_pl = c(g_qa, out), c(g_qa_pl, out_pl)
        prevb, prevb_pl = c(g_prev, post), c(g_prev_pl, post_pl)
        qb, qb_pl = torch.empty_like(q), torch.empty_like(q_pl)
        postb, postb_pl = torch.empty_like(post), torch.empty_like(post_pl)
        chb, chb_pl = torch.empty_like(ch), torch.empty_like(ch_pl)
        fctx.nctx.horizontal_limiter_thuburn_adjoint(qab, qab_pl, qb, qb_pl, prevb, prevb_pl, postb, postb_pl, out,
                                                     out_pl, q, q_pl, d, d_pl, ch, ch_pl, cm, cm_pl, post, post_pl,
                                                     mask, mask_pl, chb, chb_pl)
        return (None, qab, qab_pl, qb, qb_pl, postb, postb_pl, None, None, chb, chb_pl, None, None)


class _HaloGather(torch.autograd.Function):
    """flat[..., dst] = flat[..., src] with every source read before any
    write.  Its transpose zeroes the destinations' cotangents and adds each
    destination's to its source, in rounds whose sources are distinct, so
    the sums do not depend on the order of atomic additions."""

    @staticmethod
    def forward(fctx, flat, dst, src, rounds):
        fctx.dst, fctx.rounds = dst, rounds
        return flat.index_copy(-1, dst, flat.index_select(-1, src))

    @staticmethod
    def backward(fctx, g):
        out = g.contiguous().index_fill(-1, fctx.dst, 0.0)
        for d, s in fctx.rounds:
            out.index_add_(-1, s, g.index_select(-1, d))
        return out, None, None, None


class HaloExchange:
    """COMM_data_transfer as index operations: every halo point and pole ring
    point takes the value of its source (the library's halo map)."""

    def __init__(self, hm, lall: int, gall: int, gall_pl: int, device):
        code = lambda kind, l, g: np.where(kind == 0, l * gall + g, lall * gall + l * gall_pl + g)  # noqa: E731
        dst = code(hm["dst_kind"], hm["dst_l"], hm["dst_g"]).astype(np.int64)
        src = code(hm["src_kind"], hm["src_l"], hm["src_g"]).astype(np.int64)
        self.dst = torch.from_numpy(dst).to(device)
        self.src = torch.from_numpy(src).to(device)
        # the r-th pair of each source, in (source, destination) order, goes to round r
        order = np.lexsort((dst, src))
        ss, dd = src[order], dst[order]
        first = np.r_[True, ss[1:] != ss[:-1]]
        start = np.maximum.accumulate(np.where(first, np.arange(len(ss)), 0))
        rank = np.arange(len(ss)) - start
        self.rounds = [(torch.from_numpy(dd[rank == r]).to(device), torch.from_numpy(ss[rank == r]).to(device))
                       for r in range(int(rank.max()) + 1 if len(ss) else 0)]
        self.lall, self.gall, self.gall_pl = lall, gall, gall_pl

    def __call__(self, x, x_pl):
        """x [..., lall, kall, gall], x_pl [..., 2, kall, gall_pl]."""
        lead = x.shape[:-3]
        kall = x.shape[-2]
        a = x.reshape(-1, self.lall, kall, self.gall).permute(0, 2, 1, 3).reshape(-1, kall, self.lall * self.gall)
        b = x_pl.reshape(-1, 2, kall, self.gall_pl).permute(0, 2, 1, 3).reshape(-1, kall, 2 * self.gall_pl)
        flat = torch.cat([a, b], dim=-1)
        out = _HaloGather.apply(flat, self.dst, self.src, self.rounds)
        a2 = out[..., :self.lall * self.gall].reshape(-1, kall, self.lall, self.gall).permute(0, 2, 1, 3)
        b2 = out[..., self.lall * self.gall:].reshape(-1, kall, 2, self.gall_pl).permute(0, 2, 1, 3)
        return (a2.reshape(lead + (self.lall, kall, self.gall)).contiguous(),
                b2.reshape(lead + (2, kall, self.gall_pl)).contiguous())


class TracerStep:
    """The step of the module docstring for one grid; call it with
    (q, q_pl, rhog, rhog_pl) and get the same four back, one step on."""

    def __init__(self, ctx: Context, spec, geo, case, hm, fac, dt: float = 360.0, dt_vi: float = 60.0,
                 courant_v: float = 0.3, device="cuda:0"):
        dev = torch.device(device)
        T = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)  # noqa: E731
        self.ctx, self.spec, self.dt, self.dt_vi = ctx, spec, dt, dt_vi
        n1, kall = spec.gall_1d, spec.kall
        self.n1, self.kall = n1, kall
        self.eps = float(np.finfo(np.float64).eps)
        f = case["flux"]
        self.rv = [T(f[k]) for k in ("rhovx", "rhovx_pl", "rhovy", "rhovy_pl", "rhovz", "rhovz_pl")]
        self.d, self.d_pl = T(case["hlim"]["d"]) * 0.0, T(case["hlim"]["d_pl"]) * 0.0
        self.dv, self.dv_pl = T(case["vlim"]["d"]) * 0.0, T(case["vlim"]["d_pl"]) * 0.0
        self.cg, self.cg_pl = T(geo["coef_grad"]), T(geo["coef_grad_pl"])
        self.X, self.X_pl = T(geo["GRD_x"][:, :, 0, :]), T(geo["GRD_x_pl"][:, :, 0, :])  # [d][l][g], [d][p][v]
        self.halo = HaloExchange(hm, spec.lall, spec.gall, spec.gall_pl, dev)
        vi = case["vi"]
        self.vi = {k + s: T(vi[k + s]) for k in VI_IN for s in ("", "_pl")}
        self.rhogw_in, self.rhogw_in_pl = T(vi["rhogw"]), T(vi["rhogw_pl"])
        self.rdgz, self.afact, self.bfact = T(fac["rdgz"]), T(fac["afact"]), T(fac["bfact"])
        inner = np.zeros((n1, n1), dtype=bool)
        inner[1:-1, 1:-1] = True
        self.inner = torch.from_numpy(inner.reshape(-1)).to(dev)
        self.glo, self.ghi = n1 + 1, spec.gall - n1 - 1  # the flat range gmin..gmax
        self.goff = (0, 1, n1 + 1, n1, -1, -n1 - 1, -n1)
        kk = torch.arange(kall, device=dev)
        self.kin = ((kk >= 1) & (kk <= kall - 2)).to(torch.float64)[:, None]        # kmin..kmax
        self.kface = ((kk >= 2) & (kk <= kall - 2)).to(torch.float64)[:, None]      # interior half levels
        # the vertical mass flux is alpha * rho*w * dt, alpha fixed at set-up so that |ck| <= courant_v
        with torch.no_grad():
            w, w_pl = RhowSolve.apply(ctx, dt_vi, self.rhogw_in, self.rhogw_in_pl, *self._vi_args(T(f["rho"]),
                                                                                                  T(f["rho_pl"])))
            m = float((w.abs() * self.kface).max() * dt * self.rdgz.max() / T(f["rho"]).min())
        self.alpha = courant_v / m if m > 0 else 0.0

    # -- glue ------------------------------------------------------------------
    def _vi_args(self, rhog, rhog_pl):
        a = []
        for k in VI_IN:
            if k == "rhog0":
                a += [rhog, rhog_pl]
            else:
                a += [self.vi[k], self.vi[k + "_pl"]]
        return a

    def courant(self, flx, flx_pl, rhog, rhog_pl):
        ch = flx / rhog
        ch_pl = torch.cat([torch.zeros_like(rhog_pl[..., :1]), flx_pl[..., 1:] / rhog_pl[..., :1]], dim=-1)
        cm = (0.5 - torch.copysign(torch.full_like(ch, 0.5), ch - self.eps)).detach()
        cm_pl = (0.5 - torch.copysign(torch.full_like(ch_pl, 0.5), ch_pl - self.eps)).detach()
        cm_pl[..., 0] = 0.0
        return ch, ch_pl, cm, cm_pl

    def gradient(self, q, q_pl):
        """q [l][k][g] -> grad [3][l][k][g] on the flat range gmin..gmax (zero
        elsewhere); the pole centre from its ring.  Slices and padding, not
        index tensors: their transposes are gathers, so the gradient's sums do
        not depend on the order of atomic additions."""
        lo, hi = self.glo, self.ghi
        parts = []
        for d in range(3):
            s = None
            for m in range(7):
                o = self.goff[m]
                t = self.cg[:, d, m, lo:hi][:, None, :] * q[:, :, lo + o:hi + o]
                s = t if s is None else s + t
            parts.append(F.pad(s, (lo, self.spec.gall - hi)))
        gp = []
        for d in range(3):
            s = (self.cg_pl[:, d, :][:, None, :] * q_pl).sum(-1)
            gp.append(torch.cat([s[..., None], torch.zeros_like(q_pl[..., 1:])], dim=-1))
        return torch.stack(parts), torch.stack(gp)

    def remap(self, q, q_pl, grad, grad_pl, cm, cm_pl, xc, xc_pl):
        """Upwind face values q + grad.(GRD_xc - GRD_x) of every arc from the
        origins i, j in [gmin-1, gmax]; the faces no arc reaches stay 0.  The
        origins and their neighbours are slices of the (j, i) view."""
        n = self.n1
        sq = lambda x: x.reshape(x.shape[:-1] + (n, n))  # noqa: E731  [..., j, i]
        org = lambda x: sq(x)[..., :n - 1, :n - 1]  # noqa: E731
        # neighbour of arc a: (i+1, j), (i+1, j+1), (i, j+1); the padding that puts it back in place
        nbs = (lambda x: sq(x)[..., :n - 1, 1:], lambda x: sq(x)[..., 1:, 1:], lambda x: sq(x)[..., 1:, :n - 1])
        pads = ((1, 0, 0, 1), (1, 0, 1, 0), (0, 1, 1, 0))
        flat = lambda x: x.reshape(x.shape[:-2] + (n * n,))  # noqa: E731
        planes = [None] * 6
        for a in range(3):
            xa = xc[:, a]  # [d][l][k][g]

            def rec(sel):
                r = None
                for d in range(3):
                    t = sel(grad[d]) * (org(xa[d]) - sel(self.X[d])[:, None])
                    r = t if r is None else r + t
                return sel(q) + r

            qs, qn = rec(org), rec(nbs[a])
            cs, cn = org(cm[a]), nbs[a](cm[a + 3])
            planes[a] = flat(F.pad(cs * qn + (1.0 - cs) * qs, (0, 1, 0, 1)))
            planes[a + 3] = flat(F.pad(cn * qs + (1.0 - cn) * qn, pads[a]))
        q_a = torch.stack(planes)
        rec_pl = []
        for v in range(1, self.spec.gall_pl):
            x = xc_pl[:, :, :, v]  # [d][p][k]
            qs = q_pl[:, :, 0] + sum(grad_pl[d][:, :, 0] * (x[d] - self.X_pl[d][:, None, 0]) for d in range(3))
            qn = q_pl[:, :, v] + sum(grad_pl[d][:, :, v] * (x[d] - self.X_pl[d][:, None, v]) for d in range(3))
            c = cm_pl[:, :, v]
            rec_pl.append(c * qn + (1.0 - c) * qs)
        q_a_pl = torch.stack([torch.zeros_like(q_pl[..., 0])] + rec_pl, dim=-1)
        return q_a, q_a_pl

    def hupdate(self, rq, rq_pl, flx, flx_pl, q_a, q_a_pl):
        s = flx[0] * q_a[0]
        for m in range(1, 6):
            s = s + flx[m] * q_a[m]
        out = torch.where(self.inner, rq - s, rq)
        sp = (flx_pl[..., 1:] * q_a_pl[..., 1:]).sum(-1)
        out_pl = torch.cat([(rq_pl[..., 0] - sp)[..., None], rq_pl[..., 1:]], dim=-1)
        return out, out_pl

    def vertical(self, q, q_pl, rq, rq_pl, rhog, rhog_pl, flx_v, flx_v_pl):
        """ck, q_h, the vertical limiter, the flux-form update of rho*q and rho."""
        kall = self.kall
        rd = self.rdgz[:, None]
        nxt = lambda a: torch.cat([a[..., 1:, :], torch.zeros_like(a[..., :1, :])], dim=-2)  # noqa: E731  k+1
        prv = lambda a: torch.cat([torch.zeros_like(a[..., :1, :]), a[..., :-1, :]], dim=-2)  # noqa: E731  k-1
        ck = torch.stack([-flx_v / rhog * rd * self.kface, nxt(flx_v) / rhog * rd * nxt(self.kface) * self.kin])
        ck_pl = torch.stack([-flx_v_pl / rhog_pl * rd * self.kface,
                             nxt(flx_v_pl) / rhog_pl * rd * nxt(self.kface) * self.kin])
        q_h = (self.afact[:, None] * q + self.bfact[:, None] * prv(q)) * self.kface
        q_h_pl = (self.afact[:, None] * q_pl + self.bfact[:, None] * prv(q_pl)) * self.kface
        outs, outs_pl = [], []
        for t in range(q.shape[0]):
            a, b = VerticalLimiter.apply(self.ctx, q_h[t].contiguous(), q_h_pl[t].contiguous(), q[t].contiguous(),
                                         q_pl[t].contiguous(), self.dv, self.dv_pl, ck.contiguous(),
                                         ck_pl.contiguous())
            outs.append(a)
            outs_pl.append(b)
        q_h, q_h_pl = torch.stack(outs), torch.stack(outs_pl)
        F, F_pl = flx_v * q_h * self.kface, flx_v_pl * q_h_pl * self.kface
        rq = rq - (nxt(F) - F) * rd * self.kin
        rq_pl = rq_pl - (nxt(F_pl) - F_pl) * rd * self.kin
        fv, fv_pl = flx_v * self.kface, flx_v_pl * self.kface
        rhog = rhog - (nxt(fv) - fv) * rd * self.kin
        rhog_pl = rhog_pl - (nxt(fv_pl) - fv_pl) * rd * self.kin
        return rq, rq_pl, rhog, rhog_pl

    # -- the step ----------------------------------------------------------------
    def face_values(self, q, q_pl, xc, xc_pl, cm, cm_pl, ch, ch_pl):
        """One tracer's limited face values: gradient, exchange, remap, the
        limiter's allowable range, its exchange, the clamp."""
        ctx = self.ctx
        grad, grad_pl = self.gradient(q, q_pl)
        grad, grad_pl = self.halo(grad, grad_pl)
        q_a, q_a_pl = self.remap(q, q_pl, grad, grad_pl, cm, cm_pl, xc, xc_pl)
        zero, zero_pl = q.new_zeros((2,) + q.shape), q.new_zeros((2,) + q_pl.shape)
        q, q_pl, q_a, q_a_pl = q.contiguous(), q_pl.contiguous(), q_a.contiguous(), q_a_pl.contiguous()
        _, _, prev, prev_pl = HorizontalLimiter.apply(ctx, q_a, q_a_pl, q, q_pl, zero, zero_pl, self.d, self.d_pl,
                                                      ch, ch_pl, cm, cm_pl)
        post, post_pl = self.halo(prev, prev_pl)
        q_a, q_a_pl, _, _ = HorizontalLimiter.apply(ctx, q_a, q_a_pl, q, q_pl, post.contiguous(),
                                                    post_pl.contiguous(), self.d, self.d_pl, ch, ch_pl, cm, cm_pl)
        return q_a, q_a_pl

    def __call__(self, q, q_pl, rhog, rhog_pl, recompute_tracers: bool = False):
        """One step.  recompute_tracers: each tracer's face values are
        recomputed in the backward pass instead of kept (torch checkpointing),
        so one tracer's intermediates are alive at a time."""
        from torch.utils.checkpoint import checkpoint

        ctx, ntr = self.ctx, q.shape[0]
        rhog, rhog_pl = rhog.contiguous(), rhog_pl.contiguous()
        flx, flx_pl, xc, xc_pl = HorizontalFlux.apply(ctx, self.dt, rhog, rhog_pl, *self.rv)
        ch, ch_pl, cm, cm_pl = self.courant(flx, flx_pl, rhog, rhog_pl)
        ch, ch_pl = ch.contiguous(), ch_pl.contiguous()
        qas, qas_pl = [], []
        for t in range(ntr):
            args = (q[t], q_pl[t], xc, xc_pl, cm, cm_pl, ch, ch_pl)
            if recompute_tracers:
                q_a, q_a_pl = checkpoint(self.face_values, *args, use_reentrant=False)
            else:
                q_a, q_a_pl = self.face_values(*args)
            qas.append(q_a)
            qas_pl.append(q_a_pl)
        rq, rq_pl = self.hupdate(rhog * q, rhog_pl * q_pl, flx, flx_pl, torch.stack(qas, 1), torch.stack(qas_pl, 0))
        ones, ones_pl = torch.ones_like(flx), torch.ones_like(flx_pl)
        rhog1, rhog1_pl = self.hupdate(rhog, rhog_pl, flx, flx_pl, ones, ones_pl)
        q1, q1_pl = rq / rhog1, rq_pl / rhog1_pl
        # vertical: rho*w from the solve on the step's density, its flux fixed to a multiple of it
        w, w_pl = RhowSolve.apply(ctx, self.dt_vi, self.rhogw_in, self.rhogw_in_pl, *self._vi_args(rhog, rhog_pl))
        flx_v, flx_v_pl = self.alpha * w * self.dt, self.alpha * w_pl * self.dt
        rq2, rq2_pl, rhog2, rhog2_pl = self.vertical(q1, q1_pl, rq, rq_pl, rhog1, rhog1_pl, flx_v, flx_v_pl)
        q2, q2_pl = rq2 / rhog2, rq2_pl / rhog2_pl
        q2, q2_pl = self.halo(q2, q2_pl)
        rhog2, rhog2_pl = self.halo(rhog2, rhog2_pl)
        return q2, q2_pl, rhog2, rhog2_pl

    def cost(self, q, q_pl):
        """1/2 sum of q^2 over the inner cells, the pole centres and kmin..kmax."""
        w = self.inner.to(q.dtype) * self.kin
        return 0.5 * ((q * q) * w).sum() + 0.5 * ((q_pl[..., 0] * q_pl[..., 0]) * self.kin[:, 0]).sum()
