"""A differentiable tracer step: the adjoint test harness around the
hand-written adjoint kernels.

One step advances ntr tracers q and the density rhog of an icosahedral grid
(all regions and both poles on one GPU) over fixed winds:

    horizontal_flux(rhog, rho*v)              -> flx_h, GRD_xc    [HIP; adjoint HIP]
    ch = flx_h / rhog, cmask = upwind mask                          [HIP; adjoint HIP]
    grad = gradient(q), halo exchange; q_a = upwind remap            [HIP; adjoint HIP]
    Qout = phase 1 of the limiter, halo exchange, q_a limited        [HIP, one call; adjoint HIP]
    rhog*q and rhog updated by the face fluxes, q = rhog*q / rhog    [HIP; adjoint HIP]
    rho*w = vi_rhow_solver(..., rhog0 = rhog, ...)                   [HIP; adjoint HIP]
    flx_v = alpha * rho*w * dt                                       [torch]
    ck; q_h = half-level interpolation                               [HIP; adjoint HIP]
    q_h limited by vertical_limiter_thuburn                          [HIP; adjoint HIP]
    rhog*q and rhog updated by the vertical fluxes; halo exchange    [HIP; adjoint HIP]

The four reference kernels, the horizontal and vertical glue (Courant
numbers and masks, gradient, remap, half-level values, updates) and the
halo exchange run as torch.autograd.Functions over the context's forward
entry points (the masked forwards for the limiters) and their *_adjoint
entry points, each deterministic; the context is the library
(runtime.Context, which needs its halo map, nicam_comm_setup) or, in the
tests, the oracle.  What is left to torch's autograd is elementwise: the
products rhog*q, the quotients rhog*q / rhog and the scaling of rho*w into
flx_v.  The limiters' adjoints return the cotangents of the Courant
numbers ch and ck, through which their bounds depend on rhog; what stays
fixed is piecewise constant (cmask and the inflow flags, the signs of the
Courant numbers) or zero here (d), so the gradient in q and rhog is the
step's own wherever no switch sits exactly at its threshold.

Arrays are C arrays with the Fortran dimensions reversed: q[t][l][k][g],
q_pl[t][p][k][v]; rhog[l][k][g], rhog_pl[p][k][v].
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import numpy as np
import torch

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
    """q_a' = horizontal_limiter_thuburn(q_a, q, ch; d, cmask) (HLIM:42-430) in
    the library's exchange mode: the allowable range Qout_prev, its halo
    exchange (Qout_post) and the clamp, one masked call; d and cmask held fixed.
    Its transpose: the faces' half pass of the limiter's adjoint (q_a_b ->
    q_a's, q's and Qout_post's cotangents), the halo exchange's transpose
    (Qout_post's -> Qout_prev's) and the rows' half pass (Qout_prev's -> q's and
    ch's)."""

    @staticmethod
    def forward(fctx, nctx: Context, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cm, cm_pl):
        out, out_pl = q_a.clone(), q_a_pl.clone()
        post, post_pl = q.new_zeros((2,) + q.shape), q_pl.new_zeros((2,) + q_pl.shape)
        mask = torch.empty(q.shape, dtype=torch.uint8, device=q.device)
        mask_pl = torch.empty(q_pl.shape, dtype=torch.uint8, device=q.device)
        nctx.horizontal_limiter_thuburn_mask(out, out_pl, q, q_pl, d, d_pl, ch, ch_pl, cm, cm_pl, post, post_pl,
                                             None, None, mask, mask_pl)
        fctx.nctx = nctx
        fctx.save_for_backward(out, out_pl, q, q_pl, post, post_pl, d, d_pl, ch, ch_pl, cm, cm_pl, mask, mask_pl)
        return out, out_pl

    @staticmethod
    def backward(fctx, g_qa, g_qa_pl):
        out, out_pl, q, q_pl, post, post_pl, d, d_pl, ch, ch_pl, cm, cm_pl, mask, mask_pl = fctx.saved_tensors
        nctx = fctx.nctx
        qab = torch.zeros_like(out) if g_qa is None else g_qa.contiguous().clone()
        qab_pl = torch.zeros_like(out_pl) if g_qa_pl is None else g_qa_pl.contiguous().clone()
        qb, qb_pl = torch.empty_like(q), torch.empty_like(q_pl)
        pb, pb_pl = torch.empty_like(post), torch.empty_like(post_pl)
        fixed = (out, out_pl, q, q_pl, d, d_pl, ch, ch_pl, cm, cm_pl, post, post_pl, mask, mask_pl)
        nctx.horizontal_limiter_thuburn_adjoint(qab, qab_pl, qb, qb_pl, None, None, pb, pb_pl, *fixed)
        nctx.comm_data_transfer_adjoint(pb, pb_pl, vmax=2, kmax=q.shape[-2])
        qb2, qb2_pl = torch.empty_like(q), torch.empty_like(q_pl)
        chb, chb_pl = torch.empty_like(ch), torch.empty_like(ch_pl)
        nctx.horizontal_limiter_thuburn_adjoint(None, None, qb2, qb2_pl, pb, pb_pl, None, None, *fixed, chb, chb_pl)
        return (None, qab, qab_pl, qb + qb2, qb_pl + qb2_pl, None, None, chb, chb_pl, None, None)


def _c(g, ref):
    """A cotangent autograd may give as None, as a contiguous tensor."""
    return torch.zeros_like(ref) if g is None else g.contiguous()


class Halo(torch.autograd.Function):
    """COMM_data_transfer of x [vmax][l][k][g], x_pl [vmax][p][k][v] (the
    library's on-device halo fill, on a copy); its transpose
    (nicam_comm_data_transfer_adjoint) adds every destination's cotangent to
    its source's in a fixed order and clears the destination's."""

    @staticmethod
    def forward(fctx, nctx: Context, vmax: int, x, x_pl):
        y, y_pl = x.contiguous().clone(), x_pl.contiguous().clone()
        nctx.comm_data_transfer(y, y_pl, vmax=vmax, kmax=x.shape[-2])
        fctx.nctx, fctx.vmax = nctx, vmax
        return y, y_pl

    @staticmethod
    def backward(fctx, g, g_pl):
        b, b_pl = _c(g, g).clone(), _c(g_pl, g_pl).clone()
        fctx.nctx.comm_data_transfer_adjoint(b, b_pl, vmax=fctx.vmax, kmax=b.shape[-2])
        return None, None, b, b_pl


class Courant(torch.autograd.Function):
    """ch = flx_h / rhog and the upwind masks (nicam_tracer_courant); the masks
    are piecewise constant and carry no cotangent."""

    @staticmethod
    def forward(fctx, nctx: Context, flx, flx_pl, rhog, rhog_pl):
        ch, ch_pl = torch.empty_like(flx), torch.empty_like(flx_pl)
        cm, cm_pl = torch.empty_like(flx), torch.empty_like(flx_pl)
        nctx.tracer_courant(ch, ch_pl, cm, cm_pl, flx, flx_pl, rhog, rhog_pl)
        fctx.nctx = nctx
        fctx.save_for_backward(flx, flx_pl, rhog, rhog_pl)
        fctx.mark_non_differentiable(cm, cm_pl)
        return ch, ch_pl, cm, cm_pl

    @staticmethod
    def backward(fctx, g_ch, g_ch_pl, g_cm, g_cm_pl):
        flx, flx_pl, rhog, rhog_pl = fctx.saved_tensors
        fb, fb_pl = torch.empty_like(flx), torch.empty_like(flx_pl)
        rb, rb_pl = torch.empty_like(rhog), torch.empty_like(rhog_pl)
        fctx.nctx.tracer_courant_adjoint(fb, fb_pl, rb, rb_pl, _c(g_ch, flx), _c(g_ch_pl, flx_pl), flx, flx_pl, rhog,
                                         rhog_pl)
        return None, fb, fb_pl, rb, rb_pl


class Gradient(torch.autograd.Function):
    """grad = the 7-point stencil of coef_grad on q (nicam_oprt_gradient)."""

    @staticmethod
    def forward(fctx, nctx: Context, cg, cg_pl, q, q_pl):
        grad = q.new_empty((3,) + q.shape)
        grad_pl = q_pl.new_empty((3,) + q_pl.shape)
        nctx.oprt_gradient(grad, grad_pl, q, q_pl, cg, cg_pl)
        fctx.nctx, fctx.cg, fctx.cg_pl = nctx, cg, cg_pl
        return grad, grad_pl

    @staticmethod
    def backward(fctx, g, g_pl):
        b, b_pl = g.new_empty(g.shape[1:]), g_pl.new_empty(g_pl.shape[1:])
        fctx.nctx.oprt_gradient_adjoint(b, b_pl, _c(g, g), _c(g_pl, g_pl), fctx.cg, fctx.cg_pl)
        return None, None, None, b, b_pl


class Remap(torch.autograd.Function):
    """Upwind face values q + grad . (GRD_xc - GRD_x) (nicam_horizontal_remap)
    into zeroed arrays; cmask and GRD_x held fixed."""

    @staticmethod
    def forward(fctx, nctx: Context, cm, cm_pl, X, X_pl, q, q_pl, grad, grad_pl, xc, xc_pl):
        q_a, q_a_pl = q.new_zeros((6,) + q.shape), q_pl.new_zeros(q_pl.shape)
        nctx.horizontal_remap(q_a, q_a_pl, q, q_pl, grad, grad_pl, cm, cm_pl, xc, xc_pl, X, X_pl)
        fctx.nctx = nctx
        fctx.save_for_backward(grad, grad_pl, cm, cm_pl, xc, xc_pl, X, X_pl)
        return q_a, q_a_pl

    @staticmethod
    def backward(fctx, g, g_pl):
        grad, grad_pl, cm, cm_pl, xc, xc_pl, X, X_pl = fctx.saved_tensors
        qb, qb_pl = grad.new_empty(grad.shape[1:]), grad_pl.new_empty(grad_pl.shape[1:])
        gb, gb_pl = torch.empty_like(grad), torch.empty_like(grad_pl)
        xb, xb_pl = torch.empty_like(xc), torch.empty_like(xc_pl)
        g = qb.new_zeros((6,) + qb.shape) if g is None else g.contiguous()
        fctx.nctx.horizontal_remap_adjoint(qb, qb_pl, gb, gb_pl, xb, xb_pl, g, _c(g_pl, qb_pl), grad, grad_pl, cm,
                                           cm_pl, xc, xc_pl, X, X_pl)
        return None, None, None, None, None, qb, qb_pl, gb, gb_pl, xb, xb_pl


class Update(torch.autograd.Function):
    """rhogq - the sum over the faces of flx_h * q_a on the inner cells and the
    pole centres (nicam_tracer_update), for one tracer."""

    @staticmethod
    def forward(fctx, nctx: Context, rq, rq_pl, flx, flx_pl, q_a, q_a_pl):
        out, out_pl = rq.contiguous().clone(), rq_pl.contiguous().clone()
        nctx.tracer_update(out, out_pl, flx, flx_pl, q_a, q_a_pl)
        fctx.nctx = nctx
        fctx.save_for_backward(flx, flx_pl, q_a, q_a_pl)
        return out, out_pl

    @staticmethod
    def backward(fctx, g, g_pl):
        flx, flx_pl, q_a, q_a_pl = fctx.saved_tensors
        g, g_pl = _c(g, flx[0]), _c(g_pl, flx_pl)
        fb, fb_pl = torch.empty_like(flx), torch.empty_like(flx_pl)
        qb, qb_pl = torch.empty_like(q_a), torch.empty_like(q_a_pl)
        fctx.nctx.tracer_update_adjoint(fb, fb_pl, qb, qb_pl, g, g_pl, flx, flx_pl, q_a, q_a_pl)
        return None, g, g_pl, fb, fb_pl, qb, qb_pl


class VCourant(torch.autograd.Function):
    """ck [2][l][k][g] of the lower and upper faces (nicam_vertical_courant);
    its transpose gives the cotangents of flx_v and rhog."""

    @staticmethod
    def forward(fctx, nctx: Context, flx_v, flx_v_pl, rhog, rhog_pl):
        ck, ck_pl = flx_v.new_empty((2,) + flx_v.shape), flx_v_pl.new_empty((2,) + flx_v_pl.shape)
        nctx.vertical_courant(ck, ck_pl, flx_v, flx_v_pl, rhog, rhog_pl)
        fctx.nctx = nctx
        fctx.save_for_backward(flx_v, flx_v_pl, rhog, rhog_pl)
        return ck, ck_pl

    @staticmethod
    def backward(fctx, g, g_pl):
        flx_v, flx_v_pl, rhog, rhog_pl = fctx.saved_tensors
        fb, fb_pl = torch.empty_like(flx_v), torch.empty_like(flx_v_pl)
        rb, rb_pl = torch.empty_like(rhog), torch.empty_like(rhog_pl)
        g = flx_v.new_zeros((2,) + flx_v.shape) if g is None else g.contiguous()
        g_pl = flx_v_pl.new_zeros((2,) + flx_v_pl.shape) if g_pl is None else g_pl.contiguous()
        fctx.nctx.vertical_courant_adjoint(fb, fb_pl, rb, rb_pl, g, g_pl, flx_v, flx_v_pl, rhog, rhog_pl)
        return None, fb, fb_pl, rb, rb_pl


class VFace(torch.autograd.Function):
    """q_h = afact q(k) + bfact q(k-1) on the interior half levels, 0 on the
    others (nicam_vertical_face into a zeroed array), for one tracer."""

    @staticmethod
    def forward(fctx, nctx: Context, q, q_pl):
        q_h, q_h_pl = torch.zeros_like(q), torch.zeros_like(q_pl)
        nctx.vertical_face(q_h, q_h_pl, q, q_pl)
        fctx.nctx = nctx
        fctx.save_for_backward(q, q_pl)
        return q_h, q_h_pl

    @staticmethod
    def backward(fctx, g, g_pl):
        q, q_pl = fctx.saved_tensors
        qb, qb_pl = torch.empty_like(q), torch.empty_like(q_pl)
        fctx.nctx.vertical_face_adjoint(qb, qb_pl, _c(g, q), _c(g_pl, q_pl))
        return None, qb, qb_pl


class VUpdate(torch.autograd.Function):
    """rhogq less the vertical flux divergence of flx_v * q_h on kmin..kmax
    (nicam_tracer_update_vertical), for one tracer; with q_h = 1 the density."""

    @staticmethod
    def forward(fctx, nctx: Context, rq, rq_pl, flx_v, flx_v_pl, q_h, q_h_pl):
        out, out_pl = rq.contiguous().clone(), rq_pl.contiguous().clone()
        nctx.tracer_update_vertical(out, out_pl, flx_v, flx_v_pl, q_h, q_h_pl)
        fctx.nctx = nctx
        fctx.save_for_backward(flx_v, flx_v_pl, q_h, q_h_pl)
        return out, out_pl

    @staticmethod
    def backward(fctx, g, g_pl):
        flx_v, flx_v_pl, q_h, q_h_pl = fctx.saved_tensors
        g, g_pl = _c(g, flx_v), _c(g_pl, flx_v_pl)
        fb, fb_pl = torch.empty_like(flx_v), torch.empty_like(flx_v_pl)
        qb, qb_pl = torch.empty_like(q_h), torch.empty_like(q_h_pl)
        fctx.nctx.tracer_update_vertical_adjoint(fb, fb_pl, qb, qb_pl, g, g_pl, flx_v, flx_v_pl, q_h, q_h_pl)
        return None, g, g_pl, fb, fb_pl, qb, qb_pl


class HaloExchange:
    """COMM_data_transfer through the context: x [..., lall, kall, gall],
    x_pl [..., 2, kall, gall_pl], the leading dimensions as variables."""

    def __init__(self, nctx: Context):
        self.nctx = nctx

    def __call__(self, x, x_pl):
        lead = x.shape[:-3]
        vmax = int(np.prod(lead)) if lead else 1
        y, y_pl = Halo.apply(self.nctx, vmax, x.contiguous().reshape((vmax,) + x.shape[-3:]),
                             x_pl.contiguous().reshape((vmax,) + x_pl.shape[-3:]))
        return y.reshape(x.shape), y_pl.reshape(x_pl.shape)


class TracerStep:
    """The step of the module docstring for one grid; call it with
    (q, q_pl, rhog, rhog_pl) and get the same four back, one step on."""

    def __init__(self, ctx: Context, spec, geo, case, fac, dt: float = 360.0, dt_vi: float = 60.0,
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
        self.halo = HaloExchange(ctx)  # the context holds the halo map (nicam_comm_setup)
        vi = case["vi"]
        self.vi = {k + s: T(vi[k + s]) for k in VI_IN for s in ("", "_pl")}
        self.rhogw_in, self.rhogw_in_pl = T(vi["rhogw"]), T(vi["rhogw_pl"])
        self.rdgz = T(fac["rdgz"])
        inner = np.zeros((n1, n1), dtype=bool)
        inner[1:-1, 1:-1] = True
        self.inner = torch.from_numpy(inner.reshape(-1)).to(dev)
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
        return Courant.apply(self.ctx, flx, flx_pl, rhog, rhog_pl)

    def gradient(self, q, q_pl):
        """q [l][k][g] -> grad [3][l][k][g] on the flat range gmin..gmax (zero
        elsewhere); the pole centre from its ring."""
        return Gradient.apply(self.ctx, self.cg, self.cg_pl, q.contiguous(), q_pl.contiguous())

    def remap(self, q, q_pl, grad, grad_pl, cm, cm_pl, xc, xc_pl):
        """Upwind face values q + grad.(GRD_xc - GRD_x) of every arc from the
        origins i, j in [gmin-1, gmax]; the faces no arc reaches stay 0."""
        return Remap.apply(self.ctx, cm, cm_pl, self.X, self.X_pl, q.contiguous(), q_pl.contiguous(),
                           grad.contiguous(), grad_pl.contiguous(), xc, xc_pl)

    def hupdate(self, rq, rq_pl, flx, flx_pl, q_a, q_a_pl):
        """rq [l][k][g] (one tracer) less the sum of flx * q_a over the faces."""
        return Update.apply(self.ctx, rq, rq_pl, flx, flx_pl, q_a.contiguous(), q_a_pl.contiguous())

    def vertical(self, q, q_pl, rq, rq_pl, rhog, rhog_pl, flx_v, flx_v_pl):
        """ck, q_h, the vertical limiter, the flux-form update of rho*q and rho."""
        ctx = self.ctx
        flx_v, flx_v_pl, rhog, rhog_pl = flx_v.contiguous(), flx_v_pl.contiguous(), rhog.contiguous(), rhog_pl.contiguous()
        ck, ck_pl = VCourant.apply(ctx, flx_v, flx_v_pl, rhog, rhog_pl)
        ck, ck_pl = ck.contiguous(), ck_pl.contiguous()
        outs, outs_pl = [], []
        for t in range(q.shape[0]):
            qt, qt_pl = q[t].contiguous(), q_pl[t].contiguous()
            q_h, q_h_pl = VFace.apply(ctx, qt, qt_pl)
            a, b = VerticalLimiter.apply(ctx, q_h, q_h_pl, qt, qt_pl, self.dv, self.dv_pl, ck, ck_pl)
            a, b = VUpdate.apply(ctx, rq[t], rq_pl[t], flx_v, flx_v_pl, a.contiguous(), b.contiguous())
            outs.append(a)
            outs_pl.append(b)
        ones, ones_pl = torch.ones_like(rhog), torch.ones_like(rhog_pl)
        rhog, rhog_pl = VUpdate.apply(ctx, rhog, rhog_pl, flx_v, flx_v_pl, ones, ones_pl)
        return torch.stack(outs), torch.stack(outs_pl), rhog, rhog_pl

    # -- the step ----------------------------------------------------------------
    def face_values(self, q, q_pl, xc, xc_pl, cm, cm_pl, ch, ch_pl):
        """One tracer's limited face values: gradient, exchange, remap, the
        limiter's allowable range, its exchange, the clamp."""
        ctx = self.ctx
        grad, grad_pl = self.gradient(q, q_pl)
        grad, grad_pl = self.halo(grad, grad_pl)
        q_a, q_a_pl = self.remap(q, q_pl, grad, grad_pl, cm, cm_pl, xc, xc_pl)
        q, q_pl, q_a, q_a_pl = q.contiguous(), q_pl.contiguous(), q_a.contiguous(), q_a_pl.contiguous()
        return HorizontalLimiter.apply(ctx, q_a, q_a_pl, q, q_pl, self.d, self.d_pl, ch, ch_pl, cm, cm_pl)

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
        rqs = [self.hupdate(rhog * q[t], rhog_pl * q_pl[t], flx, flx_pl, qas[t], qas_pl[t]) for t in range(ntr)]
        rq, rq_pl = torch.stack([a for a, _ in rqs]), torch.stack([b for _, b in rqs])
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
