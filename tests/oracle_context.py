"""A stand-in for runtime.Context on CPU tensors whose kernels are the
oracle's C restatement: the flux, both limiters (masked forwards and
adjoints, with the Courant numbers' cotangents), the rho*w solve, the
horizontal and vertical glue, each with its adjoint, and the halo fill of a halo map
(halo_sim) with its transpose.  It lets the tests run tracer_ad.TracerStep, the end-to-end
gradient harness, on the CPU, and compare the GPU harness with it."""
import numpy as np
import torch

from halo_sim import apply_map, apply_map_adjoint
from test_adjoint import pack_hlim_mask, pack_vlim_mask

FLUX = ("rho", "rho_pl", "rhovx", "rhovx_pl", "rhovy", "rhovy_pl", "rhovz", "rhovz_pl")
VI = ("rhogw", "rhogw0", "preg0", "rhog0", "Srho", "Sw", "Spre")


def _np(t):
    return t.detach().numpy()


def _put(t, a):
    t.copy_(torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).reshape(t.shape))


class OracleContext:
    def __init__(self, oracle, spec, geo, vi, fac, hm):
        self.orc, self.spec, self.geo, self.hm = oracle, spec, geo, hm
        self.sgp = [int(x) for x in geo["have_sgp"]]
        self.vi = vi
        self.fac = (fac["rdgzh"], fac["afact"], fac["bfact"])
        self.rdgz = fac["rdgz"]
        self.kall, self.gall, self.lall, self.gall_pl = spec.kall, spec.gall, spec.lall, spec.gall_pl
        self.nreg, self.npl = spec.gall * spec.kall * spec.lall, spec.gall_pl * spec.kall * 2
        self.hmask = {}

    def close(self):
        pass

    # ---- halo fill and the horizontal glue ----------------------------------------
    def comm_data_transfer(self, var, var_pl, vmax=1, kmax=None):
        a, b = apply_map(self.hm, _np(var).reshape((vmax,) + var.shape[-3:]),
                         _np(var_pl).reshape((vmax,) + var_pl.shape[-3:]))
        _put(var, a)
        _put(var_pl, b)

    def comm_data_transfer_adjoint(self, var_b, var_b_pl, vmax=1, kmax=None):
        a, b = apply_map_adjoint(self.hm, _np(var_b).reshape((vmax,) + var_b.shape[-3:]),
                                 _np(var_b_pl).reshape((vmax,) + var_b_pl.shape[-3:]))
        _put(var_b, a)
        _put(var_b_pl, b)

    def tracer_courant(self, ch, ch_pl, cm, cm_pl, flx, flx_pl, rhog, rhog_pl):
        out = self.orc.tracer_courant(self.spec, _np(flx), _np(flx_pl), _np(rhog), _np(rhog_pl))
        for t, k in ((ch, "ch"), (ch_pl, "ch_pl"), (cm, "cmask"), (cm_pl, "cmask_pl")):
            _put(t, out[k])

    def tracer_courant_adjoint(self, fb, fb_pl, rb, rb_pl, chb, chb_pl, flx, flx_pl, rhog, rhog_pl):
        out = self.orc.tracer_courant_adjoint(self.spec, _np(chb), _np(chb_pl), _np(flx), _np(flx_pl), _np(rhog),
                                              _np(rhog_pl))
        for t, k in ((fb, "flx_b"), (fb_pl, "flx_b_pl"), (rb, "rhog_b"), (rb_pl, "rhog_b_pl")):
            _put(t, out[k])

    def oprt_gradient(self, grad, grad_pl, q, q_pl, cg, cg_pl):
        out = self.orc.oprt_gradient(self.spec, _np(q), _np(q_pl), _np(cg), _np(cg_pl))
        _put(grad, out["grad"])
        _put(grad_pl, out["grad_pl"])

    def oprt_gradient_adjoint(self, qb, qb_pl, gb, gb_pl, cg, cg_pl):
        out = self.orc.oprt_gradient_adjoint(self.spec, _np(gb), _np(gb_pl), _np(cg), _np(cg_pl))
        _put(qb, out["scl_b"])
        _put(qb_pl, out["scl_b_pl"])

    def horizontal_remap(self, q_a, q_a_pl, q, q_pl, grad, grad_pl, cm, cm_pl, xc, xc_pl, X, X_pl):
        out = self.orc.horizontal_remap(self.spec, _np(q_a), _np(q_a_pl), _np(q), _np(q_pl), _np(grad), _np(grad_pl),
                                        _np(cm), _np(cm_pl), _np(xc), _np(xc_pl), _np(X), _np(X_pl))
        _put(q_a, out["q_a"])
        _put(q_a_pl, out["q_a_pl"])

    def horizontal_remap_adjoint(self, qb, qb_pl, gb, gb_pl, xb, xb_pl, qab, qab_pl, grad, grad_pl, cm, cm_pl, xc,
                                 xc_pl, X, X_pl):
        out = self.orc.horizontal_remap_adjoint(self.spec, _np(qab), _np(qab_pl), _np(grad), _np(grad_pl), _np(cm),
                                                _np(cm_pl), _np(xc), _np(xc_pl), _np(X), _np(X_pl))
        for t, k in ((qb, "q_b"), (qb_pl, "q_b_pl"), (gb, "grad_b"), (gb_pl, "grad_b_pl"), (xb, "xc_b"),
                     (xb_pl, "xc_b_pl")):
            _put(t, out[k])

    def tracer_update(self, rq, rq_pl, flx, flx_pl, q_a, q_a_pl):
        out = self.orc.tracer_update(self.spec, _np(rq), _np(rq_pl), _np(flx), _np(flx_pl), _np(q_a), _np(q_a_pl))
        _put(rq, out["rhogq"])
        _put(rq_pl, out["rhogq_pl"])

    def tracer_update_adjoint(self, fb, fb_pl, qab, qab_pl, rqb, rqb_pl, flx, flx_pl, q_a, q_a_pl):
        out = self.orc.tracer_update_adjoint(self.spec, _np(rqb), _np(rqb_pl), _np(flx), _np(flx_pl), _np(q_a),
                                             _np(q_a_pl))
        for t, k in ((fb, "flx_b"), (fb_pl, "flx_b_pl"), (qab, "qa_b"), (qab_pl, "qa_b_pl")):
            _put(t, out[k])

    # ---- the vertical glue ----------------------------------------------------------
    def vertical_courant(self, ck, ck_pl, flx_v, flx_v_pl, rhog, rhog_pl):
        out = self.orc.vertical_courant(self.spec, self.rdgz, _np(flx_v), _np(flx_v_pl), _np(rhog), _np(rhog_pl))
        _put(ck, out["ck"])
        _put(ck_pl, out["ck_pl"])

    def vertical_courant_adjoint(self, fb, fb_pl, rb, rb_pl, ckb, ckb_pl, flx_v, flx_v_pl, rhog, rhog_pl):
        out = self.orc.vertical_courant_adjoint(self.spec, self.rdgz, _np(ckb), _np(ckb_pl), _np(flx_v),
                                                _np(flx_v_pl), _np(rhog), _np(rhog_pl))
        for t, k in ((fb, "flx_b"), (fb_pl, "flx_b_pl"), (rb, "rhog_b"), (rb_pl, "rhog_b_pl")):
            _put(t, out[k])

    def vertical_face(self, q_h, q_h_pl, q, q_pl):
        out = self.orc.vertical_face(self.spec, self.fac[1], self.fac[2], _np(q_h), _np(q_h_pl), _np(q), _np(q_pl))
        _put(q_h, out["q_h"])
        _put(q_h_pl, out["q_h_pl"])

    def vertical_face_adjoint(self, qb, qb_pl, qhb, qhb_pl):
        out = self.orc.vertical_face_adjoint(self.spec, self.fac[1], self.fac[2], _np(qhb), _np(qhb_pl))
        _put(qb, out["q_b"])
        _put(qb_pl, out["q_b_pl"])

    def tracer_update_vertical(self, rq, rq_pl, flx_v, flx_v_pl, q_h, q_h_pl):
        out = self.orc.tracer_update_vertical(self.spec, self.rdgz, _np(rq), _np(rq_pl), _np(flx_v), _np(flx_v_pl),
                                              _np(q_h), _np(q_h_pl))
        _put(rq, out["rhogq"])
        _put(rq_pl, out["rhogq_pl"])

    def tracer_update_vertical_adjoint(self, fb, fb_pl, qhb, qhb_pl, rqb, rqb_pl, flx_v, flx_v_pl, q_h, q_h_pl):
        out = self.orc.tracer_update_vertical_adjoint(self.spec, self.rdgz, _np(rqb), _np(rqb_pl), _np(flx_v),
                                                      _np(flx_v_pl), _np(q_h), _np(q_h_pl))
        for t, k in ((fb, "flx_b"), (fb_pl, "flx_b_pl"), (qhb, "q_h_b"), (qhb_pl, "q_h_b_pl")):
            _put(t, out[k])

    # ---- horizontal_flux --------------------------------------------------------
    def horizontal_flux(self, flx, flx_pl, xc, xc_pl, *X_dt):
        X, dt = X_dt[:-1], X_dt[-1]
        f = {k: _np(v) for k, v in zip(FLUX, X)}
        out = self.orc.horizontal_flux(self.spec, self.geo, f, dt)
        for t, k in ((flx, "flx_h"), (flx_pl, "flx_h_pl"), (xc, "GRD_xc"), (xc_pl, "GRD_xc_pl")):
            _put(t, out[k])

    def horizontal_flux_adjoint(self, fb, fb_pl, xb, xb_pl, *rest):
        Xb, X, dt = rest[:8], rest[8:16], rest[16]
        f = {k: _np(v) for k, v in zip(FLUX, X)}
        a = dict(flx_h_b=_np(fb), flx_h_b_pl=_np(fb_pl), GRD_xc_b=_np(xb), GRD_xc_b_pl=_np(xb_pl))
        out = self.orc.horizontal_flux_lin(self.spec, self.geo, f, dt, True, a)
        for t, k in zip(Xb, ("rho", "rho_pl", "rhovx", "rhovx_pl", "rhovy", "rhovy_pl", "rhovz", "rhovz_pl")):
            _put(t, out[k.replace("_pl", "") + "_b" + ("_pl" if k.endswith("_pl") else "")])

    # ---- horizontal limiter -----------------------------------------------------
    def horizontal_limiter_thuburn_mask(self, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cm, cm_pl, prev, prev_pl,
                                        post, post_pl, mask, mask_pl):
        exchange = post is None  # the library's exchange mode: Qout_prev exchanged, then the clamp on it
        z, z_pl = np.zeros((2,) + tuple(q.shape)), np.zeros((2,) + tuple(q_pl.shape))
        f = dict(q_a=_np(q_a).copy(), q_a_pl=_np(q_a_pl).copy(), q=_np(q), q_pl=_np(q_pl), d=_np(d),
                 d_pl=_np(d_pl), ch=_np(ch), ch_pl=_np(ch_pl), cmask=_np(cm), cmask_pl=_np(cm_pl),
                 Qout_post=z if exchange else _np(post), Qout_post_pl=z_pl if exchange else _np(post_pl))
        out = self.orc.horizontal_limiter_thuburn(self.spec, self.sgp, f)
        if exchange:
            f["Qout_post"], f["Qout_post_pl"] = apply_map(self.hm, out["Qout_prev"], out["Qout_prev_pl"])
            out = self.orc.horizontal_limiter_thuburn(self.spec, self.sgp, f)
            out["Qout_prev"], out["Qout_prev_pl"] = f["Qout_post"], f["Qout_post_pl"]
        m, mp = pack_hlim_mask(self.spec, f, out)
        _put(q_a, out["q_a"])
        _put(q_a_pl, out["q_a_pl"])
        _put(prev, out["Qout_prev"])
        _put(prev_pl, out["Qout_prev_pl"])
        mask.copy_(torch.from_numpy(m.reshape(mask.shape)))
        mask_pl.copy_(torch.from_numpy(mp.reshape(mask_pl.shape)))

    def horizontal_limiter_thuburn_adjoint(self, qab, qab_pl, qb, qb_pl, prevb, prevb_pl, postb, postb_pl, q_a,
                                           q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cm, cm_pl, post, post_pl, mask,
                                           mask_pl, ch_b=None, ch_b_pl=None):
        # a half pass (q_a_b or Qout_prev_b None) is the full one with that cotangent zero
        f = dict(q=_np(q), q_pl=_np(q_pl), d=_np(d), d_pl=_np(d_pl), ch=_np(ch), ch_pl=_np(ch_pl), cmask=_np(cm),
                 cmask_pl=_np(cm_pl), Qout_post=_np(post), Qout_post_pl=_np(post_pl))
        P2, P2_pl = (2,) + tuple(q.shape), (2,) + tuple(q_pl.shape)
        u = dict(q_a_b=np.zeros((6,) + tuple(q.shape)) if qab is None else _np(qab),
                 q_a_b_pl=np.zeros(tuple(q_pl.shape)) if qab_pl is None else _np(qab_pl),
                 Qprev_b=np.zeros(P2) if prevb is None else _np(prevb),
                 Qprev_b_pl=np.zeros(P2_pl) if prevb_pl is None else _np(prevb_pl))
        out = self.orc.horizontal_limiter_thuburn_lin(self.spec, self.sgp, f, dict(q_a=_np(q_a), q_a_pl=_np(q_a_pl)),
                                                      _np(mask).reshape(self.lall, self.kall, self.gall),
                                                      _np(mask_pl), True, u)
        for t, k in ((qab, "q_a_b"), (qab_pl, "q_a_b_pl"), (qb, "q_b"), (qb_pl, "q_b_pl"), (postb, "Qpost_b"),
                     (postb_pl, "Qpost_b_pl"), (ch_b, "ch_b"), (ch_b_pl, "ch_b_pl")):
            if t is not None:
                _put(t, out[k])

    # ---- vertical limiter -------------------------------------------------------
    def vlim_mask_words(self):
        nword = (self.kall - 3 + 31) // 32 if self.kall > 3 else 0
        return nword * self.gall * self.lall, nword * self.gall_pl * 2

    def vertical_limiter_thuburn_mask(self, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl, mask, mask_pl):
        f = dict(q_h=_np(q_h).copy(), q_h_pl=_np(q_h_pl).copy(), q=_np(q), q_pl=_np(q_pl), d=_np(d),
                 d_pl=_np(d_pl), ck=_np(ck), ck_pl=_np(ck_pl))
        out = self.orc.vertical_limiter_thuburn(self.spec, f)
        m, mp = pack_vlim_mask(self.spec, {"q_h": f["q_h"], "q_h_pl": f["q_h_pl"]}, out)
        _put(q_h, out["q_h"])
        _put(q_h_pl, out["q_h_pl"])
        mask.copy_(torch.from_numpy(m.reshape(-1).view(np.int32)))
        mask_pl.copy_(torch.from_numpy(mp.reshape(-1).view(np.int32)))

    def vertical_limiter_thuburn_adjoint(self, qhb, qhb_pl, qb, qb_pl, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl,
                                         mask, mask_pl, ck_b=None, ck_b_pl=None):
        f = dict(q=_np(q), q_pl=_np(q_pl), d=_np(d), d_pl=_np(d_pl), ck=_np(ck), ck_pl=_np(ck_pl))
        nword = (self.kall - 3 + 31) // 32
        m = _np(mask).view(np.uint32).reshape(self.lall, nword, self.gall)
        mp = _np(mask_pl).view(np.uint32).reshape(2, nword, self.gall_pl)
        u = dict(q_h_b=_np(qhb), q_h_b_pl=_np(qhb_pl))
        out = self.orc.vertical_limiter_thuburn_lin(self.spec, f, dict(q_h=_np(q_h), q_h_pl=_np(q_h_pl)), m, mp,
                                                    True, u)
        for t, k in ((qhb, "q_h_b"), (qhb_pl, "q_h_b_pl"), (qb, "q_b"), (qb_pl, "q_b_pl"), (ck_b, "ck_b"),
                     (ck_b_pl, "ck_b_pl")):
            if t is not None:
                _put(t, out[k])

    # ---- rho*w solve ------------------------------------------------------------
    def vi_rhow_solver(self, w, w_pl, *rest):
        ins, dt = rest[:-1], rest[-1]
        f = dict(self.vi, rhogw=_np(w), rhogw_pl=_np(w_pl))
        for i, n in enumerate(VI[1:]):
            f[n], f[n + "_pl"] = _np(ins[2 * i]), _np(ins[2 * i + 1])
        out = self.orc.vi_rhow_solver(self.spec, f, dt, *self.fac)
        _put(w, out["rhogw"])
        _put(w_pl, out["rhogw_pl"])

    def vi_rhow_solver_adjoint(self, *rest):
        bufs, dt = rest[:-1], rest[-1]
        out = self.orc.vi_rhow_solver_adjoint(self.spec, self.vi, {"rhogw": _np(bufs[0]), "rhogw_pl": _np(bufs[1])},
                                              dt, *self.fac)
        for i, n in enumerate(VI):
            _put(bufs[2 * i], out[n])
            _put(bufs[2 * i + 1], out[n + "_pl"])
