"""Kernel drivers: the GPU counterparts of the reference's snapshot-replay
programs (kernels/NICAM/dyn_*/src/main.f90).

Each ``run_*`` function sets up nothing itself: it takes a :class:`Context`
whose metrics/coefficients are already in place, moves the kernel's inputs to
the device, calls the kernel once and returns the outputs as numpy arrays.
``setup_context`` performs the set-up the reference drivers do before their
kernel loops (CONST_setup, GRD_setup, metric arrays, vi_rhow_solver_putcoef).
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np

from .runtime import PTR_DEVICE, PTR_HOST, Context
from .synth import GridSpec

FLUX_IN = ("rho", "rho_pl", "rhovx", "rhovx_pl", "rhovy", "rhovy_pl", "rhovz", "rhovz_pl")
HLIM_IN = ("q", "q_pl", "d", "d_pl", "ch", "ch_pl", "cmask", "cmask_pl", "Qout_post", "Qout_post_pl")
VLIM_IN = ("q", "q_pl", "d", "d_pl", "ck", "ck_pl")
VI_IN = ("rhogw0", "preg0", "rhog0", "Srho", "Sw", "Spre")


def setup_context(spec: GridSpec, geo: Dict[str, np.ndarray], vi: Optional[Dict[str, np.ndarray]] = None,
                  pointer_mode: int = PTR_DEVICE) -> Context:
    ctx = Context(spec.gall_1d, spec.kall, spec.lall, spec.gall_pl, spec.have_pl, list(geo["have_sgp"]),
                  pointer_mode=pointer_mode)
    ctx.const_setup()
    ctx.set_hmetrics(geo["GRD_xr"], geo["GRD_xr_pl"], geo["GMTR_p"], geo["GMTR_p_pl"], geo["GMTR_t"],
                     geo["GMTR_t_pl"], geo["GMTR_a"], geo["GMTR_a_pl"])
    if vi is not None:
        ctx.grd_setup(vi["gz"], vi["gzh"])
        ctx.set_vmetrics(vi["RGAM"], vi["RGAM_pl"], vi["RGAMH"], vi["RGAMH_pl"], vi["RGSGAM2"], vi["RGSGAM2_pl"],
                         vi["RGSGAM2H"], vi["RGSGAM2H_pl"], vi["GSGAM2H"], vi["GSGAM2H_pl"])
        ctx.vi_rhow_solver_putcoef(vi["Mc"], vi["Mc_pl"], vi["Ml"], vi["Ml_pl"], vi["Mu"], vi["Mu_pl"])
    return ctx


def _dev(x: np.ndarray, device):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(device)


def _host(t) -> np.ndarray:
    return t.detach().cpu().numpy()


def _prep(ctx: Context, arrays: Dict[str, np.ndarray], device):
    if ctx.mode == PTR_HOST:
        return {k: np.ascontiguousarray(v, dtype=np.float64).copy() for k, v in arrays.items()}
    return {k: _dev(v, device) for k, v in arrays.items()}


def _back(ctx: Context, arrays):
    if ctx.mode == PTR_HOST:
        return arrays
    import torch

    torch.cuda.synchronize()
    return {k: _host(v) for k, v in arrays.items()}


def run_flux(ctx: Context, spec: GridSpec, f: Dict[str, np.ndarray], dt: float, fill: float = 0.0,
             device="cuda:0") -> Dict[str, np.ndarray]:
    outs = dict(flx_h=np.full((6,) + spec.shape3, fill), flx_h_pl=np.full(spec.shape_pl, fill),
                GRD_xc=np.full((3, 3) + spec.shape3, fill), GRD_xc_pl=np.full((3,) + spec.shape_pl, fill))
    a = _prep(ctx, {**outs, **{k: f[k] for k in FLUX_IN}}, device)
    ctx.horizontal_flux(a["flx_h"], a["flx_h_pl"], a["GRD_xc"], a["GRD_xc_pl"], *[a[k] for k in FLUX_IN], dt)
    return _back(ctx, {k: a[k] for k in outs})


def run_hlim(ctx: Context, spec: GridSpec, f: Dict[str, np.ndarray], fill: float = 0.0,
             device="cuda:0", exchange: bool = False) -> Dict[str, np.ndarray]:
    """exchange=True passes Qout_post = NULL: the library exchanges Qout itself
    (needs ctx.comm_setup)."""
    outs = dict(q_a=f["q_a"], q_a_pl=f["q_a_pl"], Qout_prev=np.full((2,) + spec.shape3, fill),
                Qout_prev_pl=np.full((2,) + spec.shape_pl, fill))
    ins = [k for k in HLIM_IN if not (exchange and k.startswith("Qout_post"))]
    a = _prep(ctx, {**outs, **{k: f[k] for k in ins}}, device)
    ctx.horizontal_limiter_thuburn(a["q_a"], a["q_a_pl"], a["q"], a["q_pl"], a["d"], a["d_pl"], a["ch"], a["ch_pl"],
                                   a["cmask"], a["cmask_pl"], a["Qout_prev"], a["Qout_prev_pl"], a.get("Qout_post"),
                                   a.get("Qout_post_pl"))
    return _back(ctx, {k: a[k] for k in outs})


def run_vlim(ctx: Context, spec: GridSpec, f: Dict[str, np.ndarray], device="cuda:0") -> Dict[str, np.ndarray]:
    outs = dict(q_h=f["q_h"], q_h_pl=f["q_h_pl"])
    a = _prep(ctx, {**outs, **{k: f[k] for k in VLIM_IN}}, device)
    ctx.vertical_limiter_thuburn(a["q_h"], a["q_h_pl"], a["q"], a["q_pl"], a["d"], a["d_pl"], a["ck"], a["ck_pl"])
    return _back(ctx, {k: a[k] for k in outs})


def run_vi(ctx: Context, spec: GridSpec, f: Dict[str, np.ndarray], dt: float,
           device="cuda:0") -> Dict[str, np.ndarray]:
    outs = dict(rhogw=f["rhogw"], rhogw_pl=f["rhogw_pl"])
    ins = {}
    for k in VI_IN:
        ins[k] = f[k]
        ins[k + "_pl"] = f[k + "_pl"]
    a = _prep(ctx, {**outs, **ins}, device)
    args = []
    for k in VI_IN:
        args += [a[k], a[k + "_pl"]]
    ctx.vi_rhow_solver(a["rhogw"], a["rhogw_pl"], *args, dt)
    return _back(ctx, {k: a[k] for k in outs})


DIFF_IN = ("scl", "scl_pl", "kh", "kh_pl", "coef_intp", "coef_intp_pl", "coef_diff", "coef_diff_pl")


def run_diffusion(ctx: Context, spec: GridSpec, f: Dict[str, np.ndarray], fill: float = 0.0,
                  device="cuda:0") -> Dict[str, np.ndarray]:
    outs = dict(dscl=np.full(spec.shape3, fill), dscl_pl=np.full(spec.shape_pl, fill))
    a = _prep(ctx, {**outs, **{k: f[k] for k in DIFF_IN}}, device)
    ctx.oprt_diffusion(a["dscl"], a["dscl_pl"], *[a[k] for k in DIFF_IN])
    return _back(ctx, {k: a[k] for k in outs})


DDMP_IN = ("rhogvx", "rhogvx_pl", "rhogvy", "rhogvy_pl", "rhogvz", "rhogvz_pl", "rhogw", "rhogw_pl",
           "coef_intp", "coef_intp_pl", "coef_diff", "coef_diff_pl")
DDMP_OUT = ("ddivdx", "ddivdx_pl", "ddivdy", "ddivdy_pl", "ddivdz", "ddivdz_pl")


def setup_divdamp(ctx: Context, f: Dict[str, np.ndarray]) -> None:
    """The module state OPRT3D_divdamp reads: VMTR_* and GRD_rdgz."""
    ctx.set_vmetrics3d(f["RGSQRTH"], f["RGSQRTH_pl"], f["RGAM"], f["RGAM_pl"], f["RGAMH"], f["RGAMH_pl"],
                       f["C2WfactGz"], f["C2WfactGz_pl"])
    ctx.set_rdgz(f["rdgz"])


def run_divdamp(ctx: Context, spec: GridSpec, f: Dict[str, np.ndarray], fill: float = 0.0,
                device="cuda:0") -> Dict[str, np.ndarray]:
    outs = {k: np.full(spec.shape_pl if k.endswith("_pl") else spec.shape3, fill) for k in DDMP_OUT}
    a = _prep(ctx, {**outs, **{k: f[k] for k in DDMP_IN}}, device)
    ctx.oprt3d_divdamp(*[a[k] for k in DDMP_OUT], *[a[k] for k in DDMP_IN])
    return _back(ctx, {k: a[k] for k in outs})


# ---- one chained tracer step, FP64 or mixed precision ------------------------
def chain_state(ctx: Context, spec: GridSpec, case: Dict[str, Dict[str, np.ndarray]], mixed: bool, device="cuda:0"):
    """Device tensors of one chained step on the case's inputs: the flux's
    (rho, rho*v), the limiters' (q, d, ch, cmask, Qout_post, ck, pre-limit
    q_a and q_h), the vertical mass flux flx_v (from the lower-face Courant
    number: flx_v(k) = -ck(k,1) rhog(k) / rdgz(k)), the rho*w solve's, and the
    tracer mass rhogq = rho*q.  With mixed=True every input of the flux, the
    limiters and the solve's sources is rounded to FP32; rhogq and rhogw stay
    FP64 (the accumulated state)."""
    import torch

    dt = torch.float32 if mixed else torch.float64

    def up(x, typ=dt):
        return torch.from_numpy(np.ascontiguousarray(x)).to(device=device, dtype=typ)

    f, h, v, vi = case["flux"], case["hlim"], case["vlim"], case["vi"]
    S = {k: up(f[k]) for k in FLUX_IN}
    S.update({"h_" + k: up(h[k]) for k in HLIM_IN + ("q_a", "q_a_pl")})
    S.update({"v_" + k: up(v[k]) for k in VLIM_IN + ("q_h", "q_h_pl")})
    z = lambda shape, typ=dt: torch.zeros(shape, dtype=typ, device=device)  # noqa: E731
    S.update(flx_h=z((6,) + spec.shape3), flx_h_pl=z(spec.shape_pl), GRD_xc=z((3, 3) + spec.shape3),
             GRD_xc_pl=z((3,) + spec.shape_pl), Qout_prev=z((2,) + spec.shape3), Qout_prev_pl=z((2,) + spec.shape_pl))
    rdgz = 1.0 / np.diff(vi["gzh"], append=2 * vi["gzh"][-1] - vi["gzh"][-2])  # 1/(gzh(k+1)-gzh(k))
    fv = -v["ck"][0] * f["rho"] / rdgz[None, :, None]
    fv_pl = -v["ck_pl"][0] * f["rho_pl"] / rdgz[None, :, None]
    S.update(flx_v=up(fv), flx_v_pl=up(fv_pl))
    S.update({k + s: up(vi[k + s]) for k in VI_IN for s in ("", "_pl")})
    S.update(rhogw=up(vi["rhogw"], torch.float64), rhogw_pl=up(vi["rhogw_pl"], torch.float64))
    S.update(rhogq=up(f["rho"] * h["q"], torch.float64), rhogq_pl=up(f["rho_pl"] * h["q_pl"], torch.float64))
    S["mixed"] = mixed
    return S


def chain_step(ctx: Context, S, dt_large: float, dt_small: float) -> None:
    """horizontal_flux -> horizontal_limiter_thuburn -> the conservative
    horizontal update of rhogq -> vertical_limiter_thuburn -> the vertical
    update of rhogq -> vi_rhow_solver, on the tensors of chain_state.  Mixed:
    the _f32 flux and limiters, FP64 accumulation of rhogq (the _f32 updates)
    and the mixed rho*w solve."""
    m = S["mixed"]
    sfx = "_f32" if m else ""
    getattr(ctx, "horizontal_flux" + sfx)(S["flx_h"], S["flx_h_pl"], S["GRD_xc"], S["GRD_xc_pl"],
                                          *[S[k] for k in FLUX_IN], dt_large)
    getattr(ctx, "horizontal_limiter_thuburn" + sfx)(
        S["h_q_a"], S["h_q_a_pl"], S["h_q"], S["h_q_pl"], S["h_d"], S["h_d_pl"], S["h_ch"], S["h_ch_pl"],
        S["h_cmask"], S["h_cmask_pl"], S["Qout_prev"], S["Qout_prev_pl"], S["h_Qout_post"], S["h_Qout_post_pl"])
    ctx.tracer_update_conservative(S["rhogq"], S["rhogq_pl"], S["flx_h"], S["h_q_a"], single=m)
    getattr(ctx, "vertical_limiter_thuburn" + sfx)(S["v_q_h"], S["v_q_h_pl"], S["v_q"], S["v_q_pl"], S["v_d"],
                                                   S["v_d_pl"], S["v_ck"], S["v_ck_pl"])
    getattr(ctx, "tracer_update_vertical" + sfx)(S["rhogq"], S["rhogq_pl"], S["flx_v"], S["flx_v_pl"], S["v_q_h"],
                                                 S["v_q_h_pl"])
    args = []
    for k in VI_IN:
        args += [S[k], S[k + "_pl"]]
    getattr(ctx, "vi_rhow_solver_mixed" if m else "vi_rhow_solver")(S["rhogw"], S["rhogw_pl"], *args, dt_small)


def total_mass(ctx: Context, rhogq, rhogq_pl, gzh: np.ndarray) -> float:
    """The global tracer mass sum_k dz(k) * M(k) over kmin..kmax, M(k) the
    FP64 area sums of nicam_tracer_mass; the level sum exact (math.fsum)."""
    import math

    import torch

    m = torch.zeros(ctx.kall, dtype=torch.float64, device=rhogq.device)
    ctx.tracer_mass(rhogq, rhogq_pl, m)
    mk = m.cpu().numpy()
    dz = np.diff(gzh, append=2 * gzh[-1] - gzh[-2])
    return math.fsum(float(mk[k]) * float(dz[k]) for k in range(1, ctx.kall - 1))
