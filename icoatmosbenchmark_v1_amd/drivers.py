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
