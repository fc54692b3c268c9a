"""Many tracers on several grids: ctypes mirror of the batched limiters and
of the grid group (nicam_group_*), and the device arrays of a batched step.

NICAM keeps its tracers as the last Fortran dimension, q(gall,kall,lall,ntr);
as C arrays (torch tensors) that is q[ntr][lall][kall][gall], q_a
[ntr][6][lall][kall][gall], Qout [ntr][2][lall][kall][gall], and the pole
arrays [ntr][...] likewise.
"""
from __future__ import annotations

import ctypes
from ctypes import c_double, c_float, c_int, c_void_p
from typing import Dict, List, Optional, Sequence

import numpy as np

from .runtime import Context, _check, load_library

STEP_FIELDS = ("flx_h", "flx_h_pl", "GRD_xc", "GRD_xc_pl", "rho", "rho_pl", "rhovx", "rhovx_pl", "rhovy",
               "rhovy_pl", "rhovz", "rhovz_pl", "q_a", "q_a_pl", "q", "q_pl", "d", "d_pl", "ch", "ch_pl", "cmask",
               "cmask_pl", "Qout_prev", "Qout_prev_pl", "Qout_post", "Qout_post_pl", "q_h", "q_h_pl", "qv", "qv_pl",
               "dv", "dv_pl", "ck", "ck_pl", "rhogw", "rhogw_pl", "rhogw0", "rhogw0_pl", "preg0", "preg0_pl",
               "rhog0", "rhog0_pl", "Srho", "Srho_pl", "Sw", "Sw_pl", "Spre", "Spre_pl")


class TracerStepF32(ctypes.Structure):
    """struct nicam_tracer_step_f32 (include/nicam_tt.h): one grid's device arrays."""
    _fields_ = [(name, c_void_p) for name in STEP_FIELDS]


def step_struct(S: Dict[str, object]) -> TracerStepF32:
    s = TracerStepF32()
    for name in STEP_FIELDS:
        t = S.get(name)
        setattr(s, name, None if t is None else t.data_ptr())
    return s


class Group:
    """nicam_group: up to eight grids (contexts) advanced by one launch per kernel."""

    def __init__(self, ctxs: Sequence[Context]):
        self.lib = load_library()
        self.ctxs = list(ctxs)
        arr = (c_void_p * len(ctxs))(*[c.h for c in ctxs])
        h = c_void_p()
        _check(self.lib.nicam_group_create(ctypes.byref(h), len(ctxs), arr), "nicam_group_create")
        self.h = h

    def close(self) -> None:
        if self.h:
            self.lib.nicam_group_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream) -> None:
        ptr = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
        _check(self.lib.nicam_group_set_stream(self.h, c_void_p(ptr)), "nicam_group_set_stream")

    def _structs(self, grids: List[Dict[str, object]]):
        if len(grids) != len(self.ctxs):
            raise ValueError("one argument set per grid of the group")
        return (TracerStepF32 * len(grids))(*[step_struct(S) for S in grids])

    def tracer_step_f32(self, ntr: int, grids: List[Dict[str, object]], dt_large: float, dt_small: float,
                        use_graph: bool = False) -> None:
        arr = self._structs(grids)
        _check(self.lib.nicam_group_tracer_step_f32(self.h, int(ntr), arr, c_float(dt_large), c_double(dt_small),
                                                    int(bool(use_graph))), "nicam_group_tracer_step_f32")

    def autotune(self, ntr: int, grids: List[Dict[str, object]], dt_large: float, dt_small: float,
                 cache_path: Optional[str] = None) -> int:
        arr = self._structs(grids)
        n = c_int(0)
        _check(self.lib.nicam_group_autotune(self.h, int(ntr), arr, c_float(dt_large), c_double(dt_small),
                                             None if cache_path is None else str(cache_path).encode(),
                                             ctypes.byref(n)), "nicam_group_autotune")
        return n.value


def batched_state(spec, case, ntr: int, device="cuda:0") -> Dict[str, object]:
    """FP32 device arrays of one grid's batched step: the case's flux, limiter
    and rho*w inputs rounded to FP32 (rhogw FP64), and ntr tracers, tracer t
    being the case's q, q_a (horizontal) and q, q_h (vertical) shifted by 7t
    cells along each level, in the tracer-last layout."""
    import torch

    def up(x, dt=torch.float32):
        return torch.from_numpy(np.ascontiguousarray(x)).to(device=device, dtype=dt)

    f, h, v, vi = case["flux"], case["hlim"], case["vlim"], case["vi"]
    z = lambda shape, dt=torch.float32: torch.zeros(shape, dtype=dt, device=device)  # noqa: E731
    S: Dict[str, object] = {k: up(f[k]) for k in ("rho", "rho_pl", "rhovx", "rhovx_pl", "rhovy", "rhovy_pl",
                                                    "rhovz", "rhovz_pl")}
    S.update(flx_h=z((6,) + spec.shape3), flx_h_pl=z(spec.shape_pl), GRD_xc=z((3, 3) + spec.shape3),
             GRD_xc_pl=z((3,) + spec.shape_pl))
    for k in ("d", "d_pl", "ch", "ch_pl", "cmask", "cmask_pl"):
        S[k] = up(h[k])
    q0, qa0 = up(h["q"]), up(h["q_a"])
    post0 = up(h["Qout_post"])
    S["q"] = torch.stack([torch.roll(q0, 7 * t, -1) for t in range(ntr)]).contiguous()
    S["q_a"] = torch.stack([torch.roll(qa0, 7 * t, -1) for t in range(ntr)]).contiguous()
    S["Qout_post"] = torch.stack([torch.roll(post0, 7 * t, -1) for t in range(ntr)]).contiguous()
    S["q_pl"] = up(h["q_pl"]).expand((ntr,) + tuple(spec.shape_pl)).contiguous()
    S["q_a_pl"] = up(h["q_a_pl"]).expand((ntr,) + tuple(spec.shape_pl)).contiguous()
    S["Qout_post_pl"] = up(h["Qout_post_pl"]).expand((ntr, 2) + tuple(spec.shape_pl)).contiguous()
    S["Qout_prev"] = z((ntr, 2) + spec.shape3)
    S["Qout_prev_pl"] = z((ntr, 2) + spec.shape_pl)
    qv0, qh0 = up(v["q"]), up(v["q_h"])
    S["qv"] = torch.stack([torch.roll(qv0, 7 * t, -1) for t in range(ntr)]).contiguous()
    S["q_h"] = torch.stack([torch.roll(qh0, 7 * t, -1) for t in range(ntr)]).contiguous()
    S["qv_pl"] = up(v["q_pl"]).expand((ntr,) + tuple(spec.shape_pl)).contiguous()
    S["q_h_pl"] = up(v["q_h_pl"]).expand((ntr,) + tuple(spec.shape_pl)).contiguous()
    S["dv"], S["dv_pl"], S["ck"], S["ck_pl"] = up(v["d"]), up(v["d_pl"]), up(v["ck"]), up(v["ck_pl"])
    for k in ("rhogw0", "preg0", "rhog0", "Srho", "Sw", "Spre"):
        S[k], S[k + "_pl"] = up(vi[k]), up(vi[k + "_pl"])
    S["rhogw"], S["rhogw_pl"] = up(vi["rhogw"], torch.float64), up(vi["rhogw_pl"], torch.float64)
    return S
