"""The snapshot files the reference's kernel drivers read, and the vertical
grid file.

A snapshot (``snapshot.<kernel>.pe<6-digit rank>``) is what dumpio writes
(dyn_vi_rhow_solver/src/dumpio.c:255-295): the fields one after another, each
a raw array of big-endian float64 with no header or record marker, in the
order the driver reads them (paths relative to kernels/NICAM/):

* dyn_horiz_adv_flux/src/main.f90:145-164
* dyn_horiz_adv_limiter/src/main.f90:132-147
* dyn_vert_adv_limiter/src/main.f90:96-105
* dyn_vi_rhow_solver/src/main.f90:198-229
* dyn_diffusion/src/main.f90:96-105
* dyn_divdamp/src/main.f90:181-206
* dyn_metrics/src/main.f90:217-240

Fields named like the kernel's outputs (``flx_h``, ``q_a``, ``Qout_prev``,
``q_h``, ``rhogw``) hold the outputs the driver checks its result against
("point-by-point diff"); the ``*_prev`` fields of the limiters and the
solver are the inputs it restores before each call.

Arrays are numpy arrays in this package's layout (the Fortran dimensions
reversed, e.g. flx_h(gall,kall,lall,6) is flx_h[6][lall][kall][gall]), so
writing one is a byte swap of its C-order buffer.
"""
from __future__ import annotations

from pathlib import Path
from typing import Dict, List, Tuple

import numpy as np

# extents: "R" = gall*kall*lall, "P" = gall_pl*kall*2, "R1"/"P1" = one level (K0),
# "V" = the two pole regions alone (coef_diff_pl(vlink,3,2))
_R, _P, _R1, _P1 = "R", "P", "R1", "P1"

LAYOUTS: Dict[str, List[Tuple[str, str, int]]] = {
    "dyn_horiz_adv_flux": [
        ("flx_h", _R, 6), ("flx_h_pl", _P, 1), ("GRD_xc", _R, 9), ("GRD_xc_pl", _P, 3),
        ("rho", _R, 1), ("rho_pl", _P, 1), ("rhovx", _R, 1), ("rhovx_pl", _P, 1),
        ("rhovy", _R, 1), ("rhovy_pl", _P, 1), ("rhovz", _R, 1), ("rhovz_pl", _P, 1),
        ("GRD_xr", _R1, 9), ("GRD_xr_pl", _P1, 3), ("GMTR_p", _R1, 8), ("GMTR_p_pl", _P1, 8),
        ("GMTR_t", _R1, 10), ("GMTR_t_pl", _P1, 5), ("GMTR_a", _R1, 36), ("GMTR_a_pl", _P1, 18),
    ],
    "dyn_horiz_adv_limiter": [
        ("q_a_prev", _R, 6), ("q_a_prev_pl", _P, 1), ("q_a", _R, 6), ("q_a_pl", _P, 1),
        ("q", _R, 1), ("q_pl", _P, 1), ("d", _R, 1), ("d_pl", _P, 1),
        ("ch", _R, 6), ("ch_pl", _P, 1), ("cmask", _R, 6), ("cmask_pl", _P, 1),
        ("Qout_prev", _R, 2), ("Qout_prev_pl", _P, 2), ("Qout_post", _R, 2), ("Qout_post_pl", _P, 2),
    ],
    "dyn_vert_adv_limiter": [
        ("q_h_prev", _R, 1), ("q_h_prev_pl", _P, 1), ("q_h", _R, 1), ("q_h_pl", _P, 1),
        ("q", _R, 1), ("q_pl", _P, 1), ("d", _R, 1), ("d_pl", _P, 1), ("ck", _R, 2), ("ck_pl", _P, 2),
    ],
    "dyn_vi_rhow_solver": [
        ("rhogw_prev", _R, 1), ("rhogw_prev_pl", _P, 1), ("rhogw", _R, 1), ("rhogw_pl", _P, 1),
        ("rhogw0", _R, 1), ("rhogw0_pl", _P, 1), ("preg0", _R, 1), ("preg0_pl", _P, 1),
        ("rhog0", _R, 1), ("rhog0_pl", _P, 1), ("Srho", _R, 1), ("Srho_pl", _P, 1),
        ("Sw", _R, 1), ("Sw_pl", _P, 1), ("Spre", _R, 1), ("Spre_pl", _P, 1),
        ("Mc", _R, 1), ("Mc_pl", _P, 1), ("Ml", _R, 1), ("Ml_pl", _P, 1), ("Mu", _R, 1), ("Mu_pl", _P, 1),
        ("RGSGAM2", _R, 1), ("RGSGAM2_pl", _P, 1), ("RGSGAM2H", _R, 1), ("RGSGAM2H_pl", _P, 1),
        ("RGAMH", _R, 1), ("RGAMH_pl", _P, 1), ("RGAM", _R, 1), ("RGAM_pl", _P, 1),
        ("GSGAM2H", _R, 1), ("GSGAM2H_pl", _P, 1),
    ],
    "dyn_diffusion": [
        ("dscl", _R, 1), ("dscl_pl", _P, 1), ("scl", _R, 1), ("scl_pl", _P, 1), ("kh", _R, 1), ("kh_pl", _P, 1),
        ("coef_intp", _R1, 18), ("coef_intp_pl", _P1, 9), ("coef_diff", _R1, 18), ("coef_diff_pl", "V", 15),
    ],
    "dyn_divdamp": [
        ("ddivdx", _R, 1), ("ddivdx_pl", _P, 1), ("ddivdy", _R, 1), ("ddivdy_pl", _P, 1),
        ("ddivdz", _R, 1), ("ddivdz_pl", _P, 1), ("rhogvx", _R, 1), ("rhogvx_pl", _P, 1),
        ("rhogvy", _R, 1), ("rhogvy_pl", _P, 1), ("rhogvz", _R, 1), ("rhogvz_pl", _P, 1),
        ("rhogw", _R, 1), ("rhogw_pl", _P, 1),
        ("coef_intp", _R1, 18), ("coef_intp_pl", _P1, 9), ("coef_diff", _R1, 18), ("coef_diff_pl", "V", 15),
        ("RGSQRTH", _R, 1), ("RGSQRTH_pl", _P, 1), ("RGAM", _R, 1), ("RGAM_pl", _P, 1),
        ("RGAMH", _R, 1), ("RGAMH_pl", _P, 1), ("C2WfactGz", _R, 6), ("C2WfactGz_pl", _P, 6),
    ],
    # one level (k0); the pole coefficients (0:vlink, ...) span gall_pl = vlink+1 points
    "dyn_metrics": [
        ("GRD_x", _R1, 3), ("GRD_x_pl", _P1, 3), ("GRD_xt", _R1, 6), ("GRD_xt_pl", _P1, 3),
        ("GRD_s", _R1, 2), ("GRD_s_pl", _P1, 2), ("GMTR_p", _R1, 8), ("GMTR_p_pl", _P1, 8),
        ("GMTR_t", _R1, 10), ("GMTR_t_pl", _P1, 5), ("GMTR_a", _R1, 36), ("GMTR_a_pl", _P1, 18),
        ("OPRT_coef_div", _R1, 21), ("OPRT_coef_div_pl", _P1, 3), ("OPRT_coef_rot", _R1, 21),
        ("OPRT_coef_rot_pl", _P1, 3), ("OPRT_coef_grad", _R1, 21), ("OPRT_coef_grad_pl", _P1, 3),
        ("OPRT_coef_lap", _R1, 7), ("OPRT_coef_lap_pl", _P1, 1), ("OPRT_coef_intp", _R1, 18),
        ("OPRT_coef_intp_pl", _P1, 9), ("OPRT_coef_diff", _R1, 18), ("OPRT_coef_diff_pl", "V", 15),
    ],
}

KERNELS = tuple(LAYOUTS)


def snapshot_name(kernel: str, rank: int = 0) -> str:
    """dumpio_mk_fname(fname, 'snapshot.<kernel>', 'pe', rank, 6)."""
    return f"snapshot.{kernel}.pe{rank:06d}"


def field_sizes(kernel: str, gall_1d: int, kall: int, lall: int, gall_pl: int = 6) -> List[Tuple[str, int]]:
    ext = {_R: gall_1d * gall_1d * kall * lall, _P: gall_pl * kall * 2,
           _R1: gall_1d * gall_1d * lall, _P1: gall_pl * 2, "V": 2}
    return [(name, ext[kind] * n) for name, kind, n in LAYOUTS[kernel]]


def write_snapshot(path, kernel: str, fields: Dict[str, np.ndarray], gall_1d: int, kall: int, lall: int,
                   gall_pl: int = 6) -> Path:
    """Writes every field of the kernel's layout (each must be present, of the
    right size) as big-endian float64, in the driver's read order."""
    path = Path(path)
    with open(path, "wb") as f:
        for name, n in field_sizes(kernel, gall_1d, kall, lall, gall_pl):
            a = np.ascontiguousarray(fields[name], dtype=np.float64)
            if a.size != n:
                raise ValueError(f"{kernel}.{name}: {a.size} values, the driver reads {n}")
            a.astype(">f8").tofile(f)
    return path


def read_snapshot(path, kernel: str, gall_1d: int, kall: int, lall: int, gall_pl: int = 6) -> Dict[str, np.ndarray]:
    """The fields of a snapshot as flat native float64 arrays (Fortran order of
    each field kept: reshape with the reversed Fortran shape)."""
    sizes = field_sizes(kernel, gall_1d, kall, lall, gall_pl)
    total = sum(n for _, n in sizes)
    raw = np.fromfile(path, dtype=">f8")
    if raw.size != total:
        raise ValueError(f"{path}: {raw.size} values, the {kernel} driver reads {total}")
    out, o = {}, 0
    for name, n in sizes:
        out[name] = raw[o:o + n].astype(np.float64)
        o += n
    return out


def write_vgrid(path, gz: np.ndarray, gzh: np.ndarray) -> Path:
    """A vertical grid file as GRD_input_vgrid reads it (mod_misc.f90:588-610):
    Fortran unformatted sequential, big-endian, 4-byte record markers;
    records nlayer (int32, = kall-2), gz(kall), gzh(kall) (float64)."""
    gz, gzh = np.asarray(gz, dtype=np.float64), np.asarray(gzh, dtype=np.float64)
    if gz.size != gzh.size:
        raise ValueError("gz and gzh need the same length")

    def rec(payload: bytes) -> bytes:
        m = np.array([len(payload)], dtype=">i4").tobytes()
        return m + payload + m

    path = Path(path)
    path.write_bytes(rec(np.array([gz.size - 2], dtype=">i4").tobytes()) + rec(gz.astype(">f8").tobytes())
                     + rec(gzh.astype(">f8").tobytes()))
    return path
