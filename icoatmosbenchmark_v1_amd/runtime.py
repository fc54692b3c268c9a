"""ctypes binding of libnicam_tt.so and a thin Python mirror of the reference
interface.

The reference exposes Fortran module procedures (paths relative to
kernels/NICAM/ in the reference tree):

* ``horizontal_flux``            dyn_horiz_adv_flux/src/mod_src_tracer.f90:42-379
* ``horizontal_limiter_thuburn`` dyn_horiz_adv_limiter/src/mod_src_tracer.f90:42-430
* ``vertical_limiter_thuburn``   dyn_vert_adv_limiter/src/mod_src_tracer.f90:41-243
* ``vi_rhow_solver_putcoef``     dyn_vi_rhow_solver/src/mod_vi.f90:58-89
* ``vi_rhow_solver``             dyn_vi_rhow_solver/src/mod_vi.f90:93-306

:class:`Context` carries what the reference keeps in module globals and has
one method per procedure, with the same argument order.  Arrays are numpy
arrays (host pointer mode) or torch tensors on the GPU (device pointer mode),
float64 and C-contiguous with the Fortran dimensions reversed: the Fortran
array ``flx_h(gall,kall,lall,6)`` is the C array ``flx_h[6][lall][kall][gall]``.

There is no CPU fallback: every call goes to the HIP library, and a missing
library raises :class:`LibraryNotBuilt`.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_size_t, c_void_p
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libnicam_tt.so"

NICAM_OK = 0
NICAM_ERR_ARG = 1
NICAM_ERR_STATE = 2
NICAM_ERR_HIP = 3
NICAM_ERR_IO = 4
NICAM_ERR_NOMEM = 5
PTR_DEVICE = 0
PTR_HOST = 1
TUNE_HFLUX_KPER = 1
TUNE_VLIM_KCHUNK = 3
TUNE_VI_VARIANT = 4
TUNE_HLIM_VARIANT = 7
TUNE_DIFF_KPER = 5
TUNE_DDMP_KPER = 6

STATUS_NAMES = {
    NICAM_ERR_ARG: "NICAM_ERR_ARG",
    NICAM_ERR_STATE: "NICAM_ERR_STATE",
    NICAM_ERR_HIP: "NICAM_ERR_HIP",
    NICAM_ERR_IO: "NICAM_ERR_IO",
    NICAM_ERR_NOMEM: "NICAM_ERR_NOMEM",
}

# Every symbol declared in include/nicam_tt.h: (name, restype, argtypes)
_P = c_void_p
SIGNATURES = {
    "nicam_version": (c_char_p, []),
    "nicam_last_error": (c_char_p, []),
    "nicam_ctx_create": (c_int, [POINTER(c_void_p), c_int, c_int, c_int, c_int, c_int, POINTER(c_int)]),
    "nicam_ctx_destroy": (c_int, [_P]),
    "nicam_ctx_set_stream": (c_int, [_P, _P]),
    "nicam_ctx_set_pointer_mode": (c_int, [_P, c_int]),
    "nicam_ctx_set_tuning": (c_int, [_P, c_int, c_int]),
    "nicam_ctx_device_bytes": (c_size_t, [_P]),
    "nicam_const_setup": (c_int, [_P]),
    "nicam_set_const": (c_int, [_P, c_double, c_double]),
    "nicam_set_phys_const": (c_int, [_P, c_double, c_double, c_double, c_int]),
    "nicam_set_const_f32": (c_int, [_P, c_float, c_float]),
    "nicam_grd_input_vgrid": (c_int, [c_char_p, c_int, _P, _P]),
    "nicam_grd_setup": (c_int, [_P, _P, _P, _P, _P, _P]),
    "nicam_set_vgrid_factors": (c_int, [_P, _P, _P, _P]),
    "nicam_set_hmetrics": (c_int, [_P] + [_P] * 8),
    "nicam_set_vmetrics": (c_int, [_P] + [_P] * 10),
    "nicam_horizontal_flux": (c_int, [_P] + [_P] * 12 + [c_double]),
    "nicam_horizontal_limiter_thuburn": (c_int, [_P] + [_P] * 14),
    "nicam_vertical_limiter_thuburn": (c_int, [_P] + [_P] * 8),
    "nicam_vi_rhow_solver_putcoef": (c_int, [_P] + [_P] * 6),
    "nicam_vi_rhow_solver": (c_int, [_P] + [_P] * 14 + [c_double]),
    "nicam_ctx_synchronize": (c_int, [_P]),
    "nicam_halo_map": (c_int, [c_int, c_int, c_int, c_int, _P, _P, _P, _P, _P, _P, _P, _P]),
    "nicam_comm_setup": (c_int, [_P, c_int]),
    "nicam_pole_fill_map": (c_int, [c_int, c_int, _P, _P]),
    "nicam_comm_data_transfer": (c_int, [_P, _P, _P, c_int, c_int]),
    "nicam_comm_var": (c_int, [_P, _P, _P, c_int, c_int]),
    "nicam_grid_icosahedral": (c_int, [c_int, c_int, c_double, _P, _P]),
    "nicam_grid_vertices": (c_int, [c_int, c_int, c_int, c_int, _P, c_double] + [_P] * 8),
    "nicam_gmtr_setup": (c_int, [c_int, c_int, c_int, c_int, _P, c_double] + [_P] * 12),
    "nicam_oprt_diffusion_setup": (c_int, [c_int, c_int, c_int, c_int, _P] + [_P] * 10),
    "nicam_oprt_divergence_setup": (c_int, [c_int, c_int, c_int, c_int, _P] + [_P] * 8),
    "nicam_oprt_rotation_setup": (c_int, [c_int, c_int, c_int, c_int, _P] + [_P] * 8),
    "nicam_oprt_gradient_setup": (c_int, [c_int, c_int, c_int, c_int, _P] + [_P] * 8),
    "nicam_oprt_laplacian_setup": (c_int, [c_int, c_int, c_int, c_int, _P] + [_P] * 8),
    "nicam_oprt_gradient": (c_int, [_P] * 7),
    "nicam_tracer_courant": (c_int, [_P] * 9),
    "nicam_horizontal_remap": (c_int, [_P] * 13),
    "nicam_tracer_update": (c_int, [_P] * 7),
    "nicam_oprt_gradient_adjoint": (c_int, [_P] * 7),
    "nicam_tracer_courant_adjoint": (c_int, [_P] * 11),
    "nicam_horizontal_remap_adjoint": (c_int, [_P] * 17),
    "nicam_tracer_update_adjoint": (c_int, [_P] * 11),
    "nicam_comm_data_transfer_adjoint": (c_int, [_P, _P, _P, c_int, c_int]),
    "nicam_vertical_courant_adjoint": (c_int, [_P] * 11),
    "nicam_vertical_face_adjoint": (c_int, [_P] * 5),
    "nicam_tracer_update_vertical_adjoint": (c_int, [_P] * 11),
    "nicam_vertical_mass_flux": (c_int, [_P] * 11 + [c_double]),
    "nicam_vertical_courant": (c_int, [_P] * 7),
    "nicam_vertical_face": (c_int, [_P] * 5),
    "nicam_tracer_update_vertical": (c_int, [_P] * 7),
    "nicam_vi_rhow_update_matrix": (c_int, [_P] * 5 + [c_double]),
    "nicam_horizontal_flux_f32": (c_int, [_P] * 13 + [c_float]),
    "nicam_horizontal_limiter_thuburn_f32": (c_int, [_P] * 15),
    "nicam_vertical_limiter_thuburn_f32": (c_int, [_P] * 9),
    "nicam_vi_rhow_getcoef": (c_int, [_P] * 7),
    "nicam_oprt_diffusion": (c_int, [_P] + [_P] * 10),
    "nicam_set_rdgz": (c_int, [_P, _P]),
    "nicam_set_vmetrics3d": (c_int, [_P] + [_P] * 8),
    "nicam_oprt3d_divdamp": (c_int, [_P] + [_P] * 18),
    "nicam_vi_rhow_solver_adjoint": (c_int, [_P] + [_P] * 14 + [c_double]),
    "nicam_vi_rhow_solver_mixed": (c_int, [_P] + [_P] * 14 + [c_double]),
    "nicam_vertical_limiter_thuburn_mask": (c_int, [_P] * 11),
    "nicam_horizontal_limiter_thuburn_mask": (c_int, [_P] * 17),
    "nicam_horizontal_flux_adjoint": (c_int, [_P] * 21 + [c_double]),
    "nicam_horizontal_limiter_thuburn_adjoint": (c_int, [_P] * 25),
    "nicam_vertical_limiter_thuburn_adjoint": (c_int, [_P] * 17),
    "nicam_tracer_mass": (c_int, [_P] * 4),
    "nicam_tracer_mass_f32": (c_int, [_P] * 4),
    "nicam_tracer_mass_noalloc": (c_int, [_P] * 4),
    "nicam_tracer_update_conservative": (c_int, [_P] * 5),
    "nicam_tracer_update_conservative_f32": (c_int, [_P] * 5),
    "nicam_tracer_update_vertical_f32": (c_int, [_P] * 7),
    "nicam_horizontal_limiter_thuburn_batch": (c_int, [_P, c_int] + [_P] * 14),
    "nicam_horizontal_limiter_thuburn_batch_f32": (c_int, [_P, c_int] + [_P] * 14),
    "nicam_vertical_limiter_thuburn_batch": (c_int, [_P, c_int] + [_P] * 8),
    "nicam_vertical_limiter_thuburn_batch_f32": (c_int, [_P, c_int] + [_P] * 8),
    "nicam_group_create": (c_int, [_P, c_int, _P]),
    "nicam_group_destroy": (c_int, [_P]),
    "nicam_group_set_stream": (c_int, [_P, _P]),
    "nicam_group_tracer_step_f32": (c_int, [_P, c_int, _P, c_float, c_double, c_int]),
    "nicam_group_autotune": (c_int, [_P, c_int, _P, c_float, c_double, c_char_p, _P]),
}


class LibraryNotBuilt(RuntimeError):
    pass


class NicamError(RuntimeError):
    def __init__(self, status: int, func: str, msg: str):
        super().__init__(f"{func} failed with {STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


_lib: Optional[ctypes.CDLL] = None


def load_library(path: Optional[Path] = None) -> ctypes.CDLL:
    """Loads libnicam_tt.so (once) and declares every exported signature."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise LibraryNotBuilt(f"{p} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _check(rc: int, func: str) -> None:
    if rc != NICAM_OK:
        msg = load_library().nicam_last_error().decode(errors="replace")
        raise NicamError(rc, func, msg)


def _ptr(x, mode: int, name: str, numel: Optional[int] = None, single: bool = False):
    """Raw pointer of a numpy array (host mode) or torch tensor (device mode),
    float64 (float32 with single=True)."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        if mode != PTR_HOST:
            raise TypeError(f"{name}: numpy array passed in device pointer mode")
        want = np.float32 if single else np.float64
        if x.dtype != want or not x.flags.c_contiguous:
            raise TypeError(f"{name}: need a C-contiguous {np.dtype(want).name} array")
        if numel is not None and x.size != numel:
            raise ValueError(f"{name}: has {x.size} elements, expected {numel}")
        return x.ctypes.data
    # torch tensor
    try:
        import torch
    except ImportError as e:  # pragma: no cover
        raise TypeError(f"{name}: unsupported array type {type(x)}") from e
    if not isinstance(x, torch.Tensor):
        raise TypeError(f"{name}: unsupported array type {type(x)}")
    if mode != PTR_DEVICE or not x.is_cuda:
        raise TypeError(f"{name}: torch tensors must live on the GPU and use device pointer mode")
    want = torch.float32 if single else torch.float64
    if x.dtype != want or not x.is_contiguous():
        raise TypeError(f"{name}: need a contiguous {want} tensor")
    if numel is not None and x.numel() != numel:
        raise ValueError(f"{name}: has {x.numel()} elements, expected {numel}")
    return x.data_ptr()


def _ptr_u32(x, mode: int, name: str, numel: Optional[int] = None):
    """Raw pointer of a 32-bit mask array: numpy uint32 (host mode) or a
    torch int32/uint32 tensor on the GPU (device mode)."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        if mode != PTR_HOST:
            raise TypeError(f"{name}: numpy array passed in device pointer mode")
        if x.dtype != np.uint32 or not x.flags.c_contiguous:
            raise TypeError(f"{name}: need a C-contiguous uint32 array")
        if numel is not None and x.size != numel:
            raise ValueError(f"{name}: has {x.size} elements, expected {numel}")
        return x.ctypes.data
    import torch

    if not isinstance(x, torch.Tensor) or mode != PTR_DEVICE or not x.is_cuda:
        raise TypeError(f"{name}: need a GPU tensor in device pointer mode")
    if x.dtype not in (torch.int32, getattr(torch, "uint32", torch.int32)) or not x.is_contiguous():
        raise TypeError(f"{name}: need a contiguous 32-bit integer tensor")
    if numel is not None and x.numel() != numel:
        raise ValueError(f"{name}: has {x.numel()} elements, expected {numel}")
    return x.data_ptr()


def _ptr_u8(x, mode: int, name: str, numel: Optional[int] = None):
    """Raw pointer of a byte mask: numpy uint8 (host mode) or a torch uint8
    tensor on the GPU (device mode)."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        if mode != PTR_HOST:
            raise TypeError(f"{name}: numpy array passed in device pointer mode")
        if x.dtype != np.uint8 or not x.flags.c_contiguous:
            raise TypeError(f"{name}: need a C-contiguous uint8 array")
        if numel is not None and x.size != numel:
            raise ValueError(f"{name}: has {x.size} elements, expected {numel}")
        return x.ctypes.data
    import torch

    if not isinstance(x, torch.Tensor) or mode != PTR_DEVICE or not x.is_cuda:
        raise TypeError(f"{name}: need a GPU tensor in device pointer mode")
    if x.dtype != torch.uint8 or not x.is_contiguous():
        raise TypeError(f"{name}: need a contiguous uint8 tensor")
    if numel is not None and x.numel() != numel:
        raise ValueError(f"{name}: has {x.numel()} elements, expected {numel}")
    return x.data_ptr()


def _setup_ptr(x, name: str, numel: Optional[int] = None):
    """Set-up calls accept host (numpy) or device (torch) pointers in any mode."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if numel is not None and x.size != numel:
            raise ValueError(f"{name}: has {x.size} elements, expected {numel}")
        return x.ctypes.data, x
    return _ptr(x, PTR_DEVICE, name, numel), x


class Context:
    """One grid: sizes, singular-point flags, constants, metrics, coefficients.

    Mirrors the module state of the reference's mod_misc (problem_size.inc:28-59,
    mod_misc.f90:78-152) and mod_vi (mod_vi.f90:47-52).
    """

    def __init__(self, gall_1d: int, kall: int, lall: int, gall_pl: int = 6, have_pl: bool = True,
                 have_sgp: Optional[Sequence[int]] = None, pointer_mode: int = PTR_DEVICE):
        self.lib = load_library()
        if have_sgp is None:
            have_sgp = [0] * lall
        sgp = (c_int * lall)(*[1 if v else 0 for v in have_sgp])
        h = c_void_p()
        _check(self.lib.nicam_ctx_create(ctypes.byref(h), gall_1d, kall, lall, gall_pl, int(bool(have_pl)), sgp),
               "nicam_ctx_create")
        self.h = h
        self.gall_1d, self.gall, self.kall, self.lall = gall_1d, gall_1d * gall_1d, kall, lall
        self.gall_pl, self.have_pl = gall_pl, bool(have_pl)
        self.have_sgp = list(have_sgp)
        self.mode = PTR_DEVICE
        self._keep = []  # host copies kept alive during asynchronous set-up copies
        self.set_pointer_mode(pointer_mode)

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.nicam_ctx_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- sizes -------------------------------------------------------------
    @property
    def nreg(self) -> int:
        return self.gall * self.kall * self.lall

    @property
    def npl(self) -> int:
        return self.gall_pl * self.kall * 2

    # -- configuration -------------------------------------------------------
    def set_pointer_mode(self, mode: int) -> None:
        _check(self.lib.nicam_ctx_set_pointer_mode(self.h, mode), "nicam_ctx_set_pointer_mode")
        self.mode = mode

    def set_tuning(self, param: int, value: int) -> None:
        """Launch-shape choice (TUNE_* constants); results do not depend on it."""
        _check(self.lib.nicam_ctx_set_tuning(self.h, int(param), int(value)), "nicam_ctx_set_tuning")

    def set_stream(self, stream) -> None:
        """stream: a torch.cuda.Stream, a raw hipStream_t integer, or None."""
        raw = getattr(stream, "cuda_stream", stream)
        _check(self.lib.nicam_ctx_set_stream(self.h, c_void_p(raw) if raw else None), "nicam_ctx_set_stream")

    def const_setup(self) -> None:
        _check(self.lib.nicam_const_setup(self.h), "nicam_const_setup")

    def set_const(self, eps: float, huge: float) -> None:
        _check(self.lib.nicam_set_const(self.h, eps, huge), "nicam_set_const")

    def set_const_f32(self, eps: float, huge: float) -> None:
        """CONST_EPS / CONST_HUGE of the _f32 entry points (default: the -DSINGLE build's)."""
        _check(self.lib.nicam_set_const_f32(self.h, float(eps), float(huge)), "nicam_set_const_f32")

    def set_phys_const(self, grav=9.80616, rdry=287.0, cvdry=717.5, non_hydro_alpha=1) -> None:
        _check(self.lib.nicam_set_phys_const(self.h, grav, rdry, cvdry, non_hydro_alpha), "nicam_set_phys_const")

    def device_bytes(self) -> int:
        return int(self.lib.nicam_ctx_device_bytes(self.h))

    def synchronize(self) -> None:
        _check(self.lib.nicam_ctx_synchronize(self.h), "nicam_ctx_synchronize")

    # -- halo exchange ---------------------------------------------------------
    def comm_setup(self, rlevel: int) -> None:
        """COMM_setup for a context that holds all 10*4^rlevel regions and the poles."""
        _check(self.lib.nicam_comm_setup(self.h, int(rlevel)), "nicam_comm_setup")

    def _halo_args(self, var, var_pl, kmax, vmax):
        kmax = self.kall if kmax is None else int(kmax)
        n = kmax * int(vmax)
        return [_ptr(var, self.mode, "var", self.gall * self.lall * n),
                _ptr(var_pl, self.mode, "var_pl", self.gall_pl * 2 * n), kmax, int(vmax)]

    def comm_data_transfer(self, var, var_pl, vmax: int = 1, kmax: Optional[int] = None) -> None:
        """COMM_data_transfer(var(gall,kmax,lall,vmax), var_pl(gall_pl,kmax,2,vmax)); kmax defaults to kall."""
        _check(self.lib.nicam_comm_data_transfer(self.h, *self._halo_args(var, var_pl, kmax, vmax)),
               "nicam_comm_data_transfer")

    def comm_data_transfer_adjoint(self, var_b, var_b_pl, vmax: int = 1, kmax: Optional[int] = None) -> None:
        """The transpose of comm_data_transfer, in place: each halo destination's cotangent is added to its
        source's and cleared."""
        _check(self.lib.nicam_comm_data_transfer_adjoint(self.h, *self._halo_args(var_b, var_b_pl, kmax, vmax)),
               "nicam_comm_data_transfer_adjoint")

    def comm_var(self, var, var_pl, vmax: int = 1, kmax: Optional[int] = None) -> None:
        """COMM_var: pole centres from their regions, then COMM_data_transfer."""
        _check(self.lib.nicam_comm_var(self.h, *self._halo_args(var, var_pl, kmax, vmax)), "nicam_comm_var")

    # -- set-up routines -----------------------------------------------------
    def grd_setup(self, gz: np.ndarray, gzh: np.ndarray):
        """GRD_setup from gz/gzh; returns (rdgzh, afact, bfact)."""
        gz = np.ascontiguousarray(gz, dtype=np.float64)
        gzh = np.ascontiguousarray(gzh, dtype=np.float64)
        if gz.size != self.kall or gzh.size != self.kall:
            raise ValueError("gz/gzh need kall elements")
        out = [np.empty(self.kall) for _ in range(3)]
        _check(self.lib.nicam_grd_setup(self.h, gz.ctypes.data, gzh.ctypes.data, *[o.ctypes.data for o in out]),
               "nicam_grd_setup")
        return tuple(out)

    def set_vgrid_factors(self, rdgzh, afact, bfact) -> None:
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (rdgzh, afact, bfact)]
        if any(x.size != self.kall for x in a):
            raise ValueError("vertical factors need kall elements")
        _check(self.lib.nicam_set_vgrid_factors(self.h, *[x.ctypes.data for x in a]), "nicam_set_vgrid_factors")

    def set_hmetrics(self, GRD_xr, GRD_xr_pl, GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl) -> None:
        g, p = self.gall * self.lall, self.gall_pl * 2
        spec = [(GRD_xr, g * 9), (GRD_xr_pl, p * 3), (GMTR_p, g * 8), (GMTR_p_pl, p * 8),
                (GMTR_t, g * 10), (GMTR_t_pl, p * 5), (GMTR_a, g * 36), (GMTR_a_pl, p * 18)]
        names = ["GRD_xr", "GRD_xr_pl", "GMTR_p", "GMTR_p_pl", "GMTR_t", "GMTR_t_pl", "GMTR_a", "GMTR_a_pl"]
        ptrs = []
        for (x, n), nm in zip(spec, names):
            r = _setup_ptr(x if (self.have_pl or not nm.endswith("_pl")) else None, nm, n)
            ptrs.append(r[0] if r else None)
            if r:
                self._keep.append(r[1])
        _check(self.lib.nicam_set_hmetrics(self.h, *ptrs), "nicam_set_hmetrics")
        self._keep.clear()

    def _pairs(self, arrays, names):
        ptrs = []
        for x, nm in zip(arrays, names):
            pole = nm.endswith("_pl")
            if pole and not self.have_pl:
                ptrs.append(None)
                continue
            r = _setup_ptr(x, nm, self.npl if pole else self.nreg)
            ptrs.append(r[0])
            self._keep.append(r[1])
        return ptrs

    def set_vmetrics(self, RGAM, RGAM_pl, RGAMH, RGAMH_pl, RGSGAM2, RGSGAM2_pl, RGSGAM2H, RGSGAM2H_pl,
                     GSGAM2H, GSGAM2H_pl) -> None:
        names = ["RGAM", "RGAM_pl", "RGAMH", "RGAMH_pl", "RGSGAM2", "RGSGAM2_pl", "RGSGAM2H", "RGSGAM2H_pl",
                 "GSGAM2H", "GSGAM2H_pl"]
        ptrs = self._pairs([RGAM, RGAM_pl, RGAMH, RGAMH_pl, RGSGAM2, RGSGAM2_pl, RGSGAM2H, RGSGAM2H_pl,
                            GSGAM2H, GSGAM2H_pl], names)
        _check(self.lib.nicam_set_vmetrics(self.h, *ptrs), "nicam_set_vmetrics")
        self._keep.clear()

    def set_rdgz(self, rdgz) -> None:
        a = np.ascontiguousarray(rdgz, dtype=np.float64)
        if a.size != self.kall:
            raise ValueError("rdgz needs kall elements")
        _check(self.lib.nicam_set_rdgz(self.h, a.ctypes.data), "nicam_set_rdgz")

    def set_vmetrics3d(self, RGSQRTH, RGSQRTH_pl, RGAM, RGAM_pl, RGAMH, RGAMH_pl, C2WfactGz, C2WfactGz_pl) -> None:
        ptrs = self._pairs([RGSQRTH, RGSQRTH_pl, RGAM, RGAM_pl, RGAMH, RGAMH_pl],
                           ["RGSQRTH", "RGSQRTH_pl", "RGAM", "RGAM_pl", "RGAMH", "RGAMH_pl"])
        c = _setup_ptr(C2WfactGz, "C2WfactGz", 6 * self.nreg)
        cp = _setup_ptr(C2WfactGz_pl, "C2WfactGz_pl", 6 * self.npl) if self.have_pl else None
        self._keep += [c[1]] + ([cp[1]] if cp else [])
        _check(self.lib.nicam_set_vmetrics3d(self.h, *ptrs, c[0], cp[0] if cp else None), "nicam_set_vmetrics3d")
        self._keep.clear()

    def vi_rhow_solver_putcoef(self, Mc, Mc_pl, Ml, Ml_pl, Mu, Mu_pl) -> None:
        ptrs = self._pairs([Mc, Mc_pl, Ml, Ml_pl, Mu, Mu_pl], ["Mc", "Mc_pl", "Ml", "Ml_pl", "Mu", "Mu_pl"])
        _check(self.lib.nicam_vi_rhow_solver_putcoef(self.h, *ptrs), "nicam_vi_rhow_solver_putcoef")
        self._keep.clear()

    # -- hot path --------------------------------------------------------------
    def _args(self, spec, single: bool = False):
        """spec: list of (array, name, numel)."""
        out = []
        for x, nm, n in spec:
            if nm.endswith("_pl") and not self.have_pl and nm != "Qout_prev_pl":
                out.append(None)
            else:
                out.append(_ptr(x, self.mode, nm, n, single))
        return out

    def horizontal_flux(self, flx_h, flx_h_pl, GRD_xc, GRD_xc_pl, rho, rho_pl, rhovx, rhovx_pl,
                        rhovy, rhovy_pl, rhovz, rhovz_pl, dt: float) -> None:
        R, P = self.nreg, self.npl
        a = self._args([(flx_h, "flx_h", 6 * R), (flx_h_pl, "flx_h_pl", P), (GRD_xc, "GRD_xc", 9 * R),
                        (GRD_xc_pl, "GRD_xc_pl", 3 * P), (rho, "rho", R), (rho_pl, "rho_pl", P),
                        (rhovx, "rhovx", R), (rhovx_pl, "rhovx_pl", P), (rhovy, "rhovy", R),
                        (rhovy_pl, "rhovy_pl", P), (rhovz, "rhovz", R), (rhovz_pl, "rhovz_pl", P)])
        _check(self.lib.nicam_horizontal_flux(self.h, *a, float(dt)), "nicam_horizontal_flux")

    def horizontal_limiter_thuburn(self, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl,
                                   Qout_prev, Qout_prev_pl, Qout_post, Qout_post_pl) -> None:
        R, P = self.nreg, self.npl
        a = self._args([(q_a, "q_a", 6 * R), (q_a_pl, "q_a_pl", P), (q, "q", R), (q_pl, "q_pl", P),
                        (d, "d", R), (d_pl, "d_pl", P), (ch, "ch", 6 * R), (ch_pl, "ch_pl", P),
                        (cmask, "cmask", 6 * R), (cmask_pl, "cmask_pl", P), (Qout_prev, "Qout_prev", 2 * R),
                        (Qout_prev_pl, "Qout_prev_pl", 2 * P), (Qout_post, "Qout_post", 2 * R),
                        (Qout_post_pl, "Qout_post_pl", 2 * P)])
        _check(self.lib.nicam_horizontal_limiter_thuburn(self.h, *a), "nicam_horizontal_limiter_thuburn")

    # -- FP32 variants (the reference built with -DSINGLE) ---------------------
    def horizontal_flux_f32(self, flx_h, flx_h_pl, GRD_xc, GRD_xc_pl, rho, rho_pl, rhovx, rhovx_pl,
                            rhovy, rhovy_pl, rhovz, rhovz_pl, dt: float) -> None:
        R, P = self.nreg, self.npl
        a = self._args([(flx_h, "flx_h", 6 * R), (flx_h_pl, "flx_h_pl", P), (GRD_xc, "GRD_xc", 9 * R),
                        (GRD_xc_pl, "GRD_xc_pl", 3 * P), (rho, "rho", R), (rho_pl, "rho_pl", P),
                        (rhovx, "rhovx", R), (rhovx_pl, "rhovx_pl", P), (rhovy, "rhovy", R),
                        (rhovy_pl, "rhovy_pl", P), (rhovz, "rhovz", R), (rhovz_pl, "rhovz_pl", P)], single=True)
        _check(self.lib.nicam_horizontal_flux_f32(self.h, *a, float(dt)), "nicam_horizontal_flux_f32")

    def horizontal_limiter_thuburn_f32(self, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl,
                                       Qout_prev, Qout_prev_pl, Qout_post, Qout_post_pl) -> None:
        R, P = self.nreg, self.npl
        a = self._args([(q_a, "q_a", 6 * R), (q_a_pl, "q_a_pl", P), (q, "q", R), (q_pl, "q_pl", P),
                        (d, "d", R), (d_pl, "d_pl", P), (ch, "ch", 6 * R), (ch_pl, "ch_pl", P),
                        (cmask, "cmask", 6 * R), (cmask_pl, "cmask_pl", P), (Qout_prev, "Qout_prev", 2 * R),
                        (Qout_prev_pl, "Qout_prev_pl", 2 * P), (Qout_post, "Qout_post", 2 * R),
                        (Qout_post_pl, "Qout_post_pl", 2 * P)], single=True)
        _check(self.lib.nicam_horizontal_limiter_thuburn_f32(self.h, *a), "nicam_horizontal_limiter_thuburn_f32")

    def vertical_limiter_thuburn_f32(self, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl) -> None:
        R, P = self.nreg, self.npl
        a = self._args([(q_h, "q_h", R), (q_h_pl, "q_h_pl", P), (q, "q", R), (q_pl, "q_pl", P), (d, "d", R),
                        (d_pl, "d_pl", P), (ck, "ck", 2 * R), (ck_pl, "ck_pl", 2 * P)], single=True)
        _check(self.lib.nicam_vertical_limiter_thuburn_f32(self.h, *a), "nicam_vertical_limiter_thuburn_f32")

    def vertical_limiter_thuburn(self, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl) -> None:
        R, P = self.nreg, self.npl
        a = self._args([(q_h, "q_h", R), (q_h_pl, "q_h_pl", P), (q, "q", R), (q_pl, "q_pl", P), (d, "d", R),
                        (d_pl, "d_pl", P), (ck, "ck", 2 * R), (ck_pl, "ck_pl", 2 * P)])
        _check(self.lib.nicam_vertical_limiter_thuburn(self.h, *a), "nicam_vertical_limiter_thuburn")

    def oprt_diffusion(self, dscl, dscl_pl, scl, scl_pl, kh, kh_pl, coef_intp, coef_intp_pl,
                       coef_diff, coef_diff_pl) -> None:
        """OPRT_diffusion (dyn_diffusion/src/mod_oprt.f90:41-296)."""
        R, P, G = self.nreg, self.npl, self.gall * self.lall
        a = [_ptr(dscl, self.mode, "dscl", R), _ptr(dscl_pl, self.mode, "dscl_pl", P)]
        a += self._args([(scl, "scl", R), (scl_pl, "scl_pl", P), (kh, "kh", R), (kh_pl, "kh_pl", P),
                         (coef_intp, "coef_intp", 18 * G), (coef_intp_pl, "coef_intp_pl", 18 * self.gall_pl),
                         (coef_diff, "coef_diff", 18 * G), (coef_diff_pl, "coef_diff_pl", 30)])
        _check(self.lib.nicam_oprt_diffusion(self.h, *a), "nicam_oprt_diffusion")

    # ---- step glue (the inputs the reference's kernels take from "other subroutine") ----
    def oprt_gradient(self, grad, grad_pl, scl, scl_pl, coef_grad, coef_grad_pl) -> None:
        """nicam_oprt_gradient: grad(gall,kall,lall,3) of scl with OPRT_gradient_setup's coefficients."""
        R, P, G = self.nreg, self.npl, self.gall * self.lall
        a = [_ptr(grad, self.mode, "grad", 3 * R), _ptr(grad_pl, self.mode, "grad_pl", 3 * P)]
        a += self._args([(scl, "scl", R), (scl_pl, "scl_pl", P), (coef_grad, "coef_grad", 21 * G),
                         (coef_grad_pl, "coef_grad_pl", 6 * self.gall_pl)])
        _check(self.lib.nicam_oprt_gradient(self.h, *a), "nicam_oprt_gradient")

    def tracer_courant(self, ch, ch_pl, cmask, cmask_pl, flx_h, flx_h_pl, rhog, rhog_pl) -> None:
        """nicam_tracer_courant: ch = flx_h/rhog and the upwind mask of the six faces."""
        R, P = self.nreg, self.npl
        a = [_ptr(ch, self.mode, "ch", 6 * R), _ptr(ch_pl, self.mode, "ch_pl", P),
             _ptr(cmask, self.mode, "cmask", 6 * R), _ptr(cmask_pl, self.mode, "cmask_pl", P)]
        a += self._args([(flx_h, "flx_h", 6 * R), (flx_h_pl, "flx_h_pl", P), (rhog, "rhog", R),
                         (rhog_pl, "rhog_pl", P)])
        _check(self.lib.nicam_tracer_courant(self.h, *a), "nicam_tracer_courant")

    def horizontal_remap(self, q_a, q_a_pl, q, q_pl, grad, grad_pl, cmask, cmask_pl, GRD_xc, GRD_xc_pl,
                         GRD_x, GRD_x_pl) -> None:
        """nicam_horizontal_remap: upwind face values q + grad.(GRD_xc - GRD_x)."""
        R, P, G = self.nreg, self.npl, self.gall * self.lall
        a = [_ptr(q_a, self.mode, "q_a", 6 * R), _ptr(q_a_pl, self.mode, "q_a_pl", P)]
        a += self._args([(q, "q", R), (q_pl, "q_pl", P), (grad, "grad", 3 * R), (grad_pl, "grad_pl", 3 * P),
                         (cmask, "cmask", 6 * R), (cmask_pl, "cmask_pl", P), (GRD_xc, "GRD_xc", 9 * R),
                         (GRD_xc_pl, "GRD_xc_pl", 3 * P), (GRD_x, "GRD_x", 3 * G),
                         (GRD_x_pl, "GRD_x_pl", 6 * self.gall_pl)])
        _check(self.lib.nicam_horizontal_remap(self.h, *a), "nicam_horizontal_remap")

    def tracer_update(self, rhogq, rhogq_pl, flx_h, flx_h_pl, q_a, q_a_pl) -> None:
        """nicam_tracer_update: rhogq -= sum of flx_h*q_a over the six faces."""
        R, P = self.nreg, self.npl
        a = [_ptr(rhogq, self.mode, "rhogq", R), _ptr(rhogq_pl, self.mode, "rhogq_pl", P)]
        a += self._args([(flx_h, "flx_h", 6 * R), (flx_h_pl, "flx_h_pl", P), (q_a, "q_a", 6 * R),
                         (q_a_pl, "q_a_pl", P)])
        _check(self.lib.nicam_tracer_update(self.h, *a), "nicam_tracer_update")

    # adjoints of the horizontal glue: outputs (the *_b arrays first in each list) overwritten
    def oprt_gradient_adjoint(self, scl_b, scl_b_pl, grad_b, grad_b_pl, coef_grad, coef_grad_pl) -> None:
        R, P = self.nreg, self.npl
        a = [_ptr(scl_b, self.mode, "scl_b", R), _ptr(scl_b_pl, self.mode, "scl_b_pl", P)]
        a += self._args([(grad_b, "grad_b", 3 * R), (grad_b_pl, "grad_b_pl", 3 * P),
                         (coef_grad, "coef_grad", self.gall * self.lall * 21),
                         (coef_grad_pl, "coef_grad_pl", self.gall_pl * 6)])
        _check(self.lib.nicam_oprt_gradient_adjoint(self.h, *a), "nicam_oprt_gradient_adjoint")

    def tracer_courant_adjoint(self, flx_h_b, flx_h_b_pl, rhog_b, rhog_b_pl, ch_b, ch_b_pl, flx_h, flx_h_pl, rhog,
                               rhog_pl) -> None:
        R, P = self.nreg, self.npl
        a = [_ptr(flx_h_b, self.mode, "flx_h_b", 6 * R), _ptr(flx_h_b_pl, self.mode, "flx_h_b_pl", P),
             _ptr(rhog_b, self.mode, "rhog_b", R), _ptr(rhog_b_pl, self.mode, "rhog_b_pl", P)]
        a += self._args([(ch_b, "ch_b", 6 * R), (ch_b_pl, "ch_b_pl", P), (flx_h, "flx_h", 6 * R),
                         (flx_h_pl, "flx_h_pl", P), (rhog, "rhog", R), (rhog_pl, "rhog_pl", P)])
        _check(self.lib.nicam_tracer_courant_adjoint(self.h, *a), "nicam_tracer_courant_adjoint")

    def horizontal_remap_adjoint(self, q_b, q_b_pl, grad_b, grad_b_pl, GRD_xc_b, GRD_xc_b_pl, q_a_b, q_a_b_pl,
                                 grad, grad_pl, cmask, cmask_pl, GRD_xc, GRD_xc_pl, GRD_x, GRD_x_pl) -> None:
        R, P = self.nreg, self.npl
        a = [_ptr(q_b, self.mode, "q_b", R), _ptr(q_b_pl, self.mode, "q_b_pl", P),
             _ptr(grad_b, self.mode, "grad_b", 3 * R), _ptr(grad_b_pl, self.mode, "grad_b_pl", 3 * P),
             _ptr(GRD_xc_b, self.mode, "GRD_xc_b", 9 * R), _ptr(GRD_xc_b_pl, self.mode, "GRD_xc_b_pl", 3 * P)]
        a += self._args([(q_a_b, "q_a_b", 6 * R), (q_a_b_pl, "q_a_b_pl", P), (grad, "grad", 3 * R),
                         (grad_pl, "grad_pl", 3 * P), (cmask, "cmask", 6 * R), (cmask_pl, "cmask_pl", P),
                         (GRD_xc, "GRD_xc", 9 * R), (GRD_xc_pl, "GRD_xc_pl", 3 * P),
                         (GRD_x, "GRD_x", 3 * self.gall * self.lall), (GRD_x_pl, "GRD_x_pl", 6 * self.gall_pl)])
        _check(self.lib.nicam_horizontal_remap_adjoint(self.h, *a), "nicam_horizontal_remap_adjoint")

    def tracer_update_adjoint(self, flx_h_b, flx_h_b_pl, q_a_b, q_a_b_pl, rhogq_b, rhogq_b_pl, flx_h, flx_h_pl, q_a,
                              q_a_pl) -> None:
        R, P = self.nreg, self.npl
        a = [_ptr(flx_h_b, self.mode, "flx_h_b", 6 * R), _ptr(flx_h_b_pl, self.mode, "flx_h_b_pl", P),
             _ptr(q_a_b, self.mode, "q_a_b", 6 * R), _ptr(q_a_b_pl, self.mode, "q_a_b_pl", P)]
        a += self._args([(rhogq_b, "rhogq_b", R), (rhogq_b_pl, "rhogq_b_pl", P), (flx_h, "flx_h", 6 * R),
                         (flx_h_pl, "flx_h_pl", P), (q_a, "q_a", 6 * R), (q_a_pl, "q_a_pl", P)])
        _check(self.lib.nicam_tracer_update_adjoint(self.h, *a), "nicam_tracer_update_adjoint")

    def vertical_courant_adjoint(self, flx_v_b, flx_v_b_pl, rhog_b, rhog_b_pl, ck_b, ck_b_pl, flx_v, flx_v_pl, rhog,
                                 rhog_pl) -> None:
        R, P = self.nreg, self.npl
        a = [_ptr(flx_v_b, self.mode, "flx_v_b", R), _ptr(flx_v_b_pl, self.mode, "flx_v_b_pl", P),
             _ptr(rhog_b, self.mode, "rhog_b", R), _ptr(rhog_b_pl, self.mode, "rhog_b_pl", P)]
        a += self._args([(ck_b, "ck_b", 2 * R), (ck_b_pl, "ck_b_pl", 2 * P), (flx_v, "flx_v", R),
                         (flx_v_pl, "flx_v_pl", P), (rhog, "rhog", R), (rhog_pl, "rhog_pl", P)])
        _check(self.lib.nicam_vertical_courant_adjoint(self.h, *a), "nicam_vertical_courant_adjoint")

    def vertical_face_adjoint(self, q_b, q_b_pl, q_h_b, q_h_b_pl) -> None:
        R, P = self.nreg, self.npl
        a = [_ptr(q_b, self.mode, "q_b", R), _ptr(q_b_pl, self.mode, "q_b_pl", P)]
        a += self._args([(q_h_b, "q_h_b", R), (q_h_b_pl, "q_h_b_pl", P)])
        _check(self.lib.nicam_vertical_face_adjoint(self.h, *a), "nicam_vertical_face_adjoint")

    def tracer_update_vertical_adjoint(self, flx_v_b, flx_v_b_pl, q_h_b, q_h_b_pl, rhogq_b, rhogq_b_pl, flx_v,
                                       flx_v_pl, q_h, q_h_pl) -> None:
        R, P = self.nreg, self.npl
        a = [_ptr(flx_v_b, self.mode, "flx_v_b", R), _ptr(flx_v_b_pl, self.mode, "flx_v_b_pl", P),
             _ptr(q_h_b, self.mode, "q_h_b", R), _ptr(q_h_b_pl, self.mode, "q_h_b_pl", P)]
        a += self._args([(rhogq_b, "rhogq_b", R), (rhogq_b_pl, "rhogq_b_pl", P), (flx_v, "flx_v", R),
                         (flx_v_pl, "flx_v_pl", P), (q_h, "q_h", R), (q_h_pl, "q_h_pl", P)])
        _check(self.lib.nicam_tracer_update_vertical_adjoint(self.h, *a), "nicam_tracer_update_vertical_adjoint")

    def tracer_mass(self, rhogq, rhogq_pl, mass, single: bool = False) -> None:
        """nicam_tracer_mass(_f32): mass[k] = FP64 sum of AREA*rhogq over the inner cells and pole centres."""
        R, P = self.nreg, self.npl
        a = self._args([(rhogq, "rhogq", R), (rhogq_pl, "rhogq_pl", P)], single=single)
        a.append(_ptr(mass, self.mode, "mass", self.kall))
        fn = "nicam_tracer_mass_f32" if single else "nicam_tracer_mass"
        _check(getattr(self.lib, fn)(self.h, *a), fn)

    def tracer_update_conservative(self, rhogq, rhogq_pl, flx_h, q_a, single: bool = False) -> None:
        """nicam_tracer_update_conservative(_f32): the flux-form update with one transfer per arc."""
        R, P = self.nreg, self.npl
        a = [_ptr(rhogq, self.mode, "rhogq", R), None if not self.have_pl else _ptr(rhogq_pl, self.mode,
                                                                                    "rhogq_pl", P)]
        a += [_ptr(flx_h, self.mode, "flx_h", 6 * R, single), _ptr(q_a, self.mode, "q_a", 6 * R, single)]
        fn = "nicam_tracer_update_conservative_f32" if single else "nicam_tracer_update_conservative"
        _check(getattr(self.lib, fn)(self.h, *a), fn)

    def vertical_mass_flux(self, flx_v, flx_v_pl, rhogvx, rhogvx_pl, rhogvy, rhogvy_pl, rhogvz, rhogvz_pl,
                           rhogw, rhogw_pl, dt: float) -> None:
        """nicam_vertical_mass_flux: flx_v = rhogw_vm*dt on the interior half levels."""
        R, P = self.nreg, self.npl
        a = self._args([(flx_v, "flx_v", R), (flx_v_pl, "flx_v_pl", P), (rhogvx, "rhogvx", R),
                        (rhogvx_pl, "rhogvx_pl", P), (rhogvy, "rhogvy", R), (rhogvy_pl, "rhogvy_pl", P),
                        (rhogvz, "rhogvz", R), (rhogvz_pl, "rhogvz_pl", P), (rhogw, "rhogw", R),
                        (rhogw_pl, "rhogw_pl", P)])
        _check(self.lib.nicam_vertical_mass_flux(self.h, *a, float(dt)), "nicam_vertical_mass_flux")

    def vertical_courant(self, ck, ck_pl, flx_v, flx_v_pl, rhog, rhog_pl) -> None:
        """nicam_vertical_courant: ck(...,2) of the lower and upper faces."""
        R, P = self.nreg, self.npl
        a = self._args([(ck, "ck", 2 * R), (ck_pl, "ck_pl", 2 * P), (flx_v, "flx_v", R), (flx_v_pl, "flx_v_pl", P),
                        (rhog, "rhog", R), (rhog_pl, "rhog_pl", P)])
        _check(self.lib.nicam_vertical_courant(self.h, *a), "nicam_vertical_courant")

    def vertical_face(self, q_h, q_h_pl, q, q_pl) -> None:
        """nicam_vertical_face: q_h = afact q(k) + bfact q(k-1) on kmin+1..kmax."""
        R, P = self.nreg, self.npl
        a = self._args([(q_h, "q_h", R), (q_h_pl, "q_h_pl", P), (q, "q", R), (q_pl, "q_pl", P)])
        _check(self.lib.nicam_vertical_face(self.h, *a), "nicam_vertical_face")

    def tracer_update_vertical(self, rhogq, rhogq_pl, flx_v, flx_v_pl, q_h, q_h_pl) -> None:
        """nicam_tracer_update_vertical: rhogq -= the vertical flux divergence of flx_v*q_h."""
        R, P = self.nreg, self.npl
        a = self._args([(rhogq, "rhogq", R), (rhogq_pl, "rhogq_pl", P), (flx_v, "flx_v", R),
                        (flx_v_pl, "flx_v_pl", P), (q_h, "q_h", R), (q_h_pl, "q_h_pl", P)])
        _check(self.lib.nicam_tracer_update_vertical(self.h, *a), "nicam_tracer_update_vertical")

    def horizontal_limiter_thuburn_batch(self, ntr, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl,
                                         Qout_prev, Qout_prev_pl, Qout_post, Qout_post_pl, single: bool = False):
        """nicam_horizontal_limiter_thuburn_batch(_f32): ntr tracers, tracer-last arrays, device pointers."""
        R, P, n = self.nreg, self.npl, int(ntr)
        spec = [(q_a, "q_a", 6 * R * n), (q_a_pl, "q_a_pl", P * n), (q, "q", R * n), (q_pl, "q_pl", P * n),
                (d, "d", R), (d_pl, "d_pl", P), (ch, "ch", 6 * R), (ch_pl, "ch_pl", P), (cmask, "cmask", 6 * R),
                (cmask_pl, "cmask_pl", P), (Qout_prev, "Qout_prev", 2 * R * n), (Qout_prev_pl, "Qout_prev_pl", 2 * P * n),
                (Qout_post, "Qout_post", 2 * R * n), (Qout_post_pl, "Qout_post_pl", 2 * P * n)]
        a = self._args(spec, single=single)
        fn = "nicam_horizontal_limiter_thuburn_batch" + ("_f32" if single else "")
        _check(getattr(self.lib, fn)(self.h, n, *a), fn)

    def vertical_limiter_thuburn_batch(self, ntr, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl, single: bool = False):
        """nicam_vertical_limiter_thuburn_batch(_f32): ntr tracers, tracer-last arrays, device pointers."""
        R, P, n = self.nreg, self.npl, int(ntr)
        spec = [(q_h, "q_h", R * n), (q_h_pl, "q_h_pl", P * n), (q, "q", R * n), (q_pl, "q_pl", P * n),
                (d, "d", R), (d_pl, "d_pl", P), (ck, "ck", 2 * R), (ck_pl, "ck_pl", 2 * P)]
        a = self._args(spec, single=single)
        fn = "nicam_vertical_limiter_thuburn_batch" + ("_f32" if single else "")
        _check(getattr(self.lib, fn)(self.h, n, *a), fn)

    def horizontal_limiter_thuburn_batch_f32(self, ntr, *a):
        self.horizontal_limiter_thuburn_batch(ntr, *a, single=True)

    def vertical_limiter_thuburn_batch_f32(self, ntr, *a):
        self.vertical_limiter_thuburn_batch(ntr, *a, single=True)

    def tracer_update_vertical_f32(self, rhogq, rhogq_pl, flx_v, flx_v_pl, q_h, q_h_pl) -> None:
        """nicam_tracer_update_vertical_f32: FP32 flx_v and q_h, FP64 rhogq."""
        R, P = self.nreg, self.npl
        a = self._args([(rhogq, "rhogq", R), (rhogq_pl, "rhogq_pl", P)])
        a += self._args([(flx_v, "flx_v", R), (flx_v_pl, "flx_v_pl", P), (q_h, "q_h", R), (q_h_pl, "q_h_pl", P)],
                        single=True)
        _check(self.lib.nicam_tracer_update_vertical_f32(self.h, *a), "nicam_tracer_update_vertical_f32")

    def vi_rhow_update_matrix(self, eth, eth_pl, g_tilde, g_tilde_pl, dt: float) -> None:
        """nicam_vi_rhow_update_matrix: Mc/Ml/Mu of the context from h and g~ at the half levels."""
        R, P = self.nreg, self.npl
        a = self._args([(eth, "eth", R), (eth_pl, "eth_pl", P), (g_tilde, "g_tilde", R),
                        (g_tilde_pl, "g_tilde_pl", P)])
        _check(self.lib.nicam_vi_rhow_update_matrix(self.h, *a, float(dt)), "nicam_vi_rhow_update_matrix")

    def vi_rhow_getcoef(self, Mc, Mc_pl, Ml, Ml_pl, Mu, Mu_pl) -> None:
        """nicam_vi_rhow_getcoef: copies the context's Mc/Ml/Mu out."""
        R, P = self.nreg, self.npl
        a = self._args([(Mc, "Mc", R), (Mc_pl, "Mc_pl", P), (Ml, "Ml", R), (Ml_pl, "Ml_pl", P), (Mu, "Mu", R),
                        (Mu_pl, "Mu_pl", P)])
        _check(self.lib.nicam_vi_rhow_getcoef(self.h, *a), "nicam_vi_rhow_getcoef")

    def oprt3d_divdamp(self, ddivdx, ddivdx_pl, ddivdy, ddivdy_pl, ddivdz, ddivdz_pl, rhogvx, rhogvx_pl,
                       rhogvy, rhogvy_pl, rhogvz, rhogvz_pl, rhogw, rhogw_pl, coef_intp, coef_intp_pl,
                       coef_diff, coef_diff_pl) -> None:
        """OPRT3D_divdamp (dyn_divdamp/src/mod_oprt3d.f90:41-402)."""
        R, P, G = self.nreg, self.npl, self.gall * self.lall
        a = [_ptr(x, self.mode, nm, n) for x, nm, n in
             [(ddivdx, "ddivdx", R), (ddivdx_pl, "ddivdx_pl", P), (ddivdy, "ddivdy", R), (ddivdy_pl, "ddivdy_pl", P),
              (ddivdz, "ddivdz", R), (ddivdz_pl, "ddivdz_pl", P)]]
        a += self._args([(rhogvx, "rhogvx", R), (rhogvx_pl, "rhogvx_pl", P), (rhogvy, "rhogvy", R),
                         (rhogvy_pl, "rhogvy_pl", P), (rhogvz, "rhogvz", R), (rhogvz_pl, "rhogvz_pl", P),
                         (rhogw, "rhogw", R), (rhogw_pl, "rhogw_pl", P), (coef_intp, "coef_intp", 18 * G),
                         (coef_intp_pl, "coef_intp_pl", 18 * self.gall_pl), (coef_diff, "coef_diff", 18 * G),
                         (coef_diff_pl, "coef_diff_pl", 30)])
        _check(self.lib.nicam_oprt3d_divdamp(self.h, *a), "nicam_oprt3d_divdamp")

    def horizontal_flux_adjoint(self, flx_h_b, flx_h_b_pl, GRD_xc_b, GRD_xc_b_pl, rho_b, rho_b_pl, rhovx_b,
                                rhovx_b_pl, rhovy_b, rhovy_b_pl, rhovz_b, rhovz_b_pl, rho, rho_pl, rhovx, rhovx_pl,
                                rhovy, rhovy_pl, rhovz, rhovz_pl, dt: float) -> None:
        """nicam_horizontal_flux_adjoint: cotangents of rho and rho*v from those of flx_h and GRD_xc."""
        R, P = self.nreg, self.npl
        a = self._args([(flx_h_b, "flx_h_b", 6 * R), (flx_h_b_pl, "flx_h_b_pl", P), (GRD_xc_b, "GRD_xc_b", 9 * R),
                        (GRD_xc_b_pl, "GRD_xc_b_pl", 3 * P), (rho_b, "rho_b", R), (rho_b_pl, "rho_b_pl", P),
                        (rhovx_b, "rhovx_b", R), (rhovx_b_pl, "rhovx_b_pl", P), (rhovy_b, "rhovy_b", R),
                        (rhovy_b_pl, "rhovy_b_pl", P), (rhovz_b, "rhovz_b", R), (rhovz_b_pl, "rhovz_b_pl", P),
                        (rho, "rho", R), (rho_pl, "rho_pl", P), (rhovx, "rhovx", R), (rhovx_pl, "rhovx_pl", P),
                        (rhovy, "rhovy", R), (rhovy_pl, "rhovy_pl", P), (rhovz, "rhovz", R),
                        (rhovz_pl, "rhovz_pl", P)])
        _check(self.lib.nicam_horizontal_flux_adjoint(self.h, *a, float(dt)), "nicam_horizontal_flux_adjoint")

    def horizontal_limiter_thuburn_mask(self, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl,
                                        Qout_prev, Qout_prev_pl, Qout_post, Qout_post_pl, mask, mask_pl) -> None:
        """nicam_horizontal_limiter_thuburn_mask: the limiter plus its uint8 change bits per cell-level."""
        R, P = self.nreg, self.npl
        a = self._args([(q_a, "q_a", 6 * R), (q_a_pl, "q_a_pl", P), (q, "q", R), (q_pl, "q_pl", P), (d, "d", R),
                        (d_pl, "d_pl", P), (ch, "ch", 6 * R), (ch_pl, "ch_pl", P), (cmask, "cmask", 6 * R),
                        (cmask_pl, "cmask_pl", P), (Qout_prev, "Qout_prev", 2 * R),
                        (Qout_prev_pl, "Qout_prev_pl", 2 * P), (Qout_post, "Qout_post", 2 * R),
                        (Qout_post_pl, "Qout_post_pl", 2 * P)])
        a += [_ptr_u8(mask, self.mode, "mask", R), None if not self.have_pl else _ptr_u8(mask_pl, self.mode,
                                                                                         "mask_pl", P)]
        _check(self.lib.nicam_horizontal_limiter_thuburn_mask(self.h, *a), "nicam_horizontal_limiter_thuburn_mask")

    def horizontal_limiter_thuburn_adjoint(self, q_a_b, q_a_b_pl, q_b, q_b_pl, Qout_prev_b, Qout_prev_b_pl,
                                           Qout_post_b, Qout_post_b_pl, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl,
                                           cmask, cmask_pl, Qout_post, Qout_post_pl, mask, mask_pl, ch_b=None,
                                           ch_b_pl=None) -> None:
        """nicam_horizontal_limiter_thuburn_adjoint (Qout_post supplied); ch_b, ch_b_pl: the Courant
        numbers' cotangent, formed when given.  q_a_b None: the rows alone (Qout_post_b, q_a, Qout_post and
        mask may be None too); Qout_prev_b None: the faces alone (no ch_b)."""
        R, P = self.nreg, self.npl
        a = self._args([(q_a_b, "q_a_b", 6 * R), (q_a_b_pl, "q_a_b_pl", P), (q_b, "q_b", R), (q_b_pl, "q_b_pl", P),
                        (Qout_prev_b, "Qout_prev_b", 2 * R), (Qout_prev_b_pl, "Qout_prev_b_pl", 2 * P),
                        (Qout_post_b, "Qout_post_b", 2 * R), (Qout_post_b_pl, "Qout_post_b_pl", 2 * P),
                        (q_a, "q_a", 6 * R), (q_a_pl, "q_a_pl", P), (q, "q", R), (q_pl, "q_pl", P), (d, "d", R),
                        (d_pl, "d_pl", P), (ch, "ch", 6 * R), (ch_pl, "ch_pl", P), (cmask, "cmask", 6 * R),
                        (cmask_pl, "cmask_pl", P), (Qout_post, "Qout_post", 2 * R),
                        (Qout_post_pl, "Qout_post_pl", 2 * P)])
        a += [_ptr_u8(mask, self.mode, "mask", R), None if not self.have_pl else _ptr_u8(mask_pl, self.mode,
                                                                                         "mask_pl", P)]
        a += self._args([(ch_b, "ch_b", 6 * R), (ch_b_pl, "ch_b_pl", P)])
        _check(self.lib.nicam_horizontal_limiter_thuburn_adjoint(self.h, *a),
               "nicam_horizontal_limiter_thuburn_adjoint")

    def vlim_mask_words(self):
        """uint32 words of the vertical limiter's branch mask: (regions, poles)."""
        nword = (self.kall - 3 + 31) // 32 if self.kall > 3 else 0
        return nword * self.gall * self.lall, nword * self.gall_pl * 2

    def vertical_limiter_thuburn_mask(self, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl, mask, mask_pl) -> None:
        """nicam_vertical_limiter_thuburn_mask: the limiter plus its clip bits (uint32 arrays)."""
        R, P = self.nreg, self.npl
        M, MP = self.vlim_mask_words()
        a = self._args([(q_h, "q_h", R), (q_h_pl, "q_h_pl", P), (q, "q", R), (q_pl, "q_pl", P), (d, "d", R),
                        (d_pl, "d_pl", P), (ck, "ck", 2 * R), (ck_pl, "ck_pl", 2 * P)])
        a += [_ptr_u32(mask, self.mode, "mask", M), None if not self.have_pl else _ptr_u32(mask_pl, self.mode,
                                                                                           "mask_pl", MP)]
        _check(self.lib.nicam_vertical_limiter_thuburn_mask(self.h, *a), "nicam_vertical_limiter_thuburn_mask")

    def vertical_limiter_thuburn_adjoint(self, q_h_b, q_h_b_pl, q_b, q_b_pl, q_h, q_h_pl, q, q_pl, d, d_pl, ck,
                                         ck_pl, mask, mask_pl, ck_b=None, ck_b_pl=None) -> None:
        """nicam_vertical_limiter_thuburn_adjoint: cotangents of q_h's input, of q and (when ck_b is
        given) of ck."""
        R, P = self.nreg, self.npl
        M, MP = self.vlim_mask_words()
        a = self._args([(q_h_b, "q_h_b", R), (q_h_b_pl, "q_h_b_pl", P), (q_b, "q_b", R), (q_b_pl, "q_b_pl", P),
                        (q_h, "q_h", R), (q_h_pl, "q_h_pl", P), (q, "q", R), (q_pl, "q_pl", P), (d, "d", R),
                        (d_pl, "d_pl", P), (ck, "ck", 2 * R), (ck_pl, "ck_pl", 2 * P)])
        a += [_ptr_u32(mask, self.mode, "mask", M), None if not self.have_pl else _ptr_u32(mask_pl, self.mode,
                                                                                           "mask_pl", MP)]
        a += self._args([(ck_b, "ck_b", 2 * R), (ck_b_pl, "ck_b_pl", 2 * P)])
        _check(self.lib.nicam_vertical_limiter_thuburn_adjoint(self.h, *a),
               "nicam_vertical_limiter_thuburn_adjoint")

    def vi_rhow_solver_mixed(self, rhogw, rhogw_pl, rhogw0, rhogw0_pl, preg0, preg0_pl, rhog0, rhog0_pl, Srho,
                             Srho_pl, Sw, Sw_pl, Spre, Spre_pl, dt: float) -> None:
        """nicam_vi_rhow_solver_mixed: float sources and state, double rhogw, FP64 elimination."""
        R, P = self.nreg, self.npl
        a = [_ptr(rhogw, self.mode, "rhogw", R), None if not self.have_pl else _ptr(rhogw_pl, self.mode,
                                                                                    "rhogw_pl", P)]
        names = ["rhogw0", "preg0", "rhog0", "Srho", "Sw", "Spre"]
        arrs = [rhogw0, rhogw0_pl, preg0, preg0_pl, rhog0, rhog0_pl, Srho, Srho_pl, Sw, Sw_pl, Spre, Spre_pl]
        spec = []
        for i, nm in enumerate(names):
            spec.append((arrs[2 * i], nm, R))
            spec.append((arrs[2 * i + 1], nm + "_pl", P))
        a += self._args(spec, single=True)
        _check(self.lib.nicam_vi_rhow_solver_mixed(self.h, *a, float(dt)), "nicam_vi_rhow_solver_mixed")

    def vi_rhow_solver_adjoint(self, rhogw_b, rhogw_b_pl, rhogw0_b, rhogw0_b_pl, preg0_b, preg0_b_pl, rhog0_b,
                               rhog0_b_pl, Srho_b, Srho_b_pl, Sw_b, Sw_b_pl, Spre_b, Spre_b_pl, dt: float) -> None:
        """nicam_vi_rhow_solver_adjoint: rhogw_b in/out, the six source cotangents written."""
        R, P = self.nreg, self.npl
        names = ["rhogw_b", "rhogw0_b", "preg0_b", "rhog0_b", "Srho_b", "Sw_b", "Spre_b"]
        arrs = [rhogw_b, rhogw_b_pl, rhogw0_b, rhogw0_b_pl, preg0_b, preg0_b_pl, rhog0_b, rhog0_b_pl, Srho_b,
                Srho_b_pl, Sw_b, Sw_b_pl, Spre_b, Spre_b_pl]
        spec = []
        for i, nm in enumerate(names):
            spec.append((arrs[2 * i], nm, R))
            spec.append((arrs[2 * i + 1], nm + "_pl", P))
        _check(self.lib.nicam_vi_rhow_solver_adjoint(self.h, *self._args(spec), float(dt)),
               "nicam_vi_rhow_solver_adjoint")

    def vi_rhow_solver(self, rhogw, rhogw_pl, rhogw0, rhogw0_pl, preg0, preg0_pl, rhog0, rhog0_pl,
                       Srho, Srho_pl, Sw, Sw_pl, Spre, Spre_pl, dt: float) -> None:
        R, P = self.nreg, self.npl
        names = ["rhogw", "rhogw0", "preg0", "rhog0", "Srho", "Sw", "Spre"]
        arrs = [rhogw, rhogw_pl, rhogw0, rhogw0_pl, preg0, preg0_pl, rhog0, rhog0_pl, Srho, Srho_pl, Sw, Sw_pl,
                Spre, Spre_pl]
        spec = []
        for i, nm in enumerate(names):
            spec.append((arrs[2 * i], nm, R))
            spec.append((arrs[2 * i + 1], nm + "_pl", P))
        a = self._args(spec)
        _check(self.lib.nicam_vi_rhow_solver(self.h, *a, float(dt)), "nicam_vi_rhow_solver")


def grd_input_vgrid(fname: str, kall: int):
    """GRD_input_vgrid: (gz, gzh) from a Fortran unformatted big-endian file."""
    lib = load_library()
    gz, gzh = np.empty(kall), np.empty(kall)
    _check(lib.nicam_grd_input_vgrid(str(fname).encode(), kall, gz.ctypes.data, gzh.ctypes.data),
           "nicam_grd_input_vgrid")
    return gz, gzh


def halo_map(rlevel: int, gall_1d: int, gall_pl: int = 6):
    """The icosahedral halo map (nicam_halo_map): a dict of int32 arrays
    dst_kind, dst_l, dst_g, src_kind, src_l, src_g (kind 0 regular, 1 pole;
    all indices 0-based), have_sgp (one flag per region), and pole_fill_l,
    pole_fill_g (the regular point each pole centre takes in COMM_var)."""
    lib = load_library()
    n = c_int(0)
    _check(lib.nicam_halo_map(rlevel, gall_1d, gall_pl, 0, ctypes.byref(n), *([None] * 7)), "nicam_halo_map")
    keys = ["dst_kind", "dst_l", "dst_g", "src_kind", "src_l", "src_g"]
    out = {k: np.empty(n.value, dtype=np.int32) for k in keys}
    out["have_sgp"] = np.empty(10 * 4 ** rlevel, dtype=np.int32)
    m = c_int(0)
    _check(lib.nicam_halo_map(rlevel, gall_1d, gall_pl, n.value, ctypes.byref(m),
                              *[out[k].ctypes.data for k in keys + ["have_sgp"]]), "nicam_halo_map")
    if m.value != n.value:
        raise RuntimeError("nicam_halo_map changed size between calls")
    src_l, src_g = np.empty(2, dtype=np.int32), np.empty(2, dtype=np.int32)
    _check(lib.nicam_pole_fill_map(rlevel, gall_1d, src_l.ctypes.data, src_g.ctypes.data), "nicam_pole_fill_map")
    out["pole_fill_l"], out["pole_fill_g"] = src_l, src_g
    return out


def grid_icosahedral(glevel: int, rlevel: int, radius: float = 6.37122e6):
    """nicam_grid_icosahedral: (GRD_x[3][lall][1][gall], GRD_x_pl[3][2][1][6])."""
    lib = load_library()
    n1 = 2 ** (glevel - rlevel) + 2
    lall = 10 * 4 ** rlevel
    x = np.empty((3, lall, 1, n1 * n1))
    x_pl = np.empty((3, 2, 1, 6))
    _check(lib.nicam_grid_icosahedral(glevel, rlevel, float(radius), x.ctypes.data, x_pl.ctypes.data),
           "nicam_grid_icosahedral")
    return x, x_pl


def grid_vertices(x, x_pl, have_sgp, radius: float = 6.37122e6):
    """nicam_grid_vertices for GRD_x[3][lall][1][gall] and GRD_x_pl[3][2][1][gall_pl]:
    a dict xt[3][2][lall][1][gall], xt_pl[3][2][1][gall_pl], s[2][lall][1][gall],
    s_pl[2][2][1][gall_pl], xr[3][3][lall][1][gall], xr_pl[3][2][1][gall_pl]
    (the Fortran shapes reversed)."""
    lib = load_library()
    x = np.ascontiguousarray(x, dtype=np.float64)
    x_pl = np.ascontiguousarray(x_pl, dtype=np.float64)
    _, lall, _, gall = x.shape
    gall_pl = x_pl.shape[-1]
    n1 = int(round(gall ** 0.5))
    sgp = np.ascontiguousarray(have_sgp, dtype=np.int32)
    if sgp.size != lall or n1 * n1 != gall:
        raise ValueError("have_sgp needs one flag per region and gall must be a square")
    out = {"xt": np.empty((3, 2, lall, 1, gall)), "xt_pl": np.empty((3, 2, 1, gall_pl)),
           "s": np.empty((2, lall, 1, gall)), "s_pl": np.empty((2, 2, 1, gall_pl)),
           "xr": np.empty((3, 3, lall, 1, gall)), "xr_pl": np.empty((3, 2, 1, gall_pl))}
    _check(lib.nicam_grid_vertices(n1, lall, gall_pl, 1, sgp.ctypes.data, float(radius), x.ctypes.data,
                                   x_pl.ctypes.data, *[out[k].ctypes.data for k in
                                                       ("xt", "xt_pl", "s", "s_pl", "xr", "xr_pl")]),
           "nicam_grid_vertices")
    return out


def gmtr_setup(x, x_pl, xt, xt_pl, s, s_pl, have_sgp, radius: float = 6.37122e6, have_pl: bool = True):
    """nicam_gmtr_setup: a dict p[8][lall][1][gall], p_pl[8][2][1][gall_pl],
    t[5][2][lall][1][gall], t_pl[5][2][1][gall_pl], a[12][3][lall][1][gall],
    a_pl[18][2][1][gall_pl] (the Fortran shapes reversed)."""
    lib = load_library()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, x_pl, xt, xt_pl, s, s_pl)]
    _, lall, _, gall = arrs[0].shape
    gall_pl = arrs[1].shape[-1]
    n1 = int(round(gall ** 0.5))
    sgp = np.ascontiguousarray(have_sgp, dtype=np.int32)
    if sgp.size != lall or n1 * n1 != gall:
        raise ValueError("have_sgp needs one flag per region and gall must be a square")
    out = {"p": np.empty((8, lall, 1, gall)), "p_pl": np.empty((8, 2, 1, gall_pl)),
           "t": np.empty((5, 2, lall, 1, gall)), "t_pl": np.empty((5, 2, 1, gall_pl)),
           "a": np.empty((12, 3, lall, 1, gall)), "a_pl": np.empty((18, 2, 1, gall_pl))}
    _check(lib.nicam_gmtr_setup(n1, lall, gall_pl, int(have_pl), sgp.ctypes.data, float(radius),
                                *[a.ctypes.data for a in arrs],
                                *[out[k].ctypes.data for k in ("p", "p_pl", "t", "t_pl", "a", "a_pl")]),
           "nicam_gmtr_setup")
    return out


def oprt_diffusion_setup(m, have_sgp, have_pl: bool = True):
    """nicam_oprt_diffusion_setup on the metrics dict of gmtr_setup: a dict
    coef_intp[lall][2][3][3][gall], coef_intp_pl[2][3][3][gall_pl],
    coef_diff[lall][3][6][gall], coef_diff_pl[2][3][gall_pl-1]."""
    lib = load_library()
    ins = [np.ascontiguousarray(m[k], dtype=np.float64) for k in ("p", "p_pl", "t", "t_pl", "a", "a_pl")]
    _, lall, _, gall = ins[0].shape
    gall_pl = ins[1].shape[-1]
    n1 = int(round(gall ** 0.5))
    sgp = np.ascontiguousarray(have_sgp, dtype=np.int32)
    if sgp.size != lall or n1 * n1 != gall:
        raise ValueError("have_sgp needs one flag per region and gall must be a square")
    out = {"coef_intp": np.empty((lall, 2, 3, 3, gall)), "coef_intp_pl": np.empty((2, 3, 3, gall_pl)),
           "coef_diff": np.empty((lall, 3, 6, gall)), "coef_diff_pl": np.empty((2, 3, gall_pl - 1))}
    _check(lib.nicam_oprt_diffusion_setup(n1, lall, gall_pl, int(have_pl), sgp.ctypes.data,
                                          *[a.ctypes.data for a in ins],
                                          *[out[k].ctypes.data for k in ("coef_intp", "coef_intp_pl", "coef_diff",
                                                                         "coef_diff_pl")]),
           "nicam_oprt_diffusion_setup")
    return out


def oprt_coef_setup(which: str, m, have_sgp, have_pl: bool = True):
    """nicam_oprt_{divergence,rotation,gradient,laplacian}_setup ("div", "rot",
    "grad", "lap") on the metrics dict of gmtr_setup: (coef, coef_pl) as
    coef[lall][3][7][gall], coef_pl[2][3][gall_pl] (coef_lap[lall][7][gall],
    coef_lap_pl[2][gall_pl] for "lap")."""
    names = {"div": "nicam_oprt_divergence_setup", "rot": "nicam_oprt_rotation_setup",
             "grad": "nicam_oprt_gradient_setup", "lap": "nicam_oprt_laplacian_setup"}
    if which not in names:
        raise ValueError(f"which must be one of {sorted(names)}")
    lib = load_library()
    ins = [np.ascontiguousarray(m[k], dtype=np.float64) for k in ("p", "p_pl", "t", "t_pl", "a", "a_pl")]
    _, lall, _, gall = ins[0].shape
    gall_pl = ins[1].shape[-1]
    n1 = int(round(gall ** 0.5))
    sgp = np.ascontiguousarray(have_sgp, dtype=np.int32)
    if sgp.size != lall or n1 * n1 != gall:
        raise ValueError("have_sgp needs one flag per region and gall must be a square")
    if which == "lap":
        coef, coef_pl = np.empty((lall, 7, gall)), np.empty((2, gall_pl))
    else:
        coef, coef_pl = np.empty((lall, 3, 7, gall)), np.empty((2, 3, gall_pl))
    _check(getattr(lib, names[which])(n1, lall, gall_pl, int(have_pl), sgp.ctypes.data,
                                      *[a.ctypes.data for a in ins], coef.ctypes.data, coef_pl.ctypes.data),
           names[which])
    return coef, coef_pl
