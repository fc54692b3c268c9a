"""Comparison helpers shared by the parity tests.

The bar for the FP64 path is 1e-12 relative to the oracle (element-wise,
|gpu - oracle| <= 1e-12 * |oracle|); the kernels keep the reference's
operation order and are built without FMA contraction, so they are expected
to be identical to the last bit, which is checked too (-0.0 == +0.0).
"""
import numpy as np

RTOL_FP64 = 1e-12


def assert_parity(got: dict, want: dict, rtol: float = RTOL_FP64, exact: bool = True, skip=()):
    for k in want:
        if k in skip:
            continue
        a, b = np.asarray(got[k]), np.asarray(want[k])
        assert a.shape == b.shape, f"{k}: shape {a.shape} != {b.shape}"
        with np.errstate(invalid="ignore", over="ignore"):  # equal infinities are equal
            bad = ~(np.abs(a - b) <= rtol * np.abs(b))
        bad &= ~((a == b) | (np.isnan(a) & np.isnan(b)))
        if bad.any():
            idx = np.argwhere(bad)[:5]
            raise AssertionError(f"{k}: {bad.sum()} of {a.size} elements beyond rtol={rtol}; first at {idx.tolist()}: "
                                 f"got {a[tuple(idx[0])]!r} want {b[tuple(idx[0])]!r}")
        if exact:
            neq = ~((a == b) | (np.isnan(a) & np.isnan(b)))
            assert not neq.any(), f"{k}: {neq.sum()} elements differ in the last bits"


def region_slice(d: dict, regions, pole: bool, reg_axis_of=None):
    """Sub-arrays of the regions listed (the region axis is the one of size lall
    right before (kall, gall)); pole arrays are kept only if pole is True."""
    out = {}
    for k, v in d.items():
        if k.endswith("_pl"):
            if pole:
                out[k] = v
            continue
        if k == "have_sgp":
            out[k] = np.asarray(v)[list(regions)]
            continue
        if k == "points":
            out[k] = v[list(regions)]
            continue
        if k == "coef_grad":  # [lall][3][7][gall]
            out[k] = np.asarray(v)[list(regions)]
            continue
        if k == "GRD_x":  # [3][lall][1][gall]
            out[k] = np.asarray(v)[:, list(regions)]
            continue
        v = np.asarray(v)
        if v.ndim < 3:  # per-level vectors (gz, gzh, ...)
            out[k] = v
            continue
        out[k] = np.take(v, list(regions), axis=v.ndim - 3)
    return out
