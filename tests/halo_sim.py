"""Host-side application of a halo map, used by the tests as the expected
result of the on-device halo fill.

Fields are C arrays with the Fortran dimensions reversed:
var[v][l][k][g] for var(gall,kmax,lall,vmax) and var_pl[v][p][k][gp] for
var_pl(gall_pl,kmax,2,vmax).  Every destination of the map is written from
the value its source held before the call (no destination is also a source),
which is how the reference's copy lists behave.
"""
from __future__ import annotations

import numpy as np


def apply_map(m, var: np.ndarray, var_pl: np.ndarray, pole_fill: bool = False):
    """Returns copies of (var, var_pl) after COMM_data_transfer (or COMM_var
    when pole_fill) with halo map m (runtime.halo_map's dict)."""
    var, var_pl = var.copy(), var_pl.copy()
    if pole_fill:  # pole centres first: a source of the exchange below
        for p in range(2):
            var_pl[:, p, :, 0] = var[:, m["pole_fill_l"][p], :, m["pole_fill_g"][p]]
    sk, sl, sg = m["src_kind"], m["src_l"], m["src_g"]
    # var[:, L, :, G] with index arrays L, G has shape (n, vmax, kmax)
    vals = np.empty((len(sk), var.shape[0], var.shape[2]), dtype=var.dtype)
    r = sk == 0
    vals[r] = var[:, sl[r], :, sg[r]]
    vals[~r] = var_pl[:, sl[~r], :, sg[~r]]
    dk, dl, dg = m["dst_kind"], m["dst_l"], m["dst_g"]
    r = dk == 0
    var[:, dl[r], :, dg[r]] = vals[r]
    var_pl[:, dl[~r], :, dg[~r]] = vals[~r]
    return var, var_pl


def apply_map_adjoint(m, var_b: np.ndarray, var_b_pl: np.ndarray):
    """The transpose of apply_map (COMM_data_transfer) applied to cotangents:
    every destination's cotangent is added to its source's and the
    destination's set to zero.  A source's additions are made in ascending
    order of the destination's code (l*gall + g for a region cell, then
    lall*gall + p*gall_pl + v for a pole point), one at a time."""
    var_b, var_b_pl = var_b.copy(), var_b_pl.copy()
    lall, gall, gall_pl = var_b.shape[1], var_b.shape[3], var_b_pl.shape[3]
    code = lambda kind, l, g: np.where(kind == 0, l * gall + g, lall * gall + l * gall_pl + g)  # noqa: E731
    dk, dl, dg = m["dst_kind"], m["dst_l"], m["dst_g"]
    sk, sl, sg = m["src_kind"], m["src_l"], m["src_g"]
    dc, sc = code(dk, dl, dg), code(sk, sl, sg)
    vals = np.empty((len(dk), var_b.shape[0], var_b.shape[2]), dtype=var_b.dtype)
    r = dk == 0
    vals[r] = var_b[:, dl[r], :, dg[r]]
    vals[~r] = var_b_pl[:, dl[~r], :, dg[~r]]
    var_b[:, dl[r], :, dg[r]] = 0.0
    var_b_pl[:, dl[~r], :, dg[~r]] = 0.0
    for i in np.lexsort((dc, sc)):
        if sk[i] == 0:
            var_b[:, sl[i], :, sg[i]] += vals[i]
        else:
            var_b_pl[:, sl[i], :, sg[i]] += vals[i]
    return var_b, var_b_pl
