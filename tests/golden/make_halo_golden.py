"""Writes halo_gl04rl00.npz from the halo-exchange self-test logs that ship
with the reference's communication kernel:

    python tests/golden/make_halo_golden.py /root/reference

kernels/NICAM/communication/reference/msg.pe00000{0,1}.Ubuntu-gnu-ompi hold
the output of COMM_debugtest (communication/src/mod_comm.f90:2120-2481) for
glevel 4, rlevel 0 on two processes (regions 1-5 and 6-10; both poles on the
first).  The test stamps every inner point with (prc, rgnid, i, j), every
pole point with (-prc, -pole, -v, -v) and every halo point with -999 (-1 at
the first point of a singular-point region), prints the tables, runs
COMM_data_transfer, prints them again, then stamps the pole-adjacent halo
points, runs COMM_var and prints them a third time.

The fixture keeps those printed tables as data (1-based, as printed; -999
where the log shows "***"):

    reg[stage, rgn-1, j-1, i-1, c]   c = prc, rgnid, i, j  (level kmin)
    pole[stage, pole, v-1, c]        from the process that owns the poles
    stage 0 = before, 1 = after COMM_data_transfer, 2 = after COMM_var

and the halo lists each process printed (COMM_list_generate,
COMM_sortdest_pl, COMM_sortdest_singular), one integer row per entry with the
printed columns (1-based):

    pe{0,1}_{Copy,Recv,Send}_list_{r2r,p2r,r2p}, pe{0,1}_Singular_list
    cols_<list name>  the column names

The gnu, intel and pgi logs are checked to agree before anything is written.
"""
from __future__ import annotations

import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
OUT = HERE / "halo_gl04rl00.npz"
GALL_1D, GALL_PL, RALL = 18, 6, 10
STAGES = ["+++ TEST start", "+++ Communication start", "+++ pole fill start"]
CELL = re.compile(r"\(\s*([-\d*]+),\s*([-\d*]+)\)")
ROW = re.compile(r"^\s*(-?\d+)\|(.*)$")
L2R = re.compile(r"---\s*\(\s*(\d+),\s*(\d+)\)\s*=>\s*\(\s*(\d+)\)")


def _num(s: str) -> int:
    return -999 if "*" in s else int(s)


def parse(path: Path):
    """Returns (regions of this process, reg[3,5,18,18,4], pole[3,2,6,4])."""
    text = path.read_text().splitlines()
    rgns = [int(m.group(3)) for m in map(L2R.search, text) if m]
    starts = [next(n for n, ln in enumerate(text) if ln.strip() == s) for s in STAGES]
    ends = starts[1:] + [len(text)]
    nl = len(rgns)
    reg = np.full((3, nl, GALL_1D, GALL_1D, 4), -999, dtype=np.int16)
    pole = np.full((3, 2, GALL_PL, 4), -999, dtype=np.int16)
    for s, (a, b) in enumerate(zip(starts, ends)):
        half = 0  # 0: (prc,rgnid) tables, 1: (i,j) tables
        rtab = ptab = -1
        last_j = None
        for ln in text[a:b]:
            if "##### (i,j) #####" in ln:
                half, rtab, ptab = 1, -1, -1
                continue
            m = ROW.match(ln)
            if not m:
                continue
            cells = [(_num(x), _num(y)) for x, y in CELL.findall(m.group(2))]
            if len(cells) == GALL_1D:
                j = int(m.group(1))
                if j == GALL_1D and last_j != GALL_1D:
                    rtab += 1
                last_j = j
                for i, (x, y) in enumerate(cells):
                    reg[s, rtab, j - 1, i, 2 * half:2 * half + 2] = (x, y)
            elif len(cells) == GALL_PL:
                ptab += 1
                for v, (x, y) in enumerate(cells):
                    pole[s, ptab, v, 2 * half:2 * half + 2] = (x, y)
        if rtab != nl - 1 or ptab != 1:
            raise ValueError(f"{path.name}: stage {s} has {rtab + 1} region and {ptab + 1} pole tables")
    return rgns, reg, pole


LISTS = [f"{k}_list_{d}" for d in ("r2r", "p2r", "r2p") for k in ("Copy", "Recv", "Send")] + ["Singular_list"]


def parse_lists(path: Path):
    """Returns {list name: (column names, int32 rows)} for every list in LISTS."""
    text = path.read_text().splitlines()
    out = {}
    for n, ln in enumerate(text):
        name = ln.strip().removeprefix("--- ")
        if not ln.startswith(" --- ") or name not in LISTS:
            continue
        cols, rows = None, []
        for row in text[n + 1:]:
            if row.startswith("number|"):
                cols = [c.strip() for c in row.split("|")]
                continue
            fields = row.split()
            if cols is None and not fields:
                continue
            if not fields or not all(f.lstrip("-").isdigit() for f in fields):
                break
            rows.append([int(f) for f in fields])
        cols = cols or []
        arr = np.array(rows, dtype=np.int32).reshape(len(rows), len(cols))
        out[name] = (cols, arr)
    missing = set(LISTS) - set(out)
    if missing:
        raise ValueError(f"{path.name}: lists not found: {sorted(missing)}")
    return out


def main(ref_root: Path) -> None:
    logdir = ref_root / "kernels" / "NICAM" / "communication" / "reference"
    result = {}
    for comp in ("gnu", "intel", "pgi"):
        reg = np.full((3, RALL, GALL_1D, GALL_1D, 4), -999, dtype=np.int16)
        pole = None
        for pe in (0, 1):
            rgns, r, p = parse(logdir / f"msg.pe00000{pe}.Ubuntu-{comp}-ompi")
            for n, rg in enumerate(rgns):
                reg[:, rg - 1] = r[:, n]
            if pe == 0:  # the first process manages both poles
                pole = p
        result[comp] = (reg, pole)
    reg, pole = result["gnu"]
    for comp in ("intel", "pgi"):
        if not (np.array_equal(result[comp][0], reg) and np.array_equal(result[comp][1], pole)):
            raise SystemExit(f"the {comp} log disagrees with the gnu log")
    lists = {}
    for pe in (0, 1):
        per = {comp: parse_lists(logdir / f"msg.pe00000{pe}.Ubuntu-{comp}-ompi") for comp in ("gnu", "intel", "pgi")}
        for name in LISTS:
            cols, arr = per["gnu"][name]
            for comp in ("intel", "pgi"):
                if not np.array_equal(per[comp][name][1], arr):
                    raise SystemExit(f"{name} of the {comp} log disagrees with the gnu log")
            if arr.size:
                if cols and f"cols_{name}" not in lists:
                    lists[f"cols_{name}"] = np.array(cols)
            lists[f"pe{pe}_{name}"] = arr
    np.savez_compressed(OUT, reg=reg, pole=pole, gall_1d=GALL_1D, rlevel=0, glevel=4, **lists)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
