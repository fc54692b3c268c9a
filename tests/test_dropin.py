"""The reference's own kernel drivers on our snapshot files, and the same
drivers relinked against the Fortran shim over libnicam_tt.

Each driver (kernels/NICAM/<kernel>/src/main.f90, unchanged) reads
snapshot.<kernel>.pe000000 -- inputs plus the outputs it checks its result
against -- runs the kernel once and prints the reference's DEBUG_valuecheck
lines (max, min, sum of every array), then the "point-by-point diff" between
its result and the snapshot's outputs.  oracle/Makefile (target drivers)
builds two executables per kernel into oracle/_ref: drv_<k>_ref.exe from
the reference sources alone, and drv_<k>_gpu.exe with the kernel module
replaced by icoatmosbenchmark_v1_amd/fortran/shim_<k>.f90.

We write the snapshot (icoatmosbenchmark_v1_amd.snapshot) from seeded inputs
at the drivers' compiled-in size (gall_1d 130, kall 42, one region with a
singular point, poles on) with the oracle's outputs as the expected values:
* CPU: the reference driver reads our file as we wrote it (its input checks
  match the arrays) and reproduces the oracle's outputs to the bit (diff 0);
* GPU: the shim-linked driver prints exactly the reference driver's lines.
"""
import re
import resource
import shutil
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from icoatmosbenchmark_v1_amd import runtime, snapshot, synth

REPO = Path(__file__).resolve().parent.parent
REF = REPO / "oracle" / "_ref"
VGRID = REPO / "tests" / "golden" / "vgrid40_600m_24km.dat"
KERNELS = {"flux": "dyn_horiz_adv_flux", "hlim": "dyn_horiz_adv_limiter", "vlim": "dyn_vert_adv_limiter",
           "vi": "dyn_vi_rhow_solver", "diff": "dyn_diffusion", "ddmp": "dyn_divdamp"}
N1, KALL, LALL = 130, 42, 1  # problem_size.inc:28-45 of every driver
CHECK = re.compile(r"\+check\[(.{16})\] max=(.*),min=(.*),sum=(.*)$")


def _exe(short, kind):
    p = REF / f"drv_{short}_{kind}.exe"
    if not p.exists():
        pytest.skip(f"{p.name} not built (oracle/Makefile target drivers needs the reference tree)")
    return p


def _unlimited_stack():
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def run_driver(exe: Path, workdir: Path, timeout=300):
    out = subprocess.run([str(exe)], cwd=workdir, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=timeout, preexec_fn=_unlimited_stack)
    assert out.returncode == 0, f"{exe.name} failed ({out.returncode}):\n{out.stdout[-3000:]}"
    sections, cur = {}, None
    for ln in out.stdout.splitlines():
        if "### Input ###" in ln:
            cur = "input"
        elif "### Output ###" in ln:
            cur = "output"
        elif "point-by-point diff" in ln:
            cur = "diff"
        m = CHECK.search(ln)
        if m and cur:
            sections.setdefault(cur, []).append((m.group(1).strip(), m.group(2).strip(), m.group(3).strip(),
                                                 m.group(4).strip()))
    assert set(sections) == {"input", "output", "diff"}, out.stdout[-3000:]
    return sections


def make_snapshot(short, oracle, workdir: Path):
    """Writes the snapshot (and the vertical grid file) into workdir; returns
    the fields as written."""
    spec = synth.GridSpec(N1, KALL, LALL, have_pl=True, seed=91)
    gz, gzh = runtime.grd_input_vgrid(str(VGRID), KALL)
    case = synth.make_case(spec, have_sgp=[1], vgrid=(gz, gzh))
    geo = case["geo"]
    if short == "flux":
        f = case["flux"]
        out = oracle.horizontal_flux(spec, geo, f, 360.0)  # SET_dt_large
        fields = {**out, **{k: f[k] for k in ("rho", "rho_pl", "rhovx", "rhovx_pl", "rhovy", "rhovy_pl",
                                              "rhovz", "rhovz_pl")},
                  **{k: geo[k] for k in ("GRD_xr", "GRD_xr_pl", "GMTR_p", "GMTR_p_pl", "GMTR_t", "GMTR_t_pl",
                                          "GMTR_a", "GMTR_a_pl")}}
    elif short == "hlim":
        f = case["hlim"]
        out = oracle.horizontal_limiter_thuburn(spec, [1], f)
        fields = {**f, **out, "q_a_prev": f["q_a"], "q_a_prev_pl": f["q_a_pl"]}
    elif short == "vlim":
        f = case["vlim"]
        out = oracle.vertical_limiter_thuburn(spec, f)
        fields = {**f, **out, "q_h_prev": f["q_h"], "q_h_prev_pl": f["q_h_pl"]}
    elif short == "diff":
        f = synth.diffusion_inputs(spec, geo)
        fields = {**f, **oracle.oprt_diffusion(spec, [1], f)}
    elif short == "ddmp":  # the driver's GRD_setup reads GRD_rdgz from the vertical grid file
        f = dict(synth.divdamp_inputs(spec, geo), rdgz=oracle.grd_setup(gz, gzh)["rdgz"])
        fields = {**f, **oracle.oprt3d_divdamp(spec, [1], f)}
        shutil.copy(VGRID, workdir / VGRID.name)
    else:
        f = case["vi"]
        g = oracle.grd_setup(gz, gzh)
        out = oracle.vi_rhow_solver(spec, f, 60.0, g["rdgzh"], g["afact"], g["bfact"])  # SET_dt_small
        fields = {**f, **out, "rhogw_prev": f["rhogw"], "rhogw_prev_pl": f["rhogw_pl"]}
        shutil.copy(VGRID, workdir / VGRID.name)
    snapshot.write_snapshot(workdir / snapshot.snapshot_name(KERNELS[short]), KERNELS[short], fields, N1, KALL, LALL)
    return fields


def test_snapshot_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    sizes = snapshot.field_sizes("dyn_vert_adv_limiter", 6, 5, 2)
    fields = {k: rng.standard_normal(n) for k, n in sizes}
    p = snapshot.write_snapshot(tmp_path / "s", "dyn_vert_adv_limiter", fields, 6, 5, 2)
    assert p.stat().st_size == 8 * sum(n for _, n in sizes)
    raw = p.read_bytes()
    assert raw[:8] == struct.pack(">d", fields["q_h_prev"][0])  # big-endian, no header
    back = snapshot.read_snapshot(p, "dyn_vert_adv_limiter", 6, 5, 2)
    for k in fields:
        assert np.array_equal(back[k], fields[k])
    with pytest.raises(ValueError):
        snapshot.write_snapshot(tmp_path / "t", "dyn_vert_adv_limiter", {**fields, "ck": fields["q"]}, 6, 5, 2)


def test_vgrid_writer_reproduces_the_shipped_file(tmp_path):
    gz, gzh = runtime.grd_input_vgrid(str(VGRID), KALL)
    p = snapshot.write_vgrid(tmp_path / "v.dat", gz, gzh)
    assert p.read_bytes() == VGRID.read_bytes()


def _close_sum(printed: str, a: np.ndarray) -> bool:
    """A printed sum against ours: the driver sums in storage order, numpy
    pairwise, so allow the recursive-summation error bound n*eps*sum|a|."""
    return abs(float(printed) - float(np.sum(a))) <= a.size * 2.3e-16 * float(np.sum(np.abs(a)))


@pytest.mark.ref
@pytest.mark.parametrize("short", list(KERNELS))
def test_reference_driver_on_our_snapshot(short, oracle, tmp_path):
    fields = make_snapshot(short, oracle, tmp_path)
    sec = run_driver(_exe(short, "ref"), tmp_path)
    # the driver's view of what it read: max and min exact, sum to its own order
    for name, mx, mn, sm in sec["input"]:
        key = name.removeprefix("check_")
        if key.startswith("VMTR_"):
            key = key[5:]
        if key not in fields:
            continue
        a = np.asarray(fields[key], dtype=np.float64)
        assert float(mx) == a.max() and float(mn) == a.min(), name
        assert _close_sum(sm, a), name
    # the reference reproduces the oracle's outputs exactly
    for name, mx, mn, sm in sec["diff"]:
        assert float(mx) == 0.0 and float(mn) == 0.0 and float(sm) == 0.0, f"{short}: {name} {mx} {mn} {sm}"


@pytest.mark.gpu
@pytest.mark.parametrize("short", list(KERNELS))
def test_shim_driver_matches_reference_driver(short, gpu, oracle, tmp_path):
    make_snapshot(short, oracle, tmp_path)
    ref = run_driver(_exe(short, "ref"), tmp_path)
    got = run_driver(_exe(short, "gpu"), tmp_path)
    assert got == ref
    for name, mx, mn, sm in got["diff"]:
        assert float(mx) == 0.0 and float(mn) == 0.0 and float(sm) == 0.0, f"{short}: {name}"
