"""The Fortran snippets of INTEGRATION.md compile against the library's
ISO_C_BINDING module (include/rocketcombustor_iface.f90).

Each ```fortran block is wrapped in a subroutine: its own `use` and
declaration lines first, then declarations of the reference host's names it
uses (module mainvar arrays, loop counters, the reference routines it calls
as internal stubs), then its statements.  A call with the wrong number or
type of arguments fails to compile.  Needs amdflang (ROCm); skipped without
it.  Nothing is linked or run.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or (os.path.exists("/opt/rocm/bin/amdflang") and "/opt/rocm/bin/amdflang")

# names of the reference's host program that the snippets use
HOST_DECLS = {
    "h": "type(c_ptr) :: h",
    "NNT": "integer(c_int) :: NNT", "NE": "integer(c_int) :: NE", "nb": "integer(c_int) :: nb",
    "Coor": "real(c_double) :: Coor(2, 4)", "cells": "integer(c_int) :: cells(3)",
    "bnodes": "integer(c_int) :: bnodes(2)", "btype": "integer(c_int) :: btype(1)",
    "PA": "real(c_double) :: PA(7, 1)", "WA": "real(c_double) :: WA(7, 1)", "hWA": "real(c_double) :: hWA(7, 1)",
    "istep": "integer :: istep", "nstep_max": "integer :: nstep_max", "NScreenOutput": "integer :: NScreenOutput",
    "NFileBackup": "integer :: NFileBackup", "time_all": "double precision :: time_all",
    "RoRes": "double precision :: RoRes", "ttime": "double precision :: ttime",
    "ttime_n": "double precision :: ttime_n", "cs": "double precision :: cs(3)", "dt": "real(c_double) :: dt",
    "InnerLoop": "integer(c_int) :: InnerLoop", "relTol": "real(c_double) :: relTol",
    "p_out": "real(c_double) :: p_out", "Nvar": "integer :: Nvar",
}
STUBS = """contains
    subroutine step_residual(istep, time_all, RoRes)
        integer :: istep
        double precision :: time_all, RoRes
    end subroutine
    subroutine BackupData()
    end subroutine
    subroutine compute_p_out()
    end subroutine
"""
DECL = re.compile(r"^\s*(use\s|type\s*\(|integer|double precision|real|character|logical)", re.I)


def fortran_blocks():
    with open(os.path.join(ROOT, "INTEGRATION.md")) as f:
        text = f.read()
    return re.findall(r"```fortran\n(.*?)```", text, re.S)


def wrap(k: int, block: str) -> str:
    lines = [ln.split("!")[0].rstrip() for ln in block.splitlines()]
    lines = [ln for ln in lines if ln.strip() and ln.strip() != "..."]
    head, body = [], []
    for ln in lines:
        (head if not body and DECL.match(ln) and ("::" in ln or ln.strip().lower().startswith("use")) else body).append(ln)
    declared = set()
    for ln in head:
        if "::" in ln:
            for name in ln.split("::", 1)[1].split(","):
                declared.add(name.split("(")[0].strip().lower())
    code = "\n".join(body)
    extra = [d for n, d in HOST_DECLS.items()
             if n.lower() not in declared and re.search(r"\b%s\b" % re.escape(n), code, re.I)]
    uses = [ln for ln in head if ln.strip().lower().startswith("use")]
    decls = [ln for ln in head if not ln.strip().lower().startswith("use")]
    if not any("rocketcombustor_iface" in u for u in uses):
        uses.insert(0, "use rocketcombustor_iface")
    return "\n".join(["subroutine doc_block_%d()" % k, "use iso_c_binding"] + uses + ["implicit none"] + decls
                     + extra + body + [STUBS, "end subroutine", ""])


@pytest.mark.skipif(not FC, reason="amdflang not available")
def test_integration_snippets_compile(tmp_path):
    blocks = fortran_blocks()
    assert len(blocks) >= 3
    r = subprocess.run([FC, "-c", "-fdefault-real-8", "-fdefault-double-8",
                        os.path.join(ROOT, "include", "rocketcombustor_iface.f90"), "-o", str(tmp_path / "iface.o")],
                       cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    for k, b in enumerate(blocks):
        src = tmp_path / ("block%d.f90" % k)
        src.write_text(wrap(k, b))
        r = subprocess.run([FC, "-c", "-fdefault-real-8", "-fdefault-double-8", str(src), "-o",
                            str(tmp_path / ("block%d.o" % k))], cwd=tmp_path, capture_output=True, text=True)
        assert r.returncode == 0, "INTEGRATION.md fortran block %d:\n%s\n%s" % (k, src.read_text(), r.stderr)
