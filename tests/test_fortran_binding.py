"""The Fortran side of the boundary: imex_sflow2d_v2_amd/fortran/imexsw_c.f90
(the iso_c_binding module a Fortran host uses) compiled with amdflang and
linked against libimexsw.so.  A check program prints the size and field
offsets of its BIND(C) types, which must equal the C layout (ctypes mirrors
of include/imexsw.h, themselves checked against the header), and calls the
host-only entry points: the version string, a create refused by
validation with its message, and a .q_2d write and read back."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

from imex_sflow2d_v2_amd.params import MAX_VARS, ImexswParams
from imex_sflow2d_v2_amd.solver import LIB_PATH, ImexswStats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODULE = os.path.join(ROOT, "imex_sflow2d_v2_amd", "fortran", "imexsw_c.f90")
CHECK = os.path.join(ROOT, "tests", "fortran", "imexsw_c_check.f90")


def _flang():
    for c in ("/opt/rocm/llvm/bin/amdflang", shutil.which("amdflang"), shutil.which("flang"),
              shutil.which("gfortran")):
        if c and os.path.exists(c):
            return c
    return None


@pytest.fixture(scope="module")
def check_output(tmp_path_factory):
    fc = _flang()
    if fc is None:
        pytest.skip("no Fortran compiler in this image")
    assert os.path.exists(LIB_PATH), "libimexsw.so not built"
    d = tmp_path_factory.mktemp("fortran")
    libdir = os.path.dirname(LIB_PATH)
    exe = str(d / "imexsw_c_check")
    cmd = [fc, "-O1", "-J", str(d), "-o", exe, MODULE, CHECK, "-L" + libdir, "-limexsw",
           "-Wl,-rpath," + libdir]
    subprocess.check_call(cmd, cwd=str(d))
    out = subprocess.check_output([exe], cwd=str(d), timeout=120).decode()
    print(out)
    return out


def _parse(out):
    size, offs = {}, {}
    for line in out.splitlines():
        f = line.split()
        if line.startswith("sizeof "):
            size[f[1]] = int(f[2])
        elif line.startswith("offset "):
            offs[(f[1], f[2])] = int(f[3])
    return size, offs


def test_params_type_matches_c_layout(check_output):
    size, offs = _parse(check_output)
    assert size["imexsw_params"] == C.sizeof(ImexswParams)
    names = [k[1] for k in offs if k[0] == "params" and "[" not in k[1]]
    assert len(names) >= 25
    for n in names:
        assert offs[("params", n)] == getattr(ImexswParams, n).offset, n
    # bc_flag(i, side) is C bc_flag[side][i - 1]
    assert offs[("params", "bc_flag[1][2]")] == ImexswParams.bc_flag.offset + 4 * (MAX_VARS + 2)
    assert offs[("params", "bc_value[3][0]")] == ImexswParams.bc_value.offset + 8 * 3 * MAX_VARS


def test_stats_type_matches_c_layout(check_output):
    size, offs = _parse(check_output)
    assert size["imexsw_stats"] == C.sizeof(ImexswStats)
    for (kind, n), v in offs.items():
        if kind == "stats":
            assert v == getattr(ImexswStats, n).offset, n


def test_host_only_entry_points_from_fortran(check_output):
    lines = check_output.splitlines()
    assert any(ln.startswith("version imexsw") for ln in lines)
    create = [ln for ln in lines if ln.startswith("create rc")][0]
    assert create.startswith("create rc -1 message grid must have nx >= 2")
    assert "q2d_write rc 0" in lines and "q2d_read rc 0" in lines
    diff = [ln for ln in lines if ln.startswith("q2d max abs diff")][0]
    assert float(diff.split()[-1]) == 0.0


def test_quadrature_rule_from_fortran_matches_oracle(check_output):
    """imexsw_gauss_legendre called from Fortran returns the oracle's
    Gauss-Legendre rule bit for bit (the rule the profile closures use)."""
    import struct

    import numpy as np
    from oracle.pyoracle import load_oracle

    lines = check_output.splitlines()
    assert "gauss_legendre rc 0" in lines
    got = [ln.split() for ln in lines if ln.startswith("gl ")]
    assert len(got) == 5
    L = load_oracle()
    x, w = np.zeros(5), np.zeros(5)
    D = L.orc_gauss_legendre.argtypes[1]
    L.orc_gauss_legendre(5, x.ctypes.data_as(D), w.ctypes.data_as(D))
    for i, (_, _, hx, hw) in enumerate(got):
        assert struct.unpack(">d", bytes.fromhex(hx))[0] == x[i]
        assert struct.unpack(">d", bytes.fromhex(hw))[0] == w[i]

