"""The C-ABI library loads and exports every entry point include/imexsw.h
declares (no device calls), and the host-side helpers behave."""
import ctypes
import os
import re

import numpy as np

from imex_sflow2d_v2_amd import decks
from imex_sflow2d_v2_amd.params import ImexswParams, MAX_LEVELS, MAX_SOLID, MAX_VARS
from imex_sflow2d_v2_amd.solver import ABI, LIB_PATH, load_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "imexsw.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(imexsw_[a-z0-9_]+)\s*\(", src)))


def test_header_and_wrapper_agree():
    assert sorted(ABI) == declared()


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH), "libimexsw.so not built"
    lib = ctypes.CDLL(LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name
    load_library()


def test_version_string_without_device():
    lib = load_library()
    assert lib.imexsw_version().decode().startswith("imexsw")


def test_params_struct_layout_matches_header():
    # sizes of the fixed arrays mirror the #defines of the header
    src = open(HEADER).read()
    vals = dict(re.findall(r"#define (IMEXSW_MAX_\w+) (\d+)", src))
    assert int(vals["IMEXSW_MAX_SOLID"]) == MAX_SOLID
    assert int(vals["IMEXSW_MAX_LEVELS"]) == MAX_LEVELS
    assert MAX_VARS == 4 + MAX_SOLID
    # 37 scalars before the arrays etc.: compare against a C compiler's view
    import subprocess, tempfile
    prog = r'''
#include <stdio.h>
#include <stddef.h>
#include "imexsw.h"
int main(void){printf("%zu %zu %zu %zu\n", sizeof(imexsw_params),
 offsetof(imexsw_params, bc_value), offsetof(imexsw_params, rho_s),
 offsetof(imexsw_params, dyn_pres_levels)); return 0;}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(prog)
        exe = os.path.join(d, "t")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        out = subprocess.check_output([exe]).decode().split()
    sz, o1, o2, o3 = map(int, out)
    assert ctypes.sizeof(ImexswParams) == sz
    assert ImexswParams.bc_value.offset == o1
    assert ImexswParams.rho_s.offset == o2
    assert ImexswParams.dyn_pres_levels.offset == o3


def test_deck_builders_shapes():
    d = decks.dambreak_1d()
    assert d.q.shape == (5, 400, 1)
    assert d.q[0, 0, 0] > d.q[0, -1, 0] > 0
    d = decks.small_2d(n=20, ny=16, source=True)
    sc, bits, v = decks.init_source_np(d.params)
    assert sc.shape == (20, 16) and v.shape == (8, 20, 16)
    B = decks.volcano_dem(64, 20.0)
    assert B.shape == (64, 64) and B.min() >= 0.0 and B.max() > 500.0
