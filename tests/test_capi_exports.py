"""The C-ABI library loads and exports every entry point include/wx_capi.h
declares, and the Fortran interfaces bind the same names (no GPU needed:
nothing is called)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "walicxe2d_amd", "libwxhip.so")


def declared():
    txt = open(os.path.join(ROOT, "include", "wx_capi.h")).read()
    return sorted(set(re.findall(r"^(?:int|void \*)\s*(wx_\w+)\(", txt, re.M)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="libwxhip.so not built")
def test_library_exports_every_declared_symbol():
    names = declared()
    assert len(names) >= 20
    L = ctypes.CDLL(LIB)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from walicxe2d_amd import capi
    assert set(capi.EXPORTS) == set(declared())


def test_fortran_interfaces_bind_declared_symbols():
    src = open(os.path.join(ROOT, "walicxe2d_amd", "fortran", "wx_gpu.f90")).read()
    bound = set(re.findall(r'bind\(C, name="(wx_\w+)"\)', src))
    assert bound and bound <= set(declared())


def test_params_struct_layout_matches_header():
    """wx_params field offsets: the C struct has 16 int32 then doubles."""
    from walicxe2d_amd.params import WxParams
    assert WxParams.cv.offset == 64
    assert ctypes.sizeof(WxParams) == 64 + 8 * (11 + 32 + 82 + 12)
