"""C ABI checks that need no GPU: the library loads, exports every entry
point include/cfd_amd.h declares, and the parameter struct matches."""
import ctypes
import os
import re

import cfd_amd
from cfd_amd import _lib

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "cfd_amd.h")


def declared():
    src = open(HDR).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(cfd_\w+)\s*\(", src, re.M)))


def test_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_params_struct_layout():
    # 16 ints then 13 doubles, as cfd_params declares them
    assert ctypes.sizeof(_lib.CParams) == 16 * 4 + 13 * 8
    assert _lib.CParams.Re.offset == 64


def test_host_setup_errors_are_reported():
    p = cfd_amd.Params(ntilex=3)  # odd tile count: boundary_circ_rough stops on it
    try:
        cfd_amd.build_iblist(p, 0, 0)
    except _lib.CfdError as e:
        assert "ntilex is not even" in str(e)
    else:
        raise AssertionError("odd ntilex accepted")
    p = cfd_amd.Params(r=20.0)
    try:
        cfd_amd.build_iblist(p, 0, 0)
    except _lib.CfdError as e:
        assert "too close to edge" in str(e)
    else:
        raise AssertionError("oversized post accepted")
