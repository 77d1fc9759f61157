"""CPU checks of the C ABI library: it loads, exports every entry point that
include/quadesth.h declares, and refuses to run without a GPU (no fallback)."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "quadesth.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(qd_\w+|dgesv_)\s*\(", text)
    return sorted(set(names))


def test_header_and_binding_agree():
    from bem3d_elastostatic_amd import _lib
    assert header_functions() == sorted(_lib.exported_symbols())


def test_library_exports_every_declared_symbol():
    from bem3d_elastostatic_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libquadesth.so not built")
    lib = C.CDLL(_lib.LIB_PATH)
    for name in header_functions():
        assert hasattr(lib, name), name
    _lib.load()


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU behaviour")
def test_no_gpu_raises():
    from bem3d_elastostatic_amd import _lib, mesh
    from bem3d_elastostatic_amd.quadesth import Quadesth
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libquadesth.so not built")
    with pytest.raises(_lib.QuadesthError):
        Quadesth(mesh.config0())
