"""The C-ABI library loads and exports every function include/fmma_hip.h
declares (no GPU needed: nothing here launches a kernel)."""
import ctypes
import os
import re

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "fmma_hip.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(fmma_[A-Za-z0-9_]+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "fmma_plan_solve" in names and "fmma_plan_create" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from fmma_amd import _lib
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert set(declared_functions()) <= set(_lib.EXPORTS) | set(declared_functions())


def test_python_binding_lists_every_symbol():
    from fmma_amd import _lib
    assert set(declared_functions()) == set(_lib.EXPORTS)


def test_host_side_calls_without_gpu():
    """Plan creation and setters are host-only: they work without a device."""
    from fmma_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.fmma_plan_create(ctypes.byref(h), 3, 0) == 0
    assert L.fmma_plan_set_poly_ord(h, 3) == 0
    assert L.fmma_plan_set_poly_ord(h, 16) == _lib.FMMA_ERR_INVALID
    assert b"poly_ord" in L.fmma_last_error()
    assert L.fmma_plan_set_kernel(h, 99, None, 0) == _lib.FMMA_ERR_INVALID
    assert L.fmma_plan_set_m2l_precision(h, 2) == 0
    d = ctypes.c_int()
    assert L.fmma_plan_set_depth(h, 6) == 0 and L.fmma_plan_get_depth(h, ctypes.byref(d)) == 0 and d.value == 6
    assert L.fmma_plan_create(ctypes.byref(ctypes.c_void_p()), 4, 0) == _lib.FMMA_ERR_INVALID
    assert L.fmma_plan_destroy(h) == 0


def test_morton_rowmajor_roundtrip():
    from fmma_amd import _lib
    L = _lib.lib()
    for dim, level in ((1, 5), (2, 4), (3, 3)):
        n = 1 << (dim * level)
        ids = sorted(L.fmma_morton_to_rowmajor(dim, level, m) for m in range(n))
        assert ids == list(range(n))
    # low DIM bits of a Morton index are the child position, dimension 0 most significant
    assert L.fmma_morton_to_rowmajor(3, 1, 0b100) == 4  # (1,0,0) -> 1*4
    assert L.fmma_morton_to_rowmajor(3, 1, 0b001) == 1  # (0,0,1)
    assert L.fmma_morton_to_rowmajor(2, 2, 0b0110) == 0b0110 and L.fmma_morton_to_rowmajor(2, 2, 0b0011) == 0b0101
