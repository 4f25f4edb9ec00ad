"""The C-ABI library loads and exports every entry point declared in
include/epic_gpu.h (no compute calls: runs without a GPU)."""
import ctypes as C
import os

import pytest

from epic_amd import core


def test_header_declares_reference_surface():
    names = core.header_symbols()
    for required in ("epic_par2grid", "epic_grid2par", "epic_vor2vel", "epic_vorticity_tendency",
                     "epic_ls_rk_step", "epic_create", "epic_destroy", "epic_cell_sort"):
        assert required in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(core.LIB_PATH):
        pytest.skip("libepic_gpu.so not built")
    lib = C.CDLL(core.LIB_PATH)
    missing = [n for n in core.header_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_create_rejects_bad_grid_without_touching_the_gpu():
    """Argument validation happens before any HIP call."""
    if not os.path.exists(core.LIB_PATH):
        pytest.skip("libepic_gpu.so not built")
    lib = core.load_library()
    p = core.Phys(0.0, 0.0, 0.0, (C.c_double * 3)(0, 0, 0))
    f = core.Flags((C.c_int * 2)(0, 0), 4, 0.0)
    h = C.c_void_p()
    # lengths with a factor 7 (stafft.f90 factorises 2, 3, 4, 5, 6), an odd
    # nx (the x transform packs nx/2 complex points), and fewer than 8 cells
    # and transform lengths beyond 2048 points (ny = 2560, 2 nz = 2560: the
    # stages' index arithmetic stays below 2^16)
    for nx, ny, nz in ((14, 16, 16), (16, 28, 16), (16, 16, 14), (15, 16, 16), (6, 16, 16), (16, 16, 3),
                       (16, 2560, 16), (16, 16, 1280)):
        g = core.Grid(nx, ny, nz, (C.c_double * 3)(0, 0, 0), (C.c_double * 3)(1, 1, 1))
        assert lib.epic_create(C.byref(g), C.byref(p), C.byref(f), 10, C.byref(h)) == -1, (nx, ny, nz)
    g = core.Grid(16, 16, 16, (C.c_double * 3)(0, 0, 0), (C.c_double * 3)(1, 1, 1))
    f = core.Flags((C.c_int * 2)(0, 0), 2, 0.0)   # unsupported RK order
    assert lib.epic_create(C.byref(g), C.byref(p), C.byref(f), 10, C.byref(h)) == -6
    f = core.Flags((C.c_int * 2)(0, 0), 4, 0.0)   # beyond EPIC_MAX_PARCELS = 2^30 - 1
    assert lib.epic_create(C.byref(g), C.byref(p), C.byref(f), 1 << 30, C.byref(h)) == -1


def test_fortran_binding_names_declared():
    """Every C function the Fortran module binds is declared in the header."""
    import re

    src = open(os.path.join(os.path.dirname(core.LIB_PATH), "fortran", "epic_gpu_mod.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="([a-z0-9_]+)"\)', src)) - {"strlen"}
    assert bound and bound <= set(core.header_symbols())
