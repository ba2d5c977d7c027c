"""The C-ABI library loads and exports every entry point declared in include/*.h;
host-side argument checks fail cleanly without touching a GPU."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

from ramses2_amd import _lib
from ramses2_amd import build as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(rh_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_reference_counterparts():
    names = declared_functions()
    for n in ("rh_init", "rh_alloc_amr", "rh_upload_level_params", "rh_set_grid_device",
              "rh_transfer_grid_host", "rh_insert_hash_all", "rh_build_nbor", "rh_set_unew",
              "rh_set_uold", "rh_cmpdt", "rh_godunov", "rh_device_sync", "rh_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    path = B.lib_path()
    assert os.path.exists(path), "build the library first (__graft_entry__.build())"
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = declared_functions() - exported
    assert not missing, missing
    L = _lib.load()
    for n in declared_functions():
        assert hasattr(L, n)


def test_host_side_errors_without_gpu():
    L = _lib.load()
    # bad ids are rejected before any device work
    assert L.rh_set_brick_shape(5, 5, 5) == 2
    assert b"compiled kernel" in L.rh_last_error()
    assert L.rh_set_precision(7) == 2
    bt = (ctypes.c_int32 * 6)(2, 2, 9, 2, 2, 2)
    assert L.rh_set_boundaries(bt) == 2
    assert L.rh_set_brick_shape(7, 7, 8) == 0
    # a fence never recorded, and one of negative length
    assert L.rh_fence_wait(3, None) == 2
    assert b"no outstanding fence" in L.rh_last_error()
    f = ctypes.c_int32()
    assert L.rh_fence_record(None, -1, ctypes.byref(f)) == 2
    with pytest.raises(_lib.HipError):
        _lib.call("rh_set_precision", 3)


def test_fortran_binding_builds():
    """hip_interface.f90 + the Sod driver compile and link against the library."""
    if not os.path.exists(B.FLANG):
        pytest.skip("flang not installed")
    assert os.path.exists(B.fortran_driver_path()), "build() compiles the Fortran driver"
    out = subprocess.run(["nm", "-u", B.fortran_driver_path()], capture_output=True, text=True).stdout
    for n in ("rh_init", "rh_cmpdt", "rh_godunov", "rh_build_nbor"):
        assert n in out
