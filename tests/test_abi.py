"""The C-ABI libraries load and export every function include/fvgpu.h
declares (no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "fvgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s+\*?([a-z][a-z0-9_]+)\s*\(",
                       src, flags=re.M)
    return sorted(set(n for n in names if n.startswith(("fv_", "fvgpu_", "fvmlp_"))))


def test_header_lists_entry_points():
    names = declared_functions()
    for must in ("fvgpu_create", "fvgpu_internal_step", "fvgpu_external_step", "fvgpu_get_state",
                 "fvgpu_put_state", "fvgpu_diagnostics", "fvgpu_destroy", "fv_model_build"):
        assert must in names


def test_libraries_export_every_declared_symbol(built):
    from myfvcomcode_amd import build
    grid = os.path.join(ROOT, "myfvcomcode_amd", "lib", "libfvgrid.so")
    gpu = os.path.join(ROOT, "myfvcomcode_amd", "lib", "libfvgpu.so")
    if not os.path.exists(gpu):
        build.build_gpu()
    lg, lgpu = C.CDLL(grid), C.CDLL(gpu)
    missing = []
    for n in declared_functions():
        lib = lgpu if n.startswith(("fvgpu_", "fvmlp_")) else lg
        if not hasattr(lib, n):
            missing.append(n)
    assert not missing, missing
    assert hasattr(lgpu, "fvgpu_last_error") and hasattr(lg, "fv_last_error")


def test_gpu_library_fails_loudly_without_device(built):
    """Without a HIP device fvgpu_create must return an error, never fall
    back to a CPU path."""
    import myfvcomcode_amd as F
    L = F.gpu_lib()
    n = C.c_int32(0)
    rc = L.fvgpu_device_count(C.byref(n))
    if rc == 0 and n.value > 0:
        pytest.skip("a GPU is present")
    m = F.Model.build(0, nx=9, ny=5, kb=4)
    with pytest.raises(RuntimeError):
        F.DeviceModel(m)


def fortran_bindings():
    src = open(os.path.join(ROOT, "myfvcomcode_amd", "fortran", "fvgpu_mod.f90")).read()
    return sorted(set(re.findall(r'bind\(c,\s*name="([a-z0-9_]+)"\)', src)))


def test_fortran_bindings_match_header():
    """Every BIND(C) interface of the Fortran module names a declared entry
    point (a misspelt name would only fail at link time of a host model)."""
    decl = set(declared_functions())
    names = fortran_bindings()
    assert len(names) >= 20
    assert not [n for n in names if n not in decl]


def test_fortran_driver_usage(built):
    """The Fortran driver links against both libraries and reports its usage
    (exit status 2) before touching the device."""
    import subprocess
    exe = os.path.join(ROOT, "myfvcomcode_amd", "lib", "fvcom_gpu")
    if not os.path.exists(exe):
        pytest.skip("the Fortran driver is not built (no amdflang)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "usage: fvcom_gpu" in r.stdout
