"""CPU checks of the C-ABI boundary: the HIP library loads and exports every
function include/*.h declares; the struct layouts the headers restate are
the ones the Python binding uses.  No compute calls (no GPU here)."""
import ctypes as C
import os
import subprocess

import pytest

from hiplib import PKG_DIR, ROOT, package


def test_library_exports_every_declared_function():
    pkg = package()
    L = pkg.lib()
    names = pkg.header_functions()
    assert len(names) > 100
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_reference_entry_point_names_are_exported():
    """Names and callback slots of diffusion_2D.hpp:246-271 and main.cpp:536."""
    L = package().lib()
    for name in ["diffusion", "laplacian", "PSetup", "PSolve", "Initial", "dom_eig",
                 "LSRKStepCreateSTS", "ARKStepCreate", "ARKodeEvolve", "SUNLinSol_PCG",
                 "SUNDomEigEstimator_Power", "N_VNew_HipNative"]:
        assert hasattr(L, name)


def test_nvector_only_library_has_no_integrator():
    """libceda_nvhip.so is linkable next to a real SUNDIALS (no clashing
    integrator symbols)."""
    so = os.path.join(PKG_DIR, "libceda_nvhip.so")
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True).stdout
    assert " N_VNew_HipNative" in out and " diffusion" in out
    assert "ARKodeEvolve" not in out and "LSRKStepCreateSTS" not in out


def test_params_struct_matches_header():
    pkg = package()
    # int kind; int64 nx, ny; 6 doubles; int inhom; double tf -> 8+16+48+8+8
    assert C.sizeof(pkg.Params) == 88
    assert pkg.Params.tf.offset == 80


def test_ops_table_layout_is_sundials7():
    """The restated _generic_N_Vector_Ops has the SUNDIALS 7.x field count
    (11 utility + 19 standard + 3 fused + 7 array + 9 local + 2 single-buffer
    + 3 XBraid + 2 debug = 56 function pointers)."""
    src = open(os.path.join(ROOT, "include", "ceda_sundials.h")).read()
    body = src[src.index("struct _generic_N_Vector_Ops {"):src.index("};", src.index("struct _generic_N_Vector_Ops {"))]
    assert body.count("(*nv") == 56
