"""The C-ABI library loads and exports every entry point include/lsmp.h
declares; the Python binding declares the same set.  No compute calls: these
run without a GPU."""
import ctypes
import os
import re

import pytest

from conftest import ROOT, load_lsmp

HEADER = os.path.join(ROOT, "include", "lsmp.h")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lsmp_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    syms = header_symbols()
    for must in ("lsmp_init", "lsmp_rk3_advect", "lsmp_reinit", "lsmp_mass_fix", "lsmp_mass_vol",
                 "lsmp_interface_l1", "lsmp_upload_phi", "lsmp_download_phi", "lsmp_finalize"):
        assert must in syms


def test_library_exports_every_header_symbol():
    lsmp = load_lsmp()
    if not os.path.exists(lsmp.LIB_PATH):
        pytest.fail(f"{lsmp.LIB_PATH} missing: run __graft_entry__.build()")
    lib = ctypes.CDLL(lsmp.LIB_PATH)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    lsmp = load_lsmp()
    assert sorted(lsmp.EXPORTS) == header_symbols()


def test_fortran_binding_covers_header():
    """The ISO_C_BINDING module binds every entry point the header declares,
    under the header's names, and nothing else."""
    text = open(os.path.join(ROOT, "mass-preserving-level-set-method_amd", "fortran", "lsmp_mod.f90")).read()
    bound = sorted(set(re.findall(r'bind\(C,\s*name="(lsmp_[a-z0-9_]+)"\)', text)))
    assert bound == header_symbols()


def test_integration_guide_names_exist():
    """Every library call the integration guide shows exists: the Fortran
    bindings (c_lsmp_*) in lsmp_mod.f90, the C entry points in the header."""
    guide = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    mod = open(os.path.join(ROOT, "mass-preserving-level-set-method_amd", "fortran", "lsmp_mod.f90")).read()
    fortran = set(re.findall(r"function (c_lsmp_[a-z0-9_]+)", mod))
    shown = set(re.findall(r"\b(c_lsmp_[a-z0-9_]+)", guide))
    assert shown and shown <= fortran, sorted(shown - fortran)
    # C calls: lsmp_name( not preceded by c_; the Fortran helpers and types aside
    helpers = {"lsmp_check", "lsmp_params", "lsmp_setup"}
    called = set(re.findall(r"(?<![a-z_])(lsmp_[a-z0-9_]+)\(", guide)) - helpers
    assert called and called <= set(header_symbols()), sorted(called - set(header_symbols()))


def test_params_struct_layout_matches_header():
    """Field order and types of lsmp_params / lsmp_stats as the header states them."""
    lsmp = load_lsmp()
    text = open(HEADER).read()
    body = re.search(r"typedef struct lsmp_params \{(.*?)\} lsmp_params;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"(\w+)\s*[,;]", body)
    assert names == [f[0] for f in lsmp.Params._fields_]
    body = re.search(r"typedef struct lsmp_stats \{(.*?)\} lsmp_stats;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"(\w+)\s*[,;]", body)
    assert names == [f[0] for f in lsmp.Stats._fields_]


def test_error_strings_without_gpu():
    lsmp = load_lsmp()
    lib = lsmp.load_library()
    for code in (0, -1, -3, -4, -5):
        assert lib.lsmp_error_string(code).decode()


def test_invalid_params_rejected_before_any_device_call():
    """Parameter validation runs on the host, so it answers without a GPU."""
    lsmp = load_lsmp()
    lib = lsmp.load_library()
    h = ctypes.c_void_p()
    bad = lsmp.make_params(16, dim=3)
    bad.ghc = 2
    assert lib.lsmp_init(ctypes.byref(bad), ctypes.byref(h)) == -1
    bad = lsmp.make_params(3, dim=3)        # lines shorter than the CRWENO stencil
    assert lib.lsmp_init(ctypes.byref(bad), ctypes.byref(h)) == -1
    bad = lsmp.make_params(16, dim=2)
    bad.nz = 4
    assert lib.lsmp_init(ctypes.byref(bad), ctypes.byref(h)) == -1
