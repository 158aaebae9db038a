"""The C-ABI library builds for gfx950, loads, and exports every entry point
declared in include/bookleaf_hip.h (no compute calls: runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bookleaf_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bl_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def libpath():
    from bookleaf_f90_amd import build
    if not os.path.exists(build.LIB):
        build.build()
    return build.LIB


def test_header_declares_the_driver_boundary():
    names = declared_functions()
    for n in ("bl_create", "bl_hydro_getdt", "bl_hydro_getforce", "bl_hydro_getstate",
              "bl_hydro_scatteracceleration", "bl_ale_advect", "bl_ale_getfluxvolume",
              "bl_run", "bl_run_graph", "bl_totals"):
        assert n in names


def test_noalloc_limit_matches_the_python_mirror():
    """BL_NOALLOC_MAX_EL (the largest mesh of bl_hydro_getdt_noalloc) and its
    mirror in abi.py agree, and the limit is whole workgroups of 256."""
    from bookleaf_f90_amd import abi
    m = re.search(r"#define BL_NOALLOC_MAX_EL \(1 << (\d+)\)", open(HEADER).read())
    assert m and (1 << int(m.group(1))) == abi.NOALLOC_MAX_EL
    assert abi.NOALLOC_MAX_EL % 256 == 0
    assert "bl_hydro_getdt_noalloc" in declared_functions()


def test_library_exports_every_declared_symbol(libpath):
    lib = ctypes.CDLL(libpath)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_library_targets_gfx950_only(libpath):
    """The embedded offload bundle names only the gfx950 code object."""
    blob = open(libpath, "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-f]+)", blob))
    assert targets == {b"gfx950"}


def test_product_does_not_link_the_oracle(libpath):
    out = subprocess.run(["ldd", libpath], capture_output=True, text=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", libpath], capture_output=True, text=True).stdout
    assert "orc_" not in syms


def test_missing_library_fails_loudly(tmp_path):
    from bookleaf_f90_amd import device
    saved = device._LIB
    device._LIB = None
    try:
        with pytest.raises(ImportError):
            device.load_library(str(tmp_path / "absent.so"))
    finally:
        device._LIB = saved


def test_fortran_binding_declares_every_entry_point():
    """bookleaf_f90_amd/fortran/bookleaf_hip_mod.f90 binds every C entry point
    under its own name."""
    src = open(os.path.join(ROOT, "bookleaf_f90_amd", "fortran", "bookleaf_hip_mod.f90")).read()
    bound = set(re.findall(r'NAME="(bl_[a-z0-9_]+)"', src))
    assert sorted(set(declared_functions()) - bound) == []


def test_fortran_binding_and_host_build(libpath):
    from bookleaf_f90_amd import build
    if not os.path.exists(build.FLANG):
        pytest.skip("flang not installed")
    exes = build.build_fortran([os.path.join(ROOT, "tests", "fortran", "fortran_host.f90")])
    assert all(os.access(e, os.X_OK) for e in exes)
    out = subprocess.run(["ldd", exes[0]], capture_output=True, text=True).stdout
    assert "libbookleaf_hip.so" in out and "not found" not in out
