#!/usr/bin/env python3
"""Build the reference's own Fortran hot path into oracle/_ref (test oracle).

TEST INFRASTRUCTURE ONLY.  Compiles the serial (no MPI, no OpenMP, local
mesh set-up) source set of the reference *where it lies* under
/root/reference/src with AMD flang, in module-dependency order, and links it
with oracle/ref/ref_driver.f90 into oracle/_ref/bookleaf_ref, and with
oracle/ref/ref_meshgen.f90 (the reference's mesh generator on a deck's &MESH
namelist) into oracle/_ref/bookleaf_meshgen.  Nothing is
copied out of the reference tree; outputs go only to oracle/_ref/.  (The
reference's own deck reader cannot be used as shipped: under flang it aborts
at setup_IC.f90:1282, where a PACK result is assigned to a longer character
array; problems therefore come in as binary files.)

The source set follows the reference's object lists for
``NO_MPI=1 NO_OMP=1 LSETUP=1`` (src/utilities/makefile.UTILITIES:41-49,
src/packages/io/driver/makefile.IODR, src/packages/setup/makefile.SETUP):
T1 serial Typhon stubs instead of T3/MPI, no Silo/TyphonIO/METIS, no
problem-specific MOD, and our own host program in place of
inf_control_main.f90.  Flags: -O2 -ffp-contract=off (no FMA contraction),
the same association rules as the HIP build.

With --dropin it also builds oracle/_ref/bookleaf_ref_dropin: the same
reference sources and host program, with the reference's hot-path driver
modules (REPLACED below) left out and bookleaf_f90_amd/fortran/dropin/
compiled in their place (same module names), linked to libbookleaf_hip.so.
The reference's unchanged inf_solver_* then drive the GPU; the host program
is compiled with -DBL_DROPIN so that its dumps first bring the device
fields back.  Objects go to oracle/_ref/dropin_obj and dropin_mod.

usage: python oracle/build_ref.py [--force] [--dropin]
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
SRC = "/root/reference/src"
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FFLAGS = ["-O2", "-ffp-contract=off", "-DNOMPI", "-DNOOMP", "-DLSETUP"]

EXCLUDE_DIRS = ("mods", "utilities/cutils", "utilities/putils")
EXCLUDE_FILES = {
    "inf_control_main.f90",      # replaced by oracle/ref/ref_driver.f90
    "io_dr_silo.f90", "io_kn_silo.f90",  # SILO only
    "setup_metis.f90",           # METIS only
    "io_dr_tio.F90",             # TyphonIO only (not in makefile.IODR)
    "hydro_kn_getclosure.f90",   # not in makefile.HYDROKN (unbuildable stub)
}
# the driver modules the drop-in replaces (src/packages/{hydro,ale}/driver)
REPLACED = {"hydro_dr_getdt.f90", "hydro_dr_set.f90", "hydro_dr_getforce.f90",
            "hydro_dr_getstate.f90", "hydro_dr_getenergy.f90", "hydro_dr_geteos.f90",
            "hydro_dr_getacceleration.f90", "ale_dr_getmeshstatus.f90",
            "ale_dr_getfluxvolume.f90", "ale_dr_advect.f90", "ale_dr_update.f90"}
ROOT = os.path.dirname(HERE)
FORTRAN = os.path.join(ROOT, "bookleaf_f90_amd", "fortran")
LIBDIR = os.path.join(ROOT, "bookleaf_f90_amd")
INTRINSIC = {"iso_c_binding", "iso_fortran_env", "omp_lib", "mpi",
             "ieee_arithmetic", "ieee_exceptions"}


def reference_sources():
    files = []
    for root, _, names in os.walk(SRC):
        rel = os.path.relpath(root, SRC)
        if rel.startswith(EXCLUDE_DIRS):
            continue
        for n in names:
            if not n.lower().endswith(".f90"):
                continue
            if n in EXCLUDE_FILES or n.startswith("t3_"):
                continue
            files.append(os.path.join(root, n))
    return sorted(files)


MOD_RE = re.compile(r"^\s*module\s+(\w+)\s*$", re.I)
USE_RE = re.compile(r"^\s*use\s*(?:,\s*intrinsic\s*::)?\s*(\w+)", re.I)


def scan(path):
    defs, uses = set(), set()
    with open(path, errors="replace") as f:
        for line in f:
            line = line.split("!")[0]
            m = MOD_RE.match(line)
            if m and m.group(1).lower() != "procedure":
                defs.add(m.group(1).lower())
            m = USE_RE.match(line)
            if m:
                uses.add(m.group(1).lower())
    return defs, uses - INTRINSIC


def order(files):
    info = {f: scan(f) for f in files}
    provider = {}
    for f, (defs, _) in info.items():
        for d in defs:
            provider[d] = f
    done, out = set(), []

    def visit(f, stack):
        if f in done:
            return
        if f in stack:
            raise RuntimeError("module cycle at " + f)
        stack.add(f)
        for u in sorted(info[f][1]):
            p = provider.get(u)
            if p is not None and p != f:
                visit(p, stack)
        stack.discard(f)
        done.add(f)
        out.append(f)

    for f in files:
        visit(f, set())
    return out


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise SystemExit(1)


def compile_all(files, objdir, moddir, force, extra=()):
    """Compile in module order; a file is rebuilt when its object is older
    than its source or a module it uses was rebuilt in this pass."""
    info = {f: scan(f) for f in files}
    rebuilt_mods, objs = set(), []
    for f in order(files):
        base = os.path.splitext(os.path.basename(f))[0]
        obj = os.path.join(objdir, base + ".o")
        defs, uses = info[f]
        if (force or not os.path.exists(obj) or os.path.getmtime(obj) < os.path.getmtime(f)
                or uses & rebuilt_mods):
            run([FLANG, "-c"] + FFLAGS + list(extra) + ["-module-dir", moddir, "-I", moddir,
                                                      f, "-o", obj])
            rebuilt_mods |= defs
        objs.append(obj)
    return objs


def build_dropin(force):
    """oracle/_ref/bookleaf_ref_dropin: the reference with its hot-path
    driver modules replaced by the drop-in modules (GPU)."""
    lib = os.path.join(LIBDIR, "libbookleaf_hip.so")
    if not os.path.exists(lib):
        raise SystemExit("build libbookleaf_hip.so first (python -m bookleaf_f90_amd.build)")
    objdir = os.path.join(OUT, "dropin_obj")
    moddir = os.path.join(OUT, "dropin_mod")
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(moddir, exist_ok=True)
    dropin = sorted(os.path.join(FORTRAN, "dropin", n)
                    for n in os.listdir(os.path.join(FORTRAN, "dropin")) if n.endswith(".f90"))
    files = [f for f in reference_sources() if os.path.basename(f) not in REPLACED]
    files += [os.path.join(FORTRAN, "bookleaf_hip_mod.f90")] + dropin
    objs = compile_all(files, objdir, moddir, force)
    host = os.path.join(HERE, "ref", "ref_driver.f90")
    hobj = os.path.join(objdir, "ref_driver.o")
    run([FLANG, "-c", "-cpp", "-DBL_DROPIN"] + FFLAGS + ["-module-dir", moddir, "-I", moddir,
                                                         host, "-o", hobj])
    exe = os.path.join(OUT, "bookleaf_ref_dropin")
    rpath = "-Wl,-rpath,$ORIGIN/" + os.path.relpath(LIBDIR, OUT)
    run([FLANG, "-O2"] + objs + [hobj, "-L", LIBDIR, "-lbookleaf_hip", rpath, "-o", exe])
    print("built", exe, "from", len(files) - len(dropin) - 1, "reference sources and",
          len(dropin), "drop-in modules")


def main():
    force = "--force" in sys.argv
    if not os.path.isdir(SRC):
        print("reference source not present; oracle/_ref not built")
        return 0
    objdir = os.path.join(OUT, "obj")
    moddir = os.path.join(OUT, "mod")
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(moddir, exist_ok=True)
    files = order(reference_sources())
    hosts = {"bookleaf_ref": os.path.join(HERE, "ref", "ref_driver.f90"),
             "bookleaf_meshgen": os.path.join(HERE, "ref", "ref_meshgen.f90")}

    def compile_one(f, extra=()):
        base = os.path.splitext(os.path.basename(f))[0]
        obj = os.path.join(objdir, base + ".o")
        if force or not os.path.exists(obj) or os.path.getmtime(obj) < os.path.getmtime(f):
            run([FLANG, "-c"] + list(extra) + FFLAGS + ["-module-dir", moddir, "-I", moddir,
                                                        f, "-o", obj])
        return obj

    objs = [compile_one(f) for f in files]
    newest = max(os.path.getmtime(o) for o in objs)
    for name, host in hosts.items():
        hobj = compile_one(host, ("-cpp",))
        exe = os.path.join(OUT, name)
        if force or not os.path.exists(exe) or os.path.getmtime(exe) < max(
                newest, os.path.getmtime(hobj)):
            run([FLANG, "-O2"] + objs + [hobj, "-o", exe])
        print("built", exe, "from", len(files), "reference sources")
    if "--dropin" in sys.argv:
        build_dropin(force)
    return 0


if __name__ == "__main__":
    sys.exit(main())
