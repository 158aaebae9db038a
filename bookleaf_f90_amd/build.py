"""Build the gfx950 HIP library libbookleaf_hip.so in-tree.

The library is compiled with hipcc for gfx950 only, with FMA contraction off
(-ffp-contract=off) so every FP64 operation rounds like the reference's
generic x86-64 build, and with packed FP32 disabled.  Usage:

    python -m bookleaf_f90_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbookleaf_hip.so")
SOURCES = ["bl_api.hip", "bl_ale.hip", "bl_step.hip", "bl_fused.hip", "bl_mix.hip", "bl_diag.hip",
           "bl_relax.hip", "bl_ensemble.hip", "bl_mlp.hip", "bl_tiles.hip"]
HEADERS = ["bl_internal.h", "bl_math.h", "bl_kernels.h", "bl_step.h", "bl_fused.h", "bl_mix.h", "bl_ale.h",
           "bl_relax.h", "bl_tiles.h"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
    "-I" + os.path.join(ROOT, "include"),
]


def _newer(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def build(force=True, verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    deps = srcs + [os.path.join(CSRC, h) for h in HEADERS] + [
        os.path.join(ROOT, "include", "bookleaf_hip.h")]
    if not force and not _newer(LIB, deps):
        return LIB
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building " + LIB)
    os.replace(tmp, LIB)
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN = os.path.join(HERE, "fortran")
FORTRAN_BUILD = os.path.join(FORTRAN, "_build")


def build_fortran(programs=()):
    """Compile the Fortran binding bookleaf_hip_mod (fortran/bookleaf_hip_mod.f90)
    into fortran/_build and link each program source given against it and
    libbookleaf_hip.so (rpath relative to the executable).  Returns the list of
    executables, or None when flang is not installed."""
    if not os.path.exists(FLANG):
        return None
    os.makedirs(FORTRAN_BUILD, exist_ok=True)
    mod_o = os.path.join(FORTRAN_BUILD, "bookleaf_hip_mod.o")
    cmds = [[FLANG, "-O2", "-J", FORTRAN_BUILD, "-c",
             os.path.join(FORTRAN, "bookleaf_hip_mod.f90"), "-o", mod_o]]
    exes = []
    rpath = "-Wl,-rpath,$ORIGIN/" + os.path.relpath(HERE, FORTRAN_BUILD)
    for src in programs:
        src = os.path.abspath(src)
        exe = os.path.join(FORTRAN_BUILD, os.path.splitext(os.path.basename(src))[0])
        cmds.append([FLANG, "-O2", "-I", FORTRAN_BUILD, src, mod_o, "-L", HERE,
                     "-lbookleaf_hip", rpath, "-o", exe])
        exes.append(exe)
    for cmd in cmds:
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=FORTRAN_BUILD)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("flang failed: " + " ".join(cmd))
    return exes


if __name__ == "__main__":
    print(build(force=True, verbose=True))
    print(build_fortran())
