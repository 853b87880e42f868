"""Build libw3dblk.so (gfx950) in-tree with hipcc.

    python -m warp3d_amd.build

Every HIP source is compiled for gfx950 only, with packed-FP32 instruction
generation disabled, into warp3d_amd/libw3dblk.so.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libw3dblk.so")
ARCH = "gfx950"

SOURCES = ["w3d_api.cpp", "w3d_elements.cpp", "w3d_mm10_setup.cpp",
           "w3d_kernels.hip", "w3d_tangent.hip", "w3d_mm10.hip",
           "w3d_solver.hip", "w3d_serve.cpp", "w3d_asym.hip"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc",
                 shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=8):
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    common = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
              "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-fopenmp", "-I", os.path.join(HERE, "..", "include")]
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        cmd = [cc] + common + ["-x", "hip", "-c", os.path.join(CSRC, src),
                               "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out = p.communicate()[0].decode(errors="replace")
        out = "\n".join(l for l in out.splitlines()
                        if "packed-fp32-ops" not in l and
                        "is not a recognized feature" not in l)
        if out.strip() and (verbose or p.returncode):
            print(out, file=sys.stderr)
        if p.returncode:
            failed.append(src)
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fopenmp", "-o", tmp] + \
        objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


FORTRAN = [os.path.join(HERE, "..", "include", "w3d.f90"),
           os.path.join(HERE, "fortran", "w3d_newton.f90"),
           os.path.join(HERE, "fortran", "w3d_drive.f90")]
DRIVE = os.path.join(HERE, "bin", "w3d_drive")


def amdflang():
    for cand in (os.environ.get("AMDFLANG"), "/opt/rocm/bin/amdflang",
                 shutil.which("amdflang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran(verbose=False):
    """The Fortran ISO_C_BINDING module, the Newton driver and the w3d_drive
    program, linked against libw3dblk.so (rpath relative to the program).
    Skipped, with a message, when no Fortran compiler is installed."""
    fc = amdflang()
    if fc is None:
        print("amdflang not found: Fortran driver not built", file=sys.stderr)
        return None
    mod = os.path.join(HERE, "_build", "fmod")
    os.makedirs(mod, exist_ok=True)
    os.makedirs(os.path.dirname(DRIVE), exist_ok=True)
    tmp = DRIVE + ".tmp"
    cmd = [fc, "-O2", "-fopenmp", "-J", mod] + FORTRAN + [
        "-L", HERE, "-lw3dblk", "-Wl,-rpath,$ORIGIN/..", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, DRIVE)
    return DRIVE


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
