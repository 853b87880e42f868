"""The C-ABI library loads on a machine without a GPU and exports every
entry point that include/w3d.h declares; the ctypes mirror matches the
header's structure layouts."""
import ctypes as C
import os
import re

import pytest

from warp3d_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "w3d.h")
LIB = os.path.join(ROOT, "warp3d_amd", "libw3dblk.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(w3d_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libw3dblk.so is not built (python -m warp3d_amd.build)")
    return C.CDLL(LIB)


def test_header_declares_the_boundary():
    names = declared()
    for must in ("w3d_create", "w3d_recover", "w3d_recover_blocks",
                 "w3d_set_displacements", "w3d_recover_block",
                 "w3d_assemble_ifv", "w3d_tangent_block",
                 "w3d_tangent", "w3d_tangent_blocks", "w3d_pattern",
                 "w3d_assemble_vss", "w3d_commit_step", "w3d_rollback",
                 "w3d_get_block_state", "w3d_last_error"):
        assert must in names


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_no_device_is_reported_cleanly(lib):
    """Without a GPU the library reports an error instead of crashing."""
    lib.w3d_version.restype = C.c_char_p
    assert b"gfx950" in lib.w3d_version()
    lib.w3d_last_error.restype = C.c_char_p
    h = C.c_void_p()
    rc = lib.w3d_create(C.byref(h), None, 0)
    assert rc < 0 and lib.w3d_last_error()


def test_struct_layouts_match_header():
    """Offsets of the ctypes mirror equal those a C compiler gives."""
    import subprocess
    import tempfile
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "w3d.h"
#define P(T, f) printf("%s %zu\n", #T "." #f, offsetof(T, f))
int main(void) {
  P(w3d_material, model); P(w3d_material, e); P(w3d_material, yld_pt);
  P(w3d_material, slip_type); P(w3d_material, cp_e); P(w3d_material, cp_mu);
  P(w3d_material, harden_n); P(w3d_material, eps_dot_0_y);
  P(w3d_material, rtol1); P(w3d_material, miter);
  P(w3d_material, angle_convention); P(w3d_material, ncrystals);
  printf("w3d_material.size %zu\n", sizeof(w3d_material));
  P(w3d_model, c); P(w3d_model, incid); P(w3d_model, angles);
  P(w3d_model, angle_stride);
  P(w3d_model, nmat); P(w3d_model, mats); P(w3d_model, nelblk);
  P(w3d_model, elblks); P(w3d_model, auto_size); P(w3d_model, eps_bbar);
  P(w3d_model, trn); P(w3d_model, trnmat);
  printf("w3d_model.size %zu\n", sizeof(w3d_model));
  P(w3d_newton_opts, nsteps); P(w3d_newton_opts, precision);
  P(w3d_newton_opts, pcg_maxit); P(w3d_newton_opts, adaptive);
  P(w3d_newton_opts, tol);
  P(w3d_newton_opts, tol_abs); P(w3d_newton_opts, pcg_rtol);
  printf("w3d_newton_opts.size %zu\n", sizeof(w3d_newton_opts));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "t.c"), "w") as f:
            f.write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"),
                        os.path.join(d, "t.c"), "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True,
                             check=True).stdout
    for line in out.splitlines():
        key, val = line.split()
        struct, field = key.split(".")
        cls = {"w3d_material": _abi.W3DMaterial,
               "w3d_model": _abi.W3DModel,
               "w3d_newton_opts": _abi.W3DNewtonOpts}[struct]
        if field == "size":
            assert C.sizeof(cls) == int(val), key
        else:
            assert getattr(cls, field).offset == int(val), key


def test_fortran_module_constants_match_the_c_header():
    """The ISO_C_BINDING module repeats the header's ABI version and the
    integer constants a Fortran caller passes; both files must agree."""
    import re
    with open(os.path.join(ROOT, "include", "w3d.h")) as f:
        h = f.read()
    with open(os.path.join(ROOT, "include", "w3d.f90")) as f:
        fm = f.read()
    cdefs = dict(re.findall(r"#define\s+(W3D_\w+)\s+(-?\d+)\b", h))
    cenum = dict(re.findall(r"\b(W3D_\w+)\s*=\s*(-?\d+)\s*[,}]", h))
    # parameter statements, with their continuation lines joined
    stmts = re.findall(r"parameter\s*::((?:[^\n]*&\s*\n)*[^\n]*)", fm)
    fdefs = {}
    for st in stmts:
        fdefs.update(re.findall(r"(W3D_\w+)\s*=\s*(-?\d+)", st))
    assert len(fdefs) >= 20, "constants of w3d.f90 not found"
    assert fdefs["W3D_ABI_VERSION"] == cdefs["W3D_ABI_VERSION"]
    cvals = {**cenum, **cdefs}
    for name, val in fdefs.items():
        if name in cvals:
            assert cvals[name] == val, (name, val, cvals[name])


def test_fortran_module_binds_declared_functions():
    """Every bind(C, name=...) of include/w3d.f90 names a function of
    w3d.h, and the per-block interface is bound."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(root, "include", "w3d.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(w3d_[a-z0-9_]+)"\)', text))
    names = set(declared())
    assert bound <= names, sorted(bound - names)
    for must in ("w3d_set_displacements", "w3d_recover_block",
                 "w3d_recover_blocks", "w3d_assemble_ifv",
                 "w3d_get_ifv_sums", "w3d_tangent_block",
                 "w3d_tangent_blocks"):
        assert must in bound, must
