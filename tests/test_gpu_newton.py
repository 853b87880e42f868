"""The Fortran Newton driver (warp3d_amd/bin/w3d_drive: module w3d_newton
over libw3dblk through ISO_C_BINDING, with the device Jacobi-PCG) against
the same Newton loop on the oracle with a sparse direct solve.

Config: unit cube, 10x10x10 8-node bricks (2x2x2, B-bar), interior nodes
jittered +-10 % (seeded), mm01 E = 200 GPa, nu = 0.3, sigma_y = 400 MPa,
H' = E/100, uniaxial tension by imposed end displacement to 2 % in 20
steps, blocks of 128.

Bar: identical Newton iteration counts in every step; displacements after
every step within 1e-8 relative (FP64 tangent, the default cvtest
tolerance 0.01 %).  With the FP32 MFMA tangent: at most one extra Newton
iteration per step and displacements within 1e-7 of the FP64 oracle, both
converged to 1e-5 % (a relative residual of 1e-7, tight enough that the
converged displacements are fixed to better than 1e-7: at the default
0.01 % a converged state is only fixed to about the tolerance, and two
tangents that differ in the eighth digit stop at different points of that
band).
"""
import os

import numpy as np
import pytest

from tests import newton_ref as NR
from tests._util import c1_model
from warp3d_amd import _abi
from warp3d_amd import deck as D
from warp3d_amd import mesh as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    m, arr = c1_model(10)
    cst, top, val = M.uniaxial_bc(m, 0.02 / 20)
    du = np.zeros(3 * m.nonode)
    du[top] = val
    steps = [du] * 20
    ref = NR.run_history(arr, cst, steps)
    assert all(h[0] for h in ref)
    return m, arr, cst, steps, ref


TIGHT = 1e-5  # cvtest test-2 tolerance (%) of the FP32 comparison


def run_drive(tmp_path, arr, cst, steps, precision, res_tol=0.01):
    deck = os.path.join(tmp_path, f"c1_{precision}.deck")
    res = os.path.join(tmp_path, f"c1_{precision}.res")
    D.write_deck(deck, arr, cst, np.array(steps), precision=precision,
                 pcg_rtol=1e-13, res_tol=res_tol)
    out, steps_out = D.run(deck, res)
    print(out, flush=True)
    return steps_out


def test_fortran_driver_fp64_matches_oracle(c1, tmp_path):
    m, arr, cst, steps, ref = c1
    got = run_drive(str(tmp_path), arr, cst, steps, _abi.W3D_FP64)
    assert len(got) == len(ref)
    for s, (g, (conv, it, ratios, u)) in enumerate(zip(got, ref), start=1):
        assert g["converged"] and g["status"] == 0, s
        assert g["iters"] == it, (s, g["iters"], it)
        err = np.abs(g["u"] - u).max() / np.abs(u).max()
        assert err <= 1e-8, (s, err)
        assert (g["pcg_iters"] > 0).all()


def test_fortran_driver_fp32_tangent(c1, tmp_path):
    m, arr, cst, steps, _ = c1
    ref = NR.run_history(arr, cst, steps, res_tol=TIGHT)
    assert all(h[0] for h in ref)
    got = run_drive(str(tmp_path), arr, cst, steps, _abi.W3D_FP32,
                    res_tol=TIGHT)
    assert len(got) == len(ref)
    for s, (g, (conv, it, ratios, u)) in enumerate(zip(got, ref), start=1):
        assert g["converged"], s
        assert g["iters"] <= it + 1, (s, g["iters"], it)
        err = np.abs(g["u"] - u).max() / np.abs(u).max()
        assert err <= 1e-7, (s, err)
