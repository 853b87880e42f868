"""The oracle Newton loop (tests/newton_ref.py, the algorithm of the
Fortran driver) against the closed-form uniaxial response of a bilinear
Mises bar: a jittered cube under homogeneous uniaxial stress (symmetry
planes, imposed end displacement) stays homogeneous, so the end reaction
is sigma_zz times the unit area at every step."""
import numpy as np

from oracle import oracle as O
from tests import newton_ref as NR
from warp3d_amd import _abi
from warp3d_amd import mesh as M


def uniaxial_stress(eps, e, yld, hp):
    """sigma(eps) of 1-D bilinear plasticity with hardening modulus H'."""
    if e * eps <= yld:
        return e * eps
    return (yld + hp * eps) / (1.0 + hp / e)


def test_newton_uniaxial_bar_matches_closed_form():
    m = M.box_hex8(4, jitter=0.1, seed=11)
    mat = _abi.mm01_material(200e3, 0.3, 400.0, hprime=2000.0)
    arr = m.model(mat, auto_size=16)
    de = 0.0008
    cst, top, val = M.uniaxial_bc(m, de)
    du = np.zeros(3 * m.nonode)
    du[top] = val
    hist = NR.run_history(arr, cst, [du] * 6)
    e = float(np.float32(200e3))
    hp = float(O.hprime(np.float32(200e3), mat.tan_e))
    om = O.OracleModel(arr)
    om.pattern(cst)
    for s, (conv, it, ratios, u) in enumerate(hist, start=1):
        assert conv
        # reaction at the loaded end: internal forces of the top dofs
        ifv, _, _, _ = om.recover(1, 1, np.zeros_like(u), u)  # one-shot
        ref = uniaxial_stress(s * de, e, 400.0, hp)
        assert abs(ifv[top].sum() - ref) <= 1e-9 * ref, (s, ifv[top].sum(),
                                                          ref)
        # Newton iterations: 1 while elastic, 2 once plastic
        assert it == (1 if e * s * de <= 400.0 else 2)
