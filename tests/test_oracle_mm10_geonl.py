"""Pins of the oracle's mm10 under finite strains (drive_10_update with
[R n+1] from the polar decomposition, rstgp1.f:4327; mm10_setup's lattice
spin terms qc built from R n+1 Rp_n^T, mm10_a.f:889-896; the stored
rotation words, mm10_a.f:307; and mm10_calc_grads' finite-strain branch,
mm10_a.f:1611-1661).  No reference fixture covers mm10 with finite
strains, so each piece is pinned by what the mathematics fixes: identity
R reproduces the small-strain update bit for bit; with a rotation the
Jacobian is still the derivative of the residual (finite differences);
the accepted state solves the residual formed with the same R; R is live
(a rotation changes the result); the stored rotation words are R; and the
gradient fit recovers a linear field exactly and takes it to the current
axes with the inverse Jacobian."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle_mm10 import ANGLES, cubic, ry, rz
from warp3d_amd import _abi

DEPS = np.array([2e-4, -0.6e-4, -0.6e-4, 0.3e-4, 0.0, 0.1e-4])


def rotation(step):
    """A rotation growing with the step: 3 degrees about a skew axis."""
    return rz(np.radians(3.0 * step)) @ ry(np.radians(1.5 * step))


def drive(mat, cr, nsteps, rot=True):
    H = _abi.mm10_hist_size(1)
    hn = np.zeros(H)
    urcs = np.zeros(9)
    out = []
    for step in range(1, nsteps + 1):
        r = rotation(step) if rot else None
        rc, u1, h1, hused = O.mm10_gp(mat, [cr], step, 0, 1.0, DEPS, urcs,
                                      hn, rn1=r)
        assert rc == 0, step
        out.append((u1.copy(), h1.copy(), hused.copy(), r))
        hn, urcs = h1, u1
    return out


@pytest.mark.parametrize("ang", ANGLES[:2])
def test_identity_rotation_is_the_small_strain_update(ang):
    mat = cubic()
    cr = O.mm10_crystal(mat, ang)
    H = _abi.mm10_hist_size(1)
    hn_a, hn_b = np.zeros(H), np.zeros(H)
    ua, ub = np.zeros(9), np.zeros(9)
    for step in range(1, 8):
        rc, ua, ha, _ = O.mm10_gp(mat, [cr], step, 0, 1.0, DEPS, ua, hn_a)
        rc2, ub, hb, _ = O.mm10_gp(mat, [cr], step, 0, 1.0, DEPS, ub, hn_b,
                                   rn1=np.eye(3))
        assert rc == rc2 == 0
        assert np.array_equal(ua, ub) and np.array_equal(ha, hb)
        hn_a, hn_b = ha, hb


@pytest.mark.parametrize("ang", ANGLES[:3])
def test_rotation_is_live_and_stored(ang):
    mat = cubic()
    cr = O.mm10_crystal(mat, ang)
    rot = drive(mat, cr, 10)
    ref = drive(mat, cr, 10, rot=False)
    for (u1, h1, _, r), (v1, g1, _, _) in zip(rot, ref):
        # history words 64-72 hold [R n+1] column-major (mm10_a.f:307)
        assert np.array_equal(h1[63:72], r.ravel(order="F"))
        assert np.array_equal(g1[63:72], np.eye(3).ravel())
    # the lattice spin terms feel the rotation once the crystal slips
    u_last, v_last = rot[-1][0], ref[-1][0]
    assert np.abs(u_last[:6] - v_last[:6]).max() > 1e-6 * \
        np.abs(v_last[:6]).max()
    assert rot[-1][0][7] > 0.0   # plastic work: the state is plastic


@pytest.mark.parametrize("ang", ANGLES[:3])
def test_jacobian_matches_finite_differences_with_rotation(ang):
    """As test_oracle_mm10's pin, with R n+1 != I in every sub-step term:
    J11, J12, J21 within 1e-6 of central differences of the residual."""
    mat = cubic()
    cr = O.mm10_crystal(mat, ang)
    hist = drive(mat, cr, 12)
    u1, h1, hn, r = hist[-1]
    hc = hn[87:87 + 71]
    x = np.concatenate([u1[:6], [h1[87 + 42]]])
    R, J = O.mm10_residual(mat, cr, hc, DEPS, 1.0, x, rn1=r)
    Jfd = np.zeros((7, 7))
    for k in range(7):
        h = 1e-6 * max(abs(x[k]), 1.0)
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        Rp, _ = O.mm10_residual(mat, cr, hc, DEPS, 1.0, xp, rn1=r)
        Rm, _ = O.mm10_residual(mat, cr, hc, DEPS, 1.0, xm, rn1=r)
        Jfd[:, k] = (Rp - Rm) / (2 * h)
    mask = np.ones((7, 7), bool)
    mask[6, 6] = False   # the reference's sign quirk of J22
    assert np.abs(J - Jfd)[mask].max() < 1e-6 * np.abs(J).max()
    # the residual formed without the rotation is not solved by the state
    R0, _ = O.mm10_residual(mat, cr, hc, DEPS, 1.0, x)
    assert np.linalg.norm(R0) > 10 * np.linalg.norm(R)


@pytest.mark.parametrize("ang", ANGLES[:3])
def test_accepted_state_solves_the_rotated_residual(ang):
    mat = cubic()
    cr = O.mm10_crystal(mat, ang)
    for u1, h1, hn, r in drive(mat, cr, 12)[-4:]:
        x = np.concatenate([u1[:6], [h1[87 + 42]]])
        R, _ = O.mm10_residual(mat, cr, hn[87:87 + 71], DEPS, 1.0, x, rn1=r)
        assert np.linalg.norm(R) < 1e-5 + 5e-5 * np.abs(u1[:6]).max()


def test_grads_of_a_linear_field():
    """mm10_calc_grads (finite strains): with Rp R^T = A0 + xi A1 + eta A2
    + zeta A3 at the 8 points of the 2x2x2 rule, the least-squares fit
    recovers A1..A3 exactly and the stored slopes are jacinv [A1 A2 A3]
    per component, jacinv the inverse of the given Jacobian."""
    rng = np.random.default_rng(3)
    A = rng.uniform(-1, 1, (4, 3, 3))
    G, H = 8, _abi.mm10_hist_size(1)
    hel = np.zeros((G, H))
    rot = []
    for p in range(G):
        xi, eta, zeta, _ = O.getgpts(_abi.W3D_L3DISOP, 1, p + 1)
        m = A[0] + xi * A[1] + eta * A[2] + zeta * A[3]
        r = rz(0.3 * p) @ ry(0.2 + 0.1 * p)      # a different R a point
        rp = m @ r                               # so that Rp R^T = m
        hel[p, 87 + 9:87 + 18] = rp.ravel(order="F")
        rot.append(r)
    jac = np.array([[0.5, 0.1, 0.0], [0.05, 0.4, 0.02], [0.0, -0.03, 0.6]])
    bad, g = O.mm10_calc_grads_geonl(_abi.W3D_L3DISOP, 1, hel, rot, jac)
    assert bad == 0
    ji = np.linalg.inv(jac)
    want = np.zeros(27)
    for a in range(3):
        for b in range(3):
            slopes = np.array([A[1][a, b], A[2][a, b], A[3][a, b]])
            want[a + 3 * b + 9 * np.arange(3)] = ji @ slopes
    assert np.allclose(g, want, rtol=0, atol=1e-12)
    # a singular Jacobian leaves the words zero, as the reference does
    bad, g0 = O.mm10_calc_grads_geonl(_abi.W3D_L3DISOP, 1, hel, rot,
                                      np.zeros((3, 3)))
    assert bad == 1 and not g0.any()
