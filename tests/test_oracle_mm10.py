"""Pins of the oracle's mm10 crystal plasticity (Voce) restatement.

The reference's mm10 sources cannot be compiled here (mod_crystals.f and
mm10_a.f need the main_data module, whose source does not compile with the
available Fortran compiler), so each piece is pinned by what the mathematics
fixes: crystallography of the slip systems, proper rotations, the fourth-order
tensor rotation of the stiffness, the exponential map, the elastic limit,
finite differences of the local Jacobian and of the consistent tangent, the
closed-form Voce saturation curve and the converged local residual.
"""
import numpy as np
import pytest

from oracle import oracle as O
from warp3d_amd import _abi

ANGLES = [(20.0, 35.0, 50.0), (0.0, 0.0, 0.0), (123.0, 77.0, 301.0),
          (270.0, 150.0, 10.0)]


def rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def ry(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def cubic(**kw):
    base = dict(tau_y=50.0, tau_v=100.0, theta_0=1000.0, e=100e3, nu=0.3,
                mu=80e3, elastic="cubic")
    base.update(kw)
    return _abi.mm10_material(**base)


def voigt_to_tensor4(Cv):
    """Engineering-strain Voigt stiffness (11,22,33,12,23,13) -> C_ijkl."""
    idx = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (1, 0): 3, (1, 2): 4,
           (2, 1): 4, (0, 2): 5, (2, 0): 5}
    C4 = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for m in range(3):
                    C4[i, j, k, m] = Cv[idx[i, j], idx[k, m]]
    return C4


def tensor4_to_voigt(C4):
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]
    return np.array([[C4[a + b] for b in pairs] for a in pairs])


@pytest.mark.parametrize("ang", ANGLES)
def test_kocks_rotation_is_zyz(ang):
    """mm10_rotation_matrix (Kocks): the transpose of Rz(psi) Ry(theta)
    Rz(pi - phi); radians and degrees agree."""
    psi, th, phi = np.radians(ang)
    r = O.mm10_rotation_matrix(ang, 1)
    ref = (rz(psi) @ ry(th) @ rz(np.pi - phi)).T
    assert np.allclose(r, ref, atol=1e-14)
    assert np.allclose(O.mm10_rotation_matrix(np.radians(ang), 2), r,
                       atol=1e-14)
    assert abs(np.linalg.det(r) - 1) < 1e-14


@pytest.mark.parametrize("slip", ["fcc", "bcc"])
def test_slip_systems_crystallography(slip):
    """At zero angles the 12 Schmid pairs are the {111}<110> (fcc) or
    {110}<111> (bcc) families: unit b, unit n, b . n = 0, 4 (fcc) or 6
    (bcc) distinct planes, no repeated system; ms is trace-free with
    eigenvalues -1/2, 0, 1/2."""
    cr = O.mm10_crystal(cubic(slip_type=slip), [0.0, 0.0, 0.0])
    ms, qs = cr.schmid()
    planes = set()
    systems = set()
    for s in range(12):
        m = ms[s]
        A = np.array([[m[0], m[3] / 2, m[5] / 2], [m[3] / 2, m[1], m[4] / 2],
                      [m[5] / 2, m[4] / 2, m[2]]])
        w = qs[s]
        W = np.array([[0, w[2], w[1]], [-w[2], 0, w[0]], [-w[1], -w[0], 0]])
        P = A + W  # = b (x) n
        assert abs(np.trace(A)) < 1e-15
        assert np.allclose(np.linalg.eigvalsh(A), [-0.5, 0, 0.5], atol=1e-14)
        assert np.linalg.matrix_rank(P, tol=1e-12) == 1
        u, sv, vt = np.linalg.svd(P)
        b, n = u[:, 0] * sv[0], vt[0]
        assert abs(np.linalg.norm(b) - 1) < 1e-14
        assert abs(b @ n) < 1e-14
        def canon(v):  # direction up to sign: first nonzero entry positive
            v = v / np.abs(v).max()
            return tuple(np.round(v * np.sign(v[np.abs(v) > 1e-6][0]), 6))
        key_n = canon(n)
        key_b = canon(b)
        planes.add(key_n)
        systems.add((key_n, key_b))
        fam_n = sorted(np.round(np.abs(n) * (np.sqrt(3) if slip == "fcc"
                                              else np.sqrt(2)), 6))
        fam_b = sorted(np.round(np.abs(b) * (np.sqrt(2) if slip == "fcc"
                                              else np.sqrt(3)), 6))
        if slip == "fcc":
            assert fam_n == [1, 1, 1] and fam_b == [0, 1, 1]
        else:
            assert fam_n == [0, 1, 1] and fam_b == [1, 1, 1]
    assert len(systems) == 12
    assert len(planes) == (4 if slip == "fcc" else 6)


@pytest.mark.parametrize("ang", ANGLES)
def test_schmid_tensors_rotate_with_the_crystal(ang):
    """ms, qs of a crystal are the sym / skew parts of (Q b) (x) (Q n) with
    Q = g^T and (b, n) the lattice pairs; so between two orientations the
    tensors map by g1^T g0 (Kocks (0,0,0) is itself a half turn about z)."""
    mat = cubic()
    c0 = O.mm10_crystal(mat, [0.0, 0.0, 0.0])
    c1 = O.mm10_crystal(mat, ang)
    Q = c1.rotation().T @ c0.rotation()
    ms0, qs0 = c0.schmid()
    ms1, qs1 = c1.schmid()
    for s in range(12):
        def tens(m, w):
            A = np.array([[m[0], m[3] / 2, m[5] / 2],
                          [m[3] / 2, m[1], m[4] / 2],
                          [m[5] / 2, m[4] / 2, m[2]]])
            W = np.array([[0, w[2], w[1]], [-w[2], 0, w[0]],
                          [-w[1], -w[0], 0]])
            return A + W
        P0 = tens(ms0[s], qs0[s])
        P1 = tens(ms1[s], qs1[s])
        assert np.allclose(P1, Q @ P0 @ Q.T, atol=1e-14)


@pytest.mark.parametrize("ang", ANGLES)
def test_stiffness_rotation_is_tensor_rotation(ang):
    """The rotated crystal stiffness equals C'_ijkl = Q_ia Q_jb Q_kc Q_ld
    C_abcd with Q = g^T; the unrotated cubic stiffness is the inverse of the
    (e, nu, mu) flexibility."""
    mat = cubic()
    c0 = O.mm10_crystal(mat, [0.0, 0.0, 0.0])
    e, nu, mu = mat.cp_e, mat.cp_nu, mat.cp_mu
    S = np.zeros((6, 6))
    S[:3, :3] = -nu / e
    S[[0, 1, 2], [0, 1, 2]] = 1 / e
    S[[3, 4, 5], [3, 4, 5]] = 1 / mu
    assert np.allclose(c0.stiffness() @ S, np.eye(6), atol=1e-13)
    c1 = O.mm10_crystal(mat, ang)
    Q = c1.rotation().T @ c0.rotation()
    C4 = voigt_to_tensor4(c0.stiffness())
    C4r = np.einsum("ia,jb,kc,ld,abcd->ijkl", Q, Q, Q, Q, C4)
    assert np.allclose(c1.stiffness(), tensor4_to_voigt(C4r),
                       rtol=0, atol=1e-9 * abs(C4).max())


def test_isotropic_stiffness_is_rotation_invariant():
    mat = _abi.mm10_material(tau_y=50.0, tau_v=100.0, theta_0=1000.0,
                             e=200e3, nu=0.3)
    c0 = O.mm10_crystal(mat, [0.0, 0.0, 0.0])
    lam = 200e3 * 0.3 / (1.3 * 0.4)
    G = 200e3 / 2.6
    ref = np.zeros((6, 6))
    ref[:3, :3] = lam
    ref[[0, 1, 2], [0, 1, 2]] += 2 * G
    ref[[3, 4, 5], [3, 4, 5]] = G
    assert np.allclose(c0.stiffness(), ref, rtol=1e-13, atol=1e-9)
    for ang in ANGLES:
        assert np.allclose(O.mm10_crystal(mat, ang).stiffness(), ref,
                           rtol=1e-12, atol=1e-8)


def test_exponential_map():
    """mm10_expmw3x3 is the matrix exponential of the skew tensor built by
    mm10_wv2wt (Rodrigues, cross-checked against a Taylor series), and the
    identity below 1e-16."""
    rng = np.random.default_rng(1)
    for scale in (1e-3, 0.1, 1.0, 2.5):
        w = rng.normal(size=3) * scale
        W = np.array([[0, w[2], w[1]], [-w[2], 0, w[0]], [-w[1], -w[0], 0]])
        ref = np.eye(3)
        term = np.eye(3)
        for k in range(1, 60):
            term = term @ W / k
            ref = ref + term
        a = O.mm10_expmw(w)
        assert np.allclose(a, ref, atol=1e-14)
        assert np.allclose(a @ a.T, np.eye(3), atol=1e-14)
    assert np.array_equal(O.mm10_expmw([0.0, 0.0, 0.0]), np.eye(3))


def start_history(mat, cr, ncry=1):
    """n-state after step 1's initialisation (run one elastic update)."""
    H = _abi.mm10_hist_size(ncry)
    rc, u1, h1, hused = O.mm10_gp(mat, [cr] * ncry, 1, 1, 1.0, np.zeros(6),
                                  np.zeros(9), np.zeros(H))
    assert rc == 0
    return hused


def drive(mat, cr, deps, nsteps, dt=1.0, ncry=1):
    """Monotonic strain-increment loading: iteration 0 (elastic pseudo
    update, extrapolation off) then iteration 1 each step, committing."""
    H = _abi.mm10_hist_size(ncry)
    hn = np.zeros(H)
    urcs = np.zeros(9)
    out = []
    for step in range(1, nsteps + 1):
        rc, u1, h1, _ = O.mm10_gp(mat, [cr] * ncry, step, 0, dt, deps, urcs,
                                  hn)
        assert rc == 0, step
        out.append((u1.copy(), h1.copy(), hn.copy()))
        hn, urcs = h1, u1
    return out


def test_elastic_limit_is_hookes_law():
    """Far below the slip resistance (resolved shear / tau ~ 1e-3, so
    |slip| ~ 1e-57 relative) the update is sigma = C deps."""
    mat = cubic()
    cr = O.mm10_crystal(mat, ANGLES[2])
    deps = np.array([1e-7, -3e-8, 2e-8, 4e-8, -1e-8, 5e-8])
    hn = start_history(mat, cr)
    rc, u1, h1, _ = O.mm10_gp(mat, [cr], 2, 0, 1.0, deps, np.zeros(9), hn)
    assert rc == 0
    assert np.allclose(u1[:6], cr.stiffness() @ deps, rtol=1e-12, atol=0)
    # iteration 0 with extrapolation off: the tangent is the stiffness
    rc, u0, h0, _ = O.mm10_gp(mat, [cr], 2, 1, 1.0, deps, np.zeros(9), hn)
    assert np.array_equal(h0[:36], cr.stiffness().ravel(order="F"))


def plastic_state(mat, cr, nsteps=12):
    deps = np.array([2e-4, -0.6e-4, -0.6e-4, 0.3e-4, 0.0, 0.1e-4])
    hist = drive(mat, cr, deps, nsteps)
    return deps, hist


@pytest.mark.parametrize("ang", ANGLES[:3])
def test_jacobian_matches_finite_differences(ang):
    """formJ is the derivative of formR at a plastic state: central
    differences agree to 1e-6 of the Jacobian's scale in J11, J12 and J21.
    J22 = dR2/dtau carries the reference's extra sign(slip) on the rate term
    of mm10_ehard_voche (mm10_b.f:1978-1983), so it differs from the exact
    derivative by -theta_0 n/tau sgn(h)|h|^m (sum|slip| - sum slip) with the
    Voce factor h = 1 - (tau - tau_y)/tau_v."""
    mat = cubic()
    cr = O.mm10_crystal(mat, ang)
    deps, hist = plastic_state(mat, cr)
    u1, h1, hn = hist[-1]
    hc = hn[87:87 + 71]
    x = np.concatenate([u1[:6], [h1[87 + 42]]])
    R, J = O.mm10_residual(mat, cr, hc, deps, 1.0, x)
    Jfd = np.zeros((7, 7))
    for k in range(7):
        h = 1e-6 * max(abs(x[k]), 1.0)
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        Rp, _ = O.mm10_residual(mat, cr, hc, deps, 1.0, xp)
        Rm, _ = O.mm10_residual(mat, cr, hc, deps, 1.0, xm)
        Jfd[:, k] = (Rp - Rm) / (2 * h)
    scale = np.abs(J).max()
    mask = np.ones((7, 7), bool)
    mask[6, 6] = False
    assert np.abs(J - Jfd)[mask].max() < 1e-6 * scale
    tau = x[6]
    slips = h1[87 + 30:87 + 42]
    hv = 1 - (tau - mat.tau_y) / mat.tau_v
    quirk = -mat.theta_0 * mat.harden_n / tau * np.sign(hv) * \
        abs(hv) ** mat.voce_m * (np.abs(slips).sum() - slips.sum())
    assert abs(J[6, 6] - (Jfd[6, 6] + quirk)) < 1e-6 * abs(J[6, 6])
    # the state is plastic with slip of both signs, so the quirk is live
    assert abs(quirk) > 1e-2 and (slips < 0).any()


@pytest.mark.parametrize("ang", ANGLES[:3])
def test_converged_residual_is_small(ang):
    """The accepted (sigma, tau) of a step solves the local equations to the
    reference's tolerances (atol 1e-5, rtol 5e-5 of the initial norm)."""
    mat = cubic()
    cr = O.mm10_crystal(mat, ang)
    deps, hist = plastic_state(mat, cr)
    for u1, h1, hn in hist[-4:]:
        x = np.concatenate([u1[:6], [h1[87 + 42]]])
        R, _ = O.mm10_residual(mat, cr, hn[87:87 + 71], deps, 1.0, x)
        assert np.linalg.norm(R) < 1e-5 + 5e-5 * np.abs(u1[:6]).max()


def test_consistent_tangent_from_finite_difference_jacobian():
    """The stored tangent (history words 1-36) is mm10_tangent's
    sym((J11 - J12 J21 / J22)^-1 C): rebuilt here from a Jacobian taken by
    central differences of the residual at the converged state (plus the
    reference's J22 sign quirk), it agrees to 1e-6.  With tight local
    tolerances the reference's use of the Jacobian of the last Newton
    iterate, not of the converged point, is below that."""
    mat = cubic(atol=1e-10, atol1=1e-10, rtol=1e-12, rtol1=1e-12)
    cr = O.mm10_crystal(mat, ANGLES[0])
    deps, hist = plastic_state(mat, cr)
    _, _, hn = hist[-1]
    urcs_n = hist[-2][0]
    rc, u1, h1, _ = O.mm10_gp(mat, [cr], 13, 0, 1.0, deps, urcs_n, hn)
    assert rc == 0
    T = h1[:36].reshape(6, 6, order="F")
    hc = hn[87:87 + 71]
    x = np.concatenate([u1[:6], [h1[87 + 42]]])
    Jfd = np.zeros((7, 7))
    for k in range(7):
        h = 1e-6 * max(abs(x[k]), 1.0)
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        Rp, _ = O.mm10_residual(mat, cr, hc, deps, 1.0, xp)
        Rm, _ = O.mm10_residual(mat, cr, hc, deps, 1.0, xm)
        Jfd[:, k] = (Rp - Rm) / (2 * h)
    tau = x[6]
    slips = h1[87 + 30:87 + 42]
    hv = 1 - (tau - mat.tau_y) / mat.tau_v
    j22 = Jfd[6, 6] - mat.theta_0 * mat.harden_n / tau * np.sign(hv) * \
        abs(hv) ** mat.voce_m * (np.abs(slips).sum() - slips.sum())
    JJ = Jfd[:6, :6] - np.outer(Jfd[:6, 6], Jfd[6, :6]) / j22
    ref = np.linalg.solve(JJ, cr.stiffness())
    ref = 0.5 * (ref + ref.T)
    assert np.linalg.norm(T - ref) < 1e-6 * np.linalg.norm(T)
    # plastic flow has softened it well below the elastic stiffness
    assert np.linalg.norm(T - cr.stiffness()) > 0.05 * np.linalg.norm(T)


def test_voce_saturation_curve():
    """Under sustained loading tau_tilde rises monotonically to just below
    tau_y + tau_v, and every step satisfies the implicit Voce update
    tau_n+1 = tau_n + theta_0 (1 - (tau_n+1 - tau_y)/tau_v) sum|slip_n+1|
    (voce_m = 1) to the local solver's tolerance."""
    mat = cubic(theta_0=300.0, tau_v=60.0)
    cr = O.mm10_crystal(mat, ANGLES[2])
    deps = np.array([1e-3, -3e-4, -3e-4, 0, 0, 0])
    hist = drive(mat, cr, deps, 300)
    tau = np.array([50.0 + 1e-5] + [h1[87 + 42] for _, h1, _ in hist])
    gam = np.array([np.abs(h1[87 + 30:87 + 42]).sum() for _, h1, _ in hist])
    assert np.all(np.diff(tau) >= 0)
    assert 50.0 + 0.95 * 60.0 < tau[-1] < 50.0 + 60.0
    th, ty, tv = 300.0, 50.0, 60.0
    pred = (tau[:-1] + th * gam * (1 + ty / tv)) / (1 + th * gam / tv)
    assert np.abs(tau[1:] - pred).max() < 1e-4


def test_rotation_and_angles_update():
    """Rp stays a proper rotation, and the stored Euler angles reproduce
    g Rp (mm10_update_euler_angles, Kocks) through mm10_rotation_matrix."""
    mat = cubic()
    ang = ANGLES[2]
    cr = O.mm10_crystal(mat, ang)
    deps, hist = plastic_state(mat, cr, nsteps=20)
    _, h1, _ = hist[-1]
    Rp = h1[87 + 9:87 + 18].reshape(3, 3, order="F")
    assert np.allclose(Rp @ Rp.T, np.eye(3), atol=1e-13)
    assert abs(np.linalg.det(Rp) - 1) < 1e-13
    assert abs(Rp - np.eye(3)).max() > 1e-6  # lattice has rotated
    full = cr.rotation() @ Rp
    assert np.allclose(O.mm10_rotation_matrix(h1[87 + 6:87 + 9], 1), full,
                       atol=1e-12)


def test_taylor_average_of_identical_crystals():
    """Two crystals of one orientation average to the single crystal."""
    mat1 = cubic()
    mat2 = cubic(ncrystals=2)
    cr = O.mm10_crystal(mat1, ANGLES[0])
    deps = np.array([2e-4, -0.6e-4, -0.6e-4, 0.3e-4, 0.0, 0.1e-4])
    a = drive(mat1, cr, deps, 10)
    b = drive(mat2, cr, deps, 10, ncry=2)
    for (u1, h1, _), (u2, h2, _) in zip(a, b):
        assert np.allclose(u1, u2, rtol=1e-14, atol=1e-14)
        assert np.allclose(h1[:87], h2[:87], rtol=1e-14, atol=1e-14)
        assert np.array_equal(h2[87:158], h2[158:229])


def test_iteration_zero_pseudo_update():
    """Iteration 0 with extrapolation off: sigma = sigma_n - R1(sigma_n) and
    the stiffness as tangent; no lattice rotation output (Rp written as
    zeros by the reference's mm10_setup_np1)."""
    mat = cubic()
    cr = O.mm10_crystal(mat, ANGLES[1])
    deps, hist = plastic_state(mat, cr)
    u1, h1, hn = hist[-1]
    rc, u0, h0, _ = O.mm10_gp(mat, [cr], 13, 1, 1.0, deps, u1, h1)
    x = np.concatenate([h1[87:93], [h1[87 + 42]]])
    R, _ = O.mm10_residual(mat, cr, h1[87:158], deps, 1.0, x)
    assert np.allclose(u0[:6], h1[87:93] - R[:6], rtol=1e-14, atol=1e-12)
    assert np.array_equal(h0[87 + 9:87 + 18], np.zeros(9))
    assert h0[87 + 42] == h1[87 + 42]


@pytest.mark.parametrize("m", [1, 4, 8, 14, 27])
def test_least_squares_matches_numpy(m):
    """The DGELS restatement: least-squares solution for m >= 4, minimum-norm
    solution for the one-point rule (numpy's lstsq gives both)."""
    rng = np.random.default_rng(m)
    a = rng.uniform(-1, 1, (m, 4))
    a[:, 3] = 1.0
    b = rng.uniform(-2, 2, m)
    ref = np.linalg.lstsq(a, b, rcond=None)[0]
    assert np.allclose(O.lsq4(a, b), ref, rtol=1e-12, atol=1e-13)


def gauss_coords(etype, order):
    n = O.ngp(etype, order)
    return np.array([O.getgpts(etype, order, p + 1)[:3] for p in range(n)])


@pytest.mark.parametrize("etype,order,span", [(_abi.W3D_L3DISOP, 1, 128),
                                              (_abi.W3D_L3DISOP, 1, 7),
                                              (_abi.W3D_L3DISOP, 1, 1),
                                              (_abi.W3D_Q3DISOP, 1, 3),
                                              (_abi.W3D_TET10, 4, 5)])
def test_grad_fe_flat_copy_and_fit(etype, order, span):
    """rknstr_finish_cp for small strain: the 9 ngp words that
    mm10_a_copy_vector copies out of elem_hist(span, hist, ngp) -- numpy's
    own Fortran memory order here -- fitted linearly over the Gauss points
    (numpy lstsq); a word of an element other than the first and words past
    the first point's history are reached for small spans."""
    hist = 158
    xyz = gauss_coords(etype, order)
    ngp_ = len(xyz)
    rng = np.random.default_rng(span)
    hloc = rng.uniform(-1, 1, (hist, ngp_, span))
    elem_hist = np.transpose(hloc, (2, 0, 1)).copy(order="F")
    flat = elem_hist.ravel(order="F")
    start = np.ravel_multi_index((0, 96, 0), elem_hist.shape, order="F")
    rt = flat[start:start + 9 * ngp_].reshape((3, 3, ngp_), order="F")
    amat = np.column_stack([xyz, np.ones(ngp_)])
    grads = np.zeros((3, 3, 3))
    for a in range(3):
        for b in range(3):
            grads[a, b] = np.linalg.lstsq(amat, rt[a, b], rcond=None)[0][:3]
    got = O.mm10_calc_grads(etype, order, ngp_, span, hist, hloc)
    assert np.allclose(got, grads.ravel(order="F"), rtol=1e-12, atol=1e-13)


def test_grad_fe_recovers_a_linear_field():
    """Words that are an exact linear function of the Gauss-point
    coordinates give back its slopes (span 1: the copied words are those of
    one element, consecutive in the history)."""
    etype, order, hist = _abi.W3D_Q3DISOP, 1, 158
    xyz = gauss_coords(etype, order)
    ngp_ = len(xyz)
    slope = np.arange(27, dtype=float).reshape(3, 3, 3, order="F") * 0.1 - 1
    icpt = np.linspace(-1, 1, 9).reshape(3, 3, order="F")
    hloc = np.zeros((hist, ngp_, 1))
    flat = np.zeros(hist * ngp_)
    for p in range(ngp_):
        val = icpt + slope @ xyz[p]
        flat[96 + 9 * p:96 + 9 * p + 9] = val.ravel(order="F")
    hloc[:, :, 0] = flat.reshape((hist, ngp_), order="F")
    got = O.mm10_calc_grads(etype, order, ngp_, 1, hist, hloc)
    assert np.allclose(got, slope.ravel(order="F"), rtol=1e-12, atol=1e-12)
