"""Pins of the oracle's mm01 (bilinear Mises) update and cnst1 tangent.

mm01.f cannot be compiled here (its module dependencies do not build), so it
is pinned by what the mathematics fixes: the radial-return closed form for
proportional loading with linear mixed hardening, elastic unloading, the
consistency condition, the consistent tangent against finite differences of
the update, the trapezoidal energy density and the REAL-arithmetic H'.
A slip such as a dropped back-stress term, a wrong shear convention or a
transposed tangent fails one of these.
"""
import numpy as np
import pytest

from oracle import oracle as O

E, NU, SY = 200e3, 0.3, 400.0
G = E / (2 * (1 + NU))
K = E / (3 * (1 - 2 * NU))


def fresh_hist():
    return np.zeros(11)


def run(step, deps, cgn=None, hist=None, beta=1.0, H=2000.0, lnelas=False):
    cgn = np.zeros(9) if cgn is None else cgn
    hist = fresh_hist() if hist is None else hist
    return O.mm01_gp(step, E, NU, beta, H, SY, cgn, deps, hist,
                     lnelas=lnelas)


def state(h):
    return int(np.array([h[3]]).view(np.int64)[0] & 0xffffffff)


def seq_eq(s):
    """Mises equivalent stress with engineering shear words 4-6."""
    d = s[:3] - s[:3].mean()
    return np.sqrt(1.5 * (d @ d + 2 * (s[3:6] @ s[3:6])))


@pytest.mark.parametrize("beta", [1.0, 0.0, 0.4])
@pytest.mark.parametrize("H", [2000.0, 0.0, 20000.0])
def test_proportional_radial_return_closed_form(beta, H):
    """From the virgin state a deviatoric increment e gives
    dep = (sig_eq_trial - sy) / (3G + H) for any iso/kin split, and the
    stress deviator is the trial one scaled to the hardened yield radius."""
    for dirn in ([1, -0.5, -0.5, 0, 0, 0], [0.3, -0.1, -0.2, 0.4, -0.2, 0.1]):
        e = np.array(dirn, float)
        e[:3] -= e[:3].mean()
        e *= 0.01 / np.linalg.norm(e)
        cgn1, h1, cep = run(1, e, beta=beta, H=H)
        s_tr = np.concatenate([2 * G * e[:3], G * e[3:]])
        eq_tr = seq_eq(np.concatenate([s_tr, [0, 0, 0]]))
        dep = (eq_tr - SY) / (3 * G + H)
        assert state(h1) == 1
        assert abs(h1[2] - dep) <= 1e-9 * dep
        # volumetric part stays elastic (zero here)
        assert abs(cgn1[:3].sum()) < 1e-9 * SY
        # the stress deviator lies along the trial direction
        scale = (eq_tr - 3 * G * dep) / eq_tr
        assert np.allclose(cgn1[:6], scale * s_tr, rtol=1e-9,
                           atol=1e-9 * SY)
        # back stress (kinematic share) and yield radius (isotropic share)
        alpha = h1[5:11]
        assert np.allclose(alpha, (1 - beta) * H * dep * s_tr / eq_tr,
                           rtol=1e-8, atol=1e-9 * SY)
        assert abs(h1[1] * np.sqrt(3) - (SY + beta * H * dep)) < 1e-8 * SY
        # consistency: |s - alpha| equals the current yield radius
        rel = cgn1[:6] - alpha
        assert abs(seq_eq(np.concatenate([rel, [0, 0, 0]])) -
                   (SY + beta * H * dep)) < 1e-8 * SY


def test_elastic_increment_and_linear_elastic_flag():
    de = np.array([1e-4, -2e-5, 3e-5, 4e-5, -1e-5, 2e-5])
    lam = E * NU / ((1 + NU) * (1 - 2 * NU))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[[0, 1, 2], [0, 1, 2]] += 2 * G
    D[[3, 4, 5], [3, 4, 5]] = G
    for lnelas in (False, True):
        cgn1, h1, cep = run(1, de, lnelas=lnelas)
        assert np.allclose(cgn1[:6], D @ de, rtol=1e-12, atol=1e-12)
        assert state(h1) == 3
        assert np.allclose(cep, D, rtol=1e-12)
    # lnelas never yields
    cgn1, h1, _ = run(1, 10 * np.array([1e-2, -5e-3, -5e-3, 0, 0, 0]),
                      lnelas=True)
    assert state(h1) == 3 and h1[1] == 0.0


def test_unloading_is_elastic():
    e = np.array([1, -0.5, -0.5, 0, 0, 0]) * 0.006
    cgn1, h1, _ = run(1, e, beta=0.5)
    back = -0.1 * e
    cgn2, h2, cep2 = run(2, back, cgn=cgn1, hist=h1, beta=0.5)
    assert state(h2) == 3
    assert h2[2] == h1[2]  # no new plastic strain
    assert np.allclose(cgn2[:6] - cgn1[:6], cep2 @ back, rtol=1e-10)
    assert np.allclose(h2[5:11], h1[5:11])


@pytest.mark.parametrize("beta", [1.0, 0.0, 0.5])
def test_consistent_tangent_matches_finite_differences(beta):
    """cnst1's [D] is the derivative of the returned stress with respect to
    the strain increment (central differences)."""
    rng = np.random.default_rng(4)
    e0 = np.array([0.004, -0.001, -0.0015, 0.002, -0.001, 0.0005])
    cgn1, h1, _ = run(1, e0, beta=beta)
    # continue plastic loading from the converged state of step 1
    de = e0 * 0.3 + rng.normal(scale=2e-4, size=6)
    s1, hh, cep = run(2, de, cgn=cgn1, hist=h1, beta=beta)
    assert state(hh) == 1
    fd = np.zeros((6, 6))
    for k in range(6):
        h = 1e-8
        dp = de.copy()
        dm = de.copy()
        dp[k] += h
        dm[k] -= h
        sp, _, _ = run(2, dp, cgn=cgn1, hist=h1, beta=beta)
        sm, _, _ = run(2, dm, cgn=cgn1, hist=h1, beta=beta)
        fd[:, k] = (sp[:6] - sm[:6]) / (2 * h)
    assert np.allclose(cep, cep.T, rtol=1e-12, atol=1e-9)
    assert np.max(np.abs(cep - fd)) < 2e-5 * np.max(np.abs(fd))


def test_energy_density_is_trapezoidal():
    de = np.array([0.005, -0.001, -0.002, 0.001, 0.0, -0.0005])
    cgn1, h1, _ = run(1, de)
    assert abs(cgn1[6] - 0.5 * de @ cgn1[:6]) < 1e-12 * abs(cgn1[6])
    de2 = 0.5 * de
    cgn2, h2, _ = run(2, de2, cgn=cgn1, hist=h1)
    exp = cgn1[6] + 0.5 * de2 @ (cgn2[:6] + cgn1[:6])
    assert abs(cgn2[6] - exp) < 1e-12 * abs(exp)
    # plastic work density: trapezoid of the Mises stress over dep
    dep = h2[2] - h1[2]
    pw = cgn1[7] + 0.5 * (seq_eq(cgn2) + seq_eq(cgn1)) * dep
    assert abs(cgn2[7] - pw) < 1e-9 * abs(pw)
    assert abs(cgn2[8] - (cgn1[8] + dep)) < 1e-15


def test_step_one_resets_history():
    """mm01_set_history (mm01.f:235) re-initialises the n-history on every
    call during step 1, whatever is passed in."""
    de = np.array([0.003, -0.001, -0.001, 0, 0, 0])
    junk = np.full(11, 7.0)
    a = run(1, de, hist=junk)
    b = run(1, de)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_first_element_word11_quirk_is_inert_when_isothermal():
    """mm01_general reads history(1,11) of the block's first element
    (mm01.f:845); the term is multiplied by (1 - lk)^2 = 0 without
    temperature change, so the results do not depend on it."""
    e = np.array([1, -0.5, -0.5, 0.2, 0, 0]) * 0.005
    c1, h1, _ = run(1, e, beta=0.0)
    de = np.array([0.001, 0.0, -0.001, 0.0, 0.001, 0.0])
    ra = O.mm01_gp(2, E, NU, 0.0, 2000.0, SY, c1, de, h1, first_hist11=0.0)
    rb = O.mm01_gp(2, E, NU, 0.0, 2000.0, SY, c1, de, h1, first_hist11=123.0)
    for a, b in zip(ra, rb):
        assert np.array_equal(a, b)


def test_hprime_in_single_precision():
    """H' = E Et / (E - Et) evaluated in REAL (elprp.f:872)."""
    for e, et in ((200e3, 1980.19802), (70e3, 1372.549), (1.5e5, 7.5e3)):
        e32, et32 = np.float32(e), np.float32(et)
        ref = np.float32(np.float32(e32 * et32) / np.float32(e32 - et32))
        assert O.hprime(e32, et32) == ref
