"""Pins of the oracle's element kinematics.

Each check is chosen so a plausible slip (wrong point order, a transposed
Jacobian, a missing B-bar term) fails it:
  * against the reference's own leaf routines, built from its Fortran sources
    by `make -C oracle ref` (getgpts.f, derivs.f, blcmp1.f, bmod.f,
    vol_terms.f, vol_avg.f, gtlsn1.f), when they are built here;
  * against closed forms: polynomial exactness of the rules, partition of
    unity, exact Jacobians of affine maps, patch tests, rigid-body modes.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from warp3d_amd import mesh as M
from warp3d_amd._abi import W3D_L3DISOP, W3D_Q3DISOP, W3D_TET10

RULES = [(W3D_L3DISOP, 1, 8), (W3D_Q3DISOP, 1, 27), (W3D_Q3DISOP, 8, 8),
         (W3D_Q3DISOP, 9, 14), (W3D_Q3DISOP, 2, 9), (W3D_TET10, 1, 1),
         (W3D_TET10, 4, 4), (W3D_TET10, 5, 5)]
NN = {W3D_L3DISOP: 8, W3D_Q3DISOP: 20, W3D_TET10: 10}
NAT = {W3D_L3DISOP: M.HEX8_NAT, W3D_Q3DISOP: M.HEX20_NAT}
TET_NAT = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1),
                    (.5, 0, 0), (.5, .5, 0), (0, .5, 0), (0, 0, .5),
                    (.5, 0, .5), (0, .5, .5)])


def points(et, order):
    n = O.ngp(et, order)
    return np.array([O.getgpts(et, order, g + 1) for g in range(n)])


# ------------------------------------------------------------ Gauss rules
@pytest.mark.parametrize("et,order,n", RULES)
def test_rule_sizes(et, order, n):
    assert O.ngp(et, order) == n


@pytest.mark.parametrize("et,order,deg", [(W3D_L3DISOP, 1, 3),
                                          (W3D_Q3DISOP, 1, 5),
                                          (W3D_Q3DISOP, 8, 3),
                                          (W3D_Q3DISOP, 9, 3)])
def test_hex_rule_exactness(et, order, deg):
    """Integrates x^a y^b z^c over [-1,1]^3 exactly up to the rule's degree
    (the 14-point rule's published constants carry 9 digits)."""
    p = points(et, order)
    tol = 1e-8 if order == 9 else 1e-13
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                num = np.sum(p[:, 3] * p[:, 0] ** a * p[:, 1] ** b *
                             p[:, 2] ** c)
                ex = 1.0
                for k in (a, b, c):
                    ex *= 0.0 if k % 2 else 2.0 / (k + 1)
                assert abs(num - ex) < tol, (a, b, c, num, ex)


def test_hex27_point_order():
    """getgpts.f:221-355: eta runs fastest, then xi, then zeta."""
    p = points(W3D_Q3DISOP, 1)
    l1 = 0.774596669241483
    assert tuple(p[0, :3]) == (-l1, -l1, -l1)
    assert tuple(p[1, :3]) == (-l1, 0.0, -l1)
    assert tuple(p[3, :3]) == (0.0, -l1, -l1)
    assert tuple(p[9, :3]) == (-l1, -l1, 0.0)
    assert p[13, 3] == 0.7023319615912210


@pytest.mark.parametrize("order,deg", [(1, 1), (4, 2), (5, 3)])
def test_tet_rule_exactness(order, deg):
    """Weights sum to 1 (the 1/6 volume factor is applied separately by
    adjust_scalar_weights); monomials integrate to a!b!c!/(a+b+c+3)! x 6."""
    from math import factorial
    p = points(W3D_TET10, order)
    assert abs(p[:, 3].sum() - 1.0) < 1e-12
    tol = 1e-7 if order == 4 else 1e-13  # quad6 carries 8-digit constants
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                num = np.sum(p[:, 3] * p[:, 0] ** a * p[:, 1] ** b *
                             p[:, 2] ** c)
                ex = 6.0 * factorial(a) * factorial(b) * factorial(c) / \
                    factorial(a + b + c + 3)
                assert abs(num - ex) < tol


# ------------------------------------------------------- shape functions
def shape_values(et, x):
    """Independent closed-form shape functions (values, not derivatives)."""
    if et == W3D_TET10:
        s2, s3, s4 = x
        s1 = 1 - s2 - s3 - s4
        L = [s1, s2, s3, s4]
        v = [L[i] * (2 * L[i] - 1) for i in range(4)]
        for a, b in M.TET10_EDGES:
            v.append(4 * L[a] * L[b])
        return np.array(v)
    nat = NAT[et]
    out = []
    for (ri, si, ti) in nat:
        r, s, t = x
        if et == W3D_L3DISOP:
            out.append((1 + ri * r) * (1 + si * s) * (1 + ti * t) / 8)
        elif ri and si and ti:
            out.append((1 + ri * r) * (1 + si * s) * (1 + ti * t) *
                       (ri * r + si * s + ti * t - 2) / 8)
        elif ri == 0:
            out.append((1 - r * r) * (1 + si * s) * (1 + ti * t) / 4)
        elif si == 0:
            out.append((1 + ri * r) * (1 - s * s) * (1 + ti * t) / 4)
        else:
            out.append((1 + ri * r) * (1 + si * s) * (1 - t * t) / 4)
    return np.array(out)


@pytest.mark.parametrize("et", [W3D_L3DISOP, W3D_Q3DISOP, W3D_TET10])
def test_derivs_match_finite_differences(et):
    rng = np.random.default_rng(3)
    n = NN[et]
    for _ in range(5):
        x = rng.uniform(-0.8, 0.8, 3) if et != W3D_TET10 else \
            rng.dirichlet(np.ones(4))[1:]
        d = O.derivs(et, *x, n)
        h = 1e-6
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            fd = (shape_values(et, x + e) - shape_values(et, x - e)) / (2 * h)
            assert np.max(np.abs(fd - d[k])) < 1e-8
            assert abs(d[k].sum()) < 1e-13  # partition of unity


@pytest.mark.parametrize("et", [W3D_L3DISOP, W3D_Q3DISOP, W3D_TET10])
def test_shape_nodes_interpolate(et):
    """N_i(node_j) = delta_ij: the node numbering the meshes use matches
    the derivative routines."""
    nat = TET_NAT[:, 1:] if False else None
    nodes = (TET_NAT[:, 1:] if False else None)
    if et == W3D_TET10:
        pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1],
                        [.5, 0, 0], [.5, .5, 0], [0, .5, 0], [0, 0, .5],
                        [.5, 0, .5], [0, .5, .5]], float)
    else:
        pts = NAT[et].astype(float)
    for j, p in enumerate(pts):
        v = shape_values(et, p)
        assert np.allclose(v, np.eye(len(pts))[j], atol=1e-14)


# -------------------------------------------------------------- Jacobian
def element_coords(m, e):
    nodes = m.conn[e] - 1
    xyz = m.coords[nodes]
    return np.concatenate([xyz[:, 0], xyz[:, 1], xyz[:, 2]])


@pytest.mark.parametrize("maker,et,order", [
    (M.box_hex8, W3D_L3DISOP, 1), (M.box_hex20, W3D_Q3DISOP, 1),
    (M.box_tet10, W3D_TET10, 4)])
def test_jacobian_of_affine_map(maker, et, order):
    """For an affine image of the reference element J is the constant map
    matrix: det J = volume ratio, Gamma = J^-1, at every point."""
    m = maker(1)
    A = np.array([[1.3, 0.2, -0.1], [0.1, 0.9, 0.3], [-0.2, 0.1, 1.1]])
    m.coords = m.coords @ A.T + np.array([0.5, -1.0, 2.0])
    n = NN[et]
    for e in range(m.noelem):
        ce = element_coords(m, e)
        for p in points(et, order):
            d = O.derivs(et, *p[:3], n)
            jac, dj, gama = O.jacob1(*d, ce)
            assert np.allclose(gama @ jac, np.eye(3), atol=1e-13)
            assert dj > 0
            if et != W3D_TET10:
                assert np.allclose(jac, 0.5 * A.T, atol=1e-13)
                assert abs(dj - np.linalg.det(0.5 * A)) < 1e-13


# ---------------------------------------------------------------- B, B-bar
def element_b(m, e, et, order, bbar, vol=None):
    n = NN[et]
    ce = element_coords(m, e)
    out = []
    for p in points(et, order):
        d = O.derivs(et, *p[:3], n)
        _, dj, gama = O.jacob1(*d, ce)
        out.append((O.bmatrix(gama, *d, vol=vol), dj, p[3]))
    return out


def mean_dilatation(m, e):
    n = 8
    ce = element_coords(m, e)
    acc = np.zeros((8, 3))
    vol = 0.0
    for p in points(W3D_L3DISOP, 1):
        d = O.derivs(W3D_L3DISOP, *p[:3], n)
        _, dj, gama = O.jacob1(*d, ce)
        acc += (gama @ np.array(d)).T * dj
        vol += dj
    return acc / vol


def test_patch_test_uniform_strain():
    """A linear displacement field gives the exact uniform strain at every
    point of a jittered mesh, with and without B-bar."""
    m = M.box_hex8(3, jitter=0.2, seed=4)
    G = np.array([[1e-3, 2e-4, -3e-4], [5e-4, -2e-3, 1e-4],
                  [-1e-4, 3e-4, 4e-3]])
    u = m.coords @ G.T
    eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0],
                    G[1, 2] + G[2, 1], G[0, 2] + G[2, 0]])
    for e in range(m.noelem):
        nodes = m.conn[e] - 1
        ue = np.concatenate([u[nodes, 0], u[nodes, 1], u[nodes, 2]])
        for vol in (None, mean_dilatation(m, e)):
            for b, dj, w in element_b(m, e, W3D_L3DISOP, 1, vol is not None,
                                      vol):
                assert np.allclose(b @ ue, eps, atol=1e-15)


def test_bbar_mean_dilatation_property():
    """With B-bar the volumetric strain is the same at every point and
    equals the volume average of the standard one (bmod.f:16 comments)."""
    m = M.box_hex8(2, jitter=0.25, seed=5)
    rng = np.random.default_rng(1)
    for e in range(m.noelem):
        ue = rng.normal(size=24)
        vol = mean_dilatation(m, e)
        std = element_b(m, e, W3D_L3DISOP, 1, False)
        bb = element_b(m, e, W3D_L3DISOP, 1, True, vol)
        tr = [(b @ ue)[:3].sum() for b, _, _ in bb]
        avg = sum((b @ ue)[:3].sum() * dj for b, dj, _ in std) / \
            sum(dj for _, dj, _ in std)
        assert np.allclose(tr, avg, rtol=1e-12, atol=1e-14)
        # deviatoric and shear parts are untouched
        for (b0, _, _), (b1, _, _) in zip(std, bb):
            s0, s1 = b0 @ ue, b1 @ ue
            assert np.allclose(s0[3:], s1[3:], atol=1e-14)
            assert np.allclose(s0[:3] - s0[:3].mean(),
                               s1[:3] - s1[:3].mean(), atol=1e-13)


@pytest.mark.parametrize("maker,et,order", [
    (M.box_hex8, W3D_L3DISOP, 1), (M.box_hex20, W3D_Q3DISOP, 1),
    (M.box_tet10, W3D_TET10, 4)])
def test_rigid_body_modes(maker, et, order):
    m = maker(1, jitter=0.1, seed=2)
    for e in range(min(m.noelem, 3)):
        nodes = m.conn[e] - 1
        x = m.coords[nodes]
        for w in (np.array([1, 0, 0.]), np.array([0, 1, 0.]),
                  np.array([0, 0, 1.])):
            ur = np.cross(w, x) + np.array([0.3, -0.2, 0.1])
            ue = np.concatenate([ur[:, 0], ur[:, 1], ur[:, 2]])
            for b, _, _ in element_b(m, e, et, order, False):
                assert np.allclose(b @ ue, 0, atol=1e-13)


# ------------------------------------- the reference's own leaf routines
class _RefLeaf:
    """The reference's leaf library, loaded on the first use by a test
    (so that a test process that runs none of these never maps it)."""
    lib = None

    def __getattr__(self, name):
        if _RefLeaf.lib is None:
            _RefLeaf.lib = O.load_ref_leaf()
        return getattr(_RefLeaf.lib, name)


REF = _RefLeaf()
needs_ref = pytest.mark.skipif(not O.ref_leaf_built(),
                               reason="reference leaf routines not built "
                               "(make -C oracle ref)")
MXVL, MXEDOF = 128, 60


def _i(v):
    return C.byref(C.c_int32(v))


@needs_ref
@pytest.mark.parametrize("et,order,n", RULES)
def test_rules_bit_exact_vs_reference(et, order, n):
    for g in range(1, n + 1):
        d = [C.c_double() for _ in range(4)]
        REF.getgpts_(_i(et), _i(order), _i(g), *[C.byref(x) for x in d])
        mine = O.getgpts(et, order, g)
        assert tuple(x.value for x in d) == mine


@needs_ref
@pytest.mark.parametrize("et", [W3D_L3DISOP, W3D_Q3DISOP, W3D_TET10])
def test_derivs_bit_exact_vs_reference(et):
    rng = np.random.default_rng(7)
    n = NN[et]
    for _ in range(20):
        x = rng.uniform(-1, 1, 3)
        ref = [np.zeros(n) for _ in range(3)]
        REF.derivs_(_i(et), *[C.byref(C.c_double(v)) for v in x],
                    *[O.dptr(a) for a in ref])
        mine = O.derivs(et, *x, n)
        for a, b in zip(ref, mine):
            assert np.array_equal(a, b)


@needs_ref
@pytest.mark.parametrize("bbar", [False, True])
def test_strain_increment_bit_exact_vs_reference(bbar):
    """gtlsn1 (blcmp1 + bmod + B du) of the reference against the oracle's
    B-matrix and its gtlsn1 loop order, for a block of jittered elements."""
    m = M.box_hex8(3, jitter=0.2, seed=11)
    span = m.noelem
    rng = np.random.default_rng(5)
    due = rng.normal(scale=1e-3, size=(span, 24))
    ref_due = np.zeros((24, MXVL))
    ref_due[:, :span] = due.T
    for g in range(8):
        p = O.getgpts(W3D_L3DISOP, 1, g + 1)
        d = O.derivs(W3D_L3DISOP, *p[:3], 8)
        gam = np.zeros((3, 3, MXVL))  # Fortran gama(mxvl,3,3) transposed
        vol = np.zeros((3, 8, MXVL))
        volume = np.zeros(MXVL)
        mine = np.zeros((span, 6))
        for e in range(span):
            _, dj, ga = O.jacob1(*d, element_coords(m, e))
            gam[:, :, e] = ga.T
            v = mean_dilatation(m, e) if bbar else None
            if bbar:
                vol[:, :, e] = v.T
            b = O.bmatrix(ga, *d, vol=v)
            de = np.zeros(6)
            u = due[e]
            for j in range(8):
                de[0] = de[0] + b[0, j] * u[j] + b[0, 8 + j] * u[8 + j] + \
                    b[0, 16 + j] * u[16 + j]
                de[1] = de[1] + b[1, j] * u[j] + b[1, 8 + j] * u[8 + j] + \
                    b[1, 16 + j] * u[16 + j]
                de[2] = de[2] + b[2, j] * u[j] + b[2, 8 + j] * u[8 + j] + \
                    b[2, 16 + j] * u[16 + j]
                de[3] = de[3] + b[3, j] * u[j] + b[3, 8 + j] * u[8 + j]
                de[4] = de[4] + b[4, 8 + j] * u[8 + j] + b[4, 16 + j] * u[16 + j]
                de[5] = de[5] + b[5, j] * u[j] + b[5, 16 + j] * u[16 + j]
            mine[e] = de
        deps = np.zeros((6, MXVL))
        bref = np.zeros((6, MXEDOF, MXVL))
        gam_f = np.asfortranarray(np.transpose(gam, (2, 1, 0)))
        due_f = np.asfortranarray(ref_due.T)
        vol_f = np.asfortranarray(np.transpose(vol, (2, 1, 0)))
        deps_f = np.asfortranarray(deps.T)
        b_f = np.asfortranarray(np.transpose(bref, (2, 1, 0)))
        REF.gtlsn1_(_i(span), _i(8), O.dptr(due_f), O.dptr(deps_f),
                    O.dptr(gam_f), *[O.dptr(np.ascontiguousarray(a))
                                     for a in d],
                    O.dptr(vol_f), _i(1 if bbar else 0),
                    C.byref(C.c_double(0.0)), O.dptr(b_f))
        assert np.array_equal(deps_f[:span], mine), g


@needs_ref
def test_volume_terms_bit_exact_vs_reference():
    m = M.box_hex8(2, jitter=0.2, seed=12)
    span = m.noelem
    vol_f = np.zeros((MXVL, 8, 3), order="F")
    volume = np.zeros(MXVL)
    for g in range(8):
        p = O.getgpts(W3D_L3DISOP, 1, g + 1)
        d = O.derivs(W3D_L3DISOP, *p[:3], 8)
        gam_f = np.zeros((MXVL, 3, 3), order="F")
        dj = np.zeros(MXVL)
        for e in range(span):
            _, dj[e], ga = O.jacob1(*d, element_coords(m, e))
            gam_f[e] = ga
        REF.vol_terms_(O.dptr(gam_f), O.dptr(dj), O.dptr(vol_f),
                       *[O.dptr(np.ascontiguousarray(a)) for a in d],
                       O.dptr(volume), _i(span), _i(MXVL))
    REF.vol_avg_(O.dptr(vol_f), O.dptr(volume), _i(span), _i(MXVL))
    for e in range(span):
        # the oracle's element routine (same sums, same order)
        ce = element_coords(m, e)
        acc = np.zeros((8, 3))
        v = 0.0
        for g in range(8):
            p = O.getgpts(W3D_L3DISOP, 1, g + 1)
            d = O.derivs(W3D_L3DISOP, *p[:3], 8)
            _, dj, ga = O.jacob1(*d, ce)
            for k in range(3):
                for j in range(8):
                    acc[j, k] = acc[j, k] + (ga[k, 0] * d[0][j] +
                                             ga[k, 1] * d[1][j] +
                                             ga[k, 2] * d[2][j]) * dj
            v = v + dj
        acc = acc * (1.0 / v)
        assert np.array_equal(vol_f[e], acc)
