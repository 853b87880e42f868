"""Pins of the oracle's temperature loading: shape functions, the point
temperatures, the thermal strain increment and temperature-dependent mm01
properties (segmental curve sets).

Each check is chosen so that a plausible slip fails it:
  * the shape functions against the reference's own shapef.f (built by
    `make -C oracle ref`), bit for bit, and against partition of unity and
    the nodal Kronecker property;
  * free thermal expansion of a jittered body: the displacement field
    u = alpha dT x is stress free, so a sign flip of the thermal strain,
    a missing shear term or a wrong point temperature leaves stresses;
  * a fully restrained body: the closed form sigma = -E alpha dT / (1 - 2 nu)
    on every normal component;
  * a linear temperature field: the point temperature equals the field at
    the point's physical position;
  * temperature-dependent curves: the curves' clamping and interpolation
    seen through elastic stresses in closed form, and the consistency of
    the mm01 return with a hardening modulus that changes with temperature
    (the lk != 1 branch of mm01_general): the updated relative stress lies
    on the updated yield surface.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from warp3d_amd import _abi
from warp3d_amd import mesh as M
from warp3d_amd._abi import W3D_L3DISOP, W3D_Q3DISOP, W3D_TET10

NN = {W3D_L3DISOP: 8, W3D_Q3DISOP: 20, W3D_TET10: 10}
TET_NAT = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1),
                    (.5, 0, 0), (.5, .5, 0), (0, .5, 0), (0, 0, .5),
                    (.5, 0, .5), (0, .5, .5)])
E, NU = 200000.0, 0.25   # both exact in REAL


def natural_nodes(et):
    if et == W3D_L3DISOP:
        return M.HEX8_NAT.astype(float)
    if et == W3D_Q3DISOP:
        return M.HEX20_NAT.astype(float)
    return TET_NAT.astype(float)


# ------------------------------------------------------ shape functions
@pytest.mark.parametrize("et", [W3D_L3DISOP, W3D_Q3DISOP, W3D_TET10])
def test_shapef_partition_and_nodal_values(et):
    n = NN[et]
    rng = np.random.default_rng(3)
    for _ in range(20):
        x = rng.uniform(-1, 1, 3) if et != W3D_TET10 else rng.dirichlet(
            np.ones(4))[1:]
        assert abs(O.shapef(et, *x, n).sum() - 1.0) < 1e-14
    nat = natural_nodes(et)
    for k in range(n):
        v = O.shapef(et, *nat[k], n)
        assert np.allclose(v, np.eye(n)[k], atol=1e-15)


class _RefLeaf:
    lib = None

    def __getattr__(self, name):
        if _RefLeaf.lib is None:
            _RefLeaf.lib = O.load_ref_leaf()
        return getattr(_RefLeaf.lib, name)


REF = _RefLeaf()
needs_ref = pytest.mark.skipif(not O.ref_leaf_built(),
                               reason="reference leaf routines not built "
                               "(make -C oracle ref)")


@needs_ref
@pytest.mark.parametrize("et", [W3D_L3DISOP, W3D_Q3DISOP, W3D_TET10])
def test_shapef_bit_exact_vs_reference(et):
    rng = np.random.default_rng(11)
    n = NN[et]
    for _ in range(30):
        x = rng.uniform(-1, 1, 3)
        ref = np.zeros(n)
        REF.shapef_(C.byref(C.c_int32(et)),
                    *[C.byref(C.c_double(v)) for v in x], O.dptr(ref))
        assert np.array_equal(ref, O.shapef(et, *x, n))


# ---------------------------------------------------------- body tests
def _mesh(et, n=2, jitter=0.15, seed=5):
    f = {W3D_L3DISOP: M.box_hex8, W3D_Q3DISOP: M.box_hex20,
         W3D_TET10: M.box_tet10}[et]
    return f(n, jitter=jitter, seed=seed)


def _mm01(alpha, yld=1.0e4, curves=None, eps=None, alpha_zero=0.0):
    m = _abi.mm01_material(E, NU, yld, hprime=2000.0)
    _abi.set_thermal(m, alpha, alpha_zero)
    if curves is not None:
        _abi.set_curves(m, eps, curves)
    return m


def _stresses(orc):
    out = []
    for b in range(orc.nelblk):
        out.append(orc.get_block_state(b + 1, _abi.STRESS_N1)
                   .reshape(-1, 9)[:, :6])
    return np.vstack(out)


CASES = [(W3D_L3DISOP, 1), (W3D_L3DISOP, 0), (W3D_Q3DISOP, 0),
         (W3D_TET10, 0)]


@pytest.mark.parametrize("et,bbar", CASES)
@pytest.mark.parametrize("iter_", [0, 1])
def test_free_thermal_expansion_is_stress_free(et, bbar, iter_):
    """u = diag(alpha) dT x + the engineering shear of alpha_xy: the
    strain field is exactly the thermal strain, so every stress and every
    element force vanishes; the split of dT between node and element parts
    and the temperature at n must not matter."""
    alpha = [1.2e-5, 0.9e-5, 1.5e-5, 0.4e-5, 0.0, 0.0]
    mesh = _mesh(et)
    arr = mesh.model(_mm01(alpha), bbar=bbar)
    orc = O.OracleModel(arr)
    nn, ne = mesh.nonode, mesh.noelem
    t0 = np.full(nn, 20.0)
    dtn = np.full(nn, 30.0)
    dte = np.full(ne, 20.0)        # dT = 50 in all
    orc.set_temperatures(t0, dtn, None, dte, True)
    x = mesh.coords
    a32 = np.array(alpha, np.float32).astype(float)
    dT = 50.0
    du = np.zeros_like(x)
    du[:, 0] = a32[0] * dT * x[:, 0] + a32[3] * dT * x[:, 1]
    du[:, 1] = a32[1] * dT * x[:, 1]
    du[:, 2] = a32[2] * dT * x[:, 2]
    f, *_ = orc.recover(1, iter_, np.zeros(3 * nn), du.ravel())
    sig = _stresses(orc)
    scale = E * 1e-5 * dT
    assert np.max(np.abs(sig)) < 1e-9 * scale
    assert np.max(np.abs(f)) < 1e-9 * scale


@pytest.mark.parametrize("et,bbar", CASES)
def test_restrained_body_closed_form(et, bbar):
    """No displacement, uniform dT: sigma_ii = -E alpha dT / (1 - 2 nu),
    no shear, on every point, for the linear (iteration 0) and the mm01
    (iteration 1) updates; alpha_zero shifts both temperatures and drops
    out for a temperature-independent alpha."""
    alpha = 1.0e-5
    mesh = _mesh(et)
    arr = mesh.model(_mm01(alpha, alpha_zero=7.5), bbar=bbar)
    for it in (0, 1):
        orc = O.OracleModel(arr)
        orc.set_temperatures(np.full(mesh.nonode, 100.0),
                             np.full(mesh.nonode, -40.0), None, None, True)
        orc.recover(1, it, np.zeros(3 * mesh.nonode),
                    np.zeros(3 * mesh.nonode))
        sig = _stresses(orc)
        a = float(np.float32(alpha))
        want = -E * a * (-40.0) / (1 - 2 * NU)
        assert np.allclose(sig[:, :3], want, rtol=1e-12, atol=0)
        assert np.max(np.abs(sig[:, 3:])) < 1e-9 * abs(want)


def test_not_incremental_means_no_thermal_strain():
    mesh = _mesh(W3D_L3DISOP)
    arr = mesh.model(_mm01(1.0e-5))
    orc = O.OracleModel(arr)
    orc.set_temperatures(np.full(mesh.nonode, 100.0),
                         np.full(mesh.nonode, -40.0), None, None, False)
    orc.recover(1, 1, np.zeros(3 * mesh.nonode), np.zeros(3 * mesh.nonode))
    assert np.max(np.abs(_stresses(orc))) == 0.0


def _gp_positions(mesh, arr):
    """Physical positions of every point, in block state order
    (point fastest within element): x = sum N_j(gp) x_j."""
    et = int(arr.etype[0])
    order = int(arr.int_order[0])
    ng = O.ngp(et, order)
    n = NN[et]
    sf = np.array([O.shapef(et, *O.getgpts(et, order, g + 1)[:3], n)
                   for g in range(ng)])
    xe = mesh.coords[mesh.conn - 1]           # (ne, nn, 3)
    return np.einsum("gj,ejk->egk", sf, xe).reshape(-1, 3)


@pytest.mark.parametrize("et", [W3D_L3DISOP, W3D_Q3DISOP, W3D_TET10])
def test_linear_temperature_field_interpolates_exactly(et):
    """A temperature change linear in x through nodal values (with an
    element part on top): the restrained stress at every point gives back
    dT at the point's position."""
    mesh = _mesh(et, jitter=0.2)
    arr = mesh.model(_mm01(1.0e-5), bbar=0)
    orc = O.OracleModel(arr)
    x = mesh.coords
    g = np.array([30.0, -10.0, 20.0])
    dtn = 5.0 + x @ g
    dte = np.linspace(-3.0, 3.0, mesh.noelem)
    orc.set_temperatures(None, dtn, None, dte, True)
    orc.recover(1, 0, np.zeros(3 * mesh.nonode), np.zeros(3 * mesh.nonode))
    sig = _stresses(orc)
    a = float(np.float32(1.0e-5))
    dT = -sig[:, 0] * (1 - 2 * NU) / (E * a)
    ng = sig.shape[0] // mesh.noelem
    want = 5.0 + _gp_positions(mesh, arr) @ g + np.repeat(dte, ng)
    assert np.allclose(dT, want, rtol=1e-10, atol=1e-10)


# ------------------------------------- temperature-dependent mm01 curves
CURVES = [(0.0, 200000.0, 0.3, 1.0e-5, 300.0, 320.0),
          (400.0, 150000.0, 0.32, 1.4e-5, 200.0, 230.0),
          (800.0, 100000.0, 0.35, 1.8e-5, 100.0, 110.0)]
EPS = (0.0, 0.01)


def _one_hex(et=W3D_L3DISOP):
    return _mesh(et, n=1, jitter=0.0)


def _uniform_strain_du(mesh, eps):
    """du of a homogeneous strain tensor eps (3x3 symmetric)."""
    return (mesh.coords @ np.asarray(eps).T).ravel()


@pytest.mark.parametrize("temp,want_e,want_nu", [
    (-50.0, 200000.0, 0.3),            # clamped below the first curve
    (900.0, 100000.0, 0.35),           # clamped above the last
    (200.0, 175000.0, 0.31),           # half-way between curves 1 and 2
    (600.0, 125000.0, 0.335)])
def test_curves_interpolate_e_and_nu(temp, want_e, want_nu):
    """Elastic response at a constant temperature (no change over the
    step): sigma = D(e(T), nu(T)) eps for a homogeneous strain."""
    mesh = _one_hex()
    m = _mm01(0.0, curves=CURVES, eps=EPS)
    arr = mesh.model(m, bbar=0)
    orc = O.OracleModel(arr)
    orc.set_temperatures(np.full(mesh.nonode, temp), None, None, None,
                         False)
    eps = np.array([[1e-4, 2e-5, 0], [2e-5, -3e-5, 1e-5], [0, 1e-5, 5e-5]])
    orc.recover(1, 0, np.zeros(3 * mesh.nonode),
                _uniform_strain_du(mesh, eps))
    sig = _stresses(orc)
    lam = want_e * want_nu / ((1 + want_nu) * (1 - 2 * want_nu))
    mu = want_e / (2 * (1 + want_nu))
    s = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    want = np.array([s[0, 0], s[1, 1], s[2, 2], s[0, 1], s[1, 2], s[0, 2]])
    assert np.allclose(sig, want, rtol=1e-12, atol=1e-9)


def test_curve_yield_stress_and_first_yield():
    """The yield stress is the curve's first point at the point's
    temperature: a uniaxial-strain path just below and just above the
    interpolated yield."""
    mesh = _one_hex()
    m = _mm01(0.0, curves=CURVES, eps=EPS)
    arr = mesh.model(m, bbar=0)
    temp = 200.0                       # yield 250, e 175000, nu 0.31
    e, nu, sy = 175000.0, 0.31, 250.0
    for fac, yields in ((0.99, False), (1.01, True)):
        orc = O.OracleModel(arr)
        orc.set_temperatures(np.full(mesh.nonode, temp), None, None, None,
                             False)
        # uniaxial strain e_xx: von Mises of the elastic stress is
        # 2 mu (1) e_xx |1 - 0| ... = e_xx E / (1 + nu) (deviatoric part)
        exx = fac * sy * (1 + nu) / e
        eps = np.diag([exx, 0.0, 0.0])
        orc.recover(1, 1, np.zeros(3 * mesh.nonode),
                    _uniform_strain_du(mesh, eps))
        h = orc.get_block_state(1, _abi.HIST_N1).reshape(-1, 11)
        state = np.ascontiguousarray(h[:, 3]).view(np.int32)[::2]
        assert np.all(state == (1 if yields else 3))


def test_temperature_dependent_hardening_return_is_consistent():
    """Kinematic hardening whose modulus changes with temperature over the
    step (lk = H'k(n+1)/H'k(n) != 1): after the return the relative stress
    dev(sigma) - alpha lies on the yield surface of radius sqrt(2) k n+1,
    on a single element (no neighbour's history word 11 involved)."""
    mesh = _one_hex()
    m = _mm01(1.0e-5, curves=CURVES, eps=EPS)
    m.beta = 0.0                        # purely kinematic
    arr = mesh.model(m, bbar=0)
    orc = O.OracleModel(arr)
    nn = mesh.nonode
    u = np.zeros(3 * nn)
    # step 1: at 0 degrees, strain into the plastic range
    orc.set_temperatures(np.zeros(nn), None, None, None, False)
    eps1 = np.diag([3.0e-3, -1.0e-3, -1.0e-3])
    du = _uniform_strain_du(mesh, eps1)
    orc.recover(1, 1, u, du)
    orc.commit_step()
    u = u + du
    # step 2: heat to 200 (H' from 2000 to 2500 at the curve's slope) and
    # strain further
    orc.set_temperatures(np.zeros(nn), np.full(nn, 200.0), None, None, True)
    eps2 = np.array([[1.0e-3, 4e-4, 0], [4e-4, -5e-4, 0], [0, 0, -5e-4]])
    du = _uniform_strain_du(mesh, eps2)
    orc.recover(2, 1, u, du)
    h = orc.get_block_state(1, _abi.HIST_N1).reshape(-1, 11)
    s = orc.get_block_state(1, _abi.STRESS_N1).reshape(-1, 9)[:, :6]
    hn = orc.get_block_state(1, _abi.HIST_N).reshape(-1, 11)
    assert np.all(h[:, 0] > 0.0)                   # yielding
    assert not np.allclose(h[:, 4], hn[:, 4])      # H' changed: lk != 1
    mean = s[:, :3].mean(axis=1)
    dev = s.copy()
    dev[:, :3] -= mean[:, None]
    rel = dev - h[:, 5:11]
    norm = np.sqrt((rel[:, :3] ** 2).sum(1) + 2 * (rel[:, 3:] ** 2).sum(1))
    assert np.allclose(norm, np.sqrt(2.0) * h[:, 1], rtol=1e-10)
