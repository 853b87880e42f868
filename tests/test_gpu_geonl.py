"""Parity of the device finite-strain (geonl) path with the oracle, through
the C ABI: mm01 blocks of 8-node bricks (with and without B-bar), 20-node
bricks (27 and 8 points) and 10-node tetrahedra under a large deformation
(about 15 % strain with rotation) over two load steps, plastic.

Bar: stresses, strains, history, [D], element forces, volumes, [R n+1],
structure forces and energies within 1e-10 relative of the oracle; the FP64
tangent (rotated [D], Q-bar, geometric stiffness) within 1e-10, the FP32
one within 1e-6; the assembled matrix within 1e-10.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests._util import RTOL_FP64, assert_close, rel_err, smooth_field
from warp3d_amd import _abi
from warp3d_amd import mesh as M

pytestmark = pytest.mark.gpu

STATE = [(_abi.STRESS_N1, "stress"), (_abi.STRAIN_N1, "strain"),
         (_abi.HIST_N1, "history"), (_abi.CEP, "cep"),
         (_abi.EINFVEC, "einfvec"), (_abi.ELEM_VOL, "volume"),
         (_abi.ROT_N1, "rotation")]

CASES = [("hex8", M.box_hex8, 1, 1), ("hex8_nobbar", M.box_hex8, 1, 0),
         ("hex20_27", M.box_hex20, 1, None), ("hex20_8", M.box_hex20, 8, None),
         ("tet10", M.box_tet10, 4, None)]


@pytest.fixture(scope="module")
def W():
    from warp3d_amd import model
    return model


def large_field(m, scale, seed):
    """Stretch, shear and a rotation about an inclined axis, plus noise."""
    rng = np.random.default_rng(seed)
    x = m.coords
    a = np.array([0.3, 1.0, 0.5]) / np.linalg.norm([0.3, 1.0, 0.5])
    k = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    ang = 2.0 * scale
    q = np.eye(3) + np.sin(ang) * k + (1 - np.cos(ang)) * k @ k
    f = np.array([[1 + scale, 0.5 * scale, 0], [0, 1 - 0.3 * scale, 0],
                  [0.2 * scale, 0, 1 - 0.3 * scale]])
    u = x @ (q @ f - np.eye(3)).T
    u += 0.05 * scale * rng.uniform(-1, 1, x.shape) * m.h.mean()
    return u.ravel()


def compare(dev, orc, tag, iter_):
    for b in range(1, orc.nelblk + 1):
        for which, name in STATE:
            if which == _abi.ROT_N1 and iter_ == 0:
                continue
            a = dev.get_block_state(b, which)
            r = orc.get_block_state(b, which)
            if which == _abi.HIST_N1:
                a = a.reshape(-1, 11)
                r = r.reshape(-1, 11)
                assert np.array_equal(a[:, 3].view(np.int64),
                                      r[:, 3].view(np.int64)), tag
                a = np.delete(a, 3, axis=1)
                r = np.delete(r, 3, axis=1)
            assert_close(a, r, RTOL_FP64, f"{tag} blk {b} {name}")


def make(m, order, bbar):
    mat = _abi.mm01_material(70e3, 0.33, 300.0, hprime=70e3 / 50)
    kw = {"int_order": order, "auto_size": 16, "geonl": True}
    if bbar is not None:
        kw["bbar"] = bbar
    return m.model(mat, **kw)


@pytest.mark.parametrize("name,maker,order,bbar", CASES)
def test_geonl_recover_and_tangent_parity(W, name, maker, order, bbar):
    m = maker(3, jitter=0.05, seed=21)
    arr = make(m, order, bbar)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    u = np.zeros(3 * m.nonode)
    for step in (1, 2):
        du_final = large_field(m, 0.08, seed=step)
        for it in ([0, 1, 2] if step == 1 else [1, 2]):
            du = du_final * (0.5 if it == 0 else (0.9 if it == 1 else 1.0))
            print(f"{name} step {step} iter {it}", flush=True)
            fd, ed, pd, rcd = dev.recover(step, it, u, du)
            fo, eo, po, rco = orc.recover(step, it, u, du)
            assert rcd == rco == 0, (rcd, rco)
            assert_close(fd, fo, RTOL_FP64, f"{name} ifv s{step} i{it}")
            assert abs(ed - eo) <= 1e-10 * max(abs(eo), 1e-30)
            assert abs(pd - po) <= 1e-10 * max(abs(po), 1e-30) + 1e-300
            compare(dev, orc, f"{name} s{step} i{it}", it)
            orc.tangent(step, it)
            for prec, rtol in ((_abi.W3D_FP64, RTOL_FP64),
                               (_abi.W3D_FP32, 1e-6)):
                dev.tangent(step, it, prec)
                for b in range(1, orc.nelblk + 1):
                    assert_close(dev.get_block_state(b, _abi.ESTIFF),
                                 orc.get_block_state(b, _abi.ESTIFF), rtol,
                                 f"{name} Ke {prec} blk {b} s{step} i{it}")
        assert np.abs(orc.get_block_state(1, _abi.HIST_N1)
                      .reshape(-1, 11)[:, 2]).max() > 0, "no plastic points"
        dev.commit_step()
        orc.commit_step()
        u = u + du_final


def test_geonl_assembly_and_rigid_rotation(W):
    m = M.box_hex8(4, jitter=0.05, seed=22)
    arr = make(m, 1, 1)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    cst = np.zeros(3 * m.nonode, np.int32)
    cst[:3] = 1
    assert dev.pattern(cst) == orc.pattern(cst)
    u = np.zeros(3 * m.nonode)
    du = large_field(m, 0.1, seed=3)
    dev.recover(1, 1, u, du)
    orc.recover(1, 1, u, du)
    dev.tangent(1, 1, _abi.W3D_FP64)
    orc.tangent(1, 1)
    kd, kc = dev.assemble_vss()
    kdo, kco = orc.assemble_vss()
    assert_close(kd, kdo, RTOL_FP64, "k_diag")
    assert_close(kc, kco, RTOL_FP64, "k_coeffs")
    # a rigid rotation of the whole body strains nothing
    dev2 = W.BlockModel(arr)
    th = 0.9
    q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0],
                  [0, 0, 1]])
    dr = (m.coords @ (q - np.eye(3)).T).ravel()
    f, _, _, rc = dev2.recover(1, 1, u, dr)
    assert rc == 0
    for b in range(1, dev2.nelblk + 1):
        assert np.abs(dev2.get_block_state(b, _abi.STRAIN_N1)).max() < 1e-13
    assert np.abs(f).max() < 1e-6


def test_geonl_with_asymmetric_assembly_is_rejected(W):
    """The unsymmetrised crystal tangent is assembled only for small
    strains: finite-strain blocks refuse asymmetric assembly (mm10 itself
    runs with finite strains, test_gpu_mm10_geonl)."""
    m = M.box_hex8(1)
    mat = _abi.mm10_material(tau_y=50.0, tau_v=100.0, theta_0=250.0,
                             e=100e3, nu=0.3, mu=80e3, elastic="cubic")
    arr = m.model(mat, geonl=True, angles=np.zeros((1, 3)))
    dev = W.BlockModel(arr)
    with pytest.raises(W.W3DError):
        dev.set_asymmetric(True)
    dev.close()


def test_small_strain_finite_kinematics_within_their_conditioning(W):
    """At strains of a few 1e-3 the finite-strain kinematics are
    ill-conditioned: gtmat1's polar decomposition takes U from the
    eigenvalues of C = F^T F by the characteristic cubic, whose
    discriminant m^2 - 3 c1 is a difference of O(1) invariants that is
    O(strain^2) itself, so one ulp in the displacements moves the
    oracle's own stresses by about 1e-9.  The spread is measured here:
    the oracle's results under three seeded perturbations of every
    displacement word by up to two ulps.  The device result (which rounds
    differently: fused multiply-adds, its own libm) must lie within 20
    times that spread of the oracle's (forces, stresses and the internal
    energy), and within 1e-10 wherever the spread is below that."""
    m = M.box_hex8(3, jitter=0.15, seed=2)
    arr = make(m, 1, 1)
    u = np.zeros(3 * m.nonode)
    du = smooth_field(m, 0.004, seed=1)

    def run(model, d):
        f, e, _, rc = model.recover(1, 1, u, d)
        assert rc == 0
        return f, e, np.concatenate(
            [model.get_block_state(b, _abi.STRESS_N1)
             for b in range(1, model.nelblk + 1)])

    fo, eo, so = run(O.OracleModel(arr), du)
    rng = np.random.default_rng(5)
    spread = np.zeros(3)
    for _ in range(3):
        d = du * (1.0 + rng.integers(-2, 3, du.size) * 2.0 ** -52)
        fp, ep, sp = run(O.OracleModel(arr), d)
        spread = np.maximum(spread, [rel_err(fp, fo), rel_err(sp, so),
                                     abs(ep - eo) / abs(eo)])
    assert spread[1] > 1e-10          # the conditioning is real
    tol = np.maximum(20 * spread, 1e-10)
    fd, ed, sd = run(W.BlockModel(arr), du)
    assert rel_err(fd, fo) <= tol[0]
    assert rel_err(sd, so) <= tol[1]
    assert abs(ed - eo) / abs(eo) <= tol[2]
