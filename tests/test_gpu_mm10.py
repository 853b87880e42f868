"""Parity of the device mm10 (crystal plasticity, Voce) path with the
oracle, through the C ABI, on seeded meshes with seeded SO(3)-uniform
orientations.

Bar: stresses, strains, element forces, [D] and every history word within
1e-10 relative of the oracle (per history word, relative to that word's
largest magnitude over the block, at least 1 for the grad Fe slopes of
rotation components; Euler angles compared modulo 360 degrees); the
step-cut status identical.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests._util import RTOL_FP64, assert_close, smooth_field
from warp3d_amd import _abi
from warp3d_amd import mesh as M

pytestmark = pytest.mark.gpu

PLAIN = [(_abi.STRESS_N1, "stress"), (_abi.STRAIN_N1, "strain"),
         (_abi.CEP, "cep"), (_abi.EINFVEC, "einfvec"),
         (_abi.ELEM_VOL, "volume")]


@pytest.fixture(scope="module")
def W():
    from warp3d_amd import model
    return model


def cubic_fcc(**kw):
    base = dict(tau_y=50.0, tau_v=100.0, theta_0=250.0, e=100e3, nu=0.3,
                mu=80e3, elastic="cubic")
    base.update(kw)
    return _abi.mm10_material(**base)


def compare_history(a, r, ncry, tag):
    H = 87 + 71 * ncry
    a = a.reshape(-1, H)
    r = r.reshape(-1, H)
    ang = np.zeros(H, bool)
    for c in range(ncry):
        ang[87 + 71 * c + 6:87 + 71 * c + 9] = True
    d = np.abs(a - r)
    d[:, ang] = np.minimum(d[:, ang], 360.0 - d[:, ang])
    scale = np.abs(r).max(axis=0)
    # grad Fe (words 37-63) are slopes of a linear fit of rotation
    # components (magnitude <= 1): relative to the fitted data, so that
    # round-off slopes of constant data (about 1e-17) compare as zero
    scale[36:63] = np.maximum(scale[36:63], 1.0)
    bad = d > RTOL_FP64 * scale + 1e-300
    assert not bad.any(), (f"{tag}: history words "
                           f"{sorted(set(np.nonzero(bad)[1].tolist()))} "
                           f"rel {(d / np.maximum(scale, 1e-300)).max():.3e}")


def compare(dev, orc, ncry, tag):
    for b in range(1, orc.nelblk + 1):
        for which, name in PLAIN:
            assert_close(dev.get_block_state(b, which),
                         orc.get_block_state(b, which), RTOL_FP64,
                         f"{tag} blk {b} {name}")
        compare_history(dev.get_block_state(b, _abi.HIST_N1),
                        orc.get_block_state(b, _abi.HIST_N1), ncry,
                        f"{tag} blk {b}")


def run(dev, orc, m, ncry=1, steps=2, scale=0.0005, tangent=True, tag=""):
    u = np.zeros(3 * m.nonode)
    for step in range(1, steps + 1):
        du_final = smooth_field(m, scale, seed=step)
        iters = [0, 1, 2] if step == 1 else [0, 1, 2]
        for it in iters:
            du = du_final * (0.5 if it == 0 else (0.9 if it == 1 else 1.0))
            fd, ed, pd, rcd = dev.recover(step, it, u, du)
            fo, eo, po, rco = orc.recover(step, it, u, du)
            assert rcd == rco == 0, (rcd, rco)
            assert_close(fd, fo, RTOL_FP64, f"{tag} ifv s{step} i{it}")
            assert abs(ed - eo) <= 1e-10 * max(abs(eo), 1e-30)
            assert abs(pd - po) <= 1e-10 * max(abs(po), 1e-30) + 1e-300
            compare(dev, orc, ncry, f"{tag} s{step} i{it}")
            if tangent:
                dev.tangent(step, it, _abi.W3D_FP64)
                orc.tangent(step, it)
                for b in range(1, orc.nelblk + 1):
                    assert_close(dev.get_block_state(b, _abi.ESTIFF),
                                 orc.get_block_state(b, _abi.ESTIFF),
                                 RTOL_FP64, f"{tag} Ke blk {b} s{step}")
        dev.commit_step()
        orc.commit_step()
        u = u + du_final


@pytest.mark.parametrize("maker,order,n", [
    (M.box_hex8, 1, 3), (M.box_hex20, 8, 2), (M.box_hex20, 1, 2),
    (M.box_tet10, 4, 2)])
def test_mm10_parity_elements(W, maker, order, n):
    """Recovery, history, element forces and the FP64 tangent of mm10
    blocks (ragged blocks of 10) for every element type."""
    m = maker(n, jitter=0.1, seed=3)
    ang = M.random_orientations(m.noelem, seed=5)
    arr = m.model(cubic_fcc(), int_order=order, auto_size=10, angles=ang)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    run(dev, orc, m, tag=f"{m.etype}/{order}")


def test_mm10_bcc_isotropic_and_time_increment(W):
    m = M.box_hex8(3, jitter=0.1, seed=4)
    ang = M.random_orientations(m.noelem, seed=6)
    mat = _abi.mm10_material(tau_y=60.0, tau_v=80.0, theta_0=300.0,
                             voce_m=2.0, harden_n=15.0, slip_type="bcc",
                             e=200e3, nu=0.3)
    arr = m.model(mat, auto_size=7, angles=ang)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    dev.set_time_increment(0.25)
    orc.set_time_increment(0.25)
    run(dev, orc, m, tag="bcc")


def test_mm10_two_crystals_per_point(W):
    m = M.box_hex8(2, jitter=0.1, seed=5)
    ang = M.random_orientations(m.noelem, seed=7, per_element=2)
    arr = m.model(cubic_fcc(ncrystals=2), auto_size=5, angles=ang,
                  angle_stride=2)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    assert dev.block_info(1)["hist"] == 87 + 2 * 71
    run(dev, orc, m, ncry=2, tag="2 crystals")


def test_mm10_fp32_tangent(W):
    m = M.box_hex8(3, jitter=0.1, seed=3)
    ang = M.random_orientations(m.noelem, seed=5)
    arr = m.model(cubic_fcc(), auto_size=10, angles=ang)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    run(dev, orc, m, steps=1, tangent=False)
    dev.tangent(1, 2, _abi.W3D_FP32)
    orc.tangent(1, 2)
    for b in range(1, orc.nelblk + 1):
        assert_close(dev.get_block_state(b, _abi.ESTIFF),
                     orc.get_block_state(b, _abi.ESTIFF), 2e-6,
                     f"fp32 Ke blk {b}")


def test_mm10_cut_step_status(W):
    """A load increment the local solver cannot take (after its four
    sub-step cuts) asks for a step cut on both paths."""
    m = M.box_hex8(3, jitter=0.1, seed=3)
    ang = M.random_orientations(m.noelem, seed=5)
    arr = m.model(cubic_fcc(), auto_size=10, angles=ang)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    u = np.zeros(3 * m.nonode)
    du = smooth_field(m, 0.02, seed=1)
    _, _, _, rco = orc.recover(2, 1, u, du)
    _, _, _, rcd = dev.recover(2, 1, u, du)
    assert rco == rcd == _abi.W3D_CUT_STEP
