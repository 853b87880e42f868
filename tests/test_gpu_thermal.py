"""Parity of the device library's temperature loading with the oracle:
point temperatures from nodal and element values, the thermal strain
increment (mm01 and mm10), and mm01 with segmental curve sets (one
temperature-independent curve; temperature-dependent e, nu, alpha and
hardening, which drives mm01_general's lk != 1 branch).

Bar: stresses, strains, history, [D], element forces and the tangent
within 1e-10 relative of the oracle over a load history of two steps
(iterations 0-2), with non-uniform temperature changes in both steps.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests._util import RTOL_FP64, assert_close, smooth_field
from tests.test_gpu_geonl import large_field
from tests.test_gpu_parity import compare_blocks
from warp3d_amd import _abi
from warp3d_amd import mesh as M
from warp3d_amd._abi import W3D_L3DISOP, W3D_Q3DISOP, W3D_TET10

pytestmark = pytest.mark.gpu

CURVES = [(0.0, 200000.0, 0.3, 1.0e-5, 300.0, 320.0),
          (150.0, 180000.0, 0.31, 1.2e-5, 260.0, 290.0),
          (400.0, 150000.0, 0.32, 1.4e-5, 200.0, 230.0)]


@pytest.fixture(scope="module")
def W():
    from warp3d_amd import model
    return model


def _material(kind):
    m = _abi.mm01_material(200e3, 0.3, 300.0, hprime=2000.0, beta=0.4)
    _abi.set_thermal(m, [1.1e-5, 1.0e-5, 1.3e-5, 0.2e-5, 0.0, 0.1e-5],
                     alpha_zero=20.0)
    if kind == "curve":
        _abi.set_curves(m, (0.0, 0.02), [(300.0, 340.0)])
    elif kind == "temp_curves":
        _abi.set_curves(m, (0.0, 0.02), CURVES)
    return m


def _temperatures(m, step, seed, scale=1.0):
    """Totals at n (nodes + elements) and the step's changes: a field
    varying in space plus noise, element parts on top."""
    rng = np.random.default_rng(seed)
    x = m.coords
    tn = 20.0 + scale * 60.0 * (step - 1) * (1.0 + x[:, 0])
    dtn = scale * (60.0 * (1.0 + x[:, 0]) +
                   10.0 * rng.uniform(-1, 1, m.nonode))
    te = rng.uniform(-2, 2, m.noelem)
    dte = scale * rng.uniform(-5, 5, m.noelem)
    return tn, dtn, te, dte


MESHES = {
    "hex8_bbar": lambda: (M.box_hex8(3, jitter=0.15, seed=2), {}),
    # finite strains of a few % with rotation (test_gpu_geonl's field): at
    # strains of a few 1e-3 the polar decomposition's conditioning, not the
    # device, limits agreement (test_small_strain_finite_kinematics_...)
    "hex8_geonl": lambda: (M.box_hex8(3, jitter=0.15, seed=2),
                           {"geonl": True}),
    "hex20_27": lambda: (M.box_hex20(2, jitter=0.1, seed=3),
                         {"int_order": 1}),
    "tet10": lambda: (M.box_tet10(2, jitter=0.1, seed=4), {}),
}


@pytest.mark.parametrize("kind", ["plain", "curve", "temp_curves"])
@pytest.mark.parametrize("mesh", list(MESHES))
def test_thermal_mm01_parity(W, mesh, kind):
    m, kw = MESHES[mesh]()
    arr = m.model(_material(kind), auto_size=16, **kw)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    u = np.zeros(3 * m.nonode)
    for step in (1, 2):
        t = _temperatures(m, step, seed=10 + step)
        dev.set_temperatures(*t, incremental=True)
        orc.set_temperatures(*t, incremental=True)
        du_final = (large_field(m, 0.08, seed=step) if kw.get("geonl")
                    else smooth_field(m, 0.003, seed=step))
        for it in ([0, 1, 2] if step == 1 else [1, 2]):
            du = du_final * (0.5 if it == 0 else (0.9 if it == 1 else 1.0))
            fd, ed, pd, rcd = dev.recover(step, it, u, du)
            fo, eo, po, rco = orc.recover(step, it, u, du)
            assert rcd == rco == 0
            tag = f"{mesh} {kind} s{step} i{it}"
            assert_close(fd, fo, RTOL_FP64, f"{tag} ifv")
            compare_blocks(dev, orc, tag=tag)
            if it > 0:
                dev.tangent(step, it, _abi.W3D_FP64)
                orc.tangent(step, it)
                for b in range(1, orc.nelblk + 1):
                    assert_close(dev.get_block_state(b, _abi.ESTIFF),
                                 orc.get_block_state(b, _abi.ESTIFF),
                                 RTOL_FP64, f"{tag} blk {b} Ke")
        dev.commit_step()
        orc.commit_step()
        u = u + du_final
    dev.close()


def test_thermal_not_incremental_parity(W):
    """Temperatures without a change in the step (incremental off): no
    thermal strain, but temperature-dependent curves still read the point
    temperatures."""
    m = M.box_hex8(3, jitter=0.15, seed=2)
    arr = m.model(_material("temp_curves"), auto_size=16)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    t = _temperatures(m, 2, seed=5)
    dev.set_temperatures(t[0], None, t[2], None, incremental=False)
    orc.set_temperatures(t[0], None, t[2], None, incremental=False)
    u = np.zeros(3 * m.nonode)
    du = smooth_field(m, 0.004, seed=1)
    for it in (0, 1):
        fd, *_ = dev.recover(1, it, u, du)
        fo, *_ = orc.recover(1, it, u, du)
        assert_close(fd, fo, RTOL_FP64, f"i{it} ifv")
        compare_blocks(dev, orc, tag=f"i{it}")
    dev.close()


def test_thermal_mm10_parity(W):
    """mm10 sees the displacement increment less the thermal strain
    (drive_10_update); every history word against the oracle, on a
    history whose local solves all converge (no step cut)."""
    m = M.box_hex8(2, jitter=0.1, seed=6)
    mat = _abi.mm10_material(tau_y=50.0, tau_v=100.0, theta_0=250.0,
                             elastic="cubic", e=66700.0, nu=0.42, mu=75000.0)
    _abi.set_thermal(mat, 1.0e-5)
    ang = M.random_orientations(m.noelem, seed=3)
    arr = m.model(mat, auto_size=4, angles=ang)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    u = np.zeros(3 * m.nonode)
    for step in (1, 2):
        t = _temperatures(m, step, seed=20 + step, scale=0.2)
        dev.set_temperatures(*t, incremental=True)
        orc.set_temperatures(*t, incremental=True)
        du_final = smooth_field(m, 0.001, seed=step)
        for it in ([0, 1, 2] if step == 1 else [1, 2]):
            du = du_final * (0.5 if it == 0 else 1.0)
            fd, ed, pd, rcd = dev.recover(step, it, u, du)
            fo, eo, po, rco = orc.recover(step, it, u, du)
            assert rcd == rco == 0
            if it > 0:
                assert po > 0.0        # the crystals slip
            tag = f"mm10 s{step} i{it}"
            assert_close(fd, fo, RTOL_FP64, f"{tag} ifv")
            for b in range(1, orc.nelblk + 1):
                for which in (_abi.STRESS_N1, _abi.STRAIN_N1, _abi.HIST_N1):
                    assert_close(dev.get_block_state(b, which),
                                 orc.get_block_state(b, which), RTOL_FP64,
                                 f"{tag} blk {b} state {which}")
        dev.commit_step()
        orc.commit_step()
        u = u + du_final
    dev.close()
