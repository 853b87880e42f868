"""Parity of the device library (libw3dblk through its C ABI) with the
oracle, on the same seeded inputs.

Bar: plain FP64 results within 1e-10 relative (norm-wise per array) of the
oracle; integer work (blocking, sparsity pattern) bit-exact; the FP32
tangent within 1e-5 relative of the FP64 one.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests._util import RTOL_FP64, assert_close, c1_model, smooth_field
from warp3d_amd import _abi
from warp3d_amd import mesh as M

pytestmark = pytest.mark.gpu

STATE = [(_abi.STRESS_N1, "stress"), (_abi.STRAIN_N1, "strain"),
         (_abi.HIST_N1, "history"), (_abi.CEP, "cep"),
         (_abi.EINFVEC, "einfvec"), (_abi.ELEM_VOL, "volume")]


@pytest.fixture(scope="module")
def W():
    from warp3d_amd import model
    return model


def compare_blocks(dev, orc, which_list=STATE, rtol=RTOL_FP64, tag=""):
    for b in range(1, orc.nelblk + 1):
        for which, name in which_list:
            a = dev.get_block_state(b, which)
            r = orc.get_block_state(b, which)
            if which in (_abi.HIST_N, _abi.HIST_N1):
                # word 4 is the integer state packed in a double
                a = a.reshape(-1, 11)
                r = r.reshape(-1, 11)
                assert np.array_equal(a[:, 3].view(np.int64),
                                      r[:, 3].view(np.int64)), \
                    f"{tag} blk {b} state word"
                a = np.delete(a, 3, axis=1)
                r = np.delete(r, 3, axis=1)
            assert_close(a, r, rtol, f"{tag} blk {b} {name}")


def run_sequence(dev, orc, m, steps=2, tag=""):
    """Step 1: iter 0 (linear-elastic pseudo-update), iters 1 and 2;
    step 2: iters 1 and 2.  Compare after every call."""
    u = np.zeros(3 * m.nonode)
    for step in range(1, steps + 1):
        du_final = smooth_field(m, 0.004, seed=step)
        iters = [0, 1, 2] if step == 1 else [1, 2]
        for it in iters:
            du = du_final * (0.5 if it == 0 else (0.9 if it == 1 else 1.0))
            fd, ed, pd, rcd = dev.recover(step, it, u, du)
            fo, eo, po, rco = orc.recover(step, it, u, du)
            assert rcd == rco == 0
            assert_close(fd, fo, RTOL_FP64, f"{tag} ifv s{step} i{it}")
            assert abs(ed - eo) <= 1e-12 * max(abs(eo), 1e-30)
            assert abs(pd - po) <= 1e-10 * max(abs(po), 1e-30) + 1e-300
            compare_blocks(dev, orc, tag=f"{tag} s{step} i{it}")
        dev.commit_step()
        orc.commit_step()
        u = u + du_final


def test_library_loads_and_reports(W):
    assert b"gfx950" in W.lib().w3d_version()


def test_blocking_bit_exact(W):
    m, arr = c1_model(10)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    assert np.array_equal(dev.elblks, orc.elblks)
    assert dev.nelblk == 8


def test_recover_parity_hex8_c1(W):
    m, arr = c1_model(10)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    run_sequence(dev, orc, m, steps=2, tag="hex8")


def test_recover_plastic_points_present(W):
    """The sequence drives points into plasticity (state word 1)."""
    m, arr = c1_model(6)
    dev = W.BlockModel(arr)
    u = np.zeros(3 * m.nonode)
    x = m.coords
    du = np.zeros_like(x)
    du[:, 2] = 0.004 * x[:, 2] ** 2  # axial strain 0 at z=0 .. 0.008 at z=1
    du = du.ravel()
    dev.recover(1, 1, u, du)
    states = []
    for b in range(1, dev.nelblk + 1):
        h = dev.get_block_state(b, _abi.HIST_N1).reshape(-1, 11)
        states.append(h[:, 3].view(np.int64) & 0xffffffff)
    s = np.concatenate(states)
    assert (s == 1).sum() > 0 and (s == 3).sum() > 0


@pytest.mark.parametrize("maker,order,n", [
    (M.box_hex20, 1, 3), (M.box_hex20, 8, 3), (M.box_tet10, 4, 3)])
def test_recover_parity_other_elements(W, maker, order, n):
    m = maker(n, jitter=0.05, seed=3)
    mat = _abi.mm01_material(70e3, 0.33, 300.0, hprime=70e3 / 50)
    arr = m.model(mat, int_order=order, auto_size=16)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    run_sequence(dev, orc, m, steps=2, tag=f"{m.etype}/{order}")


@pytest.mark.parametrize("precision", [_abi.W3D_FP64, _abi.W3D_FP32])
def test_tangent_parity_hex8(W, precision):
    m, arr = c1_model(6)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    u = np.zeros(3 * m.nonode)
    du = smooth_field(m, 0.004, seed=2)
    dev.recover(1, 1, u, du)
    orc.recover(1, 1, u, du)
    dev.tangent(1, 1, precision)
    orc.tangent(1, 1)
    rtol = RTOL_FP64 if precision == _abi.W3D_FP64 else 1e-6
    for b in range(1, orc.nelblk + 1):
        assert_close(dev.get_block_state(b, _abi.ESTIFF),
                     orc.get_block_state(b, _abi.ESTIFF), rtol, f"Ke blk {b}")


@pytest.mark.parametrize("maker,order", [(M.box_hex20, 1), (M.box_tet10, 4)])
def test_tangent_parity_other_elements(W, maker, order):
    m = maker(2, jitter=0.05, seed=4)
    mat = _abi.mm01_material(70e3, 0.33, 300.0, hprime=70e3 / 50)
    arr = m.model(mat, int_order=order, auto_size=8)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    u = np.zeros(3 * m.nonode)
    du = smooth_field(m, 0.006, seed=2)
    dev.recover(1, 1, u, du)
    orc.recover(1, 1, u, du)
    dev.tangent(1, 1, _abi.W3D_FP64)
    orc.tangent(1, 1)
    for b in range(1, orc.nelblk + 1):
        assert_close(dev.get_block_state(b, _abi.ESTIFF),
                     orc.get_block_state(b, _abi.ESTIFF), RTOL_FP64,
                     f"Ke blk {b}")


def test_pattern_bit_exact_and_assembly(W):
    m, arr = c1_model(6)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    cst, top, val = M.uniaxial_bc(m, 0.001)
    assert dev.pattern(cst) == orc.pattern(cst)
    for a, b in zip(dev.get_pattern(), orc.get_pattern()):
        assert np.array_equal(a, b)
    u = np.zeros(3 * m.nonode)
    du = smooth_field(m, 0.004, seed=5)
    dev.recover(1, 1, u, du)
    orc.recover(1, 1, u, du)
    dev.tangent(1, 1, _abi.W3D_FP64)
    orc.tangent(1, 1)
    kd, kc = dev.assemble_vss()
    kdo, kco = orc.assemble_vss()
    assert_close(kd, kdo, RTOL_FP64, "k_diag")
    assert_close(kc, kco, RTOL_FP64, "k_coeffs")


def test_per_block_interface(W):
    """do_nleps_block form: blocks recovered one call each give the same
    element forces as the all-block call."""
    m, arr = c1_model(6, auto_size=32)
    dev = W.BlockModel(arr)
    orc = O.OracleModel(arr)
    u = np.zeros(3 * m.nonode)
    du = smooth_field(m, 0.004, seed=6)
    for b in range(1, dev.nelblk + 1):
        dev.recover_blocks(b, 1, 1, 1, u, du)
    orc.recover_blocks(1, orc.nelblk, 1, 1, u, du)
    compare_blocks(dev, orc, tag="per-block")


def test_errors_are_reported(W):
    m, arr = c1_model(2)
    dev = W.BlockModel(arr)
    with pytest.raises(W.W3DError):
        dev.get_block_state(99, _abi.STRESS_N)
    with pytest.raises(W.W3DError):
        dev.tangent(1, 1, precision=7)
    with pytest.raises(W.W3DError):
        dev.assemble_vss()  # no pattern yet
