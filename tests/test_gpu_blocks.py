"""The per-block interface of the library, called as the reference calls
do_nleps_block / do_nlek_block: from many host threads at once, one block
per call (drive_eps_sig_internal_forces.f:160-180, tanstf.f:80-87), after
one w3d_set_displacements, then w3d_assemble_ifv (addifv).

Bars against the oracle: every block's state, element forces, the block
energy and plastic work, the assembled forces, addifv's sum of |element
force| within 1e-10 relative; the block tangents (FP64 from threads, FP32
over a sub-range) within 1e-10 / 2e-6; the cut status of a block the
crystal update cannot take is reported by that block alone.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O
from tests._util import RTOL_FP64, assert_close, c1_model, smooth_field
from tests.test_gpu_mm10 import compare_history, cubic_fcc
from warp3d_amd import _abi
from warp3d_amd import mesh as M

pytestmark = pytest.mark.gpu

THREADS = 8


def _mm10_model():
    mm = M.MixedMesh([(M.box_hex8(4, jitter=0.1, seed=41), 1, 1),
                      (M.box_hex20(2, jitter=0.05, seed=42), 1, 0),
                      (M.box_tet10(2, jitter=0.05, seed=43), 4, 0)],
                     glue=True)
    mm.h = np.full(3, 0.25)
    ang = M.random_orientations(mm.noelem, seed=44)
    return mm, mm.model(cubic_fcc(), auto_size=9, angles=ang)


def _pass(dev, orc, step, it, u, du, pool, ncry=None, tag=""):
    """One recovery pass through the per-block calls, in a thread pool, in
    a shuffled block order; compared with the oracle's whole pass."""
    dev.set_displacements(u, du)
    order = np.random.default_rng(step * 10 + it).permutation(
        np.arange(1, dev.nelblk + 1))
    res = dict(zip(order, pool.map(
        lambda b: dev.recover_block(int(b), step, it), order)))
    fo, eo, po, rco = orc.recover(step, it, u, du)
    assert rco == 0
    for b in range(1, orc.nelblk + 1):
        rc, en, pw, cut = res[b]
        assert rc == 0 and cut == 0
        oe, op, of = orc.block_sums(b)
        assert abs(en - oe) <= 1e-10 * max(abs(oe), 1e-30), (tag, b)
        assert abs(pw - op) <= 1e-10 * max(abs(op), 1e-30) + 1e-300
        for which in (_abi.STRESS_N1, _abi.STRAIN_N1, _abi.EINFVEC):
            assert_close(dev.get_block_state(b, which),
                         orc.get_block_state(b, which), RTOL_FP64,
                         f"{tag} blk {b} state {which}")
        hd = dev.get_block_state(b, _abi.HIST_N1)
        ho = orc.get_block_state(b, _abi.HIST_N1)
        if ncry:
            compare_history(hd, ho, ncry, f"{tag} blk {b}")
        else:
            assert_close(hd, ho, RTOL_FP64, f"{tag} blk {b} history")
    fd, sd, nd = dev.assemble_ifv()
    assert_close(fd, fo, RTOL_FP64, f"{tag} ifv")
    so = sum(orc.block_sums(b)[2] for b in range(1, orc.nelblk + 1))
    if it > 0:
        assert abs(sd - so) <= 1e-10 * so
    assert nd == sum(orc.block_info(b)["span"] * orc.block_info(b)["totdof"]
                     for b in range(1, orc.nelblk + 1))
    # the whole-pass energy is the sum of the blocks' in block order
    en = 0.0
    for b in range(1, orc.nelblk + 1):
        en = en + res[b][1]
    assert abs(en - eo) <= 1e-10 * max(abs(eo), 1e-30)
    return res


def _tangents(dev, orc, step, it, pool, tag):
    list(pool.map(lambda b: dev.tangent_block(int(b), step, it,
                                              _abi.W3D_FP64),
                  range(1, dev.nelblk + 1)))
    orc.tangent(step, it)
    for b in range(1, orc.nelblk + 1):
        assert_close(dev.get_block_state(b, _abi.ESTIFF),
                     orc.get_block_state(b, _abi.ESTIFF), RTOL_FP64,
                     f"{tag} Ke blk {b}")


@pytest.mark.parametrize("kind", ["mm01", "mm10"])
def test_concurrent_block_calls_match_the_oracle(kind):
    from warp3d_amd.model import BlockModel
    if kind == "mm01":
        m, arr = c1_model(8, auto_size=16)
        ncry, amp = None, 0.004
    else:
        m, arr = _mm10_model()
        ncry, amp = 1, 0.0005
    dev, orc = BlockModel(arr), O.OracleModel(arr)
    assert dev.nelblk >= 12
    u = np.zeros(3 * m.nonode)
    with ThreadPoolExecutor(THREADS) as pool:
        for step in (1, 2):
            du_final = smooth_field(m, amp, seed=step)
            for it in (0, 1, 2):
                du = du_final * (0.5 if it == 0 else (0.9 if it == 1
                                                      else 1.0))
                tag = f"{kind} s{step} i{it}"
                _pass(dev, orc, step, it, u, du, pool, ncry, tag)
                _tangents(dev, orc, step, it, pool, tag)
                print(tag, "match", flush=True)
            dev.commit_step()
            orc.commit_step()
            u = u + du_final
    dev.close()


def test_block_ranges_and_fp32_subrange():
    """The range forms: recovery with and without new displacements, and
    the FP32 tangent of a sub-range of blocks next to FP64 blocks' words
    left from before (only the sub-range is compared)."""
    from warp3d_amd.model import BlockModel
    m, arr = c1_model(6, auto_size=32)
    dev, orc = BlockModel(arr), O.OracleModel(arr)
    u = np.zeros(3 * m.nonode)
    du = smooth_field(m, 0.004, seed=6)
    assert dev.recover_blocks(1, 3, 1, 1, u, du) == 0
    dev.set_displacements(u, du)
    lib = __import__("warp3d_amd.model", fromlist=["lib"]).lib()
    assert lib.w3d_recover_blocks(dev.h, 4, dev.nelblk - 3, 1, 1, 0, None,
                                  None) == 0
    orc.recover_blocks(1, orc.nelblk, 1, 1, u, du)
    for b in range(1, orc.nelblk + 1):
        for which in (_abi.STRESS_N1, _abi.EINFVEC, _abi.CEP):
            assert_close(dev.get_block_state(b, which),
                         orc.get_block_state(b, which), RTOL_FP64,
                         f"range blk {b} {which}")
    dev.tangent_blocks(2, 3, 1, 1, _abi.W3D_FP32)
    orc.tangent_blocks(2, 3, 1, 1)
    for b in (2, 3, 4):
        assert_close(dev.get_block_state(b, _abi.ESTIFF),
                     orc.get_block_state(b, _abi.ESTIFF), 2e-6,
                     f"fp32 Ke blk {b}")
    dev.close()


def test_a_cut_is_reported_by_its_block():
    """Only the blocks whose crystals cannot take the increment report the
    cut; the oracle's blocks agree one by one."""
    from warp3d_amd.model import BlockModel
    m = M.box_hex8(3, jitter=0.1, seed=3)
    ang = M.random_orientations(m.noelem, seed=5)
    arr = m.model(cubic_fcc(), auto_size=3, angles=ang)
    dev, orc = BlockModel(arr), O.OracleModel(arr)
    u = np.zeros(3 * m.nonode)
    # at this strain the crystals of four of the nine blocks need a cut
    du = smooth_field(m, 0.0017, seed=1)
    dev.set_displacements(u, du)
    with ThreadPoolExecutor(THREADS) as pool:
        res = list(pool.map(lambda b: dev.recover_block(b, 1, 1),
                            range(1, dev.nelblk + 1)))
    cuts_dev = [r[3] for r in res]
    cuts_orc = [orc.recover_blocks(b, 1, 1, 1, u, du) == _abi.W3D_CUT_STEP
                for b in range(1, orc.nelblk + 1)]
    assert cuts_dev == [int(c) for c in cuts_orc]
    assert any(cuts_dev) and not all(cuts_dev)
    for r in res:
        assert r[0] == (_abi.W3D_CUT_STEP if r[3] else 0)
    dev.close()
