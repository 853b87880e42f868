"""Restart records of the device model (warp3d_amd.restart): a state
written from the device reads back into a fresh device model bitwise, and
into the oracle, from which both continue within the parity bar (1e-10
relative on the forces)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests._util import smooth_field
from warp3d_amd import _abi
from warp3d_amd import mesh as M
from warp3d_amd import restart as R
from warp3d_amd.serve import RVE_MATERIAL

pytestmark = pytest.mark.gpu


def test_device_state_round_trip_and_into_the_oracle(tmp_path):
    from warp3d_amd.model import BlockModel
    m = M.box_hex8(3, jitter=0.05, seed=12)
    # the Voce constants of the reference's verification case: three steps
    # of this field yield without asking for a cut (a single recovery has
    # no Newton loop around it to cut)
    cp = _abi.mm10_material(**RVE_MATERIAL)
    arr = m.model(cp, auto_size=10,
                  angles=M.random_orientations(m.noelem, seed=13))
    cst = np.zeros(3 * m.nonode, np.int32)
    cst[:3] = 1
    a = BlockModel(arr)
    a.pattern(cst)
    u = np.zeros(3 * m.nonode)
    for s in (1, 2):
        du = smooth_field(m, 0.004, seed=s)
        rc = a.recover(s, 1, u, du)[3]
        assert rc == 0, f"step {s}: the crystal update asks for a cut"
        a.commit_step()
        u = u + du
    path = str(tmp_path / "dev.db")
    R.save_element_state(a, path)
    b = BlockModel(arr)
    b.pattern(cst)
    R.load_element_state(b, path)
    o = O.OracleModel(arr)
    o.pattern(cst)
    R.load_element_state(o, path)
    for blk in range(1, a.nelblk + 1):
        for w in (R.HIST_N, R.STRAIN_N, R.STRESS_N):
            # compared as bit patterns: words that hold NaN (the parked
            # corrector Jacobian of a crystal whose last local solve was
            # abandoned) must read back as the same NaN
            x = a.get_block_state(blk, w).view(np.uint64)
            y = b.get_block_state(blk, w).view(np.uint64)
            bad = np.flatnonzero(x != y)
            assert bad.size == 0, (blk, w, bad[:8], x.size)
            assert np.array_equal(x, o.get_block_state(blk, w).view(np.uint64))
    du = smooth_field(m, 0.004, seed=3)
    fa = a.recover(3, 1, u, du)[0]
    fb = b.recover(3, 1, u, du)[0]
    fo = o.recover(3, 1, u, du)[0]
    assert np.array_equal(fa, fb)
    assert np.abs(fa - fo).max() <= 1e-10 * np.abs(fo).max()
