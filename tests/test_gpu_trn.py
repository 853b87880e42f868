"""Skewed constraints on the device against the oracle (whose trnvec and
trnmtx are pinned in test_oracle_trn.py): nodes with constraint systems
u_c = T u_g in hex8 (B-bar), hex20 and tet10 models, small and finite
strains, over two steps.  Bar: forces, stresses, history, element forces
and the FP64 block tangents and VSS matrix within 1e-10 relative of the
oracle; the FP32 tangent's matrix within 2e-6."""
import numpy as np
import pytest

from oracle import oracle as O
from tests._util import RTOL_FP64, assert_close, smooth_field
from tests.test_gpu_geonl import large_field
from tests.test_oracle_trn import CASES, _rotations, _tg
from warp3d_amd import _abi
from warp3d_amd import mesh as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,geonl", CASES)
def test_skewed_constraints_match_the_oracle(kind, geonl):
    from warp3d_amd.model import BlockModel
    mk = {"hex8-bbar": (M.box_hex8, 1), "hex20": (M.box_hex20, 8),
          "tet10": (M.box_tet10, 4)}[kind]
    m = mk[0](3, jitter=0.1, seed=3)
    mat = _abi.mm01_material(200e3, 0.3, 400.0, hprime=2000.0)
    rng = np.random.default_rng(4)
    nodes = rng.choice(np.arange(3, m.nonode + 1), size=m.nonode // 3,
                       replace=False)
    base = m.model(mat, int_order=mk[1], auto_size=7, geonl=geonl)
    trn = _rotations(nodes, 5)
    arr = _abi.ModelArrays(
        base.coords, base.incid, base.incmap, base.etype, base.int_order,
        base.matnum, base.bbar, base.materials, auto_size=7,
        geonl=base.geonl, trnmat=trn)
    tg = _tg(m.nonode, trn)
    dev, orc = BlockModel(arr), O.OracleModel(arr)
    cst = np.zeros(3 * m.nonode, np.int32)
    cst[:3] = 1
    cst[5] = 1
    dev.pattern(cst)
    orc.pattern(cst)
    u = np.zeros(3 * m.nonode)
    for step in (1, 2):
        # finite strains of a few % with rotation (test_gpu_geonl's field):
        # at a few 1e-3 the polar decomposition's conditioning, not the
        # device, limits agreement
        # (the finite-strain field is given globally and handed over in
        # the nodes' constraint-compatible systems, u_c = T u_g)
        du = (tg @ large_field(m, 0.08, seed=step) if geonl
              else smooth_field(m, 0.004, seed=step))
        for it in (0, 1):
            tag = f"{kind} geonl={geonl} s{step} i{it}"
            fd, ed, _, rcd = dev.recover(step, it, u, du)
            fo, eo, _, rco = orc.recover(step, it, u, du)
            assert rcd == rco == 0
            assert_close(fd, fo, RTOL_FP64, f"{tag} ifv")
            assert abs(ed - eo) <= 1e-10 * max(abs(eo), 1e-30)
            for b in range(1, orc.nelblk + 1):
                for which in (_abi.STRESS_N1, _abi.STRAIN_N1, _abi.EINFVEC):
                    assert_close(dev.get_block_state(b, which),
                                 orc.get_block_state(b, which), RTOL_FP64,
                                 f"{tag} blk {b} {which}")
            dev.tangent(step, it, _abi.W3D_FP64)
            orc.tangent(step, it)
            for b in range(1, orc.nelblk + 1):
                assert_close(dev.get_block_state(b, _abi.ESTIFF),
                             orc.get_block_state(b, _abi.ESTIFF), RTOL_FP64,
                             f"{tag} Ke blk {b}")
            kd, kc = dev.assemble_vss()
            kdo, kco = orc.assemble_vss()
            assert_close(kd, kdo, RTOL_FP64, f"{tag} k_diag")
            assert_close(kc, kco, RTOL_FP64, f"{tag} k_coeffs")
            dev.tangent(step, it, _abi.W3D_FP32)
            kd, kc = dev.assemble_vss()
            assert_close(kc, kco, 2e-6, f"{tag} fp32 k_coeffs")
        dev.commit_step()
        orc.commit_step()
        u = u + du
    dev.close()
