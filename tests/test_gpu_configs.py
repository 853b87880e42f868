"""Parity of the device library with the oracle on the benchmark
configurations themselves:

* hex8 60x60x60 (216,000 elements, 1,728,000 points, mm01) and hex20
  40x40x40 (64,000 elements, 27-point rule, 1,728,000 points, mm01) at full
  size: a committed load step 1, then iterations 0 and 1 of step 2, every
  block's state, the structure forces, the energies, the FP64 tangent
  assembled into the VSS matrix (every block's Ke enters it) and the FP32
  tangent's matrix;
* a mixed model of hex8 (B-bar), hex20 (27 points) and tet10 under mm10
  fcc Voce in blocks of seeded spans 32-256 (at least one above 128),
  whose parts share nodes, over two steps and iterations 0-2: every
  block's state and history word, the forces, the FP64 tangent and the
  VSS matrix;
* a served RVE (n = 8, 64 grains, 10 steps to 2 %, FP64 tangent) against
  the same Newton algorithm run on the oracle with a direct solve.

Bars: 1e-10 relative (norm-wise per array; per history word for mm10),
identical integer work, identical Newton iterations and cuts, and
displacements within 1e-8 relative.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests._util import RTOL_FP64, assert_close, smooth_field
from tests.test_gpu_mm10 import compare_history
from warp3d_amd import _abi
from warp3d_amd import mesh as M

pytestmark = pytest.mark.gpu

MM01_STATE = [(_abi.STRESS_N1, "stress"), (_abi.STRAIN_N1, "strain"),
              (_abi.HIST_N1, "history"), (_abi.CEP, "cep"),
              (_abi.EINFVEC, "einfvec"), (_abi.ELEM_VOL, "volume")]


def _state_mm01(dev, orc, tag):
    for b in range(1, orc.nelblk + 1):
        for which, name in MM01_STATE:
            a = dev.get_block_state(b, which)
            r = orc.get_block_state(b, which)
            if which == _abi.HIST_N1:
                a = a.reshape(-1, 11)
                r = r.reshape(-1, 11)
                assert np.array_equal(a[:, 3].view(np.int64),
                                      r[:, 3].view(np.int64)), \
                    f"{tag} blk {b} state word"
                a = np.delete(a, 3, axis=1)
                r = np.delete(r, 3, axis=1)
            assert_close(a, r, RTOL_FP64, f"{tag} blk {b} {name}")


def _vss(dev, orc, rtol, tag):
    kd, kc = dev.assemble_vss()
    kdo, kco = orc.assemble_vss()
    assert_close(kd, kdo, rtol, f"{tag} k_diag")
    assert_close(kc, kco, rtol, f"{tag} k_coeffs")


@pytest.mark.parametrize("workload", ["hex8", "hex20"])
def test_full_size_iteration(workload):
    from warp3d_amd.model import BlockModel
    from warp3d_amd.workloads import build_workload
    m, arr, cst, du, parts = build_workload(workload)
    print(f"{workload}: {m.noelem} elements", flush=True)
    dev, orc = BlockModel(arr), O.OracleModel(arr)
    assert np.array_equal(dev.elblks, orc.elblks)
    nd = dev.pattern(cst)
    no = orc.pattern(cst)
    assert nd == no
    pd, po = dev.get_pattern(), orc.get_pattern()
    for a, b in zip(pd, po):
        assert np.array_equal(a, b)
    u = np.zeros_like(du)
    for model in (dev, orc):
        model.recover(1, 0, u, du)
        model.recover(1, 1, u, du)
        model.commit_step()
    u = du.copy()
    for it in (0, 1):
        fd, ed, wd, rcd = dev.recover(2, it, u, du)
        fo, eo, wo, rco = orc.recover(2, it, u, du)
        assert rcd == rco == 0
        assert_close(fd, fo, RTOL_FP64, f"{workload} ifv i{it}")
        assert abs(ed - eo) <= 1e-10 * max(abs(eo), 1e-30)
        assert abs(wd - wo) <= 1e-10 * max(abs(wo), 1e-30) + 1e-300
        print(f"{workload} iteration {it}: forces match", flush=True)
    # the step is plastic at the benchmark's increment
    assert wo > 0.0
    _state_mm01(dev, orc, workload)
    print(f"{workload}: block states match", flush=True)
    dev.tangent(2, 1, _abi.W3D_FP64)
    orc.tangent(2, 1)
    _vss(dev, orc, RTOL_FP64, f"{workload} fp64")
    dev.tangent(2, 1, _abi.W3D_FP32)
    _vss(dev, orc, 2e-6, f"{workload} fp32")
    dev.close()


def _glued_mixed():
    """hex8 12^3 (B-bar, 2x2x2), hex20 6^3 (27 points) and tet10 (6^3
    cells, 4 points), all on the same 1/12 lattice, glued at the faces
    where they touch."""
    mm = M.MixedMesh([(M.box_hex8(12, jitter=0.1, seed=31), 1, 1),
                      (M.box_hex20(6, jitter=0.05, seed=32), 1, 0),
                      (M.box_tet10(6, jitter=0.05, seed=33), 4, 0)],
                     glue=True)
    mm.h = np.full(3, 1.0 / 12.0)   # the lattice spacing (field noise)
    return mm


def test_mixed_types_mm10_seeded_spans():
    from tests.test_gpu_mm10 import cubic_fcc
    from warp3d_amd.model import BlockModel
    mm = _glued_mixed()
    blocks = mm.seeded_blocks(32, 256, seed=4)
    assert blocks[:, 0].max() > 128 and blocks[:, 0].min() >= 1
    # elements of different types share nodes
    starts = np.cumsum([0] + [p[0].noelem for p in mm.parts])
    nodes = [set(mm.incid[mm.incmap[starts[k]] - 1:
                          (mm.incmap[starts[k + 1] - 1] - 1 +
                           mm.parts[k][0].conn.shape[1])].tolist())
             for k in range(3)]
    assert nodes[0] & nodes[1] and nodes[1] & nodes[2]
    ang = M.random_orientations(mm.noelem, seed=35)
    arr = mm.model(cubic_fcc(), elblks=blocks, angles=ang)
    dev, orc = BlockModel(arr), O.OracleModel(arr)
    assert np.array_equal(dev.elblks, orc.elblks)
    cst = np.zeros(3 * mm.nonode, np.int32)
    base = np.abs(mm.coords[:, 2]) < 1e-9
    for k in range(3):
        cst[k::3][base] = 1
    assert dev.pattern(cst) == orc.pattern(cst)
    u = np.zeros(3 * mm.nonode)
    for step in (1, 2):
        # a homogeneous strain plus noise (the recovery does not look at
        # the constraints, which serve the assembly's pattern)
        du_final = smooth_field(mm, 0.0005, seed=step)
        for it in (0, 1, 2):
            du = du_final * (0.5 if it == 0 else (0.9 if it == 1 else 1.0))
            fd, ed, wd, rcd = dev.recover(step, it, u, du)
            fo, eo, wo, rco = orc.recover(step, it, u, du)
            tag = f"mixed s{step} i{it}"
            assert rcd == rco == 0, tag
            assert_close(fd, fo, RTOL_FP64, f"{tag} ifv")
            assert abs(ed - eo) <= 1e-10 * max(abs(eo), 1e-30)
            assert abs(wd - wo) <= 1e-10 * max(abs(wo), 1e-30) + 1e-300
            for b in range(1, orc.nelblk + 1):
                for which, name in ((_abi.STRESS_N1, "stress"),
                                    (_abi.STRAIN_N1, "strain"),
                                    (_abi.EINFVEC, "einfvec")):
                    assert_close(dev.get_block_state(b, which),
                                 orc.get_block_state(b, which), RTOL_FP64,
                                 f"{tag} blk {b} {name}")
                compare_history(dev.get_block_state(b, _abi.HIST_N1),
                                orc.get_block_state(b, _abi.HIST_N1), 1,
                                f"{tag} blk {b}")
            dev.tangent(step, it, _abi.W3D_FP64)
            orc.tangent(step, it)
            for b in range(1, orc.nelblk + 1):
                assert_close(dev.get_block_state(b, _abi.ESTIFF),
                             orc.get_block_state(b, _abi.ESTIFF), RTOL_FP64,
                             f"{tag} Ke blk {b}")
            _vss(dev, orc, RTOL_FP64, tag)
            print(f"{tag}: match", flush=True)
        dev.commit_step()
        orc.commit_step()
        u = u + du_final
    # the crystals slipped
    assert wo > 0.0
    dev.close()


def test_served_rve_matches_the_oracle_newton_loop():
    from tests.serve_ref import solve_served_oracle
    from tests.test_gpu_serve import run_alone
    from warp3d_amd import serve as S
    spec = S.RveSpec(8, seed=5, ngrains=64, nsteps=10, strain=0.02)
    opts = S.newton_opts(precision=_abi.W3D_FP64, pcg_rtol=1e-13,
                         pcg_maxit=20000)
    print("oracle", flush=True)
    ro, uo = solve_served_oracle(S.make_rve(spec), opts)
    print("served", flush=True)
    rd, ud = run_alone(spec, opts)
    assert len(rd) == len(ro) == 10
    assert [r["newton"] for r in rd] == [r["newton"] for r in ro]
    assert [r["cuts"] for r in rd] == [r["cuts"] for r in ro]
    for a, b in zip(rd, ro):
        assert abs(a["load"] - b["load"]) <= 1e-8 * abs(b["load"])
    assert np.abs(ud - uo).max() <= 1e-8 * np.abs(uo).max()
