"""Asymmetric assembly on the device against the oracle (whose full Ke and
CSR gather are pinned in test_asym.py): mm10's unsymmetrised tangent in
the history, the full element matrices and the CSR values within the FP64
bar (1e-10 relative), the CSR pattern identical."""
import numpy as np
import pytest

from oracle import oracle as O
from tests._util import RTOL_FP64, assert_close, smooth_field
from warp3d_amd import _abi
from warp3d_amd import mesh as M
from warp3d_amd.serve import RVE_MATERIAL

pytestmark = pytest.mark.gpu


def _case(kind):
    if kind.startswith("mm10"):
        m = M.box_hex8(3, jitter=0.05, seed=5)
        cp = _abi.mm10_material(**RVE_MATERIAL)
        arr = m.model(cp, auto_size=8, bbar=1 if kind == "mm10-bbar" else 0,
                      angles=M.random_orientations(m.noelem, seed=6))
        return m, arr, 0.004
    mat = _abi.mm01_material(200e3, 0.3, 400.0, hprime=2000.0)
    m = {"hex20": lambda: M.box_hex20(2, jitter=0.05, seed=3),
         "tet10": lambda: M.box_tet10(2, jitter=0.05, seed=3),
         # blocks of three element sizes (T = 24, 60, 30) in one model
         "mixed": lambda: M.MixedMesh(
             [(M.box_hex8(2, jitter=0.1, seed=3), 1, 1),
              (M.box_hex20(2, jitter=0.05, seed=4), 8, 0),
              (M.box_tet10(1, jitter=0.05, seed=5), 4, 0)]),
         # the same types glued where they touch: CSR rows of the shared
         # interface nodes gather Ke words of elements with T = 24, 60 and
         # 30 (boundary nodes are not jittered, so the lattices meet)
         "mixed_glued": lambda: M.MixedMesh(
             [(M.box_hex8(2, jitter=0.1, seed=3), 1, 1),
              (M.box_hex20(1, jitter=0.05, seed=4), 8, 0),
              (M.box_tet10(1, jitter=0.05, seed=5), 4, 0)],
             glue=True)}[kind]()
    return m, m.model(mat, auto_size=8), 0.003


def _drive(model, m, amp, cst=None):
    model.set_asymmetric(True)
    if cst is None:
        cst = np.zeros(3 * m.nonode, np.int32)
        cst[:3] = 1
        cst[5] = 1
    model.pattern(cst)
    u = np.zeros(3 * m.nonode)
    for s in (1, 2):
        du = smooth_field(m, amp, seed=s)
        assert model.recover(s, 1, u, du)[3] == 0
        model.commit_step()
        u = u + du
    du = smooth_field(m, amp, seed=3)
    ifv = model.recover(3, 1, u, du)[0]
    model.tangent(3, 1)
    return ifv


@pytest.mark.parametrize("kind", ["mm10-bbar", "mm10", "hex20", "tet10",
                                  "mixed", "mixed_glued"])
def test_full_matrices_and_csr_match_the_oracle(kind):
    from warp3d_amd.model import BlockModel
    m, arr, amp = _case(kind)
    d, o = BlockModel(arr), O.OracleModel(arr)
    fd, fo = _drive(d, m, amp), _drive(o, m, amp)
    assert_close(fd, fo, what="forces")
    for b in range(1, o.nelblk + 1):
        assert_close(d.get_block_state(b, _abi.HIST_N1),
                     o.get_block_state(b, _abi.HIST_N1), what=f"hist {b}")
        assert_close(d.get_ek_full(b), o.get_ek_full(b), what=f"Ke {b}")
    rp, col, val = d.assemble_csr()
    rpo, colo = o.csr_pattern()
    assert np.array_equal(rp, rpo) and np.array_equal(col, colo)
    _, _, valo = o.assemble_csr(rp, col)
    assert_close(val, valo, what="CSR values")
    if kind.startswith("mm10"):  # the matrix is genuinely unsymmetric
        n = d.neqns
        rows = np.repeat(np.arange(n), np.diff(rp))
        import scipy.sparse as sp
        K = sp.csr_matrix((val, col, rp), shape=(n, n))
        assert abs(K - K.T).max() > 1e-8 * abs(K).max()


def test_symmetric_paths_refuse_an_asymmetric_context():
    from warp3d_amd.model import BlockModel, W3DError
    m, arr, amp = _case("hex20")
    d = BlockModel(arr)
    _drive(d, m, amp)
    with pytest.raises(W3DError):
        d.assemble_vss()
    with pytest.raises(W3DError):
        d.tangent(3, 1, precision=_abi.W3D_FP32)


def test_bicgstab_solves_the_asymmetric_csr():
    """The device BiCGStab against a sparse LU of the same CSR."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from warp3d_amd.model import BlockModel
    m, arr, amp = _case("mm10-bbar")
    d = BlockModel(arr)
    # the bottom face held (a nonsingular system)
    cst = np.zeros(3 * m.nonode, np.int32)
    for k in range(3):
        cst[k::3][m.coords[:, 2] == 0.0] = 1
    _drive(d, m, amp, cst)
    rp, col, val = d.assemble_csr()
    n = d.neqns
    K = sp.csr_matrix((val, col, rp), shape=(n, n))
    b = np.random.default_rng(7).standard_normal(n)
    x, it, rr = d.solve_bicgstab(b, rtol=1e-12, maxit=5000)
    xr = spla.spsolve(K.tocsc(), b)
    assert rr <= 1e-12 and it > 0
    assert np.linalg.norm(K @ x - b) <= 1e-11 * np.linalg.norm(b)
    assert np.abs(x - xr).max() <= 1e-8 * np.abs(xr).max()
    # bitwise repeatable
    assert np.array_equal(d.solve_bicgstab(b, rtol=1e-12, maxit=5000)[0], x)
