"""Asymmetric assembly (the reference's asymmetric_assembly) on the oracle:
mm10's tangent kept unsymmetrised (mm10_a.f:1165), the full element
matrices of bdbt_asym (gptns1.f:1513) and their gather into full CSR rows
(assem_a_row_asym, assemble_code.f:1050).

Pins:
* with a symmetric [D] (mm01, hex8 B-bar / hex20 / tet10) the full Ke is
  symmetric, and its upper triangle is the packed Ke of bdbtgen (pinned by
  the symmetric-path tests), to rounding;
* with mm10 each element's full Ke is sum_p w|J| B^T D B formed with numpy
  matrix products from the oracle's B-matrix leaves and the stored history
  words 1-36 (the column-major [D] that drive_10_cnst reads), in
  bdbt_asym's word order (word c T + r: row r, column c);
* the unsymmetrised mm10 tangent's symmetric part is the symmetrised
  tangent;
* the CSR values are a dense scatter of element word (erow - 1) T + ecol
  into (equation of erow, equation of ecol), and the pattern holds every
  coupled pair.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests._util import smooth_field
from warp3d_amd import _abi
from warp3d_amd import mesh as M
from warp3d_amd.serve import RVE_MATERIAL


def _packed_to_full(ek, T):
    """(utsz,) packed upper triangle (pos = c(c-1)/2 + r, r <= c, 1-based)
    to a symmetric T x T."""
    K = np.zeros((T, T))
    k = 0
    for c in range(T):
        for r in range(c + 1):
            K[r, c] = K[c, r] = ek[k]
            k += 1
    return K


def _mm10_box(n=2, bbar=0, seed=5):
    m = M.box_hex8(n, jitter=0.05, seed=seed)
    cp = _abi.mm10_material(**RVE_MATERIAL)
    arr = m.model(cp, auto_size=4, bbar=bbar,
                  angles=M.random_orientations(m.noelem, seed=seed + 1))
    return m, arr


def _run(arr, m, asym, amp=0.004, steps=2):
    """Recover and commit steps of a smooth field, then one more recovery
    and the tangent."""
    o = O.OracleModel(arr)
    o.set_asymmetric(asym)
    cst = np.zeros(3 * m.nonode, np.int32)
    cst[:3] = 1
    cst[5] = 1
    o.pattern(cst)
    u = np.zeros(3 * m.nonode)
    for s in range(1, steps + 1):
        du = smooth_field(m, amp, seed=s)
        assert o.recover(s, 1, u, du)[3] == 0
        o.commit_step()
        u = u + du
    du = smooth_field(m, amp, seed=steps + 1)
    assert o.recover(steps + 1, 1, u, du)[3] == 0
    o.tangent(steps + 1, 1)
    return o


@pytest.mark.parametrize("kind", ["hex8", "hex20", "tet10", "mixed"])
def test_symmetric_material_gives_the_packed_matrix(kind):
    mat = _abi.mm01_material(200e3, 0.3, 400.0, hprime=2000.0)
    m = {"hex8": lambda: M.box_hex8(2, jitter=0.1, seed=3),
         "hex20": lambda: M.box_hex20(2, jitter=0.05, seed=3),
         "tet10": lambda: M.box_tet10(1, jitter=0.05, seed=3),
         "mixed": lambda: M.MixedMesh(
             [(M.box_hex8(2, jitter=0.1, seed=3), 1, 1),
              (M.box_hex20(2, jitter=0.05, seed=4), 8, 0),
              (M.box_tet10(1, jitter=0.05, seed=5), 4, 0)])}[kind]()
    arr = m.model(mat, auto_size=8)
    oa, os_ = _run(arr, m, True, amp=0.003), _run(arr, m, False, amp=0.003)
    for b in range(1, oa.nelblk + 1):
        info = oa.block_info(b)
        T = info["totdof"]
        kf = oa.get_ek_full(b)
        ek = os_.get_block_state(b, _abi.ESTIFF).reshape(info["span"], -1)
        for e in range(info["span"]):
            Kp = _packed_to_full(ek[e], T)
            sc = np.abs(Kp).max()
            assert np.abs(kf[e] - kf[e].T).max() <= 1e-13 * sc
            assert np.abs(kf[e] - Kp).max() <= 1e-13 * sc


def test_mm10_full_matrix_is_the_numpy_product():
    m, arr = _mm10_box()
    o = _run(arr, m, True)
    etype, order = _abi.W3D_L3DISOP, 1
    G = O.ngp(etype, order)
    info = o.block_info(1)
    span, H, T = info["span"], info["hist"], info["totdof"]
    h1 = o.get_block_state(1, _abi.HIST_N1).reshape(span, G, H)
    kf = o.get_ek_full(1)
    skew = 0.0
    for e in range(span):
        nodes = m.conn[e] - 1  # 1-based incidences
        ce = np.concatenate([m.coords[nodes, k] for k in range(3)])
        K = np.zeros((T, T))
        for g in range(G):
            xi, eta, zeta, w = O.getgpts(etype, order, g + 1)
            nxi, neta, nzeta = O.derivs(etype, xi, eta, zeta, 8)
            _, dj, gama = O.jacob1(nxi, neta, nzeta, ce)
            B = O.bmatrix(gama, nxi, neta, nzeta)
            D = h1[e, g, :36].reshape(6, 6, order="F")
            skew = max(skew, np.abs(D - D.T).max() / np.abs(D).max())
            K += w * dj * (B.T @ D @ B)
        assert np.abs(kf[e].T - K).max() <= 1e-12 * np.abs(K).max()
    # the crystal tangent is genuinely unsymmetric here
    assert skew > 1e-6


def test_mm10_symmetric_part_is_the_symmetrised_tangent():
    m, arr = _mm10_box()
    oa, os_ = _run(arr, m, True), _run(arr, m, False)
    info = oa.block_info(1)
    G, H = info["ngp"], info["hist"]
    ha = oa.get_block_state(1, _abi.HIST_N1).reshape(-1, G, H)
    hs = os_.get_block_state(1, _abi.HIST_N1).reshape(-1, G, H)
    Da = ha[..., :36].reshape(-1, 6, 6)
    Ds = hs[..., :36].reshape(-1, 6, 6)
    assert np.abs(0.5 * (Da + Da.transpose(0, 2, 1)) - Ds).max() <= \
        1e-12 * np.abs(Ds).max()
    # every other history word is the same
    assert np.array_equal(ha[..., 36:], hs[..., 36:])


def test_csr_values_are_the_dense_scatter():
    m, arr = _mm10_box(n=3, bbar=1)
    o = _run(arr, m, True)
    rp, col, val = o.assemble_csr()
    dmap = o.get_pattern()[0]
    n = o.neqns
    Kd = np.zeros((n, n))
    ncover = np.zeros((n, n), bool)
    first = 1
    for b in range(1, o.nelblk + 1):
        info = o.block_info(b)
        T = info["totdof"]
        kf = o.get_ek_full(b)
        for e in range(info["span"]):
            nodes = m.conn[first - 1 + e] - 1
            dofs = np.concatenate([3 * nodes + k for k in range(3)])
            eq = dmap[dofs]
            for r in range(T):
                if eq[r] == 0:
                    continue
                for c in range(T):
                    if eq[c] == 0:
                        continue
                    # the word (erow - 1) T + ecol of element e
                    Kd[eq[r] - 1, eq[c] - 1] += kf[e].ravel()[r * T + c]
                    ncover[eq[r] - 1, eq[c] - 1] = True
        first += info["span"]
    rows = np.repeat(np.arange(n), np.diff(rp))
    Ks = np.zeros((n, n))
    Ks[rows, col] = val
    assert np.array_equal(ncover, Ks != 0) or \
        np.all(ncover[rows, col])
    assert np.abs(Ks - Kd).max() <= 1e-13 * np.abs(Kd).max()
    assert ncover.sum() == rp[-1]
    # the matrix is not symmetric
    assert np.abs(Ks - Ks.T).max() > 1e-8 * np.abs(Ks).max()
