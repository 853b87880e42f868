"""The Newton loop of a displacement-controlled load step on the oracle, for
the tests: the same algorithm as warp3d_amd/fortran/w3d_newton.f90 (the
static path of mnralg.f:177-600 with upres.f's residual and load norm and
cvtest.f's test 2), with the equations solved by a sparse direct LU
(scipy) in place of the device PCG."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O


def vss_matrix(om, kd, kc):
    """Full symmetric scipy matrix from the VSS arrays."""
    dmap, _, kptr, kidx = om.get_pattern()
    n = om.neqns
    rows = np.repeat(np.arange(n), kptr)
    cols = kidx - 1
    up = sp.csr_matrix((kc, (rows, cols)), shape=(n, n))
    return (up + up.T + sp.diags(kd)).tocsc()


def residual(cst, ifv):
    res = np.where(cst == 0, -ifv, 0.0)
    mgload = np.sqrt(np.sum(np.where(cst != 0, ifv, 0.0) ** 2))
    return res, mgload, np.sqrt(np.sum(res ** 2))


def newton_step(om, step, cst, dmap, u, du_imposed, res_tol=0.01, mniter=1,
                mxiter=15):
    """One load step; returns (converged, iters, ratios, u_new)."""
    free = dmap > 0
    eq = dmap[free] - 1
    du = du_imposed.copy()
    ifv, _, _, rc = om.recover(step, 0, u, du)
    om.tangent(step, 0)
    res, mgload, nres = residual(cst, ifv)
    ratios = []
    it = 1
    while True:
        if it > 1:
            om.tangent(step, it)
        kd, kc = om.assemble_vss()
        K = vss_matrix(om, kd, kc)
        rhs = np.zeros(om.neqns)
        rhs[eq] = res[free]
        x = spla.spsolve(K, rhs)
        du[free] += x[eq]
        ifv, en, pw, rc = om.recover(step, it, u, du)
        res, mgload, nres = residual(cst, ifv)
        ratio = 0.0 if mgload == 0 else nres / mgload * 100.0
        ratios.append(ratio)
        conv = mgload == 0 or ratio <= res_tol or \
            (mgload <= 1e-5 and nres <= 1e-8)
        if conv and it >= mniter:
            om.commit_step()
            return True, it, ratios, u + du
        it += 1
        if it > mxiter:
            return False, it - 1, ratios, u


def run_history(arrays, cst, du_steps, **kw):
    """All steps on the oracle; returns a list of (iters, ratios, u)."""
    om = O.OracleModel(arrays)
    om.pattern(cst)
    dmap, _, _, _ = om.get_pattern()
    u = np.zeros(3 * arrays.nonode)
    out = []
    for s, dui in enumerate(du_steps, start=1):
        conv, it, ratios, u = newton_step(om, s, cst, dmap, u, dui, **kw)
        out.append((conv, it, ratios, u.copy()))
        if not conv:
            break
    return out
