"""The served RVE's Newton algorithm on the oracle, for the tests: the
algorithm of w3d_batch_advance and serve.solve_plain (iteration 0 that
starts every increment and is never accepted, the residual -ifv on the
free equations, convergence |R| <= tol max(|F_c|, tol_abs) with F_c the
internal forces on the constrained dofs, the adaptive sub-increments of
adapt_check.f with the first-sub-increment restart, and extrapolation of
the last converged increment), with the equations solved by a sparse
direct LU (scipy) in place of the device Jacobi-PCG.
"""
import numpy as np
import scipy.sparse.linalg as spla

from oracle import oracle as O
from tests.newton_ref import vss_matrix
from warp3d_amd.steps import Adaptive


def solve_served_oracle(rve, opts, threads=0):
    """Returns the per-step records (newton, cuts, load, residual,
    reaction) and the final displacements."""
    om = O.OracleModel(rve.arrays, threads=threads)
    om.pattern(rve.cstmap)
    dmap, _, _, _ = om.get_pattern()
    free = np.flatnonzero(rve.cstmap == 0)
    eqdof = free[np.argsort(dmap[free] - 1)]
    cdof = np.flatnonzero(rve.cstmap != 0)
    ldof = np.flatnonzero((rve.cstmap != 0) & (rve.du0 != 0.0))
    imp = rve.du0[cdof]
    u = np.zeros(3 * rve.mesh.nonode)
    du = rve.du0.copy()
    dt0 = 1.0
    records = []

    def start(sf, disp):
        du[eqdof] = sf * du[eqdof]
        du[cdof] = disp * imp
        om.set_time_increment(dt0 * disp)

    for step in range(1, opts.nsteps + 1):
        ad = Adaptive()
        if step > 1:
            start(1.0, 1.0)
        else:
            om.set_time_increment(dt0)
        iters = cuts = 0
        it, extrap = 0, True
        while True:
            if it >= 1:
                iters += 1
            ifv, _, _, rc = om.recover(step, it, u, du, extrapolated=extrap)
            cut = rc != 0
            r = -ifv[eqdof]
            res = float(np.sqrt(np.dot(r, r)))
            react = float(np.sqrt(np.dot(ifv[cdof], ifv[cdof])))
            if not cut and it >= 1 and \
                    res <= opts.tol * max(react, opts.tol_abs):
                om.commit_step()
                u = u + du
                it, extrap = 0, True
                if opts.adaptive and ad.advance():
                    start(ad.scaling, ad.disp_fact)
                    continue
                records.append(dict(newton=iters, cuts=cuts,
                                    load=float(ifv[ldof].sum()),
                                    residual=res, reaction=react))
                break
            if cut or it >= opts.max_iters:
                if opts.adaptive and ad.cut():
                    om.rollback()
                    start(0.0, ad.disp_fact)
                    cuts += 1
                    it, extrap = 0, False
                    continue
                raise RuntimeError(f"step {step}: no convergence")
            om.tangent(step, it)
            kd, kc = om.assemble_vss()
            K = vss_matrix(om, kd, kc)
            du[eqdof] += spla.spsolve(K, r)
            it += 1
    return records, u
