"""The Newton loop of a static load step on the oracle, for the tests: the
same algorithm as warp3d_amd/fortran/w3d_newton.f90, with the equations
solved by a sparse direct LU (scipy) in place of the device PCG.

Restates the static path of mnralg.f:
* the start of a step or sub-increment (mnralg.f:211-300): extrapolation
  of the last converged increment by the load-ratio and adaptive scale
  factors (clamped: < 0 -> 0, > 5 -> 1), none at step 1, after a change
  of load pattern or on the first sub-increment after a cut; imposed
  increments scaled by the adaptive displacement factor;
* iteration 0 (mnralg.f:358-430, uppbar.f, upres_iter_0): recovery with
  the starting increment, element stiffnesses, res = pbar - ifv on the
  free dofs with pbar = load at n + adapt_load_fact * dload;
* iterations (mnralg.f:440-560): solve K idu = res, update through the
  line search of mnralg_ls (mnralg.f:758-914; backtracking by rho from
  step length 1 while |s/s0| >= slack, s = -res . idu, back to 1 when r
  stops falling by 0.9 or the length falls below its minimum), residual
  and load norm (upres.f:148-234), the cvtest tests (cvtest.f:103-310;
  test 4's average nodal force from addifv's sum of |element force| and
  upres's sum of the nonzero load terms) and the divergence check (three
  non-decreasing residual norms, cvtest.f:94-107), new element
  stiffnesses;
* failure (iteration limit, divergence, a material's cut request): with
  adaptive on and step > 1, the state is restored and the increment is
  split by the adaptive stack of adapt_check.f (4 divisions, smallest
  factor 1/16), otherwise the step fails.
"""
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O
from warp3d_amd.steps import Adaptive


def vss_matrix(om, kd, kc):
    """Full symmetric scipy matrix from the VSS arrays."""
    dmap, _, kptr, kidx = om.get_pattern()
    n = om.neqns
    rows = np.repeat(np.arange(n), kptr)
    cols = kidx - 1
    up = sp.csr_matrix((kc, (rows, cols)), shape=(n, n))
    return (up + up.T + sp.diags(kd)).tocsc()


def residual(cst, ifv, pbar=None, sums=False):
    """upres: res = pbar - ifv on free dofs (0 on constrained ones), the
    load norm over the nonzero terms |pbar| (free) and |ifv| (constrained),
    and the residual norm; with sums also the sum and number of the
    nonzero load terms (sum_loads, num_term_loads)."""
    p = np.zeros_like(ifv) if pbar is None else pbar
    res = np.where(cst == 0, p - ifv, 0.0)
    term = np.where(cst == 0, np.abs(p), np.abs(ifv))
    nz = term[term != 0.0]
    mgload = np.sqrt(np.sum(nz ** 2))
    out = (res, mgload, np.sqrt(np.sum(res ** 2)))
    if sums:
        out = out + (float(np.sum(nz)), int(nz.size))
    return out


def ifv_sums(om):
    """addifv's sum of |element force| terms over the blocks and their
    number (num_term_ifv), after a recovery."""
    if hasattr(om, "ifv_sums"):   # the library keeps them itself
        return om.ifv_sums()
    s, n = 0.0, 0
    for b in range(1, om.nelblk + 1):
        s = s + om.block_sums(b)[2]
        info = om.block_info(b)
        n += info["span"] * info["totdof"]
    return s, n


def cvtest(tests, it, idu, res, nres, mgload, sum_ifv, nterm_ifv, sum_loads,
           nterm_loads, state):
    """cvtest.f:103-310 over the selected tests ({number: tolerance}):
    (converged, values).  state['magdu1'] keeps |idu| of the first
    iteration of the step (set at it = 1 when test 1 or 3 is selected)."""
    nidu = float(np.sqrt(np.sum(idu ** 2)))
    maxidu = float(np.max(np.abs(idu)))
    resmax = float(np.max(np.abs(res)))
    if it == 1 and (1 in tests or 3 in tests):
        state["magdu1"] = nidu
    m1 = state.get("magdu1", 0.0)
    ok, vals = True, {}
    for t, tol in tests.items():
        if t == 1:
            if it == 1:
                f = tol > 100.0
            elif m1 == 0.0:
                f = True
            else:
                vals[1] = nidu / m1 * 100.0
                f = vals[1] < tol
        elif t == 2:
            if mgload == 0.0:
                f = True
            else:
                vals[2] = nres / mgload * 100.0
                f = vals[2] <= tol or (mgload <= 1e-5 and nres <= 1e-8)
        elif t == 3:
            if m1 == 0.0:
                f = True
            else:
                vals[3] = maxidu / m1 * 100.0
                f = vals[3] < tol
        elif t == 4:
            avg = (sum_ifv + sum_loads) / float(nterm_ifv + nterm_loads)
            if avg == 0.0:
                f = True
            else:
                vals[4] = resmax / avg * 100.0
                f = vals[4] <= tol or (resmax <= 1e-8 and avg <= 1e-8)
        elif t == 5:
            vals[5] = nidu
            f = nidu <= tol
        elif t == 6:
            vals[6] = maxidu
            f = maxidu <= tol
        else:
            raise ValueError(f"cvtest: no test {t}")
        ok = ok and f
    return ok, vals


@dataclass
class NewtonOptions:
    res_tol: float = 0.01          # cvtest test 2 (%)
    tests: dict = None             # cvtest {test: tolerance}; None: test 2
    mniter: int = 1
    mxiter: int = 15
    extrapolate: bool = True
    line_search: bool = False
    ls_rho: float = 0.7
    ls_slack: float = 0.5
    ls_min: float = 0.01
    ls_max: float = 1.0
    adaptive: bool = False
    divergence_check: bool = True


def line_search(evaluate, opts):
    """mnralg_ls (mnralg.f:813-889): evaluate(alpha) updates the state to
    du0 + alpha idu and returns r = |s(alpha) / s0| (None if the recovery
    asks for a cut).  Returns the accepted step length; the state is left
    at it."""
    alpha = opts.ls_max
    rvals = []
    while True:
        r = evaluate(alpha)
        if r is None:
            return alpha
        rvals.append(r)
        if r < opts.ls_slack:
            return alpha
        if len(rvals) >= 2 and rvals[-1] > 0.9 * rvals[-2]:
            evaluate(1.0)              # r not falling: full step
            return 1.0
        alpha = opts.ls_rho * alpha    # backtrack
        if alpha < opts.ls_min:
            evaluate(1.0)              # too short: full step
            return 1.0


def _solve(om, cst, free, eq, res):
    rhs = np.zeros(om.neqns)
    rhs[eq] = res[free]
    if om.asym:  # asymmetric assembly: the full CSR rows as assembled
        rp, col, val = om.assemble_csr()
        if getattr(om, "device_solve", False):  # the library's BiCGStab
            x = om.solve_bicgstab(rhs, rtol=1e-13)[0]
        else:
            K = sp.csr_matrix((val, col, rp),
                              shape=(om.neqns, om.neqns)).tocsc()
            x = spla.spsolve(K, rhs)
    else:
        kd, kc = om.assemble_vss()
        K = vss_matrix(om, kd, kc)
        x = spla.spsolve(K, rhs)
    idu = np.zeros_like(res)
    idu[free] = x[eq]
    return idu


def newton_step(om, step, cst, dmap, u, du_imposed, res_tol=0.01, mniter=1,
                mxiter=15, opts=None, du_last=None, load_n=None, dload=None,
                scaling_load=1.0, mf_change=False, dt=1.0):
    """One load step (with its adaptive sub-increments).  Returns
    (converged, iters, ratios, u_new, du_last_new, info): iters summed over
    sub-increments, ratios of the last, du_last_new the last converged
    increment (the next step extrapolates from it), info the counts of
    sub-increments and line-search evaluations."""
    if opts is None:
        opts = NewtonOptions(res_tol=res_tol, mniter=mniter, mxiter=mxiter,
                             extrapolate=False)
    free = dmap > 0
    eq = dmap[free] - 1
    nd = u.size
    load_n = np.zeros(nd) if load_n is None else load_n
    dload = np.zeros(nd) if dload is None else dload
    du = np.zeros(nd) if du_last is None else du_last.copy()
    ad = Adaptive()
    adaptive = opts.adaptive and step > 1
    first_subinc = False
    total_iters = 0
    info = dict(subincrements=0, ls_evals=0, cuts=0)
    while True:
        # ---- start of the step or sub-increment
        sf = scaling_load * ad.scaling
        extrapolated = False
        if first_subinc or not (opts.extrapolate and step > 1):
            du = np.zeros(nd)
        elif not mf_change:
            if sf < 0.0:
                sf = 0.0
            if sf > 5.0:
                sf = 1.0
            extrapolated = True
            du = sf * du
        # a change of load pattern keeps the last increment as it is and
        # starts without extrapolation (mnralg.f:267-271)
        du[cst != 0] = du_imposed[cst != 0] * ad.disp_fact
        pbar = load_n + ad.load_fact * dload
        if adaptive:  # a sub-increment's share of the time (mnralg.f:306)
            om.set_time_increment(dt * ad.disp_fact)
        ifv, _, _, rc = om.recover(step, 0, u, du, extrapolated)
        failed = rc != 0
        ratios, norms = [], []
        tests = opts.tests if opts.tests else {2: opts.res_tol}
        cvstate = {}
        info["tests"] = []
        it = 0
        if not failed:
            om.tangent(step, 0)
            res, mgload, nres = residual(cst, ifv, pbar)
            it = 1
            while True:
                if it > 1:
                    om.tangent(step, it)
                idu = _solve(om, cst, free, eq, res)
                # mnralg_ls
                s0 = -np.dot(res[free], idu[free])
                du0 = du.copy()
                state = {}

                def evaluate(alpha):
                    d = du0 + alpha * idu
                    f, _, _, code = om.recover(step, it, u, d, extrapolated)
                    info["ls_evals"] += 1
                    state.update(du=d, ifv=f, rc=code)
                    if code != 0:
                        return None
                    rs, _, _ = residual(cst, f, pbar)
                    return abs(-np.dot(rs[free], idu[free]) / s0)

                if s0 >= 0.0 or not opts.line_search:
                    evaluate(1.0)
                else:
                    line_search(evaluate, opts)
                du, ifv, rc = state["du"], state["ifv"], state["rc"]
                if rc != 0:
                    failed = True
                    break
                res, mgload, nres, sl, nl = residual(cst, ifv, pbar,
                                                     sums=True)
                ratio = 0.0 if mgload == 0 else nres / mgload * 100.0
                ratios.append(ratio)
                norms.append(nres)
                si, ni = ifv_sums(om)
                conv, vals = cvtest(tests, it, idu, res, nres, mgload, si,
                                    ni, sl, nl, cvstate)
                info["tests"].append(vals)
                diverging = len(norms) >= 3 and norms[-1] >= norms[-2] and \
                    norms[-2] >= norms[-3]
                if conv and it >= opts.mniter:
                    break
                it += 1
                if it > opts.mxiter or (opts.divergence_check and
                                        diverging):
                    it -= 1
                    failed = True
                    break
        total_iters += it
        if failed:
            if adaptive and ad.cut():
                om.rollback()
                info["cuts"] += 1
                first_subinc = True
                continue
            if adaptive:
                om.set_time_increment(dt)
            return False, total_iters, ratios, u, du, info
        om.commit_step()
        u = u + du
        info["subincrements"] += 1
        if adaptive and ad.advance():
            first_subinc = False
            scaling_load = 1.0
            continue
        if adaptive:
            om.set_time_increment(dt)
        return True, total_iters, ratios, u, du, info


def run_history(arrays, cst, du_steps, opts=None, dloads=None, mfs=None,
                dt=1.0, om=None, **kw):
    """All steps on the oracle; returns a list of (converged, iters,
    ratios, u, info) per step.  dloads: external load increments per step;
    mfs: the steps' load multipliers (extrapolation scales by mf/mf_prev;
    None = proportional steps); dt: the time increment of a step; om: an
    oracle model to run on (its state is the result), else a new one."""
    if om is None:
        om = O.OracleModel(arrays)
    om.set_time_increment(dt)
    om.pattern(cst)
    dmap, _, _, _ = om.get_pattern()
    nd = 3 * arrays.nonode
    u = np.zeros(nd)
    du_last = np.zeros(nd)
    load = np.zeros(nd)
    out = []
    for s, dui in enumerate(du_steps, start=1):
        dl = np.zeros(nd) if dloads is None else dloads[s - 1]
        sl = 1.0
        if mfs is not None and s > 1:
            prev = mfs[s - 2]
            sl = mfs[s - 1] / (prev if abs(prev) > 1e-20 else 1.0)
        conv, it, ratios, u, du_last, info = newton_step(
            om, s, cst, dmap, u, dui, opts=opts, du_last=du_last,
            load_n=load, dload=dl, scaling_load=sl, dt=dt, **kw)
        load = load + dl
        out.append((conv, it, ratios, u.copy(), info))
        if not conv:
            break
    return out
