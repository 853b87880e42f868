"""Binary decks and results of the Fortran load-history driver
(warp3d_amd/bin/w3d_drive, built from warp3d_amd/fortran/w3d_drive.f90).

A deck holds a model (the w3d_model arrays and materials), the absolute
constraints, the Newton options and, per load step, the load multiplier,
the imposed displacement increments and the external nodal load increments;
the driver runs each step with the Fortran Newton loop over libw3dblk on the
GPU and writes, per step, its status, iteration counts, adaptive cuts, PCG
iterations, test-2 residual ratios, energy, plastic work and the
displacements after the step.
"""
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVE = os.path.join(HERE, "bin", "w3d_drive")


def write_deck(path, arrays, cstmap, du_steps, precision=0, res_tol=0.01,
               mniter=1, mxiter=15, pcg_rtol=1e-12, pcg_maxit=50000,
               dt=1.0, extrapolate=True, line_search=False, adaptive=False,
               divergence_check=True, ls_rho=0.7, ls_slack=0.5, ls_min=0.01,
               ls_max=1.0, dloads=None, mfs=None, mf_changes=None,
               asymmetric=False, tests=None, per_block=False):
    """arrays: warp3d_amd._abi.ModelArrays (single-type connectivity);
    du_steps: (nsteps, 3*nonode) imposed increments (zero off the
    constrained dofs); dloads: external load increments per step (default
    none); mfs: load multipliers of the steps (default 1: proportional
    steps); mf_changes: steps whose load pattern changes; tests: the
    cvtest tests as {number: tolerance} (default {2: res_tol});
    per_block: recoveries and tangents through the per-block calls in an
    OpenMP loop over the blocks.  The Newton options default to the
    reference's (mnralg: extrapolation on, line search and adaptive off,
    divergence check on)."""
    tests = {2: res_tol} if tests is None else dict(tests)
    convrg = [int(k in tests) for k in range(1, 7)]
    tols = [float(tests.get(k, 0.0)) for k in range(1, 7)]
    nnpe = int(np.diff(arrays.incmap).max()) if arrays.noelem > 1 else \
        arrays.incid.size
    if not np.array_equal(arrays.incmap,
                          np.arange(arrays.noelem) * nnpe + 1):
        raise ValueError("write_deck: one element type per deck")
    du_steps = np.ascontiguousarray(du_steps, dtype=np.float64)
    nsteps = du_steps.shape[0]
    nd = 3 * arrays.nonode
    dloads = (np.zeros((nsteps, nd)) if dloads is None else
              np.ascontiguousarray(dloads, dtype=np.float64))
    mfs = np.ones(nsteps) if mfs is None else np.asarray(mfs, float)
    chg = (np.zeros(nsteps) if mf_changes is None else
           np.asarray(mf_changes, float))
    with open(path, "wb") as f:
        np.array([arrays.nonode, arrays.noelem, nnpe,
                  len(arrays.materials), nsteps, precision, mniter, mxiter,
                  arrays.auto_size, pcg_maxit, int(extrapolate),
                  int(line_search), int(adaptive), int(divergence_check),
                  int(asymmetric)] + convrg + [int(per_block)],
                 np.int32).tofile(f)
        np.array(tols + [pcg_rtol, dt, ls_rho, ls_slack, ls_min, ls_max],
                 np.float64).tofile(f)
        arrays.coords.astype(np.float64).tofile(f)
        geonl = (np.zeros(arrays.noelem, np.int32) if arrays.geonl is None
                 else arrays.geonl)
        for a in (arrays.incid, arrays.etype, arrays.int_order, arrays.matnum,
                  arrays.bbar, geonl):
            np.asarray(a, np.int32).tofile(f)
        for m in arrays.materials:
            np.array([m.model, m.linear_elastic, m.slip_type, m.elastic_type,
                      m.miter, m.alter_mode, m.angle_type, m.ncrystals],
                     np.int32).tofile(f)
            np.array([m.e, m.nu, m.beta, m.tan_e, m.yld_pt],
                     np.float32).tofile(f)
            np.array([m.cp_e, m.cp_nu, m.cp_mu, m.harden_n, m.theta_0,
                      m.tau_y, m.tau_v, m.voce_m, m.iD_v, m.eps_dot_0_y,
                      m.atol, m.atol1, m.rtol, m.rtol1],
                     np.float64).tofile(f)
        has = arrays.angles is not None
        np.array([int(has), arrays.angle_stride], np.int32).tofile(f)
        if has:
            arrays.angles.astype(np.float64).tofile(f)
        has_trn = arrays.trn is not None
        np.array([int(has_trn)], np.int32).tofile(f)
        if has_trn:
            np.asarray(arrays.trn, np.int32).tofile(f)
            np.asarray(arrays.trnmat, np.float64).tofile(f)
        np.asarray(cstmap, np.int32).tofile(f)
        for k in range(nsteps):
            np.array([mfs[k], chg[k]], np.float64).tofile(f)
            du_steps[k].tofile(f)
            dloads[k].tofile(f)


def read_result(path):
    """List of per-step dicts (status, converged, iters (all attempts),
    last_iters, cuts, pcg_iters, ratio and the six cvtest values (tests)
    of the last sub-increment's iterations, energy, plastic_work, u)."""
    steps = []
    with open(path, "rb") as f:
        nsteps, nodof = np.fromfile(f, np.int32, 2)
        for _ in range(nsteps):
            head = np.fromfile(f, np.int32, 5)
            if head.size < 5:
                break
            status, conv, iters, last, cuts = (int(x) for x in head)
            pcg = np.fromfile(f, np.int32, last)
            ratio = np.fromfile(f, np.float64, last)
            tests = np.fromfile(f, np.float64, 6 * last).reshape(last, 6)
            en, pw = np.fromfile(f, np.float64, 2)
            u = np.fromfile(f, np.float64, nodof)
            steps.append(dict(status=status, converged=bool(conv),
                              iters=iters, last_iters=last, cuts=cuts,
                              pcg_iters=pcg, ratio=ratio, tests=tests,
                              energy=en,
                              plastic_work=pw, u=u))
    return steps


def run(deck_path, result_path, device=0, timeout=600):
    """Run the Fortran driver on a deck (a child process)."""
    if not os.path.exists(DRIVE):
        raise RuntimeError("w3d_drive is not built (python -m "
                           "warp3d_amd.build)")
    p = subprocess.run([DRIVE, deck_path, result_path, str(device)],
                       capture_output=True, text=True, timeout=timeout)
    if p.returncode != 0:
        raise RuntimeError(f"w3d_drive failed ({p.returncode}):\n{p.stdout}"
                           f"\n{p.stderr}")
    return p.stdout, read_result(result_path)
