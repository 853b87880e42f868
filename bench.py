#!/usr/bin/env python3
"""Element-block hot-path benchmark on one MI355X.

Workloads (all 1,728,000 Gauss points):
  hex8  (default): 60x60x60 8-node bricks (216,000 elements, 2x2x2 Gauss
        points with B-bar), interior nodes jittered +-5 % of the edge
        (seeded), mm01 bilinear Mises (E = 200 GPa, nu = 0.3, sigma_y =
        350 MPa, H' = E/100), fixed base, a seeded top-face displacement
        field spread through the height, 0.15 % mean axial strain per step
        of a 10-step ramp to 1.5 %;
  hex20: 40x40x40 20-node bricks (64,000 elements, 3x3x3 Gauss points),
        mm01 (E = 70 GPa, nu = 0.33, sigma_y = 300 MPa, H' = E/50), a
        bending-style field (axial strain varying linearly through the
        thickness, 1 % peak over 15 steps).
  mixed: 150,000 elements in blocks of seeded spans 32-256 (1,935,000 Gauss
        points): 75,000 hex8 (2x2x2, B-bar), 45,000 hex20 (3x3x3), 30,000
        tet10 (4 points), three boxes side by side; mm10 fcc crystal
        plasticity with Voce hardening (tau_y = 50 MPa, tau_v = 100 MPa, so
        the slip resistance saturates at 150 MPa; theta_0 = 250 MPa,
        m = 1, n = 20; cubic elasticity of copper: e = 66.7 GPa, nu = 0.42,
        mu = 75 GPa), seeded SO(3)-uniform orientations, 0.1 % axial strain
        per step of a 10-step ramp to 1 %.

One step = one Newton iteration of the element-block hot path of load step
2 (after a step-1 pseudo-update, update and commit, so every history word is
read), all device-resident: strain/stress recovery + material update +
element internal forces, the assembly of the element forces, the block
tangents and their assembly into the VSS sparse matrix.  The metric is
Gauss-point stress+tangent updates per second; `value` is the path with the
FP32 tangent (the library's mixed-precision path), reported beside the plain
FP64 path measured in the same run.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                    [--workload hex8|hex20|mixed|rve]

The rve workload serves a stream of RVE requests instead: --steps is the
number of requests timed, --warmup the requests served before; see
run_rve.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from warp3d_amd.workloads import WORKLOADS, build_workload  # noqa: E402

HBM_PEAK_GBS = 8000.0         # MI355X HBM3E


def kernel_name(phase, workload, fp32=True):
    """The kernel (as the rocprofv3 summaries name it) of a phase of the
    reported FP32-tangent path."""
    tmpl = {"hex8": "8, 8, true", "hex20": "20, 27, false",
            "mixed": "8, 8, true"}[workload]
    prec = "float" if fp32 else "double"
    return {"recover": f"k_recover<{tmpl}, 0, false>",
            "ifv_gather": "k_ifv_gather",
            "tangent": f"k_tangent_mfma<{tmpl}, {prec}, false, false>",
            "assemble": f"k_assemble<{prec}>",
            "energy_reduce": "k_energy_reduce",
            "mm10": "k_mm10"}[phase]
FP64_PEAK_TFS = 78.6          # MI355X FP64 (vector and matrix)
FP32_MFMA_PEAK_TFS = 157.3    # MI355X dense FP32 matrix


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic(parts, nonode, ncontrib, nout, ninv, ke_bytes):
    """Algorithmic work per launch of each phase (bytes or flops), summed
    over the element parts (count, nodes, Gauss points, mm10)."""
    rec = tan = mm10 = 0
    for ne, nn, ngp_e, is_mm10 in parts:
        T = 3 * nn
        utsz = T * (T + 1) // 2
        ngp = ne * ngp_e
        # per element: coordinates and increments gathered (2 x T words),
        # connectivity (NN ints), element forces T words and volume 1 word
        elem = ne * ((3 * T + 1) * 8 + 4 * nn)
        if is_mm10:
            # two passes: eps_n read, eps_n1 and the increments written;
            # then sigma_n1 (8 words) read for the forces (and the element
            # gathers twice)
            rec += ngp * (6 + 6 + 6 + 8) * 8 + 2 * elem
            # history 158 words read and written, sigma_n and the
            # increments read, sigma_n1 and [D] written
            mm10 += ngp * (2 * 158 + 9 + 6 + 9 + 21) * 8
        else:
            # per point: sigma_n 9 + eps_n 6 + hist_n 11 read, sigma 9 +
            # eps 6 + hist 11 + [D] 21 written = 73 words
            rec += ngp * 73 * 8 + elem
        # upper-triangle B^T D B as bdbtgen forms it: 2(36 T + 6 utsz)
        tan += ngp * 2 * (36 * T + 6 * utsz)
    out = {
        "recover": ("hbm", rec),
        # per (node, element) pair: its 8-byte offset and 3 force words
        # read; per structure node 3 words written
        "ifv_gather": ("hbm", ninv * (8 + 24) + nonode * 24),
        "tangent": ("flop", tan),
        # the unavoidable traffic: every Ke word read once, every assembled
        # word written once (the segment map's headers and indices, about
        # 1.5 bytes per Ke word, count as extra traffic)
        "assemble": ("hbm", ncontrib * ke_bytes + nout * 8),
    }
    if mm10:
        out["mm10"] = ("hbm", mm10)
    return out


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of a kernel from the committed rocprofv3
    counter summary (profiles/hbm_traffic.md): FETCH_SIZE scaled by the
    copy-kernel calibration factor stated there, plus WRITE_SIZE.  None
    when the summary has no row for the kernel and workload."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.md")
    if not os.path.exists(path):
        return None
    fetch_scale = None
    for line in open(path):
        if line.startswith("fetch_scale:"):
            fetch_scale = float(line.split(":")[1])
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if (fetch_scale and len(cells) >= 4 and cells[0] == workload and
                cells[1] == kernel):
            try:
                return (float(cells[2]) * fetch_scale + float(cells[3])) * 1024
            except ValueError:
                continue
    return None


def run_device(arr, cst, du, steps, warmup, precision, asym=False,
               graph=False):
    """asym: asymmetric assembly (full FP64 Ke, CSR rows) in place of the
    symmetric tangent and VSS assembly.  graph: also time the same
    iterations as one HIP graph launch each (w3d_iteration_dev), without
    the per-phase events; info["graph_ms"]."""
    from warp3d_amd import hip
    from warp3d_amd.model import BlockModel
    dev = BlockModel(arr)
    dev.set_asymmetric(asym)
    neq, nco = dev.pattern(cst)
    nnz = dev.csr_pattern()[1].size if asym else 0
    d_val = hip.DeviceArray(max(nnz, 1)) if asym else None
    d_du = hip.DeviceArray.from_host(du)
    d_u = hip.DeviceArray(du.size)
    d_ifv = hip.DeviceArray(du.size)
    d_kd = hip.DeviceArray(neq)
    d_kc = hip.DeviceArray(max(nco, 1))
    # load step 1 (pseudo-update, tangent, one update) committed, then the
    # timed iterations are those of load step 2 from u = du
    dev.recover_dev(1, 0, d_u.ptr, d_du.ptr, d_ifv.ptr)
    dev.tangent(1, 0, precision)
    dev.recover_dev(1, 1, d_u.ptr, d_du.ptr, d_ifv.ptr)
    dev.commit_step()
    d_u.upload(du)
    dev.recover_dev(2, 0, d_u.ptr, d_du.ptr, d_ifv.ptr)
    dev.tangent(2, 0, precision)

    def one(it):
        dev.recover_dev(2, it, d_u.ptr, d_du.ptr, d_ifv.ptr)
        dev.tangent(2, it, precision)
        if asym:
            dev.assemble_csr_dev(d_val.ptr)
        else:
            dev.assemble_vss_dev(d_kd.ptr, d_kc.ptr)

    for w in range(warmup):
        one(1 + w)
    hip.synchronize()
    dev.set_profiling(True)
    dev.reset_phase_times()
    t0 = time.perf_counter()
    for s in range(steps):
        one(1 + warmup + s)
    hip.synchronize()
    t1 = time.perf_counter()
    phases = dev.phase_times()
    dev.set_profiling(False)
    ifv = d_ifv.download()
    info = {"neqns": neq, "ncoeff": nco, "nnz": nnz,
            "ifv_norm": float(np.linalg.norm(ifv))}
    if not asym:
        info["k_coeffs"] = d_kc.download()
        info["k_diag"] = d_kd.download()
    if graph and not asym:
        it0 = 1 + warmup + steps
        for w in range(2):   # capture, then one replay
            dev.iteration_dev(2, it0 + w, d_u.ptr, d_du.ptr, d_ifv.ptr,
                              precision, d_kd.ptr, d_kc.ptr)
        hip.synchronize()
        g0 = time.perf_counter()
        for s in range(steps):
            dev.iteration_dev(2, it0 + 2 + s, d_u.ptr, d_du.ptr, d_ifv.ptr,
                              precision, d_kd.ptr, d_kc.ptr)
        hip.synchronize()
        info["graph_ms"] = (time.perf_counter() - g0) * 1e3 / steps
    return (t1 - t0) * 1e3 / steps, phases, info


def contrib_counts(arr, cst):
    """Ke words the assembly reads: per element, the pairs (erow, ecol)
    with both dofs free and scol >= srow (every free pair once)."""
    from warp3d_amd import _abi
    total = 0
    for et in np.unique(arr.etype):
        nn = _abi.NNODE[int(et)]
        sel = np.nonzero(arr.etype == et)[0]
        idx = arr.incmap[sel][:, None] - 1 + np.arange(nn)[None, :]
        conn = arr.incid[idx] - 1
        dofs = np.concatenate([conn * 3, conn * 3 + 1, conn * 3 + 2], axis=1)
        nf = (cst[dofs] == 0).sum(axis=1)
        total += int(np.sum(nf * (nf + 1) // 2))
    return total


# FP64 flops of the mm10 local solves, per unit of work counted by the
# oracle's instrumented restatement (orc_mm10_counters): the adds,
# multiplies and divides of each routine as the reference forms them (an
# FMA counts 2); pow is counted apart.  formR1: the 12 resolved shears and
# slip increments (180), form_dbarp (180), form_wp (348), symsw (18), C w1
# (66), the residual (30); formR adds the Voce sum (34); formJ11: the slips
# (180) and per system dgdt, symsw, 2 sym(sigma q) + C ms and the rank-one
# update (2,340 in all), form_wp (348), the spin and identity terms (24);
# formJ adds J12 (1,680), J21 and J22 (360); the 6 x 6 and 7 x 7 LU
# solves with J^T R (270, 400).
MM10_FLOPS = {"formR1": 822, "formR_extra": 34, "formJ11": 2892,
              "formJ_extra": 2040, "lu6": 270, "lu7": 400}
MM10_POW = {"formR1": 12, "formR_extra": 1, "formJ11": 25,
            "formJ_extra": 26}


def mm10_work(cnt):
    """(flops, pows) of the local solves from the oracle's counters."""
    r1, r, j11, j, lu6, lu7 = cnt[2], cnt[3], cnt[4], cnt[5], cnt[6], cnt[7]
    f = (r1 * MM10_FLOPS["formR1"] + r * MM10_FLOPS["formR_extra"] +
         j11 * MM10_FLOPS["formJ11"] + j * MM10_FLOPS["formJ_extra"] +
         lu6 * MM10_FLOPS["lu6"] + lu7 * MM10_FLOPS["lu7"])
    pw = (r1 * MM10_POW["formR1"] + r * MM10_POW["formR_extra"] +
          j11 * MM10_POW["formJ11"] + j * MM10_POW["formJ_extra"])
    return f, pw


def cpu_baseline(arr, du, ngp, budget_s=10.0):
    from oracle import oracle as O
    threads = len(os.sched_getaffinity(0))
    om = O.OracleModel(arr, threads=threads)
    u = np.zeros_like(du)
    om.recover(1, 0, u, du)
    om.recover(1, 1, u, du)
    om.commit_step()
    u = du.copy()
    om.recover(2, 0, u, du)
    om.tangent(2, 0)
    import ctypes as C
    L = O.lib()
    L.orc_mm10_counters.argtypes = [C.POINTER(C.c_longlong),
                                    C.POINTER(C.c_longlong), C.c_int]
    cnt = (C.c_longlong * 8)()
    L.orc_mm10_counters(cnt, None, 1)
    it = 0
    t0 = time.perf_counter()
    while True:
        it += 1
        om.recover(2, it, u, du)
        om.tangent(2, it)
        if time.perf_counter() - t0 > budget_s and (it >= 2 or
                                                    budget_s <= 0):
            break
    t1 = time.perf_counter()
    L.orc_mm10_counters(cnt, None, 1)
    out = {"value": ngp * it / (t1 - t0), "unit": "GP updates/s",
           "cores": threads, "kind": "port",
           "sample": f"oracle (C, OpenMP over blocks of 128) recovery + "
                     f"material update + element forces + FP64 tangent on "
                     f"the full workload at load step 2, {it} Newton "
                     f"iterations in {t1 - t0:.1f} s"}
    if cnt[0]:
        out["mm10_counts_per_iteration"] = [c / it for c in cnt]
    return out


def c1_precision(tmp):
    """The FP32 tangent's cost in Newton iterations: the C1 load history
    (10^3 hex8, 20 steps to 2 %) through the Fortran driver with the FP64
    and the FP32 tangent, both converged to a relative residual of 1e-7."""
    from tests._util import c1_model
    from warp3d_amd import _abi
    from warp3d_amd import deck as D
    from warp3d_amd import mesh as M
    m, arr = c1_model(10)
    cst, top, val = M.uniaxial_bc(m, 0.02 / 20)
    du = np.zeros(3 * m.nonode)
    du[top] = val
    steps = np.array([du] * 20)
    res = {}
    for prec in (_abi.W3D_FP64, _abi.W3D_FP32):
        deck = os.path.join(tmp, f"c1_{prec}.deck")
        out = os.path.join(tmp, f"c1_{prec}.res")
        D.write_deck(deck, arr, cst, steps, precision=prec, pcg_rtol=1e-13,
                     res_tol=1e-5, extrapolate=False)
        _, res[prec] = D.run(deck, out)
    r64, r32 = res[_abi.W3D_FP64], res[_abi.W3D_FP32]
    extra = [b["iters"] - a["iters"] for a, b in zip(r64, r32)]
    du_rel = max(float(np.abs(b["u"] - a["u"]).max() / np.abs(a["u"]).max())
                 for a, b in zip(r64, r32))
    return {"config": "C1: 10^3 hex8 B-bar, mm01, 20 steps to 2 %, "
                      "relative residual 1e-7",
            "newton_iters_fp64": int(sum(a["iters"] for a in r64)),
            "newton_iters_fp32": int(sum(b["iters"] for b in r32)),
            "extra_iters_per_step_max": int(max(extra)),
            "u_rel_diff_max": du_rel,
            "all_converged": all(a["converged"] and b["converged"]
                                 for a, b in zip(r64, r32))}


def run_rve(args):
    """RVE serving: --warmup requests served first (untimed), then a
    Poisson stream of --steps requests through the continuous-batching
    server; beside it --plain-sample requests of the same stream solved one
    at a time through the plain path (FP64 tangent, host-driven Newton and
    Jacobi-PCG).  value = RVE solves/s of the server over the stream."""
    from warp3d_amd import _abi
    from warp3d_amd import serve as S
    opts = S.newton_opts(precision=_abi.W3D_FP32)
    eng = S.DeviceEngine(opts)
    try:
        if args.warmup:
            warm = S.poisson_arrivals(args.warmup, rate=0.0, nmin=8,
                                      nmax=12, seed=99)
            S.Server(eng, max_slots=args.slots).run(warm)
            log(f"rve warmup: {args.warmup} requests")
        arr = S.poisson_arrivals(args.steps, rate=args.rate, nmin=8,
                                 nmax=args.nmax, seed=1)
        srv = S.Server(eng, max_slots=args.slots, hbm_budget=64 << 30)
        ph0, na0, np0 = eng.batch.phase_times()
        done, st = srv.run(arr)
        ph1, na1, np1 = eng.batch.phase_times()
    finally:
        eng.close()
    phases = {k: round(ph1[k] - ph0[k], 1) for k in ph1}
    log("server phases (device ms over the stream): " + json.dumps(phases)
        + f", advances {na1 - na0}, batched PCG iterations {np1 - np0}")
    log("server: " + json.dumps(st))
    newton = [r["newton"] for j in done if j.result
              for r in j.result["records"]]
    pcg = [r["pcg"] for j in done if j.result for r in j.result["records"]]
    # the plain path on a sample of the same stream: the requests at evenly
    # spaced size quantiles, so that its mean time per request estimates
    # the stream's
    by_size = sorted(range(len(arr)), key=lambda k: arr[k][1].n)
    k = args.plain_sample
    pick = [by_size[int((i + 0.5) * len(arr) / k)] for i in range(k)] \
        if k and arr else []
    plain_t, plain_steps = [], []
    popts = S.newton_opts(precision=_abi.W3D_FP64)
    plain_failed = 0
    for i in pick:
        spec = arr[i][1]
        t0 = time.perf_counter()
        try:
            _, _, times = S.solve_plain(S.make_rve(spec), popts)
        except RuntimeError as e:      # counted, and its time kept
            plain_failed += 1
            times = []
            log(f"plain n={spec.n}: {e}")
        plain_t.append(time.perf_counter() - t0)
        plain_steps += times
        log(f"plain n={spec.n}: {plain_t[-1]:.2f} s")
    mean_elems_stream = np.mean([s.noelem for _, s in arr])
    plain_rate = (len(plain_t) - plain_failed) / sum(plain_t) \
        if plain_t else None
    out = {
        "metric": "RVE solves/s at p99 load-step latency",
        "value": round(st["solves_per_s"], 3),
        "unit": "RVE solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * st["wall_s"] / max(args.steps, 1), 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (recovery, mm10, residual, Jacobi-PCG) + f32 (tangent)",
        "data": "synthetic (seeded Voronoi grains, orientations, sizes and "
                "arrival times)",
        "config": {"workload": WORKLOADS["rve"], "requests": args.steps,
                   "rate_per_s": args.rate, "slots": args.slots,
                   "n_range": [8, args.nmax],
                   "mean_elements": round(float(mean_elems_stream), 1)},
        "latency": {"step_p50_ms": st["step_p50_ms"],
                    "step_p99_ms": st["step_p99_ms"],
                    "request_p50_s": st["request_p50_s"],
                    "request_p99_s": st["request_p99_s"]},
        "server": {k: st[k] for k in ("solved", "failed", "wall_s",
                                      "iterations", "peak_slots",
                                      "peak_bytes", "steps")},
        "server_phase_ms": phases,
        "batched_pcg_iterations": np1 - np0,
        "newton_iters_per_step": round(float(np.mean(newton)), 3)
        if newton else None,
        "pcg_iters_per_step": round(float(np.mean(pcg)), 1) if pcg else None,
        "plain": {"sample_n": [arr[i][1].n for i in pick],
                  "failed": plain_failed,
                  "solves_per_s": plain_rate,
                  "step_p50_ms": 1e3 * float(np.percentile(plain_steps, 50))
                  if plain_steps else None,
                  "step_p99_ms": 1e3 * float(np.percentile(plain_steps, 99))
                  if plain_steps else None},
        "speedup_vs_plain": round(st["solves_per_s"] / plain_rate, 2)
        if plain_rate else None,
        "newton_opts": {"tol": opts.tol, "max_iters": opts.max_iters,
                        "pcg_rtol": opts.pcg_rtol,
                        "adaptive": bool(opts.adaptive)},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="hex8")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the hex20 line and the C1 Newton precision "
                         "runs that the hex8 line carries")
    ap.add_argument("--asym", action="store_true",
                    help="time the asymmetric-assembly iteration (full FP64 "
                         "Ke and CSR rows) on the workload instead")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run 3 device-to-device copies of MIB MiB "
                         "(calibrates HBM byte counters against a copy of "
                         "known size)")
    ap.add_argument("--rate", type=float, default=100.0,
                    help="rve: request arrivals per second")
    ap.add_argument("--slots", type=int, default=16,
                    help="rve: most requests on the device at once")
    ap.add_argument("--nmax", type=int, default=24,
                    help="rve: largest n of an n^3 request")
    ap.add_argument("--plain-sample", type=int, default=3,
                    help="rve: requests timed through the plain path")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not "
              "supported (the element-block path runs on a single device "
              "here; see DESIGN.md for how it would shard)", file=sys.stderr)
        return 2
    from warp3d_amd import _abi, hip
    if hip.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    if args.workload == "rve":
        return run_rve(args)
    if args.copy_probe:
        nb = args.copy_probe << 20
        a = hip.DeviceArray(nb, np.uint8)
        b = hip.DeviceArray(nb, np.uint8)
        for _ in range(3):
            hip.copy_d2d(b.ptr, a.ptr, nb)
        hip.synchronize()
        a.free()
        b.free()
        log(f"copy probe: 3 x {args.copy_probe} MiB device-to-device")
    t = time.perf_counter()
    m, arr, cst, du, parts = build_workload(args.workload)
    ngp = sum(ne * g for ne, _, g, _ in parts)
    log(f"workload {args.workload}: {m.noelem} elements, {ngp} points, "
        f"{m.nonode} nodes ({time.perf_counter() - t:.1f} s)")
    if args.asym:
        ms, phases, info = run_device(arr, cst, du, args.steps, args.warmup,
                                      _abi.W3D_FP64, asym=True)
        ph = {k: round(v[1] / max(v[0], 1), 4) for k, v in phases.items()
              if v[0]}
        log(f"asymmetric: {ms:.3f} ms/step  phases {ph}")
        print(json.dumps({
            "metric": "Gauss-pt stress+tangent updates/s, asymmetric "
                      "assembly", "value": round(ngp / (ms * 1e-3), 1),
            "unit": "GP updates/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 4),
            "higher_is_better": True, "dtype": "f64",
            "data": "synthetic (seeded jittered mesh and displacement field)",
            "config": {"workload": args.workload, "gauss_points": ngp,
                       "csr_nnz": int(info.get("nnz", 0))},
            "phase_ms": ph}), flush=True)
        return 0
    res = {}
    for prec, name in ((_abi.W3D_FP64, "fp64"), (_abi.W3D_FP32, "fp32")):
        t = time.perf_counter()
        ms, phases, info = run_device(arr, cst, du, args.steps, args.warmup,
                                      prec, graph=True)
        res[name] = (ms, phases, info)
        log(f"{name}: {ms:.3f} ms/step  phases "
            + ", ".join(f"{k} {v[1] / max(v[0], 1):.3f} ms" for k, v in
                        phases.items() if v[0]) +
            f"  ({time.perf_counter() - t:.1f} s)")
    ms32, ph32, info = res["fp32"]
    ms64, ph64, _ = res["fp64"]
    ninv = int(arr.incid.size)
    ncontrib = contrib_counts(arr, cst)
    nout = info["neqns"] + info["ncoeff"]
    alg = algorithmic(parts, m.nonode, ncontrib, nout, ninv, 4)
    # dominant kernel of the reported path
    avg = {k: v[1] / v[0] for k, v in ph32.items() if v[0]}
    dom = max(avg, key=avg.get)
    kind, work = alg.get(dom, ("hbm", 0))
    dur_s = avg[dom] * 1e-3
    if kind == "hbm":
        achieved = work / dur_s / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": profiled_traffic(kernel_name(dom, args.workload),
                                            args.workload)}
    else:
        achieved = work / dur_s / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 2),
                "peak": FP32_MFMA_PEAK_TFS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_MFMA_PEAK_TFS, 4),
                "traffic": profiled_traffic(kernel_name(dom, args.workload),
                                            args.workload)}
    roof["kernel"] = kernel_name(dom, args.workload)
    phase_report = {}
    for tag, ph, kb in (("fp32", ph32, 4), ("fp64", ph64, 8)):
        a = algorithmic(parts, m.nonode, ncontrib, nout, ninv, kb)
        rep = {}
        for k, (n, tot) in ph.items():
            if not n:
                continue
            d = tot / n
            entry = {"ms": round(d, 4)}
            if k in a:
                kd, w = a[k]
                if kd == "hbm":
                    entry["GB/s"] = round(w / (d * 1e-3) / 1e9, 1)
                else:
                    entry["TFLOP/s"] = round(w / (d * 1e-3) / 1e12, 2)
            rep[k] = entry
        phase_report[tag] = rep
    out = {
        "metric": "Gauss-pt stress+tangent updates/s",
        "value": round(ngp / (ms32 * 1e-3), 1),
        "unit": "GP updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (recovery, forces, assembly) + f32 (tangent)",
        "data": "synthetic (seeded jittered mesh and displacement field)",
        "config": {"workload": WORKLOADS[args.workload],
                   "elements": m.noelem, "gauss_points": ngp,
                   "equations": info["neqns"], "matrix_terms": info["ncoeff"]},
        "plain_fp64": {"value": round(ngp / (ms64 * 1e-3), 1),
                       "ms_per_step": round(ms64, 4)},
        "speedup_vs_fp64": round(ms64 / ms32, 3),
        "roofline": roof,
        "phases": phase_report,
    }
    # the FP32 tangent against the FP64 one on this workload (assembled
    # matrix, norm-wise), and the bound on the FP32/FP64 ratio: only the
    # tangent (FP32 MFMA peak 157.3 against FP64 78.6 TF/s on MI355X) and
    # the Ke words the assembly reads (4 against 8 bytes) change
    i64 = res["fp64"][2]
    kd_err = float(np.abs(info["k_diag"] - i64["k_diag"]).max() /
                   np.abs(i64["k_diag"]).max())
    kc_err = float(np.abs(info["k_coeffs"] - i64["k_coeffs"]).max() /
                   np.abs(i64["k_coeffs"]).max())
    out["fp32_tangent"] = {"k_diag_rel_err": kd_err,
                           "k_coeffs_rel_err": kc_err}
    out["speedup_bound_fp32_tangent"] = round(
        precision_bound(ph64, alg, parts, m.nonode, ncontrib, nout, ninv), 3)
    # the same iteration as one HIP graph launch, without phase events
    g32 = info.get("graph_ms")
    if g32:
        out["graph"] = {"ms_per_step": round(g32, 4),
                        "value": round(ngp / (g32 * 1e-3), 1),
                        "fp64_ms_per_step": round(
                            res["fp64"][2].get("graph_ms", 0.0), 4)}
    if args.workload == "hex8" and not args.no_extra:
        out["fp32_tangent"].update(c1_precision_safe())
        out["hex20"] = extra_workload("hex20", args)
    mm = ph32.get("mm10")
    if not args.no_cpu:
        t = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(arr, du, ngp)
        log(f"cpu baseline done ({time.perf_counter() - t:.1f} s)")
        cpi = out["cpu_baseline"].pop("mm10_counts_per_iteration", None)
        if mm and mm[0] and cpi:
            f, pw = mm10_work(cpi)
            ms = mm[1] / mm[0]
            out["mm10"] = {
                "ms": round(ms, 3),
                "fp64_flops_per_launch": round(f),
                "pow_per_launch": round(pw),
                "fp64_TFLOP/s": round(f / (ms * 1e-3) / 1e12, 3),
                "frac_fp64_peak": round(f / (ms * 1e-3) / 1e12 /
                                        FP64_PEAK_TFS, 4),
                "local_solve_counts_per_point": {
                    k: round(cpi[i] / cpi[0], 3) for i, k in enumerate(
                        ("crystal_updates", "substeps", "formR1", "formR",
                         "formJ11", "formJ", "lu6", "lu7")) if i},
                "flops_per_point": round(f / cpi[0])}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


def precision_bound(ph64, alg, parts, nonode, ncontrib, nout, ninv):
    """ms64 over the best FP32 path: the tangent at the FP32/FP64 MFMA peak
    ratio and the assembly at its 4-byte/8-byte Ke traffic ratio, the
    other phases unchanged."""
    tot = sum(v[1] / v[0] for v in ph64.values() if v[0])
    t = {k: v[1] / v[0] for k, v in ph64.items() if v[0]}
    a64 = algorithmic(parts, nonode, ncontrib, nout, ninv, 8)["assemble"][1]
    a32 = algorithmic(parts, nonode, ncontrib, nout, ninv, 4)["assemble"][1]
    best = (tot - t.get("tangent", 0) - t.get("assemble", 0) +
            t.get("tangent", 0) * FP64_PEAK_TFS / FP32_MFMA_PEAK_TFS +
            t.get("assemble", 0) * a32 / a64)
    return tot / best


def c1_precision_safe():
    import tempfile
    try:
        with tempfile.TemporaryDirectory() as tmp:
            return c1_precision(tmp)
    except Exception as e:  # reported, not fatal to the line
        return {"c1_error": str(e)[:300]}


def extra_workload(name, args):
    """Another workload of the same metric in the same run (FP64 and FP32
    paths), for the line's extra key."""
    from warp3d_amd import _abi
    m, arr, cst, du, parts = build_workload(name)
    ngp = sum(ne * g for ne, _, g, _ in parts)
    r = {}
    for prec, tag in ((_abi.W3D_FP64, "fp64"), (_abi.W3D_FP32, "fp32")):
        ms, phases, info = run_device(arr, cst, du, args.steps, args.warmup,
                                      prec)
        r[tag] = (ms, phases, info)
        log(f"{name} {tag}: {ms:.3f} ms/step")
    ms32, ph32, info = r["fp32"]
    ms64, ph64, _ = r["fp64"]
    ninv = int(arr.incid.size)
    ncontrib = contrib_counts(arr, cst)
    nout = info["neqns"] + info["ncoeff"]
    alg = algorithmic(parts, m.nonode, ncontrib, nout, ninv, 4)
    tan = ph32.get("tangent")
    out = {"workload": WORKLOADS[name], "gauss_points": ngp,
           "value": round(ngp / (ms32 * 1e-3), 1),
           "ms_per_step": round(ms32, 4),
           "plain_fp64": {"value": round(ngp / (ms64 * 1e-3), 1),
                          "ms_per_step": round(ms64, 4)},
           "speedup_vs_fp64": round(ms64 / ms32, 3),
           "speedup_bound_fp32_tangent": round(precision_bound(
               ph64, alg, parts, m.nonode, ncontrib, nout, ninv), 3),
           "k_coeffs_rel_err_fp32": float(
               np.abs(info["k_coeffs"] - r["fp64"][2]["k_coeffs"]).max() /
               np.abs(r["fp64"][2]["k_coeffs"]).max())}
    if tan and tan[0]:
        d = tan[1] / tan[0] * 1e-3
        out["tangent_TFLOP/s"] = round(alg["tangent"][1] / d / 1e12, 2)
        out["tangent_frac_fp32_mfma"] = round(
            alg["tangent"][1] / d / 1e12 / FP32_MFMA_PEAK_TFS, 4)
    return out


if __name__ == "__main__":
    sys.exit(main())
