"""Continuous-batching server of crystal-plasticity RVE solves on one device.

A request is a representative volume element: the unit cube meshed with
n^3 8-node hexahedra, its elements split among Voronoi grains of seeded
centres, each grain one FCC crystal of an SO(3)-uniform orientation (mm10,
Voce hardening with the constants of the reference's Voce verification
case), pulled along z in equal strain steps with symmetry planes
x = 0, y = 0, z = 0 (the "mixed" boundary conditions of a strain-controlled
RVE).  Its answer is the load (the z force on the top face) after every
step and the final displacements.

The server keeps every admitted RVE's mesh, state, displacements and solver
vectors on the device (one w3d_problem each) and runs Newton iterations of
all of them together (w3d_batch_advance): requests that arrive while others
are on the device are admitted as slots and HBM free up, and join the next
iteration.  Admission is first come, first served, limited by a slot count
and by an HBM budget (device bytes of the admitted problems).  Set-up of
admitted requests (mesh, blocking, sparsity pattern) runs on host worker
threads beside the device iterations.

The plain path (solve_plain) is one request at a time through the
per-call entry points, FP64 tangent, Newton bookkeeping on the host and the
host-driven Jacobi-PCG: the same equations and tolerances.

No reference counterpart: the reference solves one model per process.
"""
import ctypes as C
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from . import mesh as M

# The Voce constants of the reference's own Voce verification deck
# (verification/test82/voche_model/warp3d_voche.inp:6-10: isotropic
# elasticity E = 150 GPa, nu = 0.28, tau_y = 300 MPa, tau_v = 1200 MPa,
# theta_0 = 400 MPa, m = 5, n = 20) with fcc slip systems.  (The copper
# constants of the mixed workload make the crystal update ask for cuts at
# 1/16 of a 0.2 % step in some RVEs.)
RVE_MATERIAL = dict(tau_y=300.0, tau_v=1200.0, theta_0=400.0, voce_m=5.0,
                    harden_n=20.0, elastic="isotropic", e=150.0e3, nu=0.28)


@dataclass
class RveSpec:
    """One request: n^3 elements, seed of its grains and orientations."""
    n: int
    seed: int
    ngrains: int = 64
    strain: float = 0.02
    nsteps: int = 10

    @property
    def noelem(self):
        return self.n ** 3


@dataclass
class Rve:
    spec: RveSpec
    mesh: object
    arrays: object
    cstmap: np.ndarray
    du0: np.ndarray
    loaded: np.ndarray   # the top-face z dofs


def grains(mesh, ngrains, seed):
    """Grain of every element: the nearest of ngrains seeded centres
    (uniform in the unit cube) to the element's centroid."""
    rng = np.random.default_rng(seed)
    centres = rng.random((ngrains, 3))
    cen = mesh.coords[mesh.conn - 1].mean(axis=1)
    best = np.zeros(mesh.noelem, np.int64)
    bestd = np.full(mesh.noelem, np.inf)
    for g in range(ngrains):
        d = ((cen - centres[g]) ** 2).sum(axis=1)
        closer = d < bestd
        best[closer] = g
        bestd[closer] = d[closer]
    return best


def make_rve(spec, auto_size=128):
    """Mesh, model arrays, constraints and the first step's increment.

    The increment imposes strain/nsteps on the top face and starts the free
    dofs from the homogeneous field (-0.4 x, -0.4 y, z) times that strain,
    which meets the symmetry planes."""
    m = M.box_hex8(spec.n)
    g = grains(m, spec.ngrains, spec.seed)
    ori = M.random_orientations(spec.ngrains, seed=spec.seed + 7919)
    mat = _abi.mm10_material(**RVE_MATERIAL)
    arr = m.model(mat, auto_size=auto_size, angles=ori[g])
    de = spec.strain / spec.nsteps
    cst, top, val = M.uniaxial_bc(m, de)
    du0 = np.zeros(3 * m.nonode)
    x = m.coords
    du0[0::3] = -0.4 * de * x[:, 0]
    du0[1::3] = -0.4 * de * x[:, 1]
    du0[2::3] = de * x[:, 2]
    du0[cst != 0] = 0.0
    du0[top] = val
    return Rve(spec, m, arr, cst, du0, top)


def newton_opts(nsteps=10, precision=_abi.W3D_FP32, tol=1e-4, tol_abs=1e-8,
                max_iters=30, pcg_rtol=1e-6, pcg_maxit=5000, adaptive=True):
    """The serving defaults: the reference's default convergence test
    (0.01 % of the reaction norm), adaptive sub-increments on (the crystal
    update asks for them at 0.2 % strain per step), a symmetric tangent,
    so that convergence is linear and needs more iterations than the
    reference's default limit of 10 (up to 30 per increment)."""
    return _abi.W3DNewtonOpts(nsteps=nsteps, max_iters=max_iters,
                              precision=precision, pcg_maxit=pcg_maxit,
                              adaptive=int(bool(adaptive)), reserved=0,
                              tol=tol, tol_abs=tol_abs, pcg_rtol=pcg_rtol)


# ---------------------------------------------------------------------
# device objects
# ---------------------------------------------------------------------
class Problem:
    """One RVE resident on the device (w3d_problem over its context)."""

    def __init__(self, rve, opts, device=0):
        from .model import BlockModel, _chk, lib
        self._lib, self._chk = lib, _chk
        self.rve = rve
        self.dev = BlockModel(rve.arrays, device)
        h = C.c_void_p()
        _chk(lib().w3d_problem_create(
            self.dev.h, _abi.iptr(np.ascontiguousarray(rve.cstmap, np.int32)),
            _abi.dptr(np.ascontiguousarray(rve.du0, float)), C.byref(opts),
            C.byref(h)), "w3d_problem_create")
        self.h = h
        self.nsteps = opts.nsteps

    def status(self):
        st, step, it = C.c_int32(), C.c_int32(), C.c_int32()
        self._chk(self._lib().w3d_problem_status(self.h, C.byref(st),
                                                 C.byref(step), C.byref(it)),
                  "w3d_problem_status")
        return st.value, step.value, it.value

    def records(self):
        """Per converged step: Newton iterations (all attempts), PCG
        iterations, adaptive cuts, load, residual and reaction norms."""
        out = []
        for s in range(1, self.nsteps + 1):
            ni, pi, cu = C.c_int32(), C.c_int32(), C.c_int32()
            ld, rs, rc = C.c_double(), C.c_double(), C.c_double()
            r = self._lib().w3d_problem_step_record(
                self.h, s, C.byref(ni), C.byref(pi), C.byref(cu), C.byref(ld),
                C.byref(rs), C.byref(rc))
            if r != 0:
                break
            out.append(dict(newton=ni.value, pcg=pi.value, cuts=cu.value,
                            load=ld.value, residual=rs.value,
                            reaction=rc.value))
        return out

    def u(self):
        u = np.zeros(3 * self.rve.mesh.nonode)
        self._chk(self._lib().w3d_problem_get_u(self.h, _abi.dptr(u)),
                  "w3d_problem_get_u")
        return u

    def device_bytes(self):
        b = C.c_int64()
        self._chk(self._lib().w3d_problem_device_bytes(self.h, C.byref(b)),
                  "w3d_problem_device_bytes")
        return b.value

    def close(self):
        if getattr(self, "h", None):
            self._lib().w3d_problem_destroy(self.h)
            self.h = None
        if getattr(self, "dev", None) is not None:
            self.dev.close()
            self.dev = None

    def __del__(self):
        self.close()


class Batch:
    """A w3d_batch: one Newton iteration of many problems per advance."""

    def __init__(self, device=0):
        from .model import _chk, lib
        self._lib, self._chk = lib, _chk
        h = C.c_void_p()
        _chk(lib().w3d_batch_create(device, C.byref(h)), "w3d_batch_create")
        self.h = h

    def advance(self, problems):
        n = len(problems)
        arr = (C.c_void_p * max(n, 1))(*[p.h.value for p in problems])
        st = np.zeros(max(n, 1), np.int32)
        self._chk(self._lib().w3d_batch_advance(
            self.h, C.cast(arr, C.POINTER(C.c_void_p)), n, _abi.iptr(st)),
            "w3d_batch_advance")
        return st[:n].tolist()

    PHASES = ("recover", "mm10", "residual", "tangent_assembly", "pcg",
              "update")

    def phase_times(self):
        """Device ms per phase over all advances, the advances and the
        batched PCG iterations."""
        ms = np.zeros(6)
        na, npcg = C.c_int64(), C.c_int64()
        self._chk(self._lib().w3d_batch_phase_times(
            self.h, _abi.dptr(ms), C.byref(na), C.byref(npcg)),
            "w3d_batch_phase_times")
        return dict(zip(self.PHASES, ms.tolist())), na.value, npcg.value

    def close(self):
        if getattr(self, "h", None):
            self._lib().w3d_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class DeviceEngine:
    """What the scheduler drives: set-up, the batched iteration, teardown."""

    def __init__(self, opts, device=0):
        self.opts = opts
        self.device = device
        self.batch = Batch(device)

    def prepare(self, spec):
        return Problem(make_rve(spec), self.opts, self.device)

    def advance(self, problems):
        return self.batch.advance(problems)

    def result(self, prob):
        u = prob.u()
        return dict(records=prob.records(), u=u,
                    device_bytes=prob.device_bytes())

    def release(self, prob):
        prob.close()

    def close(self):
        self.batch.close()


# ---------------------------------------------------------------------
# scheduler
# ---------------------------------------------------------------------
@dataclass
class Job:
    rid: int
    spec: RveSpec
    t_arrive: float
    est_bytes: int = 0
    t_admit: float = 0.0
    t_start: float = 0.0
    t_done: float = 0.0
    step_start: float = 0.0
    step_lat: list = field(default_factory=list)
    status: int = _abi.W3D_PROBLEM_RUNNING
    result: dict = None
    prob: object = None
    future: object = None


class Server:
    """First-come-first-served admission by slots and HBM budget, host
    set-up on worker threads, batched device iterations.

    bytes_per_elem is the admission estimate of a request's device bytes
    per element; after every set-up it rises to the largest ratio
    observed (problem bytes / elements), so the budget is charged at least
    what the problems hold."""

    def __init__(self, engine, max_slots=32, hbm_budget=64 << 30,
                 prep_workers=4, bytes_per_elem=40e3, clock=time.perf_counter,
                 sleep=time.sleep, keep_results=False):
        self.engine = engine
        self.max_slots = max_slots
        self.hbm_budget = hbm_budget
        self.prep_workers = prep_workers
        self.bytes_per_elem = bytes_per_elem
        self.clock = clock
        self.sleep = sleep
        self.keep_results = keep_results

    def estimate(self, spec):
        return int(self.bytes_per_elem * spec.noelem)

    def run(self, arrivals):
        """arrivals: (seconds after start, RveSpec) in any order.  Returns
        the finished jobs (arrival order) and the run's statistics."""
        arrivals = sorted(arrivals, key=lambda a: a[0])
        jobs = [Job(k, spec, t) for k, (t, spec) in enumerate(arrivals)]
        future = deque(jobs)
        waiting, preparing, active, done = deque(), [], [], []
        in_use = 0
        peak_slots, peak_bytes, iterations = 0, 0, 0
        t0 = self.clock()
        pool = ThreadPoolExecutor(max_workers=self.prep_workers)
        try:
            while future or waiting or preparing or active:
                now = self.clock() - t0
                while future and future[0].t_arrive <= now:
                    waiting.append(future.popleft())
                # admission, in arrival order
                while waiting and len(active) + len(preparing) < self.max_slots:
                    j = waiting[0]
                    j.est_bytes = self.estimate(j.spec)
                    if in_use + j.est_bytes > self.hbm_budget and (
                            active or preparing):
                        break
                    waiting.popleft()
                    in_use += j.est_bytes
                    j.t_admit = now
                    j.future = pool.submit(self.engine.prepare, j.spec)
                    preparing.append(j)
                # set-ups that finished join the next iteration
                for j in [j for j in preparing if j.future.done()]:
                    preparing.remove(j)
                    j.prob = j.future.result()
                    j.future = None
                    nb = self._bytes(j)
                    if nb:
                        self.bytes_per_elem = max(self.bytes_per_elem,
                                                  nb / j.spec.noelem)
                        in_use += nb - j.est_bytes
                        j.est_bytes = nb
                    j.t_start = self.clock() - t0
                    j.step_start = j.t_start
                    active.append(j)
                peak_slots = max(peak_slots, len(active))
                peak_bytes = max(peak_bytes, in_use)
                if not active:
                    nxt = [future[0].t_arrive - now] if future else []
                    self.sleep(max(min(nxt + [0.002]), 0.0)
                               if (future or preparing) else 0.0)
                    continue
                status = self.engine.advance([j.prob for j in active])
                iterations += 1
                now = self.clock() - t0
                keep = []
                for j, st in zip(active, status):
                    if st in (_abi.W3D_PROBLEM_STEP_DONE,
                              _abi.W3D_PROBLEM_DONE):
                        j.step_lat.append(now - j.step_start)
                        j.step_start = now
                    if st in (_abi.W3D_PROBLEM_DONE, _abi.W3D_PROBLEM_FAILED):
                        j.status = st
                        j.t_done = now
                        res = self.engine.result(j.prob)
                        j.result = res if self.keep_results else {
                            "records": res["records"]}
                        self.engine.release(j.prob)
                        j.prob = None
                        in_use -= j.est_bytes
                        done.append(j)
                    else:
                        keep.append(j)
                active = keep
        finally:
            for j in active:
                self.engine.release(j.prob)
            for j in preparing:
                try:
                    self.engine.release(j.future.result())
                except Exception:
                    pass
            pool.shutdown(wait=True)
        wall = self.clock() - t0
        done.sort(key=lambda j: j.rid)
        return done, self.stats(done, wall, iterations, peak_slots,
                                peak_bytes)

    def _bytes(self, j):
        try:
            return j.prob.device_bytes()
        except AttributeError:
            return 0

    @staticmethod
    def stats(done, wall, iterations, peak_slots, peak_bytes):
        lat = np.array([x for j in done for x in j.step_lat])
        ok = [j for j in done if j.status == _abi.W3D_PROBLEM_DONE]
        e2e = np.array([j.t_done - j.t_arrive for j in ok])
        return dict(
            requests=len(done), solved=len(ok), failed=len(done) - len(ok),
            wall_s=wall, solves_per_s=len(ok) / wall if wall > 0 else 0.0,
            step_p50_ms=1e3 * float(np.percentile(lat, 50)) if lat.size else
            None,
            step_p99_ms=1e3 * float(np.percentile(lat, 99)) if lat.size else
            None,
            steps=int(lat.size),
            request_p50_s=float(np.percentile(e2e, 50)) if e2e.size else None,
            request_p99_s=float(np.percentile(e2e, 99)) if e2e.size else None,
            iterations=iterations, peak_slots=peak_slots,
            peak_bytes=int(peak_bytes))


def poisson_arrivals(nreq, rate, nmin=8, nmax=24, seed=0, **spec_kw):
    """nreq requests at Poisson rate (per second), n uniform in
    [nmin, nmax], each with its own grain seed."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, nreq)) if rate > 0 else \
        np.zeros(nreq)
    ns = rng.integers(nmin, nmax + 1, nreq)
    return [(float(t[k]), RveSpec(int(ns[k]), seed=int(1000 * seed + k),
                                  **spec_kw)) for k in range(nreq)]


# ---------------------------------------------------------------------
# plain path
# ---------------------------------------------------------------------
def solve_plain(rve, opts, device=0):
    """One RVE through the per-call entry points: w3d_recover (host
    vectors), w3d_tangent, w3d_assemble_vss_dev, w3d_solve_pcg_dev, with
    the residual, the convergence test, the adaptive sub-increments and the
    update on the host; the same algorithm and tolerances as
    w3d_batch_advance.  Returns the per-step records, the final
    displacements and the per-step wall times."""
    from . import hip
    from .model import BlockModel
    from .steps import Adaptive
    dev = BlockModel(rve.arrays, device)
    try:
        neq, nco = dev.pattern(rve.cstmap)
        d, _, _, _ = dev.get_pattern()
        free = np.flatnonzero(rve.cstmap == 0)
        eqdof = free[np.argsort(d[free] - 1)]   # dof of each equation
        cdof = np.flatnonzero(rve.cstmap != 0)
        ldof = np.flatnonzero((rve.cstmap != 0) & (rve.du0 != 0.0))
        imp = rve.du0[cdof]
        kd = hip.DeviceArray(neq)
        kc = hip.DeviceArray(max(nco, 1))
        rhs = hip.DeviceArray(neq)
        x = hip.DeviceArray(neq)
        u = np.zeros(3 * rve.mesh.nonode)
        du = rve.du0.copy()
        dt0 = 1.0
        records, times = [], []

        def start(sf, disp):
            du[eqdof] = sf * du[eqdof]
            du[cdof] = disp * imp
            dev.set_time_increment(dt0 * disp)   # the sub-increment's time

        for step in range(1, opts.nsteps + 1):
            t0 = time.perf_counter()
            ad = Adaptive()
            if step > 1:
                start(1.0, 1.0)
            else:
                dev.set_time_increment(dt0)
            iters = pcg = cuts = 0
            it, extrap = 0, True
            while True:
                # iteration 0 starts an increment: its recovery forms the
                # first residual and stiffness and is never accepted
                if it >= 1:
                    iters += 1
                ifv, _, _, rc = dev.recover(step, it, u, du,
                                            extrapolated=extrap)
                cut = rc != 0
                r = -ifv[eqdof]
                res = float(np.sqrt(np.dot(r, r)))
                react = float(np.sqrt(np.dot(ifv[cdof], ifv[cdof])))
                if not cut and it >= 1 and \
                        res <= opts.tol * max(react, opts.tol_abs):
                    dev.commit_step()
                    u = u + du
                    it, extrap = 0, True
                    if opts.adaptive and ad.advance():
                        start(ad.scaling, ad.disp_fact)
                        continue
                    records.append(dict(newton=iters, pcg=pcg, cuts=cuts,
                                        load=float(ifv[ldof].sum()),
                                        residual=res, reaction=react))
                    break
                if cut or it >= opts.max_iters:
                    if opts.adaptive and ad.cut():
                        dev.rollback()
                        start(0.0, ad.disp_fact)
                        cuts += 1
                        it, extrap = 0, False
                        continue
                    raise RuntimeError(f"step {step}: no convergence")
                dev.tangent(step, it, opts.precision)
                dev.assemble_vss_dev(kd.ptr, kc.ptr)
                rhs.upload(r)
                x.upload(np.zeros(neq))
                n_it, _ = dev.solve_pcg_dev(kd.ptr, kc.ptr, rhs.ptr, x.ptr,
                                            opts.pcg_rtol, opts.pcg_maxit)
                pcg += n_it
                du[eqdof] += x.download()
                it += 1
            times.append(time.perf_counter() - t0)
        for a in (kd, kc, rhs, x):
            a.free()
        return records, u, times
    finally:
        dev.close()
