"""Host logic of the RVE server: the RVE builder (Voronoi grains, boundary
conditions, first increment), Poisson arrivals, and the scheduler
(first-come-first-served admission by slots and HBM budget, continuous
joining, per-step latencies) driven by a deterministic stand-in engine on
a simulated clock."""
import numpy as np
import pytest

from warp3d_amd import _abi
from warp3d_amd import serve as S


def test_grains_are_the_nearest_centres():
    rve = S.make_rve(S.RveSpec(6, seed=3, ngrains=16))
    m = rve.mesh
    g = S.grains(m, 16, 3)
    rng = np.random.default_rng(3)
    centres = rng.random((16, 3))
    cen = m.coords[m.conn - 1].mean(axis=1)
    # brute force over all (element, centre) pairs
    d = ((cen[:, None, :] - centres[None]) ** 2).sum(-1)
    assert np.array_equal(g, d.argmin(axis=1))
    # the orientations handed to the model are those of the grains
    ori = rve.arrays.angles.reshape(-1, 3)
    same = g[:, None] == g[None, :]
    for e in range(0, m.noelem, 17):
        assert np.allclose(ori[same[e]], ori[e])
    assert len(np.unique(g)) > 8


def test_rve_boundary_conditions_and_first_increment():
    spec = S.RveSpec(5, seed=1, strain=0.02, nsteps=10)
    rve = S.make_rve(spec)
    x = rve.mesh.coords
    de = 0.002
    cst = rve.cstmap
    # symmetry planes and the loaded top face
    assert np.array_equal(cst[0::3] != 0, np.isclose(x[:, 0], 0.0))
    assert np.array_equal(cst[1::3] != 0, np.isclose(x[:, 1], 0.0))
    top = np.isclose(x[:, 2], 1.0)
    assert np.array_equal(cst[2::3] != 0, np.isclose(x[:, 2], 0.0) | top)
    assert np.array_equal(rve.loaded[2::3], top)
    assert rve.loaded.sum() == (spec.n + 1) ** 2
    # imposed values: zero on the planes, de on the top face
    assert np.all(rve.du0[(cst != 0) & ~rve.loaded] == 0.0)
    assert np.allclose(rve.du0[rve.loaded], de)
    # free dofs start on the homogeneous field
    free = cst == 0
    hom = np.stack([-0.4 * de * x[:, 0], -0.4 * de * x[:, 1],
                    de * x[:, 2]], axis=1).ravel()
    assert np.allclose(rve.du0[free], hom[free])
    assert rve.arrays.noelem == 125


def test_poisson_arrivals():
    arr = S.poisson_arrivals(2000, rate=100.0, nmin=8, nmax=24, seed=5)
    t = np.array([a[0] for a in arr])
    n = np.array([a[1].n for a in arr])
    assert np.all(np.diff(t) > 0)
    assert n.min() == 8 and n.max() == 24
    assert abs(2000 / t[-1] - 100.0) < 10.0
    assert len({a[1].seed for a in arr}) == 2000


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-4)


class FakeProblem:
    def __init__(self, spec):
        self.spec = spec
        self.step, self.iter = 1, 0
        self.status = _abi.W3D_PROBLEM_RUNNING
        # Newton iterations of each step: fixed by the request
        self.need = [2 + (spec.seed + s) % 3 for s in range(spec.nsteps)]
        self.records = []

    def device_bytes(self):
        return 30000 * self.spec.noelem


class FakeEngine:
    """Each advance costs 1 ms per active problem on the fake clock and
    checks the scheduler's invariants."""

    def __init__(self, clock, server_ref):
        self.clock = clock
        self.server_ref = server_ref
        self.max_active = 0
        self.order = []

    def prepare(self, spec):
        self.order.append(spec.seed)
        return FakeProblem(spec)

    def advance(self, probs):
        srv = self.server_ref[0]
        self.max_active = max(self.max_active, len(probs))
        assert len(probs) <= srv.max_slots
        assert sum(p.device_bytes() for p in probs) <= srv.hbm_budget
        self.clock.t += 1e-3 * len(probs)
        out = []
        for p in probs:
            if p.status == _abi.W3D_PROBLEM_STEP_DONE:
                p.step += 1
                p.iter = 0
            p.iter += 1
            if p.iter >= p.need[p.step - 1]:
                p.records.append(p.iter)
                p.status = (_abi.W3D_PROBLEM_DONE
                            if p.step == p.spec.nsteps
                            else _abi.W3D_PROBLEM_STEP_DONE)
            else:
                p.status = _abi.W3D_PROBLEM_RUNNING
            out.append(p.status)
        return out

    def result(self, p):
        return dict(records=list(p.records))

    def release(self, p):
        pass


@pytest.mark.parametrize("slots,budget", [(4, 1 << 40), (16, 60 << 20),
                                          (1, 1 << 40)])
def test_scheduler_admission_and_latencies(slots, budget):
    clock = FakeClock()
    ref = [None]
    eng = FakeEngine(clock, ref)
    srv = S.Server(eng, max_slots=slots, hbm_budget=budget, prep_workers=2,
                   bytes_per_elem=30000, clock=clock, sleep=clock.sleep)
    ref[0] = srv
    arr = S.poisson_arrivals(40, rate=200.0, nmin=4, nmax=12, seed=2,
                             nsteps=5)
    done, st = srv.run(arr)
    assert st["requests"] == 40 and st["solved"] == 40
    assert eng.max_active <= slots
    # first come, first served: admissions in arrival order (set-ups run
    # on worker threads, so only the admission times are ordered)
    assert sorted(eng.order) == sorted(a[1].seed for a in arr)
    admit = [j.t_admit for j in done]
    assert all(a <= b for a, b in zip(admit, admit[1:]))
    for j, (t, spec) in zip(done, arr):
        assert j.t_arrive == t and j.t_admit >= t
        # one latency per step, each covering that step's iterations
        assert len(j.step_lat) == spec.nsteps
        assert j.result["records"] == [2 + (spec.seed + s) % 3
                                       for s in range(spec.nsteps)]
        assert all(x > 0 for x in j.step_lat)
    assert st["steps"] == 40 * 5
    assert st["step_p50_ms"] <= st["step_p99_ms"]
    if budget < (1 << 30):
        assert st["peak_bytes"] <= budget
    if slots > 1:
        assert eng.max_active > 1  # requests overlapped


def test_scheduler_admits_a_request_larger_than_the_budget_alone():
    clock = FakeClock()
    big = [None]
    eng = FakeEngine(clock, big)
    # the engine checks slots only: each request alone exceeds the budget
    big[0] = S.Server(eng, max_slots=4, hbm_budget=1 << 40)
    srv = S.Server(eng, max_slots=4, hbm_budget=1 << 20, prep_workers=1,
                   bytes_per_elem=30000, clock=clock, sleep=clock.sleep)
    done, st = srv.run([(0.0, S.RveSpec(10, seed=1, nsteps=2)),
                        (0.0, S.RveSpec(10, seed=2, nsteps=2))])
    assert st["solved"] == 2
    assert eng.max_active == 1  # one at a time, none starved
