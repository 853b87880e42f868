"""Batched serving on the device (w3d_batch_advance) against the plain
path and against itself.

Bars:
* a problem's results (per-step Newton and PCG iteration counts, adaptive
  cuts, loads, norms, final displacements) are bitwise the same alone and
  in a mix of other problems that join and leave at other iterations;
* with the FP64 tangent and tight solver tolerances the batched path and
  the plain path (per-call entry points, host Newton, host-driven PCG)
  take the same Newton iterations and adaptive cuts per step (0.4 % strain
  steps, which the crystal update cannot always take whole) and give loads
  within 1e-9 relative and displacements within 1e-8 relative;
* the FP32 tangent pays at most one extra Newton iteration per step, its
  loads within 1e-6 relative and displacements within 1e-6 relative of
  the FP64 solution (both converged to |R| <= 1e-7 |F_c|);
* the server (admission, continuous joining) gives every request the
  result it has alone.
"""
import numpy as np
import pytest

from warp3d_amd import _abi
from warp3d_amd import serve as S

pytestmark = pytest.mark.gpu


def run_alone(spec, opts):
    b = S.Batch()
    p = S.Problem(S.make_rve(spec), opts)
    try:
        for _ in range(opts.nsteps * opts.max_iters * 16):
            st = b.advance([p])[0]
            if st in (_abi.W3D_PROBLEM_DONE, _abi.W3D_PROBLEM_FAILED):
                break
        assert st == _abi.W3D_PROBLEM_DONE, st
        return p.records(), p.u()
    finally:
        p.close()
        b.close()


def test_alone_and_in_the_mix_are_identical():
    opts = S.newton_opts(nsteps=4)
    kw = dict(nsteps=4, strain=0.012)
    specs = {"a": S.RveSpec(6, seed=11, **kw),
             "b": S.RveSpec(5, seed=12, **kw),
             "c": S.RveSpec(7, seed=13, **kw)}
    print("alone", flush=True)
    alone = {k: run_alone(s, opts) for k, s in specs.items()}
    # b starts first, c joins at the second advance, a at the fourth
    b = S.Batch()
    probs = {}
    joins = {0: "b", 1: "c", 3: "a"}
    done = {}
    try:
        for tick in range(200):
            if tick in joins:
                k = joins[tick]
                probs[k] = S.Problem(S.make_rve(specs[k]), opts)
            act = [k for k in probs if k not in done]
            if not act and tick > max(joins):
                break
            print("mix tick", tick, act, flush=True)
            st = b.advance([probs[k] for k in act])
            for k, s in zip(act, st):
                assert s != _abi.W3D_PROBLEM_FAILED, k
                if s == _abi.W3D_PROBLEM_DONE:
                    done[k] = (probs[k].records(), probs[k].u())
        assert set(done) == set(specs)
        for k in specs:
            ra, ua = alone[k]
            rm, um = done[k]
            assert ra == rm, k
            assert np.array_equal(ua, um), k
            assert len(ra) == 4
    finally:
        for p in probs.values():
            p.close()
        b.close()


def test_batched_fp64_matches_the_plain_path_and_fp32_tangent():
    spec = S.RveSpec(6, seed=21, nsteps=3, strain=0.012)
    tight = dict(nsteps=3, tol=1e-7, pcg_rtol=1e-11, pcg_maxit=20000,
                 max_iters=40)
    rve = S.make_rve(spec)
    print("plain fp64", flush=True)
    rp, up, _ = S.solve_plain(rve, S.newton_opts(precision=_abi.W3D_FP64,
                                                 **tight))
    print("batched fp64", flush=True)
    r64, u64 = run_alone(spec, S.newton_opts(precision=_abi.W3D_FP64,
                                             **tight))
    print("batched fp32", flush=True)
    r32, u32 = run_alone(spec, S.newton_opts(precision=_abi.W3D_FP32,
                                             **tight))
    scale = np.abs(up).max()
    assert [r["newton"] for r in r64] == [r["newton"] for r in rp]
    assert [r["cuts"] for r in r64] == [r["cuts"] for r in rp]
    assert sum(r["cuts"] for r in rp) > 0
    for a, b in zip(r64, rp):
        assert abs(a["load"] - b["load"]) <= 1e-9 * abs(b["load"])
    assert np.abs(u64 - up).max() <= 1e-8 * scale
    for a, b in zip(r32, r64):
        assert a["cuts"] == b["cuts"]
        assert a["newton"] <= b["newton"] + 1 + a["cuts"]
        assert abs(a["load"] - b["load"]) <= 1e-6 * abs(b["load"])
    assert np.abs(u32 - u64).max() <= 1e-6 * scale
    # the response is plastic: the load grows less than linearly
    loads = np.array([r["load"] for r in r64])
    assert np.all(np.diff(loads) > 0)
    assert loads[2] < 3 * loads[0]


def test_server_gives_each_request_its_alone_result():
    opts = S.newton_opts(nsteps=3)
    arr = S.poisson_arrivals(6, rate=400.0, nmin=4, nmax=7, seed=3,
                             nsteps=3, strain=0.006)
    eng = S.DeviceEngine(opts)
    try:
        srv = S.Server(eng, max_slots=3, hbm_budget=8 << 30, prep_workers=2,
                       keep_results=True)
        done, st = srv.run(arr)
    finally:
        eng.close()
    print(st, flush=True)
    assert st["solved"] == 6 and st["peak_slots"] >= 2
    assert st["steps"] == 18
    for j, (_, spec) in zip(done, arr):
        ra, ua = run_alone(spec, opts)
        assert j.result["records"] == ra
        assert np.array_equal(j.result["u"], ua)
