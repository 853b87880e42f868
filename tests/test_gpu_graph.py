"""The graph-launched Newton iteration (w3d_iteration_dev) and the
autotuner against the separate calls.

Bar: bit-identical to the eager sequence w3d_recover_dev + w3d_tangent +
w3d_assemble_vss_dev on the same inputs (NaN where NaN: the n+1 history
of a point whose crystal update asks for a step cut holds the diverged
local iterates, which the caller discards), for every replay of every
captured graph (steps 1 and 2, iterations 0-3, with eager work and commits
between the replays), for mm01 small strain, mm01 finite strain and mm10;
and the autotuned work lists give the same tangents and matrices.
"""
import numpy as np
import pytest

from tests._util import smooth_field
from warp3d_amd import _abi
from warp3d_amd import mesh as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def W():
    from warp3d_amd import model
    return model


def models():
    m8 = M.box_hex8(4, jitter=0.05, seed=31)
    mm01 = _abi.mm01_material(200e3, 0.3, 400.0, hprime=2000.0)
    m20 = M.box_hex20(2, jitter=0.05, seed=32)
    cp = _abi.mm10_material(tau_y=50.0, tau_v=100.0, theta_0=250.0,
                            e=100e3, nu=0.3, mu=80e3, elastic="cubic")
    return [("mm01", m8, m8.model(mm01, auto_size=16), 0.004),
            ("geonl", m8, m8.model(mm01, auto_size=16, geonl=True), 0.03),
            ("mm10", m20, m20.model(cp, int_order=1, auto_size=5,
                                    angles=M.random_orientations(
                                        m20.noelem, seed=33)), 0.004)]


class Run:
    """One device model with its buffers."""

    def __init__(self, W, arr, cst):
        from warp3d_amd import hip
        self.hip = hip
        self.dev = W.BlockModel(arr)
        self.neq, self.nco = self.dev.pattern(cst)
        n = 3 * arr.nonode
        self.u = hip.DeviceArray(n)
        self.du = hip.DeviceArray(n)
        self.ifv = hip.DeviceArray(n)
        self.kd = hip.DeviceArray(self.neq)
        self.kc = hip.DeviceArray(max(self.nco, 1))

    def eager(self, step, it, prec):
        en, pw, rc = self.dev.recover_dev(step, it, self.u.ptr, self.du.ptr,
                                          self.ifv.ptr)
        self.dev.tangent(step, it, prec)
        self.dev.assemble_vss_dev(self.kd.ptr, self.kc.ptr)
        return en, pw, rc

    def graph(self, step, it, prec):
        return self.dev.iteration_dev(step, it, self.u.ptr, self.du.ptr,
                                      self.ifv.ptr, prec, self.kd.ptr,
                                      self.kc.ptr)

    def outputs(self):
        out = [self.ifv.download(), self.kd.download(), self.kc.download()]
        for b in range(1, self.dev.nelblk + 1):
            for which in (_abi.STRESS_N1, _abi.HIST_N1, _abi.ESTIFF):
                out.append(self.dev.get_block_state(b, which))
        return out


@pytest.mark.parametrize("name", ["mm01", "geonl", "mm10"])
@pytest.mark.parametrize("prec", [_abi.W3D_FP64, _abi.W3D_FP32])
def test_graph_iteration_is_the_eager_sequence(W, name, prec):
    _, m, arr, scale = [x for x in models() if x[0] == name][0]
    cst = np.zeros(3 * m.nonode, np.int32)
    cst[:3] = 1
    a, b = Run(W, arr, cst), Run(W, arr, cst)
    u = np.zeros(3 * m.nonode)
    for step in (1, 2):
        duf = smooth_field(m, scale, seed=step)
        for r in (a, b):
            r.u.upload(u)
        for it in range(4):
            du = duf * (0.5 + 0.5 * it / 3)
            for r in (a, b):
                r.du.upload(du)
            print(f"{name} prec {prec} step {step} iter {it}", flush=True)
            ea = a.eager(step, it, prec)
            eb = b.graph(step, it, prec)
            assert ea == eb, (ea, eb)
            for x, y in zip(a.outputs(), b.outputs()):
                assert np.array_equal(x, y, equal_nan=True)
        for r in (a, b):
            r.dev.commit_step()
        u = u + duf


def test_autotuned_lists_give_the_same_results(W):
    _, m, arr, _ = models()[2]
    cst = np.zeros(3 * m.nonode, np.int32)
    cst[:3] = 1
    a, b = Run(W, arr, cst), Run(W, arr, cst)
    du = smooth_field(m, 0.004, seed=7)
    for r in (a, b):
        r.du.upload(du)
        r.eager(1, 1, _abi.W3D_FP32)
    choices = b.dev.autotune(_abi.W3D_FP32)
    assert all(v in (8, 16, 32, 64) for k, v in choices.items() if k >= 0)
    assert -1 in choices and choices[-1] > 0
    ea = a.eager(1, 2, _abi.W3D_FP32)
    assert b.eager(1, 2, _abi.W3D_FP32) == ea
    for x, y in zip(a.outputs(), b.outputs()):
        assert np.array_equal(x, y, equal_nan=True)
    # and the graph path with the tuned lists (the same status: this
    # increment may ask for a step cut)
    eb = b.graph(1, 2, _abi.W3D_FP32)
    for x, y in zip(a.outputs(), b.outputs()):
        assert np.array_equal(x, y, equal_nan=True)
    assert eb == ea
