"""w3d_solve_pcg_static (the solve that allocates nothing: its work
vectors are the library's static device array) against w3d_solve_pcg_dev
(the context's pattern-time workspace) on the same assembled matrices.

Two solves in a row on two models, the second smaller than the first, so
the static array is reused with stale contents from the larger solve: the
iterations, the relative residual and the solution must be bit-identical
to the allocating entry point's (the same kernels and reduction trees on
the same inputs), and a model beyond the static limit is refused.
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


def _system(W, n):
    """A C1-style hex8 model after one elastic recovery: its assembled
    matrix on the device and a right-hand side."""
    from warp3d_amd import hip
    m = M.box_hex8(n, jitter=0.1, seed=n)
    mat = _abi.mm01_material(200e3, 0.3, 400.0, hprime=2000.0)
    arr = m.model(mat, auto_size=32)
    dev = W.BlockModel(arr)
    cst, _, _ = M.uniaxial_bc(m, 0.0)
    neq, nco = dev.pattern(cst)
    du = smooth_field(m, 0.001, seed=1)
    dev.recover(1, 0, np.zeros(3 * m.nonode), du)
    dev.tangent(1, 0, _abi.W3D_FP64)
    kd = hip.DeviceArray(neq)
    kc = hip.DeviceArray(max(nco, 1))
    dev.assemble_vss_dev(kd.ptr, kc.ptr)
    rng = np.random.default_rng(n)
    b = hip.DeviceArray.from_host(rng.uniform(-1, 1, neq))
    return dev, neq, kd, kc, b


def _solve(hip, dev, neq, kd, kc, b, static):
    x = hip.DeviceArray.from_host(np.zeros(neq))
    f = dev.solve_pcg_static if static else dev.solve_pcg_dev
    it, rr = f(kd.ptr, kc.ptr, b.ptr, x.ptr, rtol=1e-10, maxit=5000)
    return it, rr, x.download()


def test_static_pcg_matches_the_allocating_solve(W):
    from warp3d_amd import hip
    for n in (8, 5):                       # the second call is smaller
        dev, neq, kd, kc, b = _system(W, n)
        ref = _solve(hip, dev, neq, kd, kc, b, static=False)
        got = _solve(hip, dev, neq, kd, kc, b, static=True)
        assert ref[0] == got[0] and ref[0] > 0
        assert ref[1] == got[1] and ref[1] <= 1e-10
        assert np.array_equal(ref[2], got[2]), f"n={n}"
        # and again, after the allocating solve has run in between
        again = _solve(hip, dev, neq, kd, kc, b, static=True)
        assert np.array_equal(again[2], got[2])
        dev.close()


def test_static_pcg_refuses_more_than_its_limit(W):
    """70^3 bricks with a fixed base: about 1.05 M equations, beyond the
    2^19 equations of the static work memory."""
    from warp3d_amd import hip
    from warp3d_amd.model import W3DError
    m = M.box_hex8(70)
    mat = _abi.mm01_material(200e3, 0.3, 400.0, hprime=2000.0)
    dev = W.BlockModel(m.model(mat))
    cst, _, _ = M.uniaxial_bc(m, 0.0)
    neq, nco = dev.pattern(cst)
    assert neq > 1 << 19
    kd = hip.DeviceArray(neq)
    kc = hip.DeviceArray(1)
    b = hip.DeviceArray(neq)
    x = hip.DeviceArray(neq)
    with pytest.raises(W3DError, match="static work memory"):
        dev.solve_pcg_static(kd.ptr, kc.ptr, b.ptr, x.ptr)
    dev.close()
