"""Shared helpers of the tests (comparison tolerances, model factories)."""
import numpy as np

from warp3d_amd import _abi
from warp3d_amd import mesh as M

# The plain FP64 device path must match the oracle within this relative
# tolerance (stresses, strains, history, [D], element forces, Ke).
RTOL_FP64 = 1e-10


def rel_err(a, b):
    """max |a - b| / max |b| (norm-wise relative error)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    den = np.max(np.abs(b)) if b.size else 0.0
    num = np.max(np.abs(a - b)) if b.size else 0.0
    return num / den if den > 0 else num


def assert_close(a, b, rtol=RTOL_FP64, what=""):
    e = rel_err(a, b)
    assert e <= rtol, f"{what}: relative error {e:.3e} > {rtol:.1e}"


def c1_model(n=10, jitter=0.1, seed=1, auto_size=128):
    """Unit cube, n^3 8-node bricks (2x2x2, B-bar), interior nodes
    jittered; mm01 with E = 200 GPa, nu = 0.3, sigma_y = 400 MPa,
    H' = E/100 (units MPa)."""
    m = M.box_hex8(n, jitter=jitter, seed=seed)
    mat = _abi.mm01_material(200e3, 0.3, 400.0, hprime=2000.0)
    return m, m.model(mat, auto_size=auto_size)


def smooth_field(m, scale, seed=0, noise=0.2):
    """A displacement field: uniaxial stretch + shear + seeded noise."""
    rng = np.random.default_rng(seed)
    x = m.coords
    u = np.zeros_like(x)
    u[:, 2] = scale * x[:, 2]
    u[:, 0] = -0.3 * scale * x[:, 0] + 0.2 * scale * x[:, 2]
    u[:, 1] = -0.3 * scale * x[:, 1]
    h = getattr(m, "h", None)  # element size (None: meshes of parts)
    u += noise * scale * rng.uniform(-1, 1, size=x.shape) * (
        h.mean() if h is not None else 1.0)
    return u.ravel()
