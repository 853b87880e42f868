"""The reference's Voce crystal verification case (test82/voche_model) with
the device element path under the oracle's Newton loop: the same Newton
iterations per step as the oracle, displacements within 1e-8 relative, and
element 64's centre sigma_xx after step 8 within the reference's tolerance
(1e-4) of its answer 476.645594 and within 1e-8 of the oracle's value."""
import json
import os

import numpy as np
import pytest

from tests.test_deck_test82 import HERE, run_case
from warp3d_amd import deck_input as DI

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("asym,device_solve",
                         [(True, False), (True, True), (False, False)])
def test_voce_case_on_the_device(asym, device_solve):
    """asym: the deck's asymmetric assembly (full Ke, CSR rows), else the
    symmetrised tangent; device_solve: the library's BiCGStab (to a
    relative residual of 1e-13) in place of the sparse LU."""
    from warp3d_amd.model import BlockModel

    def make(arrays):
        b = BlockModel(arrays)
        b.device_solve = device_solve
        return b

    want = json.load(open(os.path.join(HERE, "expected.json")))
    _, _, ho, om, blocks = run_case(asym=asym)
    _, _, hd, dm, blocks_d = run_case(make, asym=asym)
    assert blocks_d == blocks
    assert len(hd) == len(ho) == want["step"] and all(h[0] for h in hd)
    assert [h[1] for h in hd] == [h[1] for h in ho]
    for d, o in zip(hd, ho):
        assert np.abs(d[3] - o[3]).max() <= 1e-8 * np.abs(o[3]).max()
    so = DI.center_stress(om, want["element"], blocks)
    sd = DI.center_stress(dm, want["element"], blocks)
    assert abs(sd[0] - so[0]) <= 1e-8 * abs(so[0])
    assert abs(sd[0] - want["value"]) <= want["rel_tol"] * abs(want["value"])
    if asym:
        assert abs(sd[0] - want["value"]) <= 5e-7
