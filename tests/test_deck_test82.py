"""The reference's own verification case test82/voche_model, read from its
input deck (tests/golden/test82_voche: deck and expected answer) and run
with the oracle's Newton loop: element 64's centre sigma_xx after step 8
against the reference's answer 476.645594, at the reference's comparison
tolerance (1e-4 relative, verification/run_tests.py).  This pins, end to
end, the deck subset (elements, blocking, face pressures, symmetry planes,
load steps, Newton options), the Newton semantics (iteration 0,
extrapolation, test 2) and the mm10 Voce update of a bcc crystal."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import newton_ref as NR
from warp3d_amd import deck_input as DI

HERE = os.path.join(os.path.dirname(__file__), "golden", "test82_voche")


def run_case(make=None, asym=None):
    """The deck's 8 steps on a model made by make(arrays) (the oracle's by
    default), with the deck's assembly (asymmetric) unless asym says
    otherwise."""
    dk = DI.read(os.path.join(HERE, "warp3d_voche.inp"))
    pb = DI.build(dk)
    om = (make or O.OracleModel)(pb.arrays)
    om.set_asymmetric(pb.asymmetric if asym is None else asym)
    opts = NR.NewtonOptions(**pb.newton)
    hist = NR.run_history(pb.arrays, pb.cstmap,
                          [np.zeros(3 * dk.nonode)] * pb.nsteps, opts=opts,
                          dloads=pb.dloads, mfs=pb.mfs, dt=pb.dt, om=om)
    blocks = [(int(r[0]), int(r[1])) for r in om.elblks]
    return dk, pb, hist, om, blocks


def test_deck_reads_the_case():
    dk = DI.read(os.path.join(HERE, "warp3d_voche.inp"))
    pb = DI.build(dk)
    assert (dk.nonode, dk.noelem, dk.auto_size) == (342, 208, 20)
    assert pb.nsteps == 8 and pb.dt == 1.0e6
    assert pb.newton["mxiter"] == 10 and pb.newton["extrapolate"]
    assert pb.asymmetric
    # the unit pattern: +1 on the x = 1 face, -0.5 on the z = 1 face, both
    # of unit area; 50 per step
    f = pb.dloads[0].reshape(-1, 3)
    assert np.allclose(f.sum(axis=0), [50.0, 0.0, -25.0])
    x = dk.coords
    assert np.all(f[~np.isclose(x[:, 0], 1.0), 0] == 0.0)
    assert np.all(f[~np.isclose(x[:, 2], 1.0), 2] == 0.0)
    # symmetry planes
    for c in range(3):
        assert np.array_equal(pb.cstmap[c::3] != 0, x[:, c] == 0.0)


@pytest.mark.parametrize("asym", [True, False])
def test_reference_answer_of_the_voce_case(asym):
    """With the deck's asymmetric assembly, and with the symmetrised
    tangent: the converged answer does not depend on the tangent."""
    want = json.load(open(os.path.join(HERE, "expected.json")))
    dk, pb, hist, om, blocks = run_case(asym=asym)
    assert all(h[0] for h in hist) and len(hist) == want["step"]
    sig = DI.center_stress(om, want["element"], blocks)
    rel = abs(sig[0] - want["value"]) / abs(want["value"])
    assert rel <= want["rel_tol"], (sig[0], want["value"])
    if asym:  # as the reference ran it: every printed digit
        assert abs(sig[0] - want["value"]) <= 5e-7, sig[0]
