"""The reference's verification case test24 (a hollow sphere under
internal pressure with skewed constraint systems), read from its input deck
(tests/golden/test24_sphere) and run with the oracle's Newton loop:
element 1's centre Cauchy sigma_xx after step 21 against the reference's
answer -205.047360, at the reference's comparison tolerance (1e-4
relative, verification/run_tests.py).  This pins, end to end, the skewed
constraints (trnvec, trnmtx, rotate_loads), finite strains with B-bar
bricks, the bilinear mm01 update, nodal loads, line search, adaptive
stepping and extrapolation."""
import json
import os

import numpy as np

from oracle import oracle as O
from tests import newton_ref as NR
from warp3d_amd import deck_input as DI

HERE = os.path.join(os.path.dirname(__file__), "golden", "test24_sphere")


def run_case(make=None):
    dk = DI.read(os.path.join(HERE, "test_24"))
    pb = DI.build(dk)
    om = (make or O.OracleModel)(pb.arrays)
    opts = NR.NewtonOptions(**pb.newton)
    hist = NR.run_history(pb.arrays, pb.cstmap,
                          [np.zeros(3 * dk.nonode)] * pb.nsteps, opts=opts,
                          dloads=pb.dloads, mfs=pb.mfs, dt=pb.dt, om=om)
    blocks = [(int(r[0]), int(r[1])) for r in om.elblks]
    return dk, pb, hist, om, blocks


def test_deck_reads_the_case():
    dk = DI.read(os.path.join(HERE, "test_24"))
    pb = DI.build(dk)
    assert (dk.nonode, dk.noelem, pb.nsteps) == (96, 45, 21)
    assert pb.arrays.geonl.all() and pb.arrays.bbar.all()
    assert len(dk.trn) == 42
    for t in dk.trn.values():   # proper rotations
        assert np.allclose(t @ t.T, np.eye(3), atol=1e-6)
    assert pb.newton["line_search"] and pb.newton["adaptive"]
    assert pb.newton["mxiter"] == 20 and pb.newton["res_tol"] == 0.001
    assert pb.mfs[0] == 1.0 and pb.mfs[1] == 0.2
    # step 1's loads: the deck's 16 nodal forces, those of skewed nodes
    # in their constraint systems (T f), none on held dofs
    f = pb.dloads[0].reshape(-1, 3)
    g = DI.nodal_loads(dk, "unit_press").reshape(-1, 3)
    assert np.count_nonzero(np.abs(g).sum(axis=1)) == 16
    want = g.copy()
    for n, t in dk.trn.items():
        want[n - 1] = t @ g[n - 1]
    want = want.ravel()
    want[pb.cstmap != 0] = 0.0
    assert np.allclose(f.ravel(), want, rtol=1e-15, atol=0)
    assert np.count_nonzero(f.ravel() - g.ravel() * (pb.cstmap == 0)) > 0
    # node 10: local v and w held, y = 0 and z symmetry planes
    assert list(pb.cstmap[27:30]) == [0, 1, 1]


def test_reference_answer_of_the_sphere_case():
    want = json.load(open(os.path.join(HERE, "expected.json")))
    dk, pb, hist, om, blocks = run_case()
    assert all(h[0] for h in hist) and len(hist) == want["step"]
    sig = DI.center_stress(om, want["element"], blocks, geonl=True)
    rel = abs(sig[0] - want["value"]) / abs(want["value"])
    assert rel <= want["rel_tol"], (sig[0], want["value"])
