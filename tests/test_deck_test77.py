"""The reference's verification case test77 (a 20-node brick cantilever
driven to very large displacements and rotations), read from its input
deck (tests/golden/test77_cantilever) and run with the oracle's Newton
loop: element 1, Gauss point 5, eps_xx after step 100 against the
reference's answer 0.231553, at the reference's comparison tolerance
(1e-4 relative, verification/run_tests.py).  This pins, end to end, the
20-node brick with 2x2x2 integration under finite strains (the polar
decomposition, the mid-step strain increments, the rotated Cauchy
stresses in the internal forces and the tangent's initial-stress terms),
nodal loads, a fixed plane and the strain output's rotation."""
import json
import os

import numpy as np

from oracle import oracle as O
from tests import newton_ref as NR
from warp3d_amd import _abi
from warp3d_amd import deck_input as DI

HERE = os.path.join(os.path.dirname(__file__), "golden", "test77_cantilever")


def run_case(make=None):
    dk = DI.read(os.path.join(HERE, "test_77"))
    pb = DI.build(dk)
    om = (make or O.OracleModel)(pb.arrays)
    opts = NR.NewtonOptions(**pb.newton)
    hist = NR.run_history(pb.arrays, pb.cstmap,
                          [np.zeros(3 * dk.nonode)] * pb.nsteps, opts=opts,
                          dloads=pb.dloads, mfs=pb.mfs, dt=pb.dt, om=om)
    blocks = [(int(r[0]), int(r[1])) for r in om.elblks]
    return dk, pb, hist, om, blocks


def answer(model, blocks, want):
    return DI.point_strain(model, want["element"], want["gauss_point"],
                           blocks, 8, geonl=True)[0]


def test_deck_reads_the_case():
    dk = DI.read(os.path.join(HERE, "test_77"))
    pb = DI.build(dk)
    assert (dk.nonode, dk.noelem, pb.nsteps) == (171, 20, 100)
    assert dk.etype == _abi.W3D_Q3DISOP and dk.incid.shape == (20, 20)
    assert (pb.arrays.int_order == 8).all() and pb.arrays.geonl.all()
    assert not pb.arrays.bbar.any()
    # the plane x = 0 fixed in all three components: 21 nodes
    held = pb.cstmap.reshape(-1, 3)
    at0 = dk.coords[:, 0] == 0.0
    assert at0.sum() == 21 and held[at0].all() and not held[~at0].any()
    # 21 end nodes loaded in y and z, 0.01 of the unit pattern a step
    f = pb.dloads[0].reshape(-1, 3)
    assert np.count_nonzero(f[:, 1]) == 21
    assert np.allclose(f[:, 1].sum(), 0.01 * 21 * 0.476190e-01)
    assert np.array_equal(f[:, 1], f[:, 2]) and not f[:, 0].any()
    nw = pb.newton
    assert (nw["res_tol"], nw["mxiter"], nw["mniter"]) == (0.05, 20, 1)
    assert not nw["extrapolate"] and not nw["adaptive"]
    assert pb.dt == 1.0e6


def test_reference_answer_of_the_cantilever_case():
    want = json.load(open(os.path.join(HERE, "expected.json")))
    dk, pb, hist, om, blocks = run_case()
    assert all(h[0] for h in hist) and len(hist) == want["step"]
    eps = answer(om, blocks, want)
    rel = abs(eps - want["value"]) / abs(want["value"])
    assert rel <= want["rel_tol"], (eps, want["value"])
    # the large rotations the case is about: the free end moves by more
    # than half the beam's length in y and z
    u = hist[-1][3].reshape(-1, 3)
    assert u[170, 1] > 2.0 and u[170, 2] > 2.0
