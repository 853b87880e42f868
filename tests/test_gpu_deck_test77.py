"""The reference's large-rotation verification case (test77, a 20-node
brick cantilever) on the device: the element path under the oracle's
Newton loop, and the Fortran driver (w3d_drive over the ISO_C_BINDING
module) running the same deck with the device PCG.  Bars: the same Newton
iterations per step as the oracle, displacements within 1e-8 relative,
and element 1's Gauss point 5 eps_xx after step 100 within the
reference's tolerance (1e-4) of its answer 0.231553 and within 1e-8 of
the oracle's value."""
import json
import os

import numpy as np
import pytest

from tests.test_deck_test77 import HERE, answer, run_case
from warp3d_amd import _abi
from warp3d_amd import deck as D

pytestmark = pytest.mark.gpu


def test_cantilever_case_on_the_device():
    from warp3d_amd.model import BlockModel
    want = json.load(open(os.path.join(HERE, "expected.json")))
    _, _, ho, om, blocks = run_case()
    _, _, hd, dm, blocks_d = run_case(BlockModel)
    assert blocks_d == blocks
    assert len(hd) == len(ho) == want["step"] and all(h[0] for h in hd)
    assert [h[1] for h in hd] == [h[1] for h in ho]
    for d, o in zip(hd, ho):
        assert np.abs(d[3] - o[3]).max() <= 1e-8 * np.abs(o[3]).max()
    eo, ed = answer(om, blocks, want), answer(dm, blocks, want)
    assert abs(ed - eo) <= 1e-8 * abs(eo)
    assert abs(ed - want["value"]) <= want["rel_tol"] * abs(want["value"])


def test_cantilever_case_through_the_fortran_driver(tmp_path):
    dk, pb, ho, om, _ = run_case()
    deck = os.path.join(str(tmp_path), "cantilever.deck")
    res = os.path.join(str(tmp_path), "cantilever.res")
    nw = pb.newton
    D.write_deck(deck, pb.arrays, pb.cstmap,
                 np.zeros((pb.nsteps, 3 * dk.nonode)),
                 precision=_abi.W3D_FP64, res_tol=nw["res_tol"],
                 mniter=nw["mniter"], mxiter=nw["mxiter"], pcg_rtol=1e-13,
                 dt=pb.dt, extrapolate=nw["extrapolate"],
                 line_search=nw["line_search"], adaptive=nw["adaptive"],
                 divergence_check=nw["divergence_check"], dloads=pb.dloads,
                 mfs=pb.mfs)
    out, got = D.run(deck, res)
    print(out[-2000:], flush=True)
    assert len(got) == len(ho)
    for s, (g, (conv, it, _, u, info)) in enumerate(zip(got, ho), start=1):
        assert g["converged"] and conv, s
        assert g["iters"] == it, (s, g["iters"], it)
        assert np.abs(g["u"] - u).max() <= 1e-8 * np.abs(u).max(), s
