"""The reference's skewed-constraint verification case (test24, a hollow
sphere under internal pressure) on the device: the element path under the
oracle's Newton loop, and the Fortran driver (w3d_drive over the ISO_C_
BINDING module) running the same deck with the device PCG.  Bars: the same
Newton iterations per step as the oracle, displacements within 1e-8
relative, and element 1's centre Cauchy sigma_xx after step 21 within the
reference's tolerance (1e-4) of its answer -205.047360 and within 1e-8 of
the oracle's value."""
import json
import os

import numpy as np
import pytest

from tests.test_deck_test24 import HERE, run_case
from warp3d_amd import _abi
from warp3d_amd import deck as D
from warp3d_amd import deck_input as DI

pytestmark = pytest.mark.gpu


def test_sphere_case_on_the_device():
    from warp3d_amd.model import BlockModel
    want = json.load(open(os.path.join(HERE, "expected.json")))
    _, _, ho, om, blocks = run_case()
    _, _, hd, dm, blocks_d = run_case(BlockModel)
    assert blocks_d == blocks
    assert len(hd) == len(ho) == want["step"] and all(h[0] for h in hd)
    assert [h[1] for h in hd] == [h[1] for h in ho]
    for d, o in zip(hd, ho):
        assert np.abs(d[3] - o[3]).max() <= 1e-8 * np.abs(o[3]).max()
    so = DI.center_stress(om, want["element"], blocks, geonl=True)
    sd = DI.center_stress(dm, want["element"], blocks, geonl=True)
    assert abs(sd[0] - so[0]) <= 1e-8 * abs(so[0])
    assert abs(sd[0] - want["value"]) <= want["rel_tol"] * abs(want["value"])


def test_sphere_case_through_the_fortran_driver(tmp_path):
    dk, pb, ho, om, _ = run_case()
    deck = os.path.join(str(tmp_path), "sphere.deck")
    res = os.path.join(str(tmp_path), "sphere.res")
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
    print(out, flush=True)
    assert len(got) == len(ho)
    for s, (g, (conv, it, _, u, info)) in enumerate(zip(got, ho), start=1):
        assert g["converged"] and conv, s
        assert g["iters"] == it, (s, g["iters"], it)
        assert g["cuts"] == info["cuts"], s
        assert np.abs(g["u"] - u).max() <= 1e-8 * np.abs(u).max(), s
