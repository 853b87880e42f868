"""The synthetic workloads of the element-block path (seeded meshes,
materials, constraints and a load-step increment), shared by bench.py and
the full-size parity tests.

  hex8:  60x60x60 8-node bricks (216,000 elements, 2x2x2 Gauss points with
         B-bar), interior nodes jittered +-5 % of the edge (seeded), mm01
         bilinear Mises (E = 200 GPa, nu = 0.3, sigma_y = 350 MPa,
         H' = E/100), fixed base, a seeded top-face displacement field
         spread through the height, 0.15 % mean axial strain per step of a
         10-step ramp to 1.5 %;
  hex20: 40x40x40 20-node bricks (64,000 elements, 3x3x3 Gauss points),
         mm01 (E = 70 GPa, nu = 0.33, sigma_y = 300 MPa, H' = E/50), a
         bending-style field (axial strain varying linearly through the
         thickness, 1 % peak over 15 steps);
  mixed: 150,000 elements in blocks of seeded spans 32-256 (1,935,000
         Gauss points): 75,000 hex8 (2x2x2, B-bar), 45,000 hex20 (3x3x3),
         30,000 tet10 (4 points), three boxes side by side; mm10 fcc
         crystal plasticity with Voce hardening (tau_y = 50 MPa,
         tau_v = 100 MPa, theta_0 = 250 MPa, m = 1, n = 20; cubic
         elasticity of copper: e = 66.7 GPa, nu = 0.42, mu = 75 GPa),
         seeded SO(3)-uniform orientations, 0.1 % axial strain per step of
         a 10-step ramp to 1 %.
"""
import numpy as np

WORKLOADS = {
    "hex8": "hex8 60x60x60 box, 216000 elements, 1728000 Gauss points "
            "(2x2x2, B-bar), mm01, blocks of 128, one Newton iteration of "
            "the element-block path per step",
    "hex20": "hex20 40x40x40 box, 64000 elements, 1728000 Gauss points "
             "(3x3x3), mm01, blocks of 128, one Newton iteration of the "
             "element-block path per step",
    "mixed": "150000 elements (75000 hex8 2x2x2 B-bar, 45000 hex20 3x3x3, "
             "30000 tet10 4-point), 1935000 Gauss points, blocks of seeded "
             "spans 32-256, mm10 fcc Voce crystal plasticity with seeded "
             "SO(3)-uniform orientations, one Newton iteration of the "
             "element-block path per step",
    "rve": "Poisson stream of crystal-plasticity RVE requests, each an n^3 "
           "hex8 unit cube (n uniform in [8, 24]) of 64 seeded Voronoi "
           "grains with SO(3)-uniform fcc orientations, mm10 Voce (the "
           "test82 Voce constants), FP32 "
           "tangent, symmetry planes and a strain-controlled top face, 10 "
           "steps to 2 %; served by continuous batching",
}


def build_workload(name, seed=2024):
    """Mesh, model arrays, constraints, the step increment du (the
    displacement increment of one load step) and the element parts
    (count, nodes, Gauss points, mm10)."""
    from warp3d_amd import _abi
    from warp3d_amd import mesh as M
    rng = np.random.default_rng(seed)
    if name == "mixed":
        h8 = M.box_hex8(50, 50, 30, size=(1.0, 1.0, 0.6), jitter=0.05,
                        seed=seed)
        h20 = M.box_hex20(30, 30, 50, size=(0.6, 0.6, 1.0), jitter=0.05,
                          seed=seed + 1)
        t10 = M.box_tet10(20, 25, 10, size=(0.4, 0.5, 0.2), jitter=0.05,
                          seed=seed + 2)
        mm = M.MixedMesh([(h8, 1, 1), (h20, 1, 0), (t10, 4, 0)])
        mat = _abi.mm10_material(tau_y=50.0, tau_v=100.0, theta_0=250.0,
                                 voce_m=1.0, harden_n=20.0, elastic="cubic",
                                 e=66.7e3, nu=0.42, mu=75.0e3)
        ang = M.random_orientations(mm.noelem, seed=seed)
        arr = mm.model(mat, elblks=mm.seeded_blocks(32, 256, seed=seed),
                       angles=ang)
        x = mm.coords
        base = np.abs(x[:, 2]) < 1e-9
        cst = np.zeros(3 * mm.nonode, np.int32)
        for k in range(3):
            cst[k::3][base] = 1
        eps = 0.01 / 10
        du = np.zeros_like(x)
        du[:, 2] = eps * x[:, 2]
        du[:, 0] = -0.3 * eps * x[:, 2] * (x[:, 0] - x[:, 0].mean()) * 0.1
        du += 0.05 * eps * rng.uniform(-1, 1, size=du.shape) * \
            (x[:, 2:3] > 0) * 0.02
        du[base] = 0.0
        parts = [(h8.noelem, 8, 8, True), (h20.noelem, 20, 27, True),
                 (t10.noelem, 10, 4, True)]
        return mm, arr, cst, du.ravel(), parts
    if name == "hex8":
        m = M.box_hex8(60, jitter=0.05, seed=seed)
        mat = _abi.mm01_material(200e3, 0.3, 350.0, hprime=2000.0)
        arr = m.model(mat, auto_size=128)
        eps = 0.015 / 10
    else:
        m = M.box_hex20(40, jitter=0.05, seed=seed)
        mat = _abi.mm01_material(70e3, 0.33, 300.0, hprime=70e3 / 50)
        arr = m.model(mat, int_order=1, auto_size=128)
        eps = 0.01 / 15
    x = m.coords
    nd = 3 * m.nonode
    base = np.abs(x[:, 2]) < 1e-9
    cst = np.zeros(nd, np.int32)
    for k in range(3):
        cst[k::3][base] = 1
    du = np.zeros_like(x)
    if name == "hex8":
        # seeded top-face field spread through the height as the
        # linear-elastic solution would
        top = rng.uniform(0.5, 1.5, size=m.nonode)
        du[:, 2] = eps * x[:, 2] * top
        du[:, 0] = -0.3 * eps * (x[:, 0] - 0.5) * x[:, 2]
        du[:, 1] = -0.3 * eps * (x[:, 1] - 0.5) * x[:, 2]
    else:
        # bending about the y axis: axial strain linear through x
        du[:, 2] = eps * 2.0 * (x[:, 0] - 0.5) * x[:, 2] * 2.0
        du[:, 0] = -eps * x[:, 2] ** 2
        du[:, 1] = -0.33 * eps * (x[:, 1] - 0.5) * x[:, 2]
        du += 0.05 * eps * rng.uniform(-1, 1, size=du.shape) * \
            (x[:, 2:3] > 0)
    du[base] = 0.0
    parts = [(m.noelem, m.conn.shape[1], 8 if name == "hex8" else 27, False)]
    return m, arr, cst, du.ravel(), parts
