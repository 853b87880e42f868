"""A subset of the reference's input language, enough to drive the
library from the reference's own verification decks
(verification/test82/voche_model/warp3d_voche.inp, the Voce crystal case;
verification/test24/test_24, a hollow sphere with skewed constraints;
verification/test77/test_77, a 20-node brick cantilever at large
rotations).

Commands read (main_program.f:121-369 dispatches them in the reference):
  structure, crystal (incrystal.f: slip_type, elastic_type isotropic,
  e, nu, hardening voce, harden_n, voce_m, theta_0, tau_y, tau_v),
  list (named lists for output, read and ignored),
  material ... properties cp (inmat.f: angle_convention, angle_type,
  n_crystals 1, crystal_input single, crystal_type, orientation_input
  single angles), number of nodes/elements, elements (inelem.f: type
  l3disop or q3disop, one type per deck, linear or nonlinear (finite
  strains), material, order 2x2x2, bbar, center_output, short), material
  ... properties bilinear (inmat.f: e, nu, yld_pt, tan_e, beta),
  coordinates, incidences, blocking automatic [display] [size = N]
  (inelbk.f:208), loading <pattern> with element face pressures or nodal
  loads force_x/y/z (inlod.f), constraints (incon.f) with "plane x|y|z=0
  symmetry|fixed", node lists with u/v/w values and "transformation
  matrix <nodes> row_1 .. row_2 .. row_3 .." (skewed constraint systems),
  loading <name> nonlinear with
  "step a-b <pattern> <factor>", initial conditions temperatures (read and
  ignored: the build is isothermal), nonlinear analysis parameters
  (indypm.f: time step, maximum/minimum iterations, divergence check,
  convergence test norm res tol, adaptive, extrapolate, line search,
  bbar stabilization factor; the rest read and ignored), compute
  displacements for loading <name> step N, output ... (recorded).
Anything else raises DeckError.

Face pressures become consistent nodal forces (elem_load_b.f: the
pressure acts along the inward normal of the face, eqnrmvh; integrated
with the face's bilinear shape functions at 2x2 points); the face node
lists of the 8-node brick are those of eqelfn (elem_load_c.f:886).  Nodal
loads at nodes with a skewed constraint system are rotated into it
(rotate_loads, eqivld.f:788).
"""
import re
from dataclasses import dataclass, field

import numpy as np

from . import _abi

# eqelfn f8nod: the nodes of each face of an 8-node brick, 1-based
HEX8_FACES = {1: (4, 3, 2, 1), 2: (5, 6, 7, 8), 3: (1, 2, 6, 5),
              4: (3, 4, 8, 7), 5: (2, 3, 7, 6), 6: (4, 1, 5, 8)}


class DeckError(ValueError):
    pass


@dataclass
class Deck:
    nonode: int = 0
    noelem: int = 0
    coords: np.ndarray = None
    incid: np.ndarray = None           # (noelem, nnode) 1-based
    etype: int = _abi.W3D_L3DISOP
    nnode: int = 8
    crystals: dict = field(default_factory=dict)
    materials: dict = field(default_factory=dict)
    elem_material: np.ndarray = None   # material name index per element
    elem_opts: dict = field(default_factory=dict)
    auto_size: int = 128
    patterns: dict = field(default_factory=dict)   # name -> [(e, face, p)]
    nodal: dict = field(default_factory=dict)      # name -> {node: f(3)}
    symmetry: list = field(default_factory=list)   # planes: 'x', 'y', 'z'
    fixed_planes: list = field(default_factory=list)  # planes held in u, v, w
    trn: dict = field(default_factory=dict)        # node -> 3x3 rows
    fixed: dict = field(default_factory=dict)      # (node, dof) -> value
    steps: dict = field(default_factory=dict)      # step -> [(pat, fac)]
    nonlinear: dict = field(default_factory=dict)
    compute_to: int = 0
    outputs: list = field(default_factory=list)


def _logical_lines(text):
    """Lines without comments, a trailing comma joining the next line."""
    out, cur = [], ""
    for raw in text.splitlines():
        s = raw.split("!")[0].split("$")[0].rstrip()
        if not s.strip() or s.strip() in ("c", "C") or \
                s.lstrip().startswith(("c ", "C ", "*echo")):
            continue
        cur = (cur + " " + s.strip()) if cur else s.strip()
        if cur.endswith(","):
            cur = cur[:-1]
            continue
        out.append(cur)
        cur = ""
    if cur:
        out.append(cur)
    return out


def _kv(tokens):
    d = {}
    i = 0
    while i < len(tokens):
        k = tokens[i].lower()
        if i + 1 < len(tokens):
            d[k] = tokens[i + 1]
        i += 2
    return d


def read(path):
    text = open(path).read()
    lines = _logical_lines(text)
    dk = Deck()
    mat_names = {}
    i = 0
    mode = None
    cur_pattern = None
    while i < len(lines):
        ln = lines[i]
        low = ln.lower().strip()
        tok = low.replace("=", " = ").split()
        i += 1
        # the parameter block ends at the next command
        if mode == "nlparams" and not low.startswith(
                ("compute", "output", "stop", "loading", "constraints",
                 "list ")):
            dk.nonlinear[low] = tok
            continue
        if low.startswith("structure"):
            continue
        if low.startswith("crystal "):
            props = lines[i].lower().replace("properties", "").strip()
            i += 1
            if props.endswith("."):      # the deck's closing period
                props = props[:-1]
            d = _kv(props.split())
            dk.crystals[int(tok[1])] = d
            continue
        if low.startswith("material "):
            name = ln.split()[1]
            props = lines[i].lower().replace("properties", "").split()
            i += 1
            if not props or props[0] not in ("cp", "bilinear"):
                raise DeckError("only cp and bilinear materials")
            mat_names[name.lower()] = len(mat_names) + 1
            dk.materials[name.lower()] = props
            continue
        if low.startswith(("number of nodes", "number of elements")):
            # "number of nodes N elements M", or one per line
            for k in range(len(tok) - 1):
                if tok[k] == "nodes":
                    dk.nonode = int(tok[k + 1])
                if tok[k] == "elements":
                    dk.noelem = int(tok[k + 1])
            if dk.nonode and dk.coords is None:
                dk.coords = np.zeros((dk.nonode, 3))
            if dk.noelem and dk.incid is None:
                # sized for the largest element; trimmed at the end
                dk.incid = np.zeros((dk.noelem, 20), np.int32)
                dk.elem_material = np.zeros(dk.noelem, np.int32)
            continue
        if low == "elements":
            mode = "elements"
            continue
        if low == "coordinates":
            mode = "coordinates"
            continue
        if low == "incidences":
            mode = "incidences"
            continue
        if low.startswith("blocking automatic"):
            if "size" in tok:
                dk.auto_size = int(tok[-1])
            mode = None
            continue
        if low.startswith("loading ") and len(tok) == 2:
            # a pattern or a step definition follows
            cur_pattern = tok[1]
            nxt = lines[i].lower().strip()
            if nxt == "element loads":
                i += 1
                mode = "element_loads"
                dk.patterns[cur_pattern] = []
            elif nxt == "nodal loads":
                i += 1
                mode = "nodal_loads"
                dk.nodal[cur_pattern] = {}
            elif nxt == "nonlinear":
                i += 1
                mode = "steps"
            else:
                raise DeckError(f"loading {cur_pattern}: {nxt}")
            continue
        if low == "constraints":
            mode = "constraints"
            continue
        if low == "initial conditions":
            mode = "ignore"
            continue
        if low == "nonlinear analysis parameters":
            mode = "nlparams"
            continue
        if low.startswith("compute displacements"):
            dk.compute_to = int(tok[-1])
            mode = None
            continue
        if low.startswith("output"):
            dk.outputs.append(low)
            mode = None
            continue
        if low.startswith("list "):
            # a named node or element list for output commands (inlist)
            mode = None
            continue
        if low == "stop":
            break
        # data lines of the current mode
        if mode == "elements":
            m = re.match(r"(\d+)\s*-\s*(\d+)\s+(.*)", low)
            if not m:
                raise DeckError(f"elements: {ln}")
            a, b, rest = int(m.group(1)), int(m.group(2)), m.group(3).split()
            opts = dict(zip(rest[0::1], rest[1:] + [""]))
            if rest[:2] not in (["type", "l3disop"], ["type", "q3disop"]):
                raise DeckError("only l3disop and q3disop elements")
            dk.etype = (_abi.W3D_L3DISOP if rest[1] == "l3disop"
                        else _abi.W3D_Q3DISOP)
            dk.nnode = 8 if rest[1] == "l3disop" else 20
            mname = rest[rest.index("material") + 1]
            dk.elem_material[a - 1:b] = mat_names[mname]
            dk.elem_opts = dict(order=rest[rest.index("order") + 1]
                                if "order" in rest else "2x2x2",
                                bbar="bbar" in rest,
                                center_output="center_output" in rest,
                                linear="linear" in rest,
                                geonl="nonlinear" in rest)
            del opts
            continue
        if mode == "coordinates":
            v = ln.split()
            dk.coords[int(v[0]) - 1] = [float(x) for x in v[1:4]]
            continue
        if mode == "incidences":
            v = [int(x) for x in ln.split()]
            if len(v) != dk.nnode + 1:
                raise DeckError(f"incidences: {ln}")
            dk.incid[v[0] - 1, :dk.nnode] = v[1:]
            continue
        if mode == "element_loads":
            v = low.split()
            # <elem> face <f> pressure <p>
            if v[1] != "face" or v[3] != "pressure":
                raise DeckError(f"element loads: {ln}")
            dk.patterns[cur_pattern].append((int(v[0]), int(v[2]),
                                             float(v[4])))
            continue
        if mode == "nodal_loads":
            v = low.split()
            f = dk.nodal[cur_pattern].setdefault(int(v[0]), np.zeros(3))
            for k in range(1, len(v) - 1, 2):
                if v[k] not in ("force_x", "force_y", "force_z"):
                    raise DeckError(f"nodal loads: {ln}")
                f["xyz".index(v[k][-1])] += float(v[k + 1])
            continue
        if mode == "constraints":
            m = re.match(r"plane\s+([xyz])\s*=\s*0(\.0*)?\s+"
                         r"(symmetry|fixed)$", low)
            if m:
                # symmetry: the plane's normal component held; fixed: all
                # three components of the plane's nodes (incon.f)
                if m.group(3) == "symmetry":
                    dk.symmetry.append(m.group(1))
                else:
                    dk.fixed_planes.append(m.group(1))
                continue
            if low.startswith("transformation matrix"):
                v = low.split()[2:]
                k = v.index("row_1")
                nodes = [int(x) for x in v[:k]]
                rows = []
                for r in ("row_1", "row_2", "row_3"):
                    j = v.index(r)
                    rows.append([float(x) for x in v[j + 1:j + 4]])
                for n in nodes:
                    dk.trn[n] = np.array(rows)
                continue
            v = low.split()
            k = 0
            while k < len(v) and v[k].isdigit():
                k += 1
            if k == 0 or (len(v) - k) % 2:
                raise DeckError(f"constraints: {ln}")
            for n in (int(x) for x in v[:k]):
                for j in range(k, len(v), 2):
                    if v[j] not in ("u", "v", "w"):
                        raise DeckError(f"constraints: {ln}")
                    dk.fixed[(n, "uvw".index(v[j]))] = float(v[j + 1])
            continue
        if mode == "steps":
            m = re.match(r"step\s+(\d+)\s*(-\s*(\d+))?\s+(\S+)\s+(\S+)", low)
            if not m:
                raise DeckError(f"steps: {ln}")
            a = int(m.group(1))
            b = int(m.group(3)) if m.group(3) else a
            for s in range(a, b + 1):
                dk.steps.setdefault(s, []).append((m.group(4),
                                                   float(m.group(5))))
            continue
        if mode == "nlparams":
            dk.nonlinear[low] = tok
            continue
        if mode == "ignore":
            continue
        raise DeckError(f"unsupported line: {ln}")
    if dk.incid is not None:
        dk.incid = np.ascontiguousarray(dk.incid[:, :dk.nnode])
    return dk


def _bilinear_material(props):
    """material ... properties bilinear (inmat.f): e, nu, yld_pt, tan_e,
    beta (1: isotropic hardening)."""
    d = _kv(props[1:])
    return _abi.mm01_material(float(d["e"]), float(d["nu"]),
                              float(d["yld_pt"]),
                              tan_e=float(d.get("tan_e", 0.0)),
                              beta=float(d.get("beta", 1.0)))


def _material(dk):
    """The deck's one material: (w3d_material, Euler angles or None)."""
    if len(dk.materials) != 1:
        raise DeckError("one material")
    (props,) = dk.materials.values()
    if props[0] == "bilinear":
        return _bilinear_material(props), None
    return _crystal_material(props[1:], dk)


def _crystal_material(props, dk):
    p = props
    d = {}
    k = 0
    while k < len(p):
        if p[k] in ("angles",):
            d["angles"] = [float(x) for x in p[k + 1:k + 4]]
            k += 4
            continue
        d[p[k]] = p[k + 1] if k + 1 < len(p) else ""
        k += 2
    if d.get("n_crystals", "1") != "1" or d.get("crystal_input") != \
            "single" or d.get("orientation_input") != "single":
        raise DeckError("one crystal with a single orientation")
    c = dk.crystals[int(d["crystal_type"])]
    if c.get("hardening") != "voce" or c.get("elastic_type") != "isotropic":
        raise DeckError("Voce hardening with isotropic elasticity")
    mat = _abi.mm10_material(
        tau_y=float(c["tau_y"]), tau_v=float(c["tau_v"]),
        theta_0=float(c["theta_0"]), voce_m=float(c["voce_m"]),
        harden_n=float(c["harden_n"]), slip_type=c["slip_type"],
        elastic="isotropic", e=float(c["e"]), nu=float(c["nu"]),
        angle_type=d.get("angle_type", "degrees"))
    if d.get("angle_convention", "kocks") != "kocks":
        raise DeckError("Kocks angles")
    return mat, d["angles"]


def face_loads(dk, pattern):
    """Consistent nodal forces (3*nonode) of a pattern's face pressures."""
    f = np.zeros((dk.nonode, 3))
    g = 1.0 / np.sqrt(3.0)
    for e, face, p in dk.patterns[pattern]:
        nodes = dk.incid[e - 1]
        fn = [nodes[k - 1] for k in HEX8_FACES[face]]
        x = dk.coords[np.array(fn) - 1]
        cen = dk.coords[nodes - 1].mean(axis=0)
        for a in (-g, g):
            for b in (-g, g):
                N = 0.25 * np.array([(1 - a) * (1 - b), (1 + a) * (1 - b),
                                     (1 + a) * (1 + b), (1 - a) * (1 + b)])
                da = 0.25 * np.array([-(1 - b), (1 - b), (1 + b), -(1 + b)])
                db = 0.25 * np.array([-(1 - a), -(1 + a), (1 + a), (1 - a)])
                ta, tb = da @ x, db @ x
                nv = np.cross(ta, tb)
                pt = N @ x
                if np.dot(nv, cen - pt) < 0.0:   # inward
                    nv = -nv
                for k in range(4):
                    f[fn[k] - 1] += N[k] * p * nv
    return f.ravel()


@dataclass
class Problem:
    arrays: object
    cstmap: np.ndarray
    dloads: list
    mfs: list
    nsteps: int
    dt: float
    newton: dict
    asymmetric: bool = False    # assembly asymmetric on


def nodal_loads(dk, pattern):
    """The nodal forces (3*nonode) of a pattern, global components."""
    f = np.zeros((dk.nonode, 3))
    for n, v in dk.nodal[pattern].items():
        f[n - 1] += v
    return f.ravel()


def build(dk):
    """Model arrays, constraints, the steps' load increments and the
    Newton options of a deck."""
    mat, angles = _material(dk)
    if dk.elem_opts.get("order") != "2x2x2":
        raise DeckError("2x2x2 integration")
    ne = dk.noelem
    nn = dk.nnode
    # 2x2x2 is order 1 of the 8-node brick and order 8 of the 20-node one
    order = 1 if dk.etype == _abi.W3D_L3DISOP else 8
    eps_bbar = 0.0
    nl = dk.nonlinear
    newton = dict(res_tol=0.01, mxiter=5, mniter=1, extrapolate=True,
                  line_search=False, adaptive=False, divergence_check=True)
    dt = 1.0
    asym = False
    for k, t in nl.items():
        if k.startswith("time step"):
            dt = float(t[-1])
        elif k.startswith("maximum iterations"):
            newton["mxiter"] = int(t[-1])
        elif k.startswith("minimum iterations"):
            newton["mniter"] = int(t[-1])
        elif k.startswith("divergence check"):
            newton["divergence_check"] = t[-1] == "on"
        elif k.startswith("convergence test norm res tol"):
            newton["res_tol"] = float(t[-1])
        elif k.startswith("adaptive"):
            newton["adaptive"] = t[1] == "on"
        elif k.startswith("extrapolate"):
            newton["extrapolate"] = t[1] == "on"
        elif k.startswith("line search"):
            newton["line_search"] = t[2] == "on"
        elif k.startswith("assembly asymmetric"):
            asym = t[-1] == "on"
        elif k.startswith("bbar stabilization factor"):
            eps_bbar = float(t[-1])
    arr = _abi.ModelArrays(
        dk.coords, dk.incid.ravel(), np.arange(ne, dtype=np.int32) * nn + 1,
        np.full(ne, dk.etype, np.int32), np.full(ne, order, np.int32),
        dk.elem_material, np.full(ne, 1 if dk.elem_opts.get("bbar") else 0,
                                  np.int32),
        [mat], auto_size=dk.auto_size, eps_bbar=eps_bbar,
        angles=(None if angles is None else
                np.tile(np.asarray(angles, float), (ne, 1))),
        geonl=np.full(ne, 1 if dk.elem_opts.get("geonl") else 0, np.int32),
        trnmat=dk.trn or None)
    cst = np.zeros(3 * dk.nonode, np.int32)
    for ax in dk.symmetry:
        c = "xyz".index(ax)
        cst[c::3][dk.coords[:, c] == 0.0] = 1
    for ax in dk.fixed_planes:
        on = dk.coords[:, "xyz".index(ax)] == 0.0
        for c in range(3):
            cst[c::3][on] = 1
    for (n, c), val in dk.fixed.items():
        if val != 0.0:
            raise DeckError("nonzero imposed displacements")
        cst[3 * (n - 1) + c] = 1
    nsteps = dk.compute_to or max(dk.steps)
    dloads, mfs = [], []
    for s in range(1, nsteps + 1):
        d = np.zeros(3 * dk.nonode)
        mf = 0.0
        for pat, fac in dk.steps[s]:
            if pat in dk.nodal:
                d += fac * nodal_loads(dk, pat)
            else:
                d += fac * face_loads(dk, pat)
            mf += fac
        # rotate_loads: forces at skewed nodes in their constraint system
        for n, t in dk.trn.items():
            d[3 * (n - 1):3 * n] = t @ d[3 * (n - 1):3 * n]
        d[cst != 0] = 0.0
        dloads.append(d)
        mfs.append(mf)
    return Problem(arr, cst, dloads, mfs, nsteps, dt, newton, asym)


def center_stress(model, elem, auto_size_blocks, geonl=False):
    """The element-centre stresses of an element (the mean of its 8 point
    values, oumkcv.f), from the model's n-state; auto_size_blocks: the
    model's blocking (rows of span, first element).  With finite strains
    the point values are the Cauchy stresses {T} = [qn1] {t} of the
    unrotated ones with [R n+1] (ougts1.f:117-126), here R t R^T."""
    for b, (span, first) in enumerate(auto_size_blocks, start=1):
        if first <= elem < first + span:
            s = model.get_block_state(b, _abi.STRESS_N).reshape(span, 8, 9)
            s = s[elem - first][:, :6]
            if not geonl:
                return s.mean(axis=0)
            r = model.get_block_state(b, _abi.ROT_N1).reshape(span, 8, 9)
            out = np.zeros((8, 6))
            for g in range(8):
                rr = r[elem - first, g].reshape(3, 3, order="F")
                t = s[g]
                tm = np.array([[t[0], t[3], t[5]], [t[3], t[1], t[4]],
                               [t[5], t[4], t[2]]])
                c = rr @ tm @ rr.T
                out[g] = [c[0, 0], c[1, 1], c[2, 2], c[0, 1], c[1, 2],
                          c[0, 2]]
            return out.mean(axis=0)
    raise DeckError(f"element {elem} not found")


def point_strain(model, elem, gp, auto_size_blocks, ngp, geonl=False):
    """The strains the reference prints for a Gauss point (ougts1.f:51-103):
    eps_xx, eps_yy, eps_zz, gam_xy, gam_yz, gam_xz from the model's
    n-state; with finite strains the accumulated unrotated strains d are
    turned to the current spatial axes, D = R d R^T with [R n+1]
    (ou_get_spatial_from_material, ougts1.f:306), shears kept as
    engineering strains.  gp is 1-based."""
    for b, (span, first) in enumerate(auto_size_blocks, start=1):
        if first <= elem < first + span:
            e = model.get_block_state(b, _abi.STRAIN_N).reshape(span, ngp, 6)
            d = e[elem - first, gp - 1].copy()
            if not geonl:
                return d
            r = model.get_block_state(b, _abi.ROT_N1).reshape(span, ngp, 9)
            rr = r[elem - first, gp - 1].reshape(3, 3, order="F")
            t = np.array([[d[0], 0.5 * d[3], 0.5 * d[5]],
                          [0.5 * d[3], d[1], 0.5 * d[4]],
                          [0.5 * d[5], 0.5 * d[4], d[2]]])
            c = rr @ t @ rr.T
            return np.array([c[0, 0], c[1, 1], c[2, 2], 2.0 * c[0, 1],
                             2.0 * c[1, 2], 2.0 * c[0, 2]])
    raise DeckError(f"element {elem} not found")
