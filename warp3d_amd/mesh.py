"""Synthetic structured meshes in WARP3D's element conventions.

Node numbering inside each element follows the reference's shape functions:
  * 8-node brick (l3disop, derivs.f:283): node k at natural coordinates
    (-,-,+) (-,-,-) (-,+,-) (-,+,+) (+,-,+) (+,-,-) (+,+,-) (+,+,+);
  * 20-node brick (q3disop, derivs.f:182): the same corners, then mid-edge
    nodes 9..20 at (-,-,0) (-,0,-) (-,+,0) (-,0,+) (+,-,0) (+,0,-) (+,+,0)
    (+,0,+) (0,-,+) (0,-,-) (0,+,-) (0,+,+);
  * 10-node tetrahedron (tet10, derivs.f:606): corners 1..4, then mid-edge
    nodes on edges 1-2, 2-3, 1-3, 1-4, 2-4, 3-4.
The natural axes map onto +x, +y, +z so every element has det J > 0.
"""
import numpy as np

from ._abi import (W3D_L3DISOP, W3D_Q3DISOP, W3D_TET10, ModelArrays)

HEX8_NAT = np.array([(-1, -1, 1), (-1, -1, -1), (-1, 1, -1), (-1, 1, 1),
                     (1, -1, 1), (1, -1, -1), (1, 1, -1), (1, 1, 1)])
HEX20_NAT = np.vstack([HEX8_NAT, np.array([
    (-1, -1, 0), (-1, 0, -1), (-1, 1, 0), (-1, 0, 1),
    (1, -1, 0), (1, 0, -1), (1, 1, 0), (1, 0, 1),
    (0, -1, 1), (0, -1, -1), (0, 1, -1), (0, 1, 1)])])
TET10_EDGES = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]


class Mesh:
    """coords (nonode, 3); conn (noelem, nn) 1-based; etype per element."""

    def __init__(self, coords, conn, etype, lattice=None, h=None):
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32)
        self.etype = etype
        self.lattice = lattice  # integer lattice position of every node
        self.h = h

    @property
    def nonode(self):
        return self.coords.shape[0]

    @property
    def noelem(self):
        return self.conn.shape[0]

    def model(self, material, int_order=None, bbar=None, elblks=None,
              auto_size=128, eps_bbar=0.0, angles=None, angle_stride=1,
              geonl=False):
        """ModelArrays for this mesh with one material for all elements
        (geonl: finite strains and rotations for every element)."""
        ne, nn = self.conn.shape
        if int_order is None:
            int_order = {W3D_L3DISOP: 1, W3D_Q3DISOP: 8, W3D_TET10: 4}[
                self.etype]
        if bbar is None:
            bbar = 1 if self.etype == W3D_L3DISOP else 0
        mats = material if isinstance(material, (list, tuple)) else [material]
        return ModelArrays(
            self.coords, self.conn.ravel(),
            np.arange(ne, dtype=np.int32) * nn + 1,
            np.full(ne, self.etype, np.int32),
            np.full(ne, int_order, np.int32),
            np.ones(ne, np.int32), np.full(ne, bbar, np.int32), mats,
            elblks=elblks, auto_size=auto_size, eps_bbar=eps_bbar,
            angles=angles, angle_stride=angle_stride,
            geonl=np.full(ne, 1 if geonl else 0, np.int32))


def random_orientations(n, seed=0, per_element=1):
    """Seeded SO(3)-uniform crystal orientations as Kocks Euler angles in
    degrees, (n * per_element, 3): uniform unit quaternions turned into
    rotations r, then psi = atan2(r32, r31), theta = acos(r33),
    phi = atan2(r23, r13) (the inverse of mm10_rotation_matrix)."""
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n * per_element, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    w, x, y, z = q.T
    r = np.empty((q.shape[0], 3, 3))
    r[:, 0, 0] = 1 - 2 * (y * y + z * z)
    r[:, 0, 1] = 2 * (x * y - z * w)
    r[:, 0, 2] = 2 * (x * z + y * w)
    r[:, 1, 0] = 2 * (x * y + z * w)
    r[:, 1, 1] = 1 - 2 * (x * x + z * z)
    r[:, 1, 2] = 2 * (y * z - x * w)
    r[:, 2, 0] = 2 * (x * z - y * w)
    r[:, 2, 1] = 2 * (y * z + x * w)
    r[:, 2, 2] = 1 - 2 * (x * x + y * y)
    psi = np.arctan2(r[:, 2, 1], r[:, 2, 0])
    theta = np.arccos(np.clip(r[:, 2, 2], -1.0, 1.0))
    phi = np.arctan2(r[:, 1, 2], r[:, 0, 2])
    return np.degrees(np.stack([psi, theta, phi], axis=1)) % 360.0


def _jitter(coords, lattice, top, h, frac, seed):
    """Move nodes strictly inside the box by U(-frac, frac) * h per axis."""
    if frac <= 0:
        return coords
    rng = np.random.default_rng(seed)
    inner = np.all((lattice > 0) & (lattice < top), axis=1)
    d = rng.uniform(-frac, frac, size=coords.shape) * h
    out = coords.copy()
    out[inner] += d[inner]
    return out


def box_hex8(nx, ny=None, nz=None, size=(1.0, 1.0, 1.0), jitter=0.0,
             seed=0):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    n = np.array([nx, ny, nz])
    h = np.asarray(size, float) / n
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1),
                          np.arange(nz + 1), indexing="ij")
    lat = np.stack([i.ravel(order="F"), j.ravel(order="F"),
                    k.ravel(order="F")], axis=1)
    coords = lat * h
    coords = _jitter(coords, lat, n, h, jitter, seed)

    def nid(a, b, c):
        return 1 + a + (nx + 1) * (b + (ny + 1) * c)

    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                             indexing="ij")
    ex, ey, ez = (v.ravel(order="F") for v in (ex, ey, ez))
    conn = np.empty((ex.size, 8), np.int32)
    for k_, (r, s, t) in enumerate(HEX8_NAT):
        conn[:, k_] = nid(ex + (r + 1) // 2, ey + (s + 1) // 2,
                          ez + (t + 1) // 2)
    return Mesh(coords, conn, W3D_L3DISOP, lattice=lat, h=h)


def box_hex20(nx, ny=None, nz=None, size=(1.0, 1.0, 1.0), jitter=0.0,
              seed=0):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    n = np.array([nx, ny, nz])
    h = np.asarray(size, float) / n
    # half lattice; keep corners and mid-edge points (at most one odd index)
    i, j, k = np.meshgrid(np.arange(2 * nx + 1), np.arange(2 * ny + 1),
                          np.arange(2 * nz + 1), indexing="ij")
    lat2 = np.stack([i.ravel(order="F"), j.ravel(order="F"),
                     k.ravel(order="F")], axis=1)
    keep = (lat2 % 2).sum(axis=1) <= 1
    ids = np.full(lat2.shape[0], -1, np.int64)
    ids[keep] = np.arange(keep.sum())
    lat2 = lat2[keep]
    coords = lat2 * (h / 2)
    coords = _jitter(coords, lat2, 2 * n, h / 2, jitter, seed)

    def nid(a, b, c):
        return 1 + ids[a + (2 * nx + 1) * (b + (2 * ny + 1) * c)]

    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                             indexing="ij")
    ex, ey, ez = (v.ravel(order="F") for v in (ex, ey, ez))
    conn = np.empty((ex.size, 20), np.int32)
    for k_, (r, s, t) in enumerate(HEX20_NAT):
        conn[:, k_] = nid(2 * ex + r + 1, 2 * ey + s + 1, 2 * ez + t + 1)
    return Mesh(coords, conn, W3D_Q3DISOP, lattice=lat2, h=h / 2)


def box_tet10(nx, ny=None, nz=None, size=(1.0, 1.0, 1.0), jitter=0.0,
              seed=0):
    """Each cube cell split into the 6 Kuhn tetrahedra (conforming)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    n = np.array([nx, ny, nz])
    h = np.asarray(size, float) / n
    i, j, k = np.meshgrid(np.arange(2 * nx + 1), np.arange(2 * ny + 1),
                          np.arange(2 * nz + 1), indexing="ij")
    lat2 = np.stack([i.ravel(order="F"), j.ravel(order="F"),
                     k.ravel(order="F")], axis=1)
    coords = lat2 * (h / 2)
    coords = _jitter(coords, lat2, 2 * n, h / 2, jitter, seed)

    def nid(p):
        return 1 + p[..., 0] + (2 * nx + 1) * (p[..., 1] + (2 * ny + 1) *
                                               p[..., 2])

    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1),
             (2, 1, 0)]
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                             indexing="ij")
    base = 2 * np.stack([ex.ravel(order="F"), ey.ravel(order="F"),
                         ez.ravel(order="F")], axis=1)
    conns = []
    for p in perms:
        v = [np.zeros(3, int)]
        for ax in p:
            e = np.zeros(3, int)
            e[ax] = 2
            v.append(v[-1] + e)
        v = np.array(v)
        d = np.linalg.det(np.array([v[1] - v[0], v[2] - v[0], v[3] - v[0]],
                                   float))
        if d < 0:
            v[[1, 2]] = v[[2, 1]]
        corners = base[:, None, :] + v[None, :, :]
        mids = [(corners[:, a] + corners[:, b]) // 2 for a, b in TET10_EDGES]
        allp = np.concatenate([corners, np.stack(mids, axis=1)], axis=1)
        conns.append(nid(allp))
    conn = np.stack(conns, axis=1).reshape(-1, 10).astype(np.int32)
    return Mesh(coords, conn, W3D_TET10, lattice=lat2, h=h / 2)


def uniaxial_bc(mesh, strain, size=1.0):
    """Symmetry planes x=0, y=0, z=0 and an imposed z-displacement
    strain*size on z=size.  Returns (cstmap, u_top_dofs mask, value)."""
    nd = 3 * mesh.nonode
    cst = np.zeros(nd, np.int32)
    x = mesh.coords
    tol = 1e-9 * size
    cst[0::3][np.abs(x[:, 0]) < tol] = 1
    cst[1::3][np.abs(x[:, 1]) < tol] = 1
    cst[2::3][np.abs(x[:, 2]) < tol] = 1
    top = np.zeros(nd, bool)
    top[2::3] = np.abs(x[:, 2] - size) < tol
    cst[top] = 1
    return cst, top, strain * size


class MixedMesh:
    """Meshes of different element types side by side along x, elements
    numbered part after part (blocks are contiguous element ranges of one
    type).  parts: (mesh, int_order, bbar).  With glue, the parts touch
    (no gap) and the nodes of a part that coincide with nodes of earlier
    parts become those nodes, so elements of different types share nodes
    (the mesh need not be conforming); otherwise the node sets are
    disjoint."""

    def __init__(self, parts, gap=0.1, glue=False):
        coords, incid, incmap, etype, order, bbar = [], [], [], [], [], []
        noff, x0, pos = 0, 0.0, 1
        known = {}
        for mesh, int_order, bb in parts:
            c = mesh.coords.copy()
            c[:, 0] += x0 - c[:, 0].min()
            x0 = c[:, 0].max() + (0.0 if glue else gap)
            ne, nn = mesh.conn.shape
            ids = np.arange(1, mesh.nonode + 1, dtype=np.int64)
            if glue:
                # lattice-rounded positions identify coincident nodes
                key = np.round(c / 1e-9).astype(np.int64)
                new = []
                for k in range(mesh.nonode):
                    t = tuple(key[k])
                    if t in known:
                        ids[k] = known[t]
                    else:
                        new.append(k)
                        ids[k] = noff + len(new)
                        known[t] = ids[k]
                c = c[new]
                ids_conn = ids[mesh.conn - 1]
                noff += len(new)
            else:
                ids_conn = mesh.conn + noff
                noff += mesh.nonode
            coords.append(c)
            incid.append(np.asarray(ids_conn).ravel())
            incmap.append(pos + nn * np.arange(ne, dtype=np.int32))
            pos += ne * nn
            etype.append(np.full(ne, mesh.etype, np.int32))
            order.append(np.full(ne, int_order, np.int32))
            bbar.append(np.full(ne, bb, np.int32))
        self.coords = np.concatenate(coords)
        self.incid = np.concatenate(incid).astype(np.int32)
        self.incmap = np.concatenate(incmap).astype(np.int32)
        self.etype = np.concatenate(etype)
        self.int_order = np.concatenate(order)
        self.bbar = np.concatenate(bbar)
        self.parts = parts

    @property
    def nonode(self):
        return self.coords.shape[0]

    @property
    def noelem(self):
        return self.etype.size

    def seeded_blocks(self, lo=32, hi=256, seed=0):
        """Blocking table (span, felem, 0, 0) with seeded spans uniform in
        [lo, hi], each block inside one element type."""
        rng = np.random.default_rng(seed)
        rows, e = [], 0
        while e < self.noelem:
            end = e + 1
            while end < self.noelem and self.etype[end] == self.etype[e]:
                end += 1
            while e < end:
                span = min(int(rng.integers(lo, hi + 1)), end - e)
                rows.append((span, e + 1, 0, 0))
                e += span
        return np.array(rows, np.int32)

    def model(self, material, elblks=None, auto_size=128, angles=None,
              angle_stride=1, geonl=False):
        mats = material if isinstance(material, (list, tuple)) else [material]
        return ModelArrays(
            self.coords, self.incid, self.incmap, self.etype, self.int_order,
            np.ones(self.noelem, np.int32), self.bbar, mats, elblks=elblks,
            auto_size=auto_size, angles=angles, angle_stride=angle_stride,
            geonl=np.full(self.noelem, 1 if geonl else 0, np.int32))
