"""Structured tetrahedral meshes for the benchmark and test workloads.

gmsh is not available, so every workload is meshed here: a box of nx*ny*nz
hexahedra, each split into 6 tetrahedra around its main diagonal (Kuhn
split, conforming across hexes), interior vertices jittered, all tets
positively oriented (as gmsh's are).  Region tags come from tet centroids;
the six box faces carry face tags 1..6 (xmin, xmax, ymin, ymax, zmin, zmax).
"""
from __future__ import annotations

from itertools import permutations

import numpy as np

from .mesh import Mesh3D

XMIN, XMAX, YMIN, YMAX, ZMIN, ZMAX = 1, 2, 3, 4, 5, 6


def box_nodes(nx: int, ny: int, nz: int, size=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0)):
    """(3, nv) grid vertices, x fastest; returns nodes and the (nx+1,ny+1,nz+1) id grid."""
    xs = origin[0] + np.linspace(0, size[0], nx + 1)
    ys = origin[1] + np.linspace(0, size[1], ny + 1)
    zs = origin[2] + np.linspace(0, size[2], nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()])
    ids = np.arange(nodes.shape[1]).reshape(nz + 1, ny + 1, nx + 1).transpose(2, 1, 0)
    return nodes, ids


def kuhn_tets(ids: np.ndarray) -> np.ndarray:
    """(4, 6*nhex) tets of the Kuhn split of every hex of an id grid."""
    nx, ny, nz = (s - 1 for s in ids.shape)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.transpose(2, 1, 0).ravel(), j.transpose(2, 1, 0).ravel(), k.transpose(2, 1, 0).ravel()
    out = []
    unit = np.eye(3, dtype=np.int64)
    for perm in permutations(range(3)):
        p = np.zeros(3, dtype=np.int64)
        verts = [ids[i, j, k]]
        for ax in perm:
            p = p + unit[ax]
            verts.append(ids[i + p[0], j + p[1], k + p[2]])
        out.append(np.stack(verts))
    # interleave so the 6 tets of a hex are consecutive
    return np.stack(out, axis=1).reshape(4, -1)


def orient_positive(nodes: np.ndarray, tets: np.ndarray) -> np.ndarray:
    """Swap the last two vertices of negatively oriented tets."""
    v = nodes[:, tets]
    det = np.einsum("in,in->n", v[:, 1] - v[:, 0], np.cross((v[:, 2] - v[:, 0]).T, (v[:, 3] - v[:, 0]).T).T)
    t = tets.copy()
    neg = det < 0
    t[2, neg], t[3, neg] = tets[3, neg], tets[2, neg]
    return t


def jitter_interior(nodes: np.ndarray, ids: np.ndarray, h: np.ndarray, frac: float, rng) -> np.ndarray:
    """Move interior vertices by U(-frac, frac)*h per axis (boundary vertices stay)."""
    nodes = nodes.copy()
    inner = ids[1:-1, 1:-1, 1:-1].ravel()
    d = rng.uniform(-frac, frac, size=(3, inner.shape[0])) * np.asarray(h)[:, None]
    nodes[:, inner] += d
    return nodes


def box_face_tags(nodes: np.ndarray, tets: np.ndarray, size, origin=(0.0, 0.0, 0.0), tol=1e-9) -> dict[int, np.ndarray]:
    """Boundary triangles of the box grouped by side (tags 1..6)."""
    faces = np.concatenate([tets[[a, b, c]] for a, b, c in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))], axis=1)
    s = np.sort(faces, axis=0)
    key = np.ascontiguousarray(s.T).view([("", np.int64)] * 3).ravel()
    _, idx, cnt = np.unique(key, return_index=True, return_counts=True)
    bnd = s[:, idx[cnt == 1]]
    out = {}
    for tag, axis, val in ((XMIN, 0, origin[0]), (XMAX, 0, origin[0] + size[0]), (YMIN, 1, origin[1]),
                           (YMAX, 1, origin[1] + size[1]), (ZMIN, 2, origin[2]), (ZMAX, 2, origin[2] + size[2])):
        on = np.all(np.abs(nodes[axis, bnd] - val) < tol * max(size), axis=0)
        out[tag] = bnd[:, on]
    return out


def structured_box(nx, ny, nz, size=(1.0, 1.0, 1.0), jitter=0.0, seed=0, origin=(0.0, 0.0, 0.0),
                   keep=None, tag_fn=None) -> Mesh3D:
    """Jittered Kuhn box mesh.

    keep: optional predicate on tet centroids (3, nt) -> bool mask of tets to keep
          (used to carve PEC holes).
    tag_fn: optional centroid -> (nt,) int volume tags.
    """
    rng = np.random.default_rng(seed)
    nodes, ids = box_nodes(nx, ny, nz, size, origin)
    h = np.array([size[0] / nx, size[1] / ny, size[2] / nz])
    if jitter > 0:
        nodes = jitter_interior(nodes, ids, h, jitter, rng)
    tets = orient_positive(nodes, kuhn_tets(ids))
    cent = nodes[:, tets].mean(axis=1)
    if keep is not None:
        tets = tets[:, keep(cent)]
        cent = nodes[:, tets].mean(axis=1)
        used = np.unique(tets)
        remap = -np.ones(nodes.shape[1], dtype=np.int64)
        remap[used] = np.arange(used.shape[0])
        nodes, tets = nodes[:, used], remap[tets]
    tags = np.ones(tets.shape[1], dtype=np.int64) if tag_fn is None else tag_fn(cent)
    faces = box_face_tags(nodes, tets, size, origin)
    return Mesh3D(nodes, tets, tet_tags=tags, face_tags=faces)


# ------------------------------------------------------------------ workloads
def assembly_box(nx=50, ny=50, nz=67, seed=1):
    """Assembly workload: unit box, nx*ny*nz hexes x 6 tets, interior vertices
    jittered U(-0.1, 0.1)*h per axis; per tet eps_r' ~ U[1,10], tan(delta) ~
    U[0, 0.02] (eps_r = eps_r'(1 - j tan(delta))), mu_r = 1.  Random numbers
    are drawn in that order from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    nodes, ids = box_nodes(nx, ny, nz, (1.0, 1.0, 1.0))
    h = np.array([1.0 / nx, 1.0 / ny, 1.0 / nz])
    nodes = jitter_interior(nodes, ids, h, 0.1, rng)
    tets = orient_positive(nodes, kuhn_tets(ids))
    nt = tets.shape[1]
    er = rng.uniform(1.0, 10.0, nt)
    tand = rng.uniform(0.0, 0.02, nt)
    eps = er * (1 - 1j * tand)
    return Mesh3D(nodes, tets), eps


def waveguide_wr90(nx=15, ny=7, nz=32, seed=0, length=50e-3):
    """Air-filled WR-90 section 22.86 x 10.16 x `length` mm, ports on z faces."""
    size = (22.86e-3, 10.16e-3, length)
    return structured_box(nx, ny, nz, size, jitter=0.05, seed=seed)


POST_TAG, PORT1_TAG, PORT2_TAG = 7, 8, 9
PORT_WIDTH, PORT_HEIGHT = 2e-3, 3.75e-3


def posts_box(nx=80, ny=40, nz=16, seed=2, size=(40e-3, 20e-3, 10e-3), radius=1.5e-3, height=8e-3,
              centers_x=(8e-3, 16e-3, 24e-3, 32e-3), port_width=PORT_WIDTH, port_height=PORT_HEIGHT):
    """Sweep workload: box with PEC cylindrical posts on the centre line and
    two lumped-port rectangles on the x end faces.

    The posts (radius, height from z = 0) are carved out by tet centroid, so
    their staircase surfaces become exterior faces (PEC by default); they carry
    face tag POST_TAG.  On x = 0 and x = size[0] the triangles with centroid in
    |y - yc| < port_width / 2, z < port_height (the strip from the floor up)
    carry PORT1_TAG and PORT2_TAG; with the default 80 x 40 x 16 hexes these
    edges fall on grid lines.  80 x 40 x 16 hexes give 307,200 tets before
    carving, 298,447 after.
    """
    yc = size[1] / 2

    def keep(c):
        inside = np.zeros(c.shape[1], dtype=bool)
        for xc in centers_x:
            inside |= ((c[0] - xc) ** 2 + (c[1] - yc) ** 2 < radius ** 2) & (c[2] < height)
        return ~inside

    mesh = structured_box(nx, ny, nz, size=size, jitter=0.05, seed=seed, keep=keep)
    # tag the remaining exterior faces (the post walls)
    bnd = mesh.boundary_triangles()
    tagged = np.unique(np.concatenate(list(mesh.face_tags.values()))) if mesh.face_tags else np.zeros(0, int)
    post = np.setdiff1d(bnd, tagged)
    mesh.face_tags[POST_TAG] = post
    for side, tag in ((XMIN, PORT1_TAG), (XMAX, PORT2_TAG)):
        tri = mesh.face_tags[side]
        c = mesh.nodes[:, mesh.tris[:, tri]].mean(axis=1)
        inside = (np.abs(c[1] - yc) < port_width / 2) & (c[2] < port_height)
        mesh.face_tags[tag] = tri[inside]
        mesh.face_tags[side] = tri[~inside]
    return mesh


BLOCK_TAG = 2


def cavity_block(nx=48, ny=32, nz=24, seed=3, size=(30e-3, 20e-3, 15e-3), block=((3e-3, 13e-3), (3e-3, 13e-3), (3e-3, 8e-3))):
    """Eigenmode workload: PEC box 30 x 20 x 15 mm with a dielectric block
    10 x 10 x 5 mm placed 3 mm from one corner; tets with centroid inside the
    block carry volume tag BLOCK_TAG (the rest 1).  48 x 32 x 24 hexes give
    221,184 tets."""
    (x0, x1), (y0, y1), (z0, z1) = block

    def tag(c):
        inside = (c[0] > x0) & (c[0] < x1) & (c[1] > y0) & (c[1] < y1) & (c[2] > z0) & (c[2] < z1)
        return np.where(inside, BLOCK_TAG, 1)

    return structured_box(nx, ny, nz, size=size, jitter=0.05, seed=seed, tag_fn=tag)


PATCH_TAG, GROUND_TAG, FEED_PORT_TAG = 10, 11, 12
SUBSTRATE_TAG = 2


def tensor_nodes(xs, ys, zs):
    """(3, nv) vertices of a tensor grid of coordinate arrays, x fastest, and the id grid."""
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()])
    ids = np.arange(nodes.shape[1]).reshape(len(zs), len(ys), len(xs)).transpose(2, 1, 0)
    return nodes, ids


def patch_antenna(cell=1.25e-3, half=45e-3, below=12.5e-3, above=25e-3, h_sub=1.6e-3, sub_layers=2, ground=60e-3,
                  patch_w=38e-3, patch_l=29e-3, line_w=2.5e-3, inset=8e-3, gap=1.0e-3, port_y=-27.5e-3,
                  jitter=0.05, seed=4, er_sub=4.4, scale_w=1.0, scale_l=1.0):
    """Microstrip patch antenna in an air box: a tensor grid of Kuhn tets over
    [-half, half]^2 x [-h_sub - below, above] with a uniform `cell` in x, y and above
    the substrate, `sub_layers` layers in the h_sub substrate (|x|, |y| <= ground / 2,
    -h_sub < z < 0, volume tag SUBSTRATE_TAG).  Interior triangles carry the conductors:
    GROUND_TAG (z = -h_sub, |x|, |y| <= ground / 2), PATCH_TAG (z = 0: the patch_w x
    patch_l patch, centred, with an inset feed notch of depth `inset` and gap `gap` beside a
    line_w wide line that runs to y = port_y) and FEED_PORT_TAG (the line_w x h_sub strip
    at y = port_y between ground and line).  The outer faces keep tags 1..6 (for an
    absorbing boundary).  scale_w / scale_l perturb the patch size (training variants).
    Vertices not on the conductor planes are jittered by `jitter` of the local spacing."""
    rng = np.random.default_rng(seed)
    n_xy = int(round(2 * half / cell))
    xs = np.linspace(-half, half, n_xy + 1)
    ys = xs.copy()
    n_below = int(round(below / cell))
    n_above = int(round(above / cell))
    zs = np.concatenate([np.linspace(-h_sub - below, -h_sub, n_below + 1)[:-1],
                         np.linspace(-h_sub, 0.0, sub_layers + 1)[:-1], np.linspace(0.0, above, n_above + 1)])
    nodes, ids = tensor_nodes(xs, ys, zs)
    if jitter > 0:
        inner = ids[1:-1, 1:-1, 1:-1].ravel()
        z = nodes[2, inner]
        y = nodes[1, inner]
        free = (np.abs(z + h_sub) > 1e-12) & (np.abs(z) > 1e-12) & (np.abs(y - port_y) > 1e-12)
        inner = inner[free]
        dz = np.diff(zs)
        kz = np.clip(np.searchsorted(zs, nodes[2, inner]) - 1, 0, len(dz) - 1)
        hz = np.minimum(dz[kz], dz[np.minimum(kz + 1, len(dz) - 1)])
        d = rng.uniform(-jitter, jitter, size=(3, inner.shape[0]))
        nodes = nodes.copy()
        nodes[0, inner] += d[0] * cell
        nodes[1, inner] += d[1] * cell
        nodes[2, inner] += d[2] * hz
    tets = orient_positive(nodes, kuhn_tets(ids))
    cent = nodes[:, tets].mean(axis=1)
    g2 = ground / 2
    sub = (np.abs(cent[0]) < g2) & (np.abs(cent[1]) < g2) & (cent[2] > -h_sub) & (cent[2] < 0)
    tags = np.where(sub, SUBSTRATE_TAG, 1)
    size = (2 * half, 2 * half, zs[-1] - zs[0])
    origin = (-half, -half, zs[0])
    mesh = Mesh3D(nodes, tets, tet_tags=tags, face_tags=box_face_tags(nodes, tets, size, origin))
    # interior triangles of the conductor planes (grid planes: their vertices are not jittered)
    s = mesh.tris
    fz = nodes[2, s]
    fc = mesh.tri_centers
    pw, pl = patch_w * scale_w / 2, patch_l * scale_l / 2
    lw = line_w / 2
    on_z0 = np.all(np.abs(fz) < 1e-12, axis=0)
    in_patch = (np.abs(fc[0]) < pw) & (np.abs(fc[1]) < pl)
    notch = (np.abs(fc[0]) < lw + gap) & (fc[1] < -pl + inset) & (fc[1] > -pl)
    line = (np.abs(fc[0]) < lw) & (fc[1] > port_y) & (fc[1] < -pl + inset)
    mesh.face_tags[PATCH_TAG] = np.flatnonzero(on_z0 & ((in_patch & ~notch) | line))
    on_g = np.all(np.abs(fz + h_sub) < 1e-12, axis=0)
    mesh.face_tags[GROUND_TAG] = np.flatnonzero(on_g & (np.abs(fc[0]) < g2) & (np.abs(fc[1]) < g2))
    on_p = np.all(np.abs(nodes[1, s] - port_y) < 1e-12, axis=0)
    mesh.face_tags[FEED_PORT_TAG] = np.flatnonzero(on_p & (np.abs(fc[0]) < lw) & (fc[2] > -h_sub) & (fc[2] < 0))
    return mesh
