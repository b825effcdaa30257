"""Mesh topology for the Nedelec-2 path (reference: fem/mesh3d.py, fem/elements/nedelec2.py).

``Mesh3D`` keeps the reference's attribute names and index conventions
(fem/mesh3d.py:224-355) so callers can use it where they used the reference's
``Mesh3D``, but it is built from plain arrays instead of a gmsh session, and
it carries its own face-tag -> triangle and volume-tag -> tet maps in place
of the gmsh queries of ``get_triangles`` / ``get_tetrahedra``
(fem/mesh3d.py:179-207).

Differences, all deliberate:
  * edges and triangles are numbered in lexicographic order of their sorted
    vertex ids (the reference's order is Python set iteration order,
    fem/mesh3d.py:252-271, i.e. arbitrary); every result that depends on the
    numbering is compared by DOF key.
  * the kernels use *canonical* tets (vertices sorted by global id) and the
    per-tet entity table ``tet_ent`` (6 edges + 4 faces in the canonical local
    order (0,1)(0,2)(0,3)(1,2)(1,3)(2,3) / (0,1,2)(0,2,3)(0,1,3)(1,2,3)).
"""
from __future__ import annotations

import time

import numpy as np

CANON_EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))
CANON_FACES = ((0, 1, 2), (0, 2, 3), (0, 1, 3), (1, 2, 3))


def _tri_ordering(i1, i2, i3):
    """fem/mesh3d.py:37-51 (as written, including its `i2-1`)."""
    return np.sign(np.sign(i2 - 1) + np.sign(i3 - i2) + np.sign(i1 - i3))


class Mesh3D:
    """Tetrahedral mesh with the reference's topology arrays.

    Args:
        nodes: (3, nv) float64 vertex coordinates in metres.
        tets: (4, nt) int vertex ids per tet (any vertex order).
        tet_tags: optional (nt,) int volume tag per tet.
        face_tags: optional dict {tag: (3, k) int vertex triples} of tagged
            boundary/interface triangles (the stand-in for gmsh surface tags).
        topology: "host" (numpy) or "device" (the GPU build of emg_topology_build,
            identical arrays; ``topology_seconds`` records either's time).
    """

    def __init__(self, nodes: np.ndarray, tets: np.ndarray, tet_tags: np.ndarray | None = None,
                 face_tags: dict[int, np.ndarray] | None = None, topology: str = "host"):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.tets = np.ascontiguousarray(tets, dtype=np.int64)
        nv, nt = self.nodes.shape[1], self.tets.shape[1]
        if nt == 0:
            raise ValueError("mesh has no tetrahedra")
        if self.tets.min() < 0 or self.tets.max() >= nv:
            raise ValueError("tet vertex id out of range")
        self.tet_tags = np.zeros(nt, dtype=np.int64) if tet_tags is None else np.asarray(tet_tags, dtype=np.int64)

        t0 = time.perf_counter()
        self._nv = nv
        if topology == "device":
            self._topology_device()
        elif topology == "host":
            self._topology_host()
        else:
            raise ValueError(f"topology is 'host' or 'device', not {topology!r}")
        self.topology_seconds = time.perf_counter() - t0
        t = self.tris
        T = self.tets
        # fem/mesh3d.py:249, 273, 351-353
        self.centers = self.nodes[:, T].mean(axis=1)
        self.tri_centers = self.nodes[:, t].mean(axis=1)
        self.edge_centers = self.nodes[:, self.edges].mean(axis=1)
        if topology == "host":
            self.edge_lengths = np.sqrt(np.sum((self.nodes[:, self.edges[0]] - self.nodes[:, self.edges[1]]) ** 2, axis=0))
        a = self.nodes[:, t[1]] - self.nodes[:, t[0]]
        b = self.nodes[:, t[2]] - self.nodes[:, t[0]]
        self.areas = 0.5 * np.linalg.norm(np.cross(a.T, b.T), axis=1)

        self.face_tags: dict[int, np.ndarray] = {}
        for tag, tri in (face_tags or {}).items():
            tri = np.asarray(tri, dtype=np.int64).reshape(3, -1)
            self.face_tags[int(tag)] = np.unique(self.get_tri(tri[0], tri[1], tri[2]))
        self.defined = True

    def _topology_host(self):
        """Mesh3D.update's topology (fem/mesh3d.py:252-336) with numpy sorts and uniques."""
        nv, nt = self._nv, self.tets.shape[1]
        canon = np.sort(self.tets, axis=0)                                    # (4, nt)
        self.tets_canonical = np.ascontiguousarray(canon.T.astype(np.int32))  # (nt, 4)
        e_all = np.stack([np.stack([canon[a], canon[b]]) for a, b in CANON_EDGES], axis=1)      # (2, 6, nt)
        f_all = np.stack([np.stack([canon[a], canon[b], canon[c]]) for a, b, c in CANON_FACES], axis=1)  # (3, 4, nt)
        ekeys = e_all[0] * nv + e_all[1]                                      # (6, nt)
        uek, einv = np.unique(ekeys.ravel(), return_inverse=True)
        self.edges = np.stack([uek // nv, uek % nv]).astype(np.int64)
        fkeys = self._face_key(f_all[0], f_all[1], f_all[2])
        ufk, finv = np.unique(fkeys.ravel(), return_inverse=True)
        self.tris = self._face_unkey(ufk)
        self._ekeys, self._fkeys = uek, ufk
        ne, nf = self.edges.shape[1], self.tris.shape[1]
        tet_edge_c = einv.reshape(6, nt)
        tet_face_c = finv.reshape(4, nt)
        self.tet_ent = np.ascontiguousarray(np.concatenate([tet_edge_c, tet_face_c + ne]).T.astype(np.int32))  # (nt, 10)

        T = self.tets
        # fem/mesh3d.py:292 local pairs (1,2),(1,3),(1,4),(2,3),(4,2),(3,4) in the tet's own order
        self.tet_to_edge = np.stack([self.get_edge(T[i - 1], T[j - 1]) for i, j in zip([1, 1, 1, 2, 4, 3], [2, 3, 4, 3, 2, 4])])
        self.tet_to_edge_sign = np.stack([np.where(T[i - 1] > T[j - 1], -1, 1) for i, j in zip([1, 1, 1, 2, 4, 3], [2, 3, 4, 3, 2, 4])])
        # fem/mesh3d.py:296 local triples (1,2,3),(1,3,4),(1,4,2),(2,3,4)
        self.tet_to_tri = np.stack([self.get_tri(T[i - 1], T[j - 1], T[k - 1]) for i, j, k in zip([1, 1, 1, 2], [2, 3, 4, 3], [3, 4, 2, 4])])
        self.tet_to_tri_sign = np.stack([_tri_ordering(T[0], T[1], T[2]), _tri_ordering(T[0], T[2], T[3]),
                                         _tri_ordering(T[0], T[3], T[1]), _tri_ordering(T[1], T[2], T[3])])
        # fem/mesh3d.py:310-313 tri_to_tet, -1 padded, tets in ascending order
        flat_tri = self.tet_to_tri.T.reshape(-1)
        flat_tet = np.repeat(np.arange(nt), 4)
        o = np.argsort(flat_tri, kind="stable")
        ft, fk = flat_tri[o], flat_tet[o]
        first = np.ones(ft.shape[0], dtype=bool)
        first[1:] = ft[1:] != ft[:-1]
        starts = np.flatnonzero(first)
        rank = np.arange(ft.shape[0]) - np.repeat(starts, np.diff(np.append(starts, ft.shape[0])))
        if rank.max() > 1:
            raise ValueError("non-manifold mesh: a triangle is shared by more than two tets")
        self.tri_to_tet = np.full((2, nf), -1, dtype=np.int64)
        self.tri_to_tet[rank, ft] = fk
        # fem/mesh3d.py:330-336
        t = self.tris
        self.tri_to_edge = np.stack([self.get_edge(t[0], t[1]), self.get_edge(t[1], t[2]), self.get_edge(t[0], t[2])])
        self.tri_to_edge_sign = np.stack([np.ones(nf, np.int64), np.ones(nf, np.int64), -np.ones(nf, np.int64)])

    def _topology_device(self):
        """The same arrays from the GPU (emg_topology_build): one stable radix sort of edge
        keys and one of face keys, then binary searches; copied back to host numpy."""
        from .topology import device_topology
        out = device_topology(self.nodes, self.tets)
        for k, v in out.items():
            setattr(self, k, v)
        nv = self._nv
        self._ekeys = self.edges[0] * nv + self.edges[1]
        self._fkeys = self._face_key(self.tris[0], self.tris[1], self.tris[2])
        nf = self.tris.shape[1]
        self.tri_to_edge_sign = np.stack([np.ones(nf, np.int64), np.ones(nf, np.int64), -np.ones(nf, np.int64)])

    # ------------------------------------------------------------ keys
    def _face_key(self, a, b, c):
        nv = self._nv
        return (np.asarray(a, dtype=np.int64) * nv + b) * nv + c

    def _face_unkey(self, k):
        nv = self._nv
        return np.stack([k // (nv * nv), (k // nv) % nv, k % nv]).astype(np.int64)

    # ------------------------------------------------------------ sizes (fem/mesh3d.py:116-134)
    @property
    def n_edges(self) -> int:
        return self.edges.shape[1]

    @property
    def n_tets(self) -> int:
        return self.tets.shape[1]

    @property
    def n_tris(self) -> int:
        return self.tris.shape[1]

    @property
    def n_nodes(self) -> int:
        return self.nodes.shape[1]

    @property
    def n_entities(self) -> int:
        return self.n_edges + self.n_tris

    # ------------------------------------------------------------ lookups (fem/mesh3d.py:136-158)
    def get_edge(self, i1, i2):
        """Edge id(s) of vertex pair(s); -1 where absent (reference returns None)."""
        i1 = np.asarray(i1, dtype=np.int64)
        i2 = np.asarray(i2, dtype=np.int64)
        if np.any(i1 == i2):
            raise ValueError("Edge cannot be formed by the same node.")
        k = np.minimum(i1, i2) * self._nv + np.maximum(i1, i2)
        pos = np.searchsorted(self._ekeys, k)
        pos = np.minimum(pos, self._ekeys.shape[0] - 1)
        return np.where(self._ekeys[pos] == k, pos, -1)

    def get_tri(self, i1, i2, i3):
        """Triangle id(s) of vertex triple(s) in any order; -1 where absent."""
        s = np.sort(np.stack([np.asarray(i1, np.int64), np.asarray(i2, np.int64), np.asarray(i3, np.int64)]), axis=0)
        k = self._face_key(s[0], s[1], s[2])
        pos = np.minimum(np.searchsorted(self._fkeys, k), self._fkeys.shape[0] - 1)
        return np.where(self._fkeys[pos] == k, pos, -1)

    def get_tet(self, i1, i2, i3, i4):
        """Tet id of a vertex quadruple (any order), or None."""
        q = np.sort([int(i1), int(i2), int(i3), int(i4)])
        hit = np.flatnonzero(np.all(self.tets_canonical == q, axis=1))
        return int(hit[0]) if hit.size else None

    # ------------------------------------------------------------ selections (stand-in for gmsh)
    def get_triangles(self, face_tags) -> np.ndarray:
        """Triangle ids of the given face tags (fem/mesh3d.py:191-207)."""
        if isinstance(face_tags, (int, np.integer)):
            face_tags = [int(face_tags)]
        out = [self.face_tags[int(t)] for t in face_tags]
        return np.unique(np.concatenate(out)) if out else np.zeros(0, dtype=np.int64)

    def get_tetrahedra(self, vol_tags) -> np.ndarray:
        """Tet ids of the given volume tags (fem/mesh3d.py:179-189)."""
        if isinstance(vol_tags, (int, np.integer)):
            vol_tags = [int(vol_tags)]
        return np.flatnonzero(np.isin(self.tet_tags, np.asarray(vol_tags)))

    def boundary_surface(self, face_tags, origin):
        """The tagged triangles as a SurfaceMesh with normals pointing away from origin
        (fem/mesh3d.py:450-456)."""
        from .farfield import SurfaceMesh
        return SurfaceMesh(self, self.get_triangles(face_tags), origin)

    def boundary_triangles(self) -> np.ndarray:
        """Triangles with a single adjacent tet (the exterior surface)."""
        return np.flatnonzero(self.tri_to_tet[1] < 0)

    def retreive(self, material_selector, volumes) -> np.ndarray:
        """(3,3,nt) material tensors (fem/mesh3d.py:358-378); ``volumes`` is a
        list of objects with ``.tags`` (volume tags) and ``.material``."""
        arry = np.zeros((3, 3, self.n_tets), dtype=np.complex128)
        xs, ys, zs = self.centers
        for volume in volumes:
            ids = self.get_tetrahedra(volume.tags)
            arry[:, :, ids] = material_selector(volume.material, xs[ids], ys[ids], zs[ids])
        return arry
