"""CPU ORACLE -- test infrastructure only.

Restatement of the reference's CPU path for the Nedelec-2 hot path:
mesh topology and DOF numbering, the element loop, COO->CSR, and (in
``oracle/fd.py``) the frequency-domain pieces.  It is the checker for the
HIP path in ``emerge_amd`` and the host-core baseline timed by ``bench.py``.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it; the product path never does.

Every function cites the reference lines it follows.  Topology is built
with vectorised numpy instead of the reference's Python loops over sets and
dicts, but with the same conventions; the edge/face *order* of the
reference comes from Python set iteration (fem/mesh3d.py:252-271) and is
therefore arbitrary -- ``order_seed`` reproduces that by permuting them.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    """Compile oracle/liboracle.so (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I64 = ctypes.c_int64
        for sfx in ("", "_ld"):
            fn = getattr(L, "oracle_matrix_builder" + sfx)
            fn.argtypes = [P, I64, P, I64, P, I64, P, I64, P, P, P, P, P, I64, I64, P, P, P, P]
            fn.restype = None
            fn = getattr(L, "oracle_ned2_tet_stiff_mass" + sfx)
            fn.argtypes = [P] * 8
            fn.restype = None
        L.oracle_matinv.argtypes = [P, P]
        L.oracle_matinv.restype = None
        D = ctypes.c_double
        L.oracle_ned2_tri_stiff_force.argtypes = [P, D, D, P, P, ctypes.c_int, P, P]
        L.oracle_ned2_tri_stiff_force.restype = None
        L.oracle_ned2_tri_stiff.argtypes = [P, P, D, D, P]
        L.oracle_ned2_tri_stiff.restype = None
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- topology
class OracleMesh:
    """fem/mesh3d.py:224-355 (Mesh3D.update) restated on given nodes/tets.

    nodes: (3, nv) float64; tets: (4, nt) int64 in the caller's vertex order
    (the reference keeps gmsh's order).
    """

    def __init__(self, nodes: np.ndarray, tets: np.ndarray, order_seed: int | None = None):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.tets = np.ascontiguousarray(tets, dtype=np.int64)
        nv = self.nodes.shape[1]
        nt = self.tets.shape[1]
        st = np.sort(self.tets, axis=0)                       # mesh3d.py:255 sorted(...)
        i1, i2, i3, i4 = st
        e = np.concatenate([np.stack(p) for p in [(i1, i2), (i1, i3), (i1, i4), (i2, i3), (i2, i4), (i3, i4)]], axis=1)
        f = np.concatenate([np.stack(p) for p in [(i1, i2, i3), (i1, i2, i4), (i1, i3, i4), (i2, i3, i4)]], axis=1)
        edges = np.unique(e.T, axis=0).T                      # mesh3d.py:252-270 (set -> array)
        tris = np.unique(f.T, axis=0).T
        if order_seed is not None:                            # arbitrary set order, :267
            rng = np.random.default_rng(order_seed)
            edges = edges[:, rng.permutation(edges.shape[1])]
            tris = tris[:, rng.permutation(tris.shape[1])]
        self.edges = np.ascontiguousarray(edges)
        self.tris = np.ascontiguousarray(tris)
        ne, nf = edges.shape[1], tris.shape[1]
        ekey = edges[0] * nv + edges[1]
        eorder = np.argsort(ekey)
        esorted = ekey[eorder]

        def edge_id(a, b):                                    # mesh3d.py:136-142 get_edge
            k = np.minimum(a, b) * nv + np.maximum(a, b)
            return eorder[np.searchsorted(esorted, k)]

        tkey = (tris[0] * nv + tris[1]) * nv + tris[2]
        torder = np.argsort(tkey)
        tsorted = tkey[torder]

        def tri_id(a, b, c):                                  # mesh3d.py:152-154 get_tri
            s = np.sort(np.stack([a, b, c]), axis=0)
            k = (s[0] * nv + s[1]) * nv + s[2]
            return torder[np.searchsorted(tsorted, k)]

        T = self.tets
        # mesh3d.py:292 local pairs (1,2),(1,3),(1,4),(2,3),(4,2),(3,4)
        self.tet_to_edge = np.stack([edge_id(T[i - 1], T[j - 1]) for i, j in zip([1, 1, 1, 2, 4, 3], [2, 3, 4, 3, 2, 4])])
        # mesh3d.py:296 local triples (1,2,3),(1,3,4),(1,4,2),(2,3,4)
        self.tet_to_tri = np.stack([tri_id(T[i - 1], T[j - 1], T[k - 1]) for i, j, k in zip([1, 1, 1, 2], [2, 3, 4, 3], [3, 4, 2, 4])])
        # mesh3d.py:310-313 tri_to_tet (first two tets in tet order, -1 padded)
        tt = np.full((2, nf), -1, dtype=np.int64)
        flat_tri = self.tet_to_tri.T.reshape(-1)
        flat_tet = np.repeat(np.arange(nt), 4)
        o = np.argsort(flat_tri, kind="stable")
        ft, fk = flat_tri[o], flat_tet[o]
        first = np.ones_like(ft, dtype=bool)
        first[1:] = ft[1:] != ft[:-1]
        starts = np.flatnonzero(first)
        rank = np.arange(ft.shape[0]) - np.repeat(starts, np.diff(np.append(starts, ft.shape[0])))
        tt[rank, ft] = fk
        self.tri_to_tet = tt
        # mesh3d.py:330-335 tri_to_edge [(i1,i2),(i2,i3),(i1,i3)]
        self.tri_to_edge = np.stack([edge_id(tris[0], tris[1]), edge_id(tris[1], tris[2]), edge_id(tris[0], tris[2])])
        # mesh3d.py:351-353
        self.edge_lengths = np.sqrt(np.sum((self.nodes[:, edges[0]] - self.nodes[:, edges[1]]) ** 2, axis=0))
        a = self.nodes[:, tris[1]] - self.nodes[:, tris[0]]
        b = self.nodes[:, tris[2]] - self.nodes[:, tris[0]]
        self.areas = 0.5 * np.linalg.norm(np.cross(a.T, b.T), axis=1)
        self.centers = self.nodes[:, T].mean(axis=1)
        self.n_edges, self.n_tris, self.n_tets = ne, nf, nt
        self._edge_search = (eorder, esorted)
        self._tri_search = (torder, tsorted)

    def get_edge_ids(self, a, b):
        eorder, esorted = self._edge_search
        nv = self.nodes.shape[1]
        k = np.minimum(a, b) * nv + np.maximum(a, b)
        return eorder[np.searchsorted(esorted, k)]

    def get_tri_ids(self, tri_vertices: np.ndarray):
        torder, tsorted = self._tri_search
        nv = self.nodes.shape[1]
        s = np.sort(np.asarray(tri_vertices, dtype=np.int64), axis=0)
        k = (s[0] * nv + s[1]) * nv + s[2]
        return torder[np.searchsorted(tsorted, k)]


class OracleNedelec2:
    """fem/elements/nedelec2.py:32-70 DOF maps."""

    def __init__(self, mesh: OracleMesh):
        self.mesh = mesh
        ne, nf = mesh.n_edges, mesh.n_tris
        self.n_field = 2 * ne + 2 * nf
        t2f = np.zeros((20, mesh.n_tets), dtype=np.int64)
        t2f[:6] = mesh.tet_to_edge
        t2f[6:10] = mesh.tet_to_tri + ne
        t2f[10:16] = mesh.tet_to_edge + (nf + ne)
        t2f[16:20] = mesh.tet_to_tri + (nf + 2 * ne)
        self.tet_to_field = t2f
        self.edge_to_field = np.stack([np.arange(ne), np.arange(ne) + nf + ne])
        tri2f = np.zeros((8, nf), dtype=np.int64)
        tri2f[:3] = mesh.tri_to_edge
        tri2f[3] = np.arange(nf) + ne
        tri2f[4:7] = mesh.tri_to_edge + ne + nf
        tri2f[7] = np.arange(nf) + 2 * ne + nf
        self.tri_to_field = tri2f

    def dof_keys(self) -> np.ndarray:
        """(n_field, 4) int64 key per DOF: sorted vertex ids (-1 padded) + mode.

        Edge DOFs carry (v0, v1, -1, mode), face DOFs (v0, v1, v2, mode); the key
        is independent of the arbitrary edge/face numbering.
        """
        m = self.mesh
        ne, nf = m.n_edges, m.n_tris
        ek = np.stack([m.edges[0], m.edges[1], -np.ones(ne, dtype=np.int64)], axis=1)
        fk = m.tris.T.copy()
        ent = np.concatenate([ek, fk])
        keys = np.concatenate([np.concatenate([ent, np.zeros((ne + nf, 1), np.int64)], 1),
                               np.concatenate([ent, np.ones((ne + nf, 1), np.int64)], 1)])
        return keys


# ---------------------------------------------------------------- assembly
def material_tensors(eps: np.ndarray, mu: np.ndarray | None = None) -> tuple[np.ndarray, np.ndarray]:
    """Per-tet scalars -> (3,3,nt) complex tensors, as fem/material.py:44-55."""
    nt = eps.shape[0]
    er = np.zeros((3, 3, nt), dtype=np.complex128)
    ur = np.zeros((3, 3, nt), dtype=np.complex128)
    for k in range(3):
        er[k, k] = eps
        ur[k, k] = 1.0 if mu is None else mu
    return er, ur


def matrix_builder(field: OracleNedelec2, er: np.ndarray, ur: np.ndarray, t0: int = 0, t1: int | None = None,
                   extended: bool = False):
    """fem/physics/edm/optimized_assembly.py:66-116 over tets [t0, t1).

    extended=True evaluates the same algorithm in long double (the reference's
    absolute-coordinate volume/cofactor formulas lose digits to cancellation
    in fp64 on fine meshes away from the origin)."""
    m = field.mesh
    nt = m.n_tets
    t1 = nt if t1 is None else t1
    n = (t1 - t0) * 400
    dataE = np.empty(n, dtype=np.complex128)
    dataB = np.empty(n, dtype=np.complex128)
    rows = np.empty(n, dtype=np.int64)
    cols = np.empty(n, dtype=np.int64)
    er = np.ascontiguousarray(er, dtype=np.complex128)
    ur = np.ascontiguousarray(ur, dtype=np.complex128)
    fn = lib().oracle_matrix_builder_ld if extended else lib().oracle_matrix_builder
    fn(_p(m.nodes), m.nodes.shape[1], _p(m.tets), nt, _p(m.tris), m.n_tris,
                                _p(m.edges), m.n_edges, _p(m.edge_lengths), _p(field.tet_to_field),
                                _p(np.ascontiguousarray(m.tet_to_edge)), _p(ur), _p(er), t0, t1,
                                _p(dataE), _p(dataB), _p(rows), _p(cols))
    return dataE, dataB, rows, cols


def tet_mass_stiffness_matrices(field: OracleNedelec2, er: np.ndarray, ur: np.ndarray, extended: bool = False):
    """fem/physics/edm/optimized_assembly.py:43-64: E (curl-curl) and B (mass) CSR."""
    dataE, dataB, rows, cols = matrix_builder(field, er, ur, extended=extended)
    N = field.n_field
    E = sp.coo_matrix((dataE, (rows, cols)), shape=(N, N)).tocsr()
    B = sp.coo_matrix((dataB, (rows, cols)), shape=(N, N)).tocsr()
    return E, B


def element_matrices(verts: np.ndarray, edge_lengths: np.ndarray, lem: np.ndarray, ltm: np.ndarray,
                     Ms: np.ndarray, Mm: np.ndarray):
    """fem/mth/tet.py:854-1088 for one tet (verts (3,4), lem (2,6), ltm (3,4))."""
    D = np.empty((20, 20), dtype=np.complex128)
    F = np.empty((20, 20), dtype=np.complex128)
    lib().oracle_ned2_tet_stiff_mass(_p(np.ascontiguousarray(verts, dtype=np.float64)),
                                     _p(np.ascontiguousarray(edge_lengths, dtype=np.float64)),
                                     _p(np.ascontiguousarray(lem, dtype=np.int64)),
                                     _p(np.ascontiguousarray(ltm, dtype=np.int64)),
                                     _p(np.ascontiguousarray(Ms, dtype=np.complex128)),
                                     _p(np.ascontiguousarray(Mm, dtype=np.complex128)), _p(D), _p(F))
    return D, F


def matinv(s: np.ndarray) -> np.ndarray:
    """fem/mth/optimized.py:440-461 (with its det-multiplying branch)."""
    out = np.empty((3, 3), dtype=np.complex128)
    lib().oracle_matinv(_p(np.ascontiguousarray(s, dtype=np.complex128)), _p(out))
    return out


def local_mapping(vertex_ids: np.ndarray, ids: np.ndarray) -> np.ndarray:
    """fem/mth/optimized.py:273-307."""
    out = np.zeros(ids.shape, dtype=np.int64)
    for t in range(ids.shape[1]):
        for j in range(ids.shape[0]):
            for k in range(4):
                if vertex_ids[k] == ids[j, t]:
                    out[j, t] = k
                    break
    return out


def permutation_to(keys_src: np.ndarray, keys_dst: np.ndarray) -> np.ndarray:
    """Index array p with keys_dst[p[i]] == keys_src[i] (rows are unique keys)."""
    def pack(k):
        k = np.ascontiguousarray(k.astype(np.int64))
        return k.view([("", np.int64)] * k.shape[1]).reshape(-1)
    a, b = pack(keys_src), pack(keys_dst)
    ob = np.argsort(b)
    pos = np.searchsorted(b[ob], a)
    p = ob[pos]
    assert np.array_equal(b[p], a), "DOF key sets differ"
    return p
