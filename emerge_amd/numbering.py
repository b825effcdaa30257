"""A caller's Nedelec-2 DOF numbering and the map to this library's.

The reference numbers DOFs [edge mode 1 | face mode 1 | edge mode 2 | face
mode 2] (fem/elements/nedelec2.py:46-62) over its own edge and face lists,
whose order is Python set iteration order (fem/mesh3d.py:252-271).  This
library uses the same block layout over edges and faces sorted by vertex ids,
so the two numberings differ by one entity permutation applied to both modes.
``CallerNumbering`` finds it from the caller's field (``field.mesh.edges``,
``field.mesh.tris``) and moves matrices and vectors between the numberings on
the GPU (emg_csr_permute, emg_csr_map, emg_gather_c128, emg_scatter_c128).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import _lib
from .mesh import Mesh3D


def _caller_mesh(field):
    return field.mesh if hasattr(field, "mesh") else field


class CallerNumbering:
    """p[i] = this library's index of the caller's DOF i (and pinv, its inverse).

    field: the caller's Nedelec-2 field (or its mesh): ``mesh.nodes`` (3, nv),
    ``mesh.tets`` (4, nt), ``mesh.edges`` (2, ne), ``mesh.tris`` (3, nf).
    mesh: this library's Mesh3D of the same nodes and tets (built if None).
    """

    def __init__(self, field, mesh: Mesh3D | None = None, device: str = "cuda"):
        import torch
        cm = _caller_mesh(field)
        self.torch = torch
        self.dev = torch.device(device)
        self.mesh = mesh if mesh is not None else Mesh3D(np.asarray(cm.nodes), np.asarray(cm.tets),
                                                         tet_tags=getattr(cm, "tet_tags", None))
        m = self.mesh
        ce = np.sort(np.asarray(cm.edges, dtype=np.int64), axis=0)
        ct = np.sort(np.asarray(cm.tris, dtype=np.int64), axis=0)
        if ce.shape[1] != m.n_edges or ct.shape[1] != m.n_tris:
            raise ValueError(f"caller field has {ce.shape[1]} edges / {ct.shape[1]} faces, the mesh "
                             f"{m.n_edges} / {m.n_tris}")
        e_ours = m.get_edge(ce[0], ce[1])
        f_ours = m.get_tri(ct[0], ct[1], ct[2])
        if np.any(e_ours < 0) or np.any(f_ours < 0):
            raise ValueError("caller edges/faces are not those of the mesh")
        ent = np.concatenate([e_ours, f_ours + m.n_edges])            # caller entity -> our entity
        nent = m.n_entities
        if np.unique(ent).shape[0] != nent:
            raise ValueError("caller edge/face lists repeat an entity")
        self.n = 2 * nent
        self.p = np.concatenate([ent, ent + nent]).astype(np.int64)
        self.pinv = np.empty_like(self.p)
        self.pinv[self.p] = np.arange(self.n)
        self.p_d = torch.from_numpy(self.p.astype(np.int32)).to(self.dev)
        self.pinv_d = torch.from_numpy(self.pinv.astype(np.int32)).to(self.dev)
        self._perm = None        # (pattern key, caller rowptr, cols (host), map (device))

    # ------------------------------------------------------------ vectors
    def to_caller(self, x):
        """Vector in this library's numbering (device, complex128) -> caller's numbering (device)."""
        out = self.torch.empty(self.n, dtype=self.torch.complex128, device=self.dev)
        _lib.check(_lib.load().emg_gather_c128(self.n, _lib.ptr(self.p_d), self.n, _lib.ptr(x), _lib.ptr(out),
                                               _lib.stream_ptr()), "emg_gather_c128")
        return out

    def from_caller(self, x):
        """Vector in the caller's numbering (numpy or device) -> this library's numbering (device)."""
        torch = self.torch
        xd = torch.as_tensor(np.ascontiguousarray(x, dtype=np.complex128) if isinstance(x, np.ndarray) else x,
                             dtype=torch.complex128, device=self.dev).contiguous()
        out = torch.empty(self.n, dtype=torch.complex128, device=self.dev)
        _lib.check(_lib.load().emg_scatter_c128(self.n, _lib.ptr(self.p_d), self.n, _lib.ptr(xd), _lib.ptr(out), None,
                                                _lib.stream_ptr()), "emg_scatter_c128")
        return out

    # ------------------------------------------------------------ matrices on this library's pattern
    def permuted_pattern(self, pattern):
        """(rowptr, cols) of the caller-order CSR of `pattern` (host arrays) and the
        device map from each caller entry to the pattern's entry (built once)."""
        key = id(pattern)
        if self._perm is None or self._perm[0] != key:
            torch = self.torch
            rp = torch.empty(self.n + 1, dtype=torch.int32, device=self.dev)
            cols = torch.empty(pattern.nnz, dtype=torch.int32, device=self.dev)
            mp = torch.empty(pattern.nnz, dtype=torch.int32, device=self.dev)
            _lib.check(_lib.load().emg_csr_permute(self.n, _lib.ptr(pattern.rowptr), _lib.ptr(pattern.cols),
                                                   _lib.ptr(self.p_d), _lib.ptr(self.pinv_d), _lib.ptr(rp),
                                                   _lib.ptr(cols), _lib.ptr(mp), _lib.stream_ptr()), "emg_csr_permute")
            self._perm = (key, rp.cpu().numpy(), cols.cpu().numpy(), mp)
        return self._perm[1], self._perm[2], self._perm[3]

    def to_caller_csr(self, pattern, vals) -> sp.csr_matrix:
        """Values on this library's pattern (device) -> caller-numbered scipy CSR (host)."""
        rp, cols, mp = self.permuted_pattern(pattern)
        out = self.torch.empty(mp.shape[0], dtype=self.torch.complex128, device=self.dev)
        _lib.check(_lib.load().emg_gather_c128(mp.shape[0], _lib.ptr(mp), pattern.nnz, _lib.ptr(vals), _lib.ptr(out),
                                               _lib.stream_ptr()), "emg_gather_c128")
        return sp.csr_matrix((out.cpu().numpy(), cols, rp), shape=(self.n, self.n))

    def positions(self, pattern, A: sp.csr_matrix):
        """Device int32: the pattern position of every entry of the caller-order CSR A (-1 if absent)."""
        torch = self.torch
        rp = torch.from_numpy(np.ascontiguousarray(A.indptr, dtype=np.int32)).to(self.dev)
        ci = torch.from_numpy(np.ascontiguousarray(A.indices, dtype=np.int32)).to(self.dev)
        pos = torch.empty(A.nnz, dtype=torch.int32, device=self.dev)
        _lib.check(_lib.load().emg_csr_map(self.n, _lib.ptr(rp), _lib.ptr(ci), _lib.ptr(self.p_d),
                                           _lib.ptr(pattern.rowptr), _lib.ptr(pattern.cols), _lib.ptr(pos),
                                           _lib.stream_ptr()), "emg_csr_map")
        return pos

    def from_caller_csr(self, pattern, A: sp.csr_matrix, pos=None, out=None):
        """Caller-order CSR A (host) -> values on this library's pattern (device).  Raises if A
        has a nonzero outside the pattern (a matrix that is not a Nedelec-2 operator of this mesh)."""
        torch = self.torch
        if pos is None:
            pos = self.positions(pattern, A)
        vals = torch.from_numpy(np.ascontiguousarray(A.data, dtype=np.complex128)).to(self.dev)
        if out is None:
            out = torch.zeros(pattern.nnz, dtype=torch.complex128, device=self.dev)
        else:
            out.zero_()
        dropped = torch.zeros(1, dtype=torch.int64, device=self.dev)
        _lib.check(_lib.load().emg_scatter_c128(A.nnz, _lib.ptr(pos), pattern.nnz, _lib.ptr(vals), _lib.ptr(out),
                                                _lib.ptr(dropped), _lib.stream_ptr()), "emg_scatter_c128")
        nd = int(dropped.item())
        if nd:
            raise ValueError(f"{nd} nonzero entries of the matrix lie outside the mesh's Nedelec-2 pattern")
        return out
