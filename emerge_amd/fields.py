"""Field reconstruction at arbitrary points on the GPU.

``EMDataSet.interpolate(xs, ys, zs)`` (fem/physics/edm/emdata.py:181-199)
evaluates E and H = curl E / (-j omega mu) of a solution at any points.  The
reference searches every tet for every point (fem/mth/tet.py:371-497,
499-626).  Here a uniform grid of cells lists the tets overlapping each cell
(built once per mesh on the host), one kernel locates every point in its
cell's tets (emg_locate_points), and the field and its curl come from
emg_interp_points / emg_interp_curl_points.
"""
from __future__ import annotations

import numpy as np

from . import _lib


class PointLocator:
    """Uniform-grid tet lookup for a Mesh3D (about `tets_per_cell` tets per cell)."""

    def __init__(self, mesh, device="cuda", tets_per_cell: float = 2.0):
        import torch
        self.torch = torch
        self.mesh = mesh
        self.dev = torch.device(device)
        nodes, tets = mesh.nodes, mesh.tets_canonical.T
        lo, hi = nodes.min(axis=1), nodes.max(axis=1)
        span = np.maximum(hi - lo, 1e-300)
        ncell = max(1.0, mesh.n_tets / tets_per_cell)
        h = (np.prod(span) / ncell) ** (1.0 / 3.0)
        dims = np.maximum(1, np.ceil(span / h)).astype(np.int64)
        inv = dims / span
        v = nodes[:, tets]                                               # (3, 4, nt)
        c0 = np.clip(np.floor((v.min(axis=1) - lo[:, None]) * inv[:, None]), 0, dims[:, None] - 1).astype(np.int64)
        c1 = np.clip(np.floor((v.max(axis=1) - lo[:, None]) * inv[:, None]), 0, dims[:, None] - 1).astype(np.int64)
        ext = c1 - c0 + 1                                                # (3, nt)
        cnt = ext.prod(axis=0)
        tid = np.repeat(np.arange(mesh.n_tets), cnt)
        k = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)  # index within the tet's cell box
        ex, ey = ext[0][tid], ext[1][tid]
        ix = c0[0][tid] + k % ex
        iy = c0[1][tid] + (k // ex) % ey
        iz = c0[2][tid] + k // (ex * ey)
        cell = (iz * dims[1] + iy) * dims[0] + ix
        order = np.argsort(cell, kind="stable")                          # tet ids stay ascending per cell
        ptr = np.zeros(int(dims.prod()) + 1, dtype=np.int64)
        np.cumsum(np.bincount(cell, minlength=int(dims.prod())), out=ptr[1:])
        if ptr[-1] >= 2 ** 31:
            raise ValueError("point-location grid too large for int32 indices")
        d = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(self.dev)
        self.grid = d(np.concatenate([lo, inv, dims.astype(np.float64)]), np.float64)
        self.cell_ptr = d(ptr, np.int32)
        self.cell_tets = d(tid[order], np.int32)
        self.nodes = d(nodes.T, np.float64)
        self.tets = d(mesh.tets_canonical, np.int32)
        self.tet_ent = d(mesh.tet_ent, np.int32)

    def locate(self, pts: np.ndarray):
        """(tet ids (device int32, -1 outside), barycentric coordinates (device, npts x 4))."""
        torch = self.torch
        P = torch.from_numpy(np.ascontiguousarray(np.asarray(pts, dtype=np.float64).T)).to(self.dev)
        n = P.shape[0]
        tet = torch.empty(n, dtype=torch.int32, device=self.dev)
        bary = torch.empty((n, 4), dtype=torch.float64, device=self.dev)
        _lib.check(_lib.load().emg_locate_points(n, _lib.ptr(P), _lib.ptr(self.grid), _lib.ptr(self.cell_ptr),
                                                 _lib.ptr(self.cell_tets), _lib.ptr(self.nodes), _lib.ptr(self.tets),
                                                 _lib.ptr(tet), _lib.ptr(bary), _lib.stream_ptr()), "emg_locate_points")
        return tet, bary

    def fields(self, x, pts: np.ndarray, curl_scale=None):
        """E (3, npts) and, with curl_scale (nt complex per tet, or a scalar), curl_scale * curl E
        (3, npts) of the solution x (N complex128, numpy or device) at points (3, npts)."""
        torch = self.torch
        L = _lib.load()
        xd = torch.as_tensor(np.ascontiguousarray(x) if isinstance(x, np.ndarray) else x,
                             dtype=torch.complex128, device=self.dev).contiguous()
        tet, bary = self.locate(pts)
        n = tet.shape[0]
        nent = self.mesh.n_entities
        E = torch.empty((n, 3), dtype=torch.complex128, device=self.dev)
        _lib.check(L.emg_interp_points(n, _lib.ptr(tet), _lib.ptr(bary), _lib.ptr(self.nodes), _lib.ptr(self.tets),
                                       _lib.ptr(self.tet_ent), nent, 1, _lib.ptr(xd), _lib.ptr(E), _lib.stream_ptr()),
                   "emg_interp_points")
        H = None
        if curl_scale is not None:
            cs = np.broadcast_to(np.asarray(curl_scale, dtype=np.complex128), (self.mesh.n_tets,))
            csd = torch.from_numpy(np.ascontiguousarray(cs)).to(self.dev)
            Hd = torch.empty((n, 3), dtype=torch.complex128, device=self.dev)
            _lib.check(L.emg_interp_curl_points(n, _lib.ptr(tet), _lib.ptr(bary), _lib.ptr(self.nodes),
                                                _lib.ptr(self.tets), _lib.ptr(self.tet_ent), nent, 1, _lib.ptr(xd),
                                                _lib.ptr(csd), _lib.ptr(Hd), _lib.stream_ptr()),
                       "emg_interp_curl_points")
            H = Hd.cpu().numpy().T
        return E.cpu().numpy().T, H


class LazyLocator:
    """A PointLocator built on first use (a sweep that never interpolates pays nothing)."""

    def __init__(self, mesh, device="cuda"):
        self.mesh = mesh
        self.device = device
        self._loc = None

    def _get(self) -> PointLocator:
        if self._loc is None:
            self._loc = PointLocator(self.mesh, self.device)
        return self._loc

    def locate(self, pts):
        return self._get().locate(pts)

    def fields(self, x, pts, curl_scale=None):
        return self._get().fields(x, pts, curl_scale=curl_scale)
