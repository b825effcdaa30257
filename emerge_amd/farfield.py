"""Far field from surface fields (fem/physics/edm/sc.py:27-180) on the GPU.

Usage as in the reference (fem/simmodel.py:184-190):

    topsurf = mesh.boundary_surface(tags, origin)
    Ein, Hin = data.item(0).interpolate(*topsurf.exyz).EH
    E, H = stratton_chu(Ein, Hin, topsurf, theta, phi, k0)

The surface currents are sampled at the surface's edge centres with the
area-weighted outward normals (each triangle gives a third of its area times
its normal to each of its edges, sc.py:163-173); points whose |E| is below
1e-3 of the largest are dropped (sc.py:59-75); the radiation sum runs in one
HIP kernel (emg_farfield), one workgroup per direction, in FP64.
"""
from __future__ import annotations

import numpy as np

from . import _lib

LR = 0.001      # sc.py:26, relative |E| below which surface points are skipped


class SurfaceMesh:
    """Triangles of some face tags as a surface (fem/mesh3d.py:458-592): nodes, tris (local
    vertex ids, ascending), edges (i<j), tri_to_edge [(i1,i2),(i2,i3),(i1,i3)], edge_centers,
    areas and unit normals pointing away from `origin`."""

    def __init__(self, mesh, tri_ids: np.ndarray, origin):
        tris = mesh.tris[:, tri_ids]
        un = np.unique(tris)
        self.nodes = mesh.nodes[:, un]
        self.tris = np.searchsorted(un, tris)
        t = self.tris
        e = np.concatenate([np.stack([t[0], t[1]]), np.stack([t[1], t[2]]), np.stack([t[0], t[2]])], axis=1)
        e = np.sort(e, axis=0)
        keys = e[0] * un.shape[0] + e[1]
        uk, inv = np.unique(keys, return_inverse=True)
        self.edges = np.stack([uk // un.shape[0], uk % un.shape[0]])
        self.tri_to_edge = inv.reshape(3, -1)
        self.n_nodes, self.n_tris, self.n_edges = self.nodes.shape[1], t.shape[1], self.edges.shape[1]
        self.edge_centers = 0.5 * (self.nodes[:, self.edges[0]] + self.nodes[:, self.edges[1]])
        p1, p2, p3 = (self.nodes[:, t[k]] for k in range(3))
        cr = np.cross((p2 - p1).T, (p3 - p1).T).T
        nrm = np.linalg.norm(cr, axis=0)
        self.areas = nrm / 2                                                  # fem/mesh3d.py:30-34
        n = cr / nrm
        cen = (p1 + p2 + p3) / 3.0
        sgn = np.where(np.sum(n * (cen - np.asarray(origin, dtype=np.float64)[:, None]), axis=0) < 0, -1.0, 1.0)
        self.normals = n * sgn                                                # fem/mth/optimized.py:228-236

    @property
    def exyz(self):
        return self.edge_centers[0], self.edge_centers[1], self.edge_centers[2]

    def weighted_normals(self) -> np.ndarray:
        """(3, n_edges): sum over the triangles of each edge of normal * area / 3 (sc.py:163-173)."""
        w = np.zeros((3, self.n_edges))
        for k in range(3):
            for c in range(3):
                np.add.at(w[c], self.tri_to_edge[k], self.normals[c] * self.areas / 3)
        return w


def stratton_chu(Ein, Hin, mesh: SurfaceMesh, theta, phi, k0: float, device: str = "cuda"):
    """Far-field E and H (3, ndir) complex128 in the directions (cos t cos p, cos t sin p, sin t)
    from E and H (3, n_edges) at the surface's edge centres (sc.py:143-180)."""
    import torch
    Ein = np.asarray(Ein, dtype=np.complex128)
    Hin = np.asarray(Hin, dtype=np.complex128)
    theta = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    phi = np.broadcast_to(np.asarray(phi, dtype=np.float64), theta.shape)
    emag = np.sqrt(np.sum(np.abs(Ein) ** 2, axis=0))
    keep = emag > np.max(emag) * LR
    w = mesh.weighted_normals()
    dev = torch.device(device)
    d = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(dev)
    pts, wn = d(mesh.edge_centers[:, keep].T, np.float64), d(w[:, keep].T, np.float64)
    E, H = d(Ein[:, keep].T, np.complex128), d(Hin[:, keep].T, np.complex128)
    th, ph = d(theta.ravel(), np.float64), d(np.ascontiguousarray(phi).ravel(), np.float64)
    nd = th.shape[0]
    Eo = torch.empty((nd, 3), dtype=torch.complex128, device=dev)
    Ho = torch.empty_like(Eo)
    _lib.check(_lib.load().emg_farfield(int(keep.sum()), _lib.ptr(pts), _lib.ptr(wn), _lib.ptr(E), _lib.ptr(H), nd,
                                        _lib.ptr(th), _lib.ptr(ph), float(k0), _lib.ptr(Eo), _lib.ptr(Ho),
                                        _lib.stream_ptr()), "emg_farfield")
    return Eo.cpu().numpy().T, Ho.cpu().numpy().T
