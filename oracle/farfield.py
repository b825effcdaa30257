"""CPU ORACLE -- test infrastructure only.

Restatement of the reference's far-field computation: the surface sampling of
SurfaceMesh (fem/mesh3d.py:458-592: edge centres, areas, outward normals) and
stratton_chu / stratton_chu_ff (fem/physics/edm/sc.py:27-180), term for term,
in complex128 (the reference evaluates it in complex64).
"""
from __future__ import annotations

import numpy as np

LR = 0.001


def surface(nodes: np.ndarray, tris: np.ndarray, origin):
    """Edge centres, per-edge weighted normals (sum over the edge's triangles of n A / 3)."""
    un = np.unique(tris)
    loc = {int(v): i for i, v in enumerate(un)}
    t = np.vectorize(loc.get)(tris)
    pts = nodes[:, un]
    edges = set()
    for i in range(t.shape[1]):                                  # fem/mesh3d.py:534-541
        i1, i2, i3 = t[:, i]
        for a, b in ((i1, i2), (i2, i3), (i1, i3)):
            edges.add((min(a, b), max(a, b)))
    edges = sorted(edges)
    eid = {e: k for k, e in enumerate(edges)}
    E = np.array(edges).T
    centers = (pts[:, E[0]] + pts[:, E[1]]) / 2
    wns = np.zeros((3, E.shape[1]))
    o = np.asarray(origin, dtype=np.float64)
    for i in range(t.shape[1]):
        n1, n2, n3 = pts[:, t[0, i]], pts[:, t[1, i]], pts[:, t[2, i]]
        c = np.cross(n2 - n1, n3 - n1)
        area = np.sqrt(np.sum(c ** 2)) / 2                        # fem/mesh3d.py:30-34
        n = c / np.sqrt(np.sum(c ** 2))
        if np.dot(n, (n1 + n2 + n3) / 3.0 - o) < 0:               # fem/mth/optimized.py:228-236
            n = -n
        i1, i2, i3 = t[:, i]
        for a, b in ((i1, i2), (i2, i3), (i1, i3)):               # sc.py:166-173
            wns[:, eid[(min(a, b), max(a, b))]] += n * area / 3
    return centers, wns


def stratton_chu(Ein, Hin, centers, wns, theta, phi, k0):
    """sc.py:27-141 in complex128."""
    Ein, Hin = np.asarray(Ein), np.asarray(Hin)
    Emag = np.sqrt(np.abs(Ein[0]) ** 2 + np.abs(Ein[1]) ** 2 + np.abs(Ein[2]) ** 2)
    keep = Emag > np.max(Emag) * LR
    Ex, Ey, Ez = Ein[:, keep]
    Hx, Hy, Hz = Hin[:, keep]
    vx, vy, vz = centers[:, keep]
    nx, ny, nz = wns[:, keep]
    rx = np.cos(theta) * np.cos(phi)
    ry = np.cos(theta) * np.sin(phi)
    rz = np.sin(theta)
    Z0 = 376.73031366857
    Q = -1j * k0 / (4 * np.pi)
    NxHx, NxHy, NxHz = ny * Hz - nz * Hy, nz * Hx - nx * Hz, nx * Hy - ny * Hx
    NxEx, NxEy, NxEz = ny * Ez - nz * Ey, nz * Ex - nx * Ez, nx * Ey - ny * Ex
    Eo = np.zeros((3, np.size(theta)), dtype=np.complex128)
    for j in range(vx.shape[0]):
        G = np.exp(1j * k0 * (rx * vx[j] + ry * vy[j] + rz * vz[j]))
        RxNxHx = ry * NxHz[j] - rz * NxHy[j]
        RxNxHy = rz * NxHx[j] - rx * NxHz[j]
        RxNxHz = rx * NxHy[j] - ry * NxHx[j]
        ie1x = (NxEx[j] - Z0 * RxNxHx) * G
        ie1y = (NxEy[j] - Z0 * RxNxHy) * G
        ie1z = (NxEz[j] - Z0 * RxNxHz) * G
        Eo[0] += Q * (ry * ie1z - rz * ie1y)
        Eo[1] += Q * (rz * ie1x - rx * ie1z)
        Eo[2] += Q * (rx * ie1y - ry * ie1x)
    Ho = np.stack([(ry * Eo[2] - rz * Eo[1]) / Z0, (rz * Eo[0] - rx * Eo[2]) / Z0, (rx * Eo[1] - ry * Eo[0]) / Z0])
    return Eo, Ho
