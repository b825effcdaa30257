"""S-parameter extraction (fem/physics/edm/emfreq3d.py:734-779, fem/mth/sparam.py:72-139,
fem/mth/integrals.py:24-70).

Every port quadrature point (6-point Dunavant per port triangle) or voltage
integration point is tagged once with the tet that holds it and its
barycentric coordinates; the field of all solutions is then evaluated on the
GPU in one launch (emg_interp_points).  The reference instead searches all
tets touching a port for every point on every call (fem/mth/tet.py:393-428).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .bc import LumpedPort, PortBC


def locate_points(mesh, pts: np.ndarray, candidates: np.ndarray):
    """(tet id, barycentric coords in canonical vertex order) of each point.

    Like the reference's inside test (fem/mth/tet.py:426: lambda >= -1e-6), the
    last candidate tet that holds a point wins.
    """
    tc = mesh.tets_canonical[candidates]                        # (nc, 4)
    v = mesh.nodes[:, tc]                                       # (3, nc, 4)
    T = np.stack([v[:, :, 1] - v[:, :, 0], v[:, :, 2] - v[:, :, 0], v[:, :, 3] - v[:, :, 0]], axis=-1)  # (3,nc,3)
    Tinv = np.linalg.inv(np.transpose(T, (1, 0, 2)))            # (nc, 3, 3)
    tet = np.full(pts.shape[1], -1, dtype=np.int64)
    bary = np.zeros((pts.shape[1], 4))
    for k in range(pts.shape[1]):
        loc = np.einsum("cij,cj->ci", Tinv, pts[:, k][None, :] - v[:, :, 0].T)     # (nc, 3)
        lam = np.concatenate([1 - loc.sum(1, keepdims=True), loc], axis=1)
        inside = np.flatnonzero(np.all(lam >= -1e-6, axis=1) & (loc.sum(1) <= 1.00000001))
        if inside.size == 0:
            raise ValueError(f"point {pts[:, k]} is not inside any candidate tet")
        c = inside[-1]
        tet[k] = candidates[c]
        bary[k] = lam[c]
    return tet, bary


def port_tets(mesh, ports) -> np.ndarray:
    """Tets touching any vertex of any port face (emfreq3d.py:639-653)."""
    verts = np.unique(np.concatenate([mesh.tris[:, mesh.get_triangles(p.tags)].ravel() for p in ports]))
    return np.flatnonzero(np.isin(mesh.tets, verts).any(axis=0))


def lumped_port_line(mesh, port: LumpedPort):
    """Voltage integration line of a lumped port (emfreq3d.py:366-389)."""
    from .bc import Line
    tri = mesh.get_triangles(port.tags)
    points = np.unique(mesh.tris[:, tri].ravel())
    xs, ys, zs = mesh.nodes[:, points]
    ax = port.direction
    dotprod = xs * ax[0] + ys * ax[1] + zs * ax[2]
    start_id = points[np.argwhere(dotprod == np.min(dotprod))]
    start = np.squeeze(np.mean(mesh.nodes[:, start_id], axis=1))
    end = start + ax * port.height
    port.vintline = Line.from_points(start, end, 11)
    port.voltage_integration_points = (start, end)
    port.v_integration = True


class PortSampler:
    """Field sampling points of every port, with the per-port S-parameter formulas."""

    def __init__(self, problem):
        import torch
        self.torch = torch
        self.p = problem
        mesh = problem.mesh
        self.ports = problem.ports
        self.dpts = problem.dpts
        cand = None
        segs = []
        tets, barys = [], []
        off = 0
        self.meta = {}
        for port in self.ports:
            if port.v_integration:
                if port.vintline is None:
                    lumped_port_line(mesh, port)
                if cand is None:
                    cand = port_tets(mesh, self.ports)
                pts = np.array(port.vintline.cmid)
                t, b = locate_points(mesh, pts, cand)
                self.meta[port.port_number] = ("line", off, pts.shape[1], None)
            else:
                tri = mesh.get_triangles(port.tags)
                tet = mesh.tri_to_tet[0, tri]
                tv = mesh.tris[:, tri]                              # sorted vertex ids (3, nt)
                tc = mesh.tets_canonical[tet]                      # (nt, 4)
                nq = self.dpts.shape[1]
                b = np.zeros((tri.shape[0], nq, 4))
                for k in range(3):
                    pos = np.argmax(tc == tv[k][:, None], axis=1)  # canonical slot of tri vertex k
                    b[np.arange(tri.shape[0]), :, pos] = self.dpts[1 + k][None, :]
                t = np.repeat(tet, nq)
                b = b.reshape(-1, 4)
                xyz = np.einsum("ctv,vq->ctq", mesh.nodes[:, tv].transpose(0, 2, 1), self.dpts[1:]).reshape(3, -1)
                er, ur = problem.er[:, :, tet], problem.ur[:, :, tet]
                const = 1 / np.sqrt((ur[0, 0] + ur[1, 1] + ur[2, 2]) / (er[0, 0] + er[1, 1] + er[2, 2]))
                self.meta[port.port_number] = ("face", off, t.shape[0], (tri, xyz, const))
            tets.append(t)
            barys.append(b)
            off += t.shape[0]
        self.npts = off
        dev = problem.dev
        self.ptet = torch.from_numpy(np.concatenate(tets).astype(np.int32)).to(dev)
        self.bary = torch.from_numpy(np.ascontiguousarray(np.concatenate(barys))).to(dev)

    def fields(self, X) -> np.ndarray:
        """E at all sample points for solutions X (nsol, N) -> (nsol, npts, 3) numpy."""
        torch, p, L = self.torch, self.p, _lib.load()
        X = X.contiguous()
        nsol = X.shape[0]
        E = torch.empty((nsol, self.npts, 3), dtype=torch.complex128, device=p.dev)
        _lib.check(L.emg_interp_points(self.npts, _lib.ptr(self.ptet), _lib.ptr(self.bary), _lib.ptr(p.pat.nodes),
                                       _lib.ptr(p.pat.tets), _lib.ptr(p.pat.tet_ent), p.mesh.n_entities, nsol,
                                       _lib.ptr(X), _lib.ptr(E), _lib.stream_ptr()), "emg_interp_points")
        return E.cpu().numpy()

    def s_data(self, port: PortBC, E: np.ndarray, k0: float, active: bool):
        """(field power, mode power) of `port` for one solution's sample fields E (npts, 3)."""
        kind, off, n, extra = self.meta[port.port_number]
        Ep = E[off:off + n]
        if kind == "line":                                          # emfreq3d.py:753-768
            ln = port.vintline
            V = np.sum(Ep[:, 0] * ln.dxs + Ep[:, 1] * ln.dys + Ep[:, 2] * ln.dzs)
            a, b = (port.voltage, V - port.voltage) if active else (0, V)
            return np.sqrt(b ** 2 / (2 * port.Z0)), np.sqrt(a ** 2 / (2 * port.Z0))
        tri, xyz, const = extra                                      # emfreq3d.py:770-779, sparam.py
        mesh = self.p.mesh
        nq = self.dpts.shape[1]
        Em = port.port_mode_3d_global(xyz[0], xyz[1], xyz[2], k0)    # (3, nt*nq)
        Ef = Ep.T
        Q = 1 if active else 0
        Zm = port.Zmode(k0)
        fp = (np.sum((Ef - Q * Em) * np.conj(Em), axis=0) / (2 * Zm)).reshape(-1, nq)
        mp = (np.sum(Em * np.conj(Em), axis=0) / (2 * Zm)).reshape(-1, nq)
        A = mesh.areas[tri]
        w = self.dpts[0]
        return np.sum(const * (fp @ w) * A), np.sum(const * (mp @ w) * A)

    def s_matrix(self, E_by_active: dict[int, np.ndarray], k0: float) -> np.ndarray:
        """S[j, i] for solutions excited at each port i (emfreq3d.py:688-727)."""
        nums = [p.port_number for p in self.ports]
        S = np.zeros((len(nums), len(nums)), dtype=np.complex128)
        for ap in self.ports:
            E = E_by_active[ap.port_number]
            _, pmode = self.s_data(ap, E, k0, True)
            for bc in self.ports:
                pf, _ = self.s_data(bc, E, k0, bc is ap)
                S[nums.index(bc.port_number), nums.index(ap.port_number)] = pf / pmode
        return S
