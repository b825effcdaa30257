"""CPU ORACLE -- test infrastructure only.

Restatement of the reference's port modal analysis (a ModalPort's mode from
its face): Electrodynamics3D.modal_analysis (fem/physics/edm/emfreq3d.py:201-364),
Assembler.assemble_bma_matrices (fem/physics/edm/assembler.py:246-308), the
port element matrices generalized_matrix_GQ / _matrix_builder
(fem/physics/edm/nedeleclegrange2.py:185-417), the NedelecLegrange2 DOF map
(fem/elements/nedleg2.py:30-126), SolverLAPACK.eig with filter_real_modes
(fem/solver.py:311-341, 34-68), the 2D field and curl interpolation
(fem/mth/tri.py:228-401) and compute_avg_power_flux
(fem/physics/edm/port_functions.py:28-43), plus the ModalPort formulas the
frequency-domain oracle needs (fem/bc.py:329-500).

The symbolic expressions of the reference are restated in terms of the
barycentric coordinates l_i and their gradients, with its quirks: the
face-face entries of Dtt are never filled (nedeleclegrange2.py:350-358 fill
only Att and Btt), the quadrature table has 8 digits, and the reference's Ez
(tri.py:303) is l1 (2 l1 - 1)(e9 + e10 + e11) + 2 l1 l2 (e12 + e13 + e14) and its
"curl" (tri.py:380-396) is (dEz/dy + j b Ey, dEz/dx - j b Ex, curl_z Et), both
reproduced as written.  tests/test_golden_ref.py holds them to the reference's
own outputs.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

from . import oracle as O

C0 = 299792458.0
MU0 = 4 * np.pi * 1e-7
EPS0 = 8.854187818814e-12

# fem/physics/edm/nedeleclegrange2.py:201-204: the order-4 Dunavant rule to 8 digits (as written there)
DPTS_GQ = np.array([[0.22338159, 0.22338159, 0.22338159, 0.10995174, 0.10995174, 0.10995174],
                    [0.10810302, 0.44594849, 0.44594849, 0.81684757, 0.09157621, 0.09157621],
                    [0.44594849, 0.44594849, 0.10810302, 0.09157621, 0.09157621, 0.81684757],
                    [0.44594849, 0.10810302, 0.44594849, 0.09157621, 0.81684757, 0.09157621]])

EDGES = ((0, 1), (1, 2), (0, 2))        # local edges of a sorted triangle (fem/mesh3d.py:534-541)


def tri_coefficients(xs, ys):
    """fem/physics/edm/nedeleclegrange2.py:180-199: a, b, c (signed) and |area|."""
    x1, x2, x3 = xs
    y1, y2, y3 = ys
    a = np.array([x2 * y3 - y2 * x3, x3 * y1 - y3 * x1, x1 * y2 - y1 * x2])
    b = np.array([y2 - y3, y3 - y1, y1 - y2])
    c = np.array([x3 - x2, x1 - x3, x2 - x1])
    sA = 0.5 * ((x1 - x3) * (y2 - y1) - (x1 - x2) * (y3 - y1))
    s = np.sign(sA)
    return a * s, b * s, c * s, abs(sA)


class _Basis2D:
    """l_i and grad l_i of one triangle at points (xs, ys)."""

    def __init__(self, a, b, c, A, xs, ys):
        k = 1.0 / (2 * A)
        self.l = np.array([(a[i] + b[i] * xs + c[i] * ys) * k for i in range(3)])
        self.g = np.array([[b[i] * k, c[i] * k] for i in range(3)])              # (3, 2)

    def w(self, p, q, r):
        """l_p (l_q grad l_r - l_r grad l_q): (2, npts)."""
        l, g = self.l, self.g
        return l[p] * (l[q] * g[r][:, None] - l[r] * g[q][:, None])

    def wcurl(self, p, q, r):
        """z-curl of w(p, q, r)."""
        l, g = self.l, self.g
        cr = lambda u, v: g[u][0] * g[v][1] - g[u][1] * g[v][0]
        return l[q] * cr(p, r) - l[r] * cr(p, q) + 2 * l[p] * cr(q, r)

    def tangential(self, k):
        """The 8 tangential functions in DOF order [e1(01), e1(12), e1(02), f1, e2(01), e2(12), e2(02), f2]."""
        if k < 3:
            i, j = EDGES[k]
            return self.w(i, j, i), self.wcurl(i, j, i)              # l_i (l_j grad l_i - l_i grad l_j)
        if k == 3:
            return -self.w(1, 2, 0), -self.wcurl(1, 2, 0)            # -l2 (l3 grad l1 - l1 grad l3)
        if k < 7:
            i, j = EDGES[k - 4]
            return self.w(j, j, i), self.wcurl(j, j, i)              # l_j (l_j grad l_i - l_i grad l_j)
        return self.w(2, 1, 0), self.wcurl(2, 1, 0)                  # l3 (l2 grad l1 - l1 grad l2)

    def nodal(self, k):
        """The 6 Lagrange-2 functions and gradients: vertices 0..2, then edges (01), (12), (02)."""
        l, g = self.l, self.g
        if k < 3:
            return l[k] * (2 * l[k] - 1), g[k][:, None] * (4 * l[k] - 1)
        i, j = EDGES[k - 3]
        return 4 * l[i] * l[j], 4 * (g[i][:, None] * l[j] + l[i] * g[j][:, None])


def generalized_matrix(verts, Ms, Mm, k0, face_block=False):
    """The 14 x 14 A, B of one triangle (nedeleclegrange2.py:206-374), verts (2, 3) local.

    face_block=True fills the face-face entries of Dtt that the reference leaves zero
    (a departure from the reference, used to pin the rest against the closed form)."""
    xs, ys = verts
    a, b, c, A = tri_coefficients(xs, ys)
    W = DPTS_GQ[0]
    px = xs[0] * DPTS_GQ[1] + xs[1] * DPTS_GQ[2] + xs[2] * DPTS_GQ[3]
    py = ys[0] * DPTS_GQ[1] + ys[1] * DPTS_GQ[2] + ys[2] * DPTS_GQ[3]
    bs = _Basis2D(a, b, c, A, px, py)
    Ds = lambda i, j: np.hypot(xs[i] - xs[j], ys[i] - ys[j])
    Msz, Mmz, Ms2, Mm2 = Ms[2, 2], Mm[2, 2], Ms[:2, :2], Mm[:2, :2]
    T = [bs.tangential(k) for k in range(8)]
    Nd = [bs.nodal(k) for k in range(6)]
    q1 = lambda u, v: np.sum(u * v * W)                                     # _gqi
    q2 = lambda u, v: np.sum(W * np.sum(u * v, axis=0))                     # _gqi2
    Att = np.array([[q1(T[i][1], Msz * T[j][1]) for j in range(8)] for i in range(8)])
    Btt = np.array([[q2(T[i][0], Mm2 @ T[j][0]) for j in range(8)] for i in range(8)])
    Dtt = np.array([[q2(T[i][0], Ms2 @ T[j][0]) for j in range(8)] for i in range(8)])
    Dzt = np.array([[q2(Nd[i][1], Ms2 @ T[j][0]) for j in range(8)] for i in range(6)])
    Dzz1 = np.array([[q2(Nd[i][1], Ms2 @ Nd[j][1]) for j in range(6)] for i in range(6)])
    Dzz2 = np.array([[q1(Nd[i][0], Mmz * Nd[j][0]) for j in range(6)] for i in range(6)])
    # the reference fills the face rows / columns of Att, Btt, Dtt by symmetry (nedeleclegrange2.py:300-348)
    for M in (Att, Btt, Dtt):
        for ei in (0, 1, 2, 4, 5, 6):
            for f in (3, 7):
                M[f, ei] = M[ei, f]
    # ... and never fills the face-face entries of Dtt (only those of Att and Btt, :350-358): zero
    if not face_block:
        Dtt[3, 3] = Dtt[3, 7] = Dtt[7, 3] = Dtt[7, 7] = 0.0
    Ls = np.ones((14, 14))
    Ls[3, :] *= Ds(0, 2)
    Ls[7, :] *= Ds(0, 1)
    Ls[:, 3] *= Ds(0, 2)
    Ls[:, 7] *= Ds(0, 1)
    for ei, (i, j) in enumerate(EDGES):
        Le = Ds(i, j)
        Ls[ei, :] *= Le
        Ls[:, ei] *= Le
        Ls[ei + 4, :] *= Le
        Ls[:, ei + 4] *= Le
    Am = np.zeros((14, 14), dtype=np.complex128)
    Bm = np.zeros((14, 14), dtype=np.complex128)
    Am[:8, :8] = Att - k0 ** 2 * Btt
    Bm[:8, :8] = Dtt
    Bm[8:, :8] = Dzt
    Bm[:8, 8:] = Dzt.T
    Bm[8:, 8:] = Dzz1 - k0 ** 2 * Dzz2
    return Ls * Am * abs(A), Ls * Bm * abs(A)


def interp_full(pts, field, tris, nodes, t2f):
    """fem/mth/tri.py:228-306: Ex, Ey, Ez at points (2, n) from the 14-DOF field (last tri wins)."""
    n = pts.shape[1]
    out = np.zeros((3, n), dtype=np.complex128)
    for t in range(tris.shape[1]):
        v = nodes[:2, tris[:, t]]
        T = np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]], axis=1)
        loc = np.linalg.pinv(T) @ (pts - v[:, :1])
        inside = (loc.sum(0) <= 1.00000001) & (loc[0] >= -1e-6) & (loc[1] >= -1e-6)
        if not inside.any():
            continue
        out[:, inside] = _eval_tri(v, field[t2f[:, t]], pts[:, inside])[0]
    return out


def interp_curl(pts, field, tris, nodes, t2f, diadic, beta):
    """fem/mth/tri.py:308-401: the reference's (dEz/dy + j b Ey, dEz/dx - j b Ex, curl_z Et), times diag(diadic)."""
    n = pts.shape[1]
    out = np.zeros((3, n), dtype=np.complex128)
    for t in range(tris.shape[1]):
        v = nodes[:2, tris[:, t]]
        T = np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]], axis=1)
        loc = np.linalg.pinv(T) @ (pts - v[:, :1])
        inside = (loc.sum(0) <= 1.00000001) & (loc[0] >= -1e-6) & (loc[1] >= -1e-6)
        if not inside.any():
            continue
        H = _eval_tri(v, field[t2f[:, t]], pts[:, inside], beta)[1]
        out[:, inside] = H * np.diag(diadic[:, :, t])[:, None]
    return out


def _eval_tri(v, e, pts, beta=0.0):
    a, b, c, A = tri_coefficients(v[0], v[1])
    bs = _Basis2D(a, b, c, A, pts[0], pts[1])
    D = lambda i, j: np.hypot(v[0, i] - v[0, j], v[1, i] - v[1, j])
    mult = np.array([D(0, 1), D(1, 2), D(0, 2), D(0, 2), D(0, 1), D(1, 2), D(0, 2), D(0, 1), 1, 1, 1, 1, 1, 1])
    e = e * mult
    Et = np.zeros((2, pts.shape[1]), dtype=np.complex128)
    cz = np.zeros(pts.shape[1], dtype=np.complex128)
    for k in range(8):
        w, wc = bs.tangential(k)
        Et += e[k] * w
        cz += e[k] * wc
    l, g = bs.l, bs.g
    S9, S12 = e[8] + e[9] + e[10], e[11] + e[12] + e[13]
    Ez = l[0] * (2 * l[0] - 1) * S9 + 2 * l[0] * l[1] * S12                       # tri.py:303 as written
    dEz = (g[0][:, None] * ((2 * l[0] - 1) * S9 + 2 * l[1] * S12)
           + l[0] * (2 * g[0][:, None] * S9 + 2 * g[1][:, None] * S12))
    jb = 1j * beta
    H = np.array([dEz[1] + jb * Et[1], dEz[0] - jb * Et[0], cz])
    return np.vstack([Et, Ez[None]]), H


def filter_real_modes(vals, vecs, k0, ermax, urmax):
    """fem/solver.py:34-68."""
    upper = -(k0 ** 2) * ermax * urmax * 2
    mask = (vals <= -1) & (vals >= upper)
    fv, fw = vals[mask], vecs[:, mask]
    order = np.argsort(np.abs(np.sqrt(-fv) - k0))
    return fv[order], fw[:, order]


class OraclePortMode:
    """The port face of a 3D oracle problem as a NedelecLegrange2 field and its modes."""

    def __init__(self, om, tri_ids, frame_basis, frame_origin, origin_shift):
        tris3 = om.tris[:, tri_ids]
        un = np.unique(tris3)
        self.node_ids = un
        self.tris = np.searchsorted(un, tris3)
        self.nodes = om.nodes[:, un]
        edges = set()
        for t in range(self.tris.shape[1]):
            i1, i2, i3 = self.tris[:, t]
            edges.update([(i1, i2), (i2, i3), (i1, i3)])
        self.edges = np.array(sorted(edges)).T
        eid = {(int(a), int(b)): k for k, (a, b) in enumerate(self.edges.T)}
        self.tri_to_edge = np.array([[eid[(int(t[0]), int(t[1]))], eid[(int(t[1]), int(t[2]))],
                                      eid[(int(t[0]), int(t[2]))]] for t in self.tris.T]).T
        Nn, Ne, Nt = self.nodes.shape[1], self.edges.shape[1], self.tris.shape[1]
        self.n_xy = 2 * Ne + 2 * Nt
        self.n_field = self.n_xy + Nn + Ne
        t2f = np.zeros((14, Nt), dtype=np.int64)                              # fem/elements/nedleg2.py:54-61
        t2f[:3] = self.tri_to_edge
        t2f[3] = np.arange(Nt) + Ne
        t2f[4:7] = self.tri_to_edge + Ne + Nt
        t2f[7] = np.arange(Nt) + 2 * Ne + Nt
        t2f[8:11] = self.tris + 2 * Ne + 2 * Nt
        t2f[11:14] = self.tri_to_edge + 2 * Ne + 2 * Nt + Nn
        self.tri_to_field = t2f
        e2f = np.zeros((5, Ne), dtype=np.int64)                               # nedleg2.py:63-68
        e2f[0] = np.arange(Ne)
        e2f[1] = np.arange(Ne) + Nt + Ne
        e2f[2] = np.arange(Ne) + 2 * Ne + 2 * Nt + Nn
        e2f[3:] = self.edges + 2 * Ne + 2 * Nt
        self.edge_to_field = e2f
        self.basis = np.asarray(frame_basis)
        self.origin = np.asarray(frame_origin)
        self.local_nodes = np.linalg.pinv(self.basis) @ (self.nodes - self.origin[:, None])
        # surface normals (fem/mesh3d.py:580-584) away from origin - zhat (assembler.py:258)
        o = self.origin - origin_shift
        p1, p2, p3 = (self.nodes[:, self.tris[k]] for k in range(3))
        cr = np.cross((p2 - p1).T, (p3 - p1).T).T
        n = cr / np.linalg.norm(cr, axis=0)
        sgn = np.where(np.sum(n * ((p1 + p2 + p3) / 3 - o[:, None]), axis=0) < 0, -1.0, 1.0)
        self.normals = n * sgn

    def matrices(self, er_tri, ur_tri, k0, face_block=False):
        """assemble_bma_matrices' E, B (assembler.py:258-266; nedeleclegrange2.py:31-53, 376-417)."""
        nodes = np.linalg.pinv(self.basis) @ self.nodes                     # nedeleclegrange2.py:42 (no origin)
        rows, cols, dA, dB = [], [], [], []
        for t in range(self.tris.shape[1]):
            A, B = generalized_matrix(nodes[:2, self.tris[:, t]], O.matinv(ur_tri[:, :, t]), er_tri[:, :, t], k0,
                                      face_block)
            idx = self.tri_to_field[:, t]
            rows.append(np.repeat(idx, 14))
            cols.append(np.tile(idx, 14))
            dA.append(A.ravel())
            dB.append(B.ravel())
        n = self.n_field
        rows, cols = np.concatenate(rows), np.concatenate(cols)
        E = sp.coo_matrix((np.concatenate(dA), (rows, cols)), shape=(n, n)).tocsr()
        Bm = sp.coo_matrix((np.concatenate(dB), (rows, cols)), shape=(n, n)).tocsr()
        return E, Bm

    def pec_ids(self, om, pec_tri_ids):
        """DOFs of port-surface edges that are edges of PEC triangles (assembler.py:268-296)."""
        nmap = {int(v): i for i, v in enumerate(self.node_ids)}
        eid = {(int(a), int(b)): k for k, (a, b) in enumerate(self.edges.T)}
        out = []
        for e in np.unique(om.tri_to_edge[:, pec_tri_ids].ravel()):
            i1, i2 = nmap.get(int(om.edges[0, e])), nmap.get(int(om.edges[1, e]))
            if i1 is None or i2 is None:
                continue
            k = eid.get((min(i1, i2), max(i1, i2)))
            if k is None:
                continue
            out.extend(self.edge_to_field[:, k].tolist())
        return set(out)

    def interpolate_E(self, mode, pts_global):
        """NedelecLegrange2.interpolate_Ef (nedleg2.py:82-90): global points -> global E."""
        loc = np.linalg.pinv(self.basis) @ (pts_global - self.origin[:, None])
        El = interp_full(loc[:2], mode, self.tris, self.local_nodes, self.tri_to_field)
        return self.basis @ El

    def interpolate_H(self, mode, pts_global, k0, ur_tri, beta):
        """NedelecLegrange2.interpolate_Hf (nedleg2.py:92-105)."""
        const = 1j / ((k0 * C0) * MU0)
        urinv = np.stack([O.matinv(ur_tri[:, :, t]) for t in range(ur_tri.shape[2])], axis=2)
        loc = np.linalg.pinv(self.basis) @ (pts_global - self.origin[:, None])
        Hl = interp_curl(loc[:2], mode, self.tris, self.local_nodes, self.tri_to_field, urinv, beta)
        return (self.basis @ Hl) * const


def surface_points(nodes, tris, dpts):
    """fem/mth/optimized.py:140-171: quadrature points (3, nq * ntri), point-major per triangle."""
    P = np.einsum("cvt,vq->cqt", nodes[:, tris], dpts[1:])
    return P.reshape(3, -1), P.shape[1]


def avg_power_flux(pm: OraclePortMode, mode, k0, ur_tri, beta, dpts):
    """compute_avg_power_flux (port_functions.py:28-43) with surface_integral order 4
    (fem/mth/integrals.py:51-70): the normal of the first triangle for the whole face."""
    pts, nq = surface_points(pm.nodes, pm.tris, dpts)
    E = pm.interpolate_E(mode, pts)
    H = pm.interpolate_H(mode, pts, k0, ur_tri, beta)
    nx, ny, nz = pm.normals[:, 0]
    Sx = 0.5 * np.real(E[1] * np.conj(H[2]) - E[2] * np.conj(H[1]))
    Sy = 0.5 * np.real(E[2] * np.conj(H[0]) - E[0] * np.conj(H[2]))
    Sz = 0.5 * np.real(E[0] * np.conj(H[1]) - E[1] * np.conj(H[0]))
    S = (nx * Sx + ny * Sy + nz * Sz).reshape(nq, -1)
    p1, p2, p3 = (pm.nodes[:, pm.tris[k]] for k in range(3))
    A = np.linalg.norm(np.cross((p2 - p1).T, (p3 - p1).T), axis=1) / 2
    return float(np.sum(A * (dpts[0] @ S)))


def modal_analysis(orc, port, freq, nmodes=1, TEM=False, face_block=False):
    """Electrodynamics3D.modal_analysis (emfreq3d.py:201-364) on the oracle problem `orc` for the
    oracle port `port` (kind "modal"): fills port.modes (list of dicts) and returns them."""
    om = orc.om
    k0 = 2 * np.pi * freq / C0
    tri_ids = orc.tri_ids(port.tags)
    # er, ur per 3D-mesh triangle from its first tet (emfreq3d.py:237-241)
    er_all = orc.er[:, :, om.tri_to_tet[0]]
    ur_all = orc.ur[:, :, om.tri_to_tet[0]]
    ermean = np.mean(er_all[er_all > 0].flatten())
    urmean = np.mean(ur_all[ur_all > 0].flatten())
    ermax, urmax = np.max(er_all.flatten()), np.max(ur_all.flatten())
    kmax = k0 * np.sqrt(ermax * urmax)
    fr = port.frame
    pm = OraclePortMode(om, tri_ids, fr.basis, fr.origin, fr.basis[:, 2])
    A, B = pm.matrices(er_all[:, :, tri_ids], ur_all[:, :, tri_ids], k0, face_block)
    pec = set()
    for bc in orc.bcs:
        if bc.kind == "pec":
            pec |= pm.pec_ids(om, orc.tri_ids(bc.tags))
    solve_ids = np.array([i for i in range(pm.n_field) if i not in pec])
    target = ermean * urmean * (1.1 if TEM else 0.65) * k0
    Ar = A[solve_ids][:, solve_ids].toarray()
    Br = B[solve_ids][:, solve_ids].toarray()
    lam, vecs = sla.eig(Ar, Br, right=True, left=False)                       # SolverLAPACK (solver.py:337-340)
    lam, vecs = filter_real_modes(lam, vecs, target, 2, 2)
    dpts = _dunavant4()
    port.port_field = pm
    port.modes = []
    for i in range(lam.shape[0]):
        Em = np.zeros(pm.n_field, dtype=np.complex128)
        Em[solve_ids] = vecs[:, i]
        Em = Em * np.exp(-1j * np.angle(np.max(Em)))
        beta = min(np.emath.sqrt(-lam[i]).real, kmax.real)
        P = avg_power_flux(pm, Em, k0, ur_all, beta, dpts)
        energy = np.mean(np.abs(Em) ** 2)
        if energy < 1e-4:                                                     # fem/bc.py:433-435
            continue
        Ez = np.max(np.abs(Em[pm.n_xy:]))
        Exy = np.max(np.abs(Em[:pm.n_xy]))
        modetype = "TEM" if TEM else ("TE" if Ez / Exy < 1e-5 else "TM")
        port.modes.append(dict(field=Em, beta=beta, k0=k0, freq=freq, norm=np.sqrt(1 / np.abs(P)), power=P,
                               modetype=modetype, TEM=TEM, eig=lam[i], ez_ratio=Ez / Exy))
    return port.modes


def _dunavant4():
    from .fd import gaus_quad_tri4
    return gaus_quad_tri4()
