"""Port-face modal analysis on the GPU: the mode of a ModalPort.

Mirrors Electrodynamics3D.modal_analysis (fem/physics/edm/emfreq3d.py:201-364)
with Assembler.assemble_bma_matrices (fem/physics/edm/assembler.py:246-308):

* ``PortFace``: the port's triangles (``Mesh3D.boundary_surface``) carrying the
  14-DOF NedelecLegrange2 field of fem/elements/nedleg2.py:30-126 (tangential
  second-order Nedelec plus quadratic Lagrange normal component), its DOF maps
  and the PEC DOFs of assembler.py:268-296;
* the pencil A e = lambda B e (lambda = -beta^2) assembled densely on the GPU
  by ``emg_port_assemble`` (the reference's generalized_matrix_GQ, quirks
  included), PEC DOFs removed by the reduced numbering;
* the reference factorises the whole pencil with LAPACK (SolverLAPACK.eig,
  fem/solver.py:311-341) and keeps every eigenvalue in the window
  -8 t^2 <= lambda <= -1 of filter_real_modes (fem/solver.py:34-68, called with
  ermax = urmax = 2 and t the target propagation constant).  Here the window is
  searched by shift-invert Arnoldi around its centre (rocSOLVER LU of
  A - sigma B, the operator (A - sigma B)^{-1} B formed once by
  ``emg_port_operator``, one matrix-vector product per step of
  ``emg_port_krylov``), followed by an FP64 Rayleigh-Ritz of the
  pencil on the Krylov basis; the basis grows until every Ritz value in the
  window has a relative residual below 1e-6, and each kept pair is then
  refined by two steps of inverse iteration at its own Ritz value (a second
  LU) to a residual below ``tol`` (1e-10), as accurate as a dense LAPACK
  eigenvector, which the reference's TE / TM test on tiny Ez needs;
* each mode: phase rule of emfreq3d.py:330, beta = min(Re sqrt(-lambda), kmax),
  the average power flux of port_functions.py:28-43 through the GPU field
  kernel ``emg_port_field``, the energy threshold of fem/bc.py:433-435 and the
  TE / TM test of emfreq3d.py:342-352.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

from . import _lib
from .bc import C0, ModalPort, PEC, PortMode

MU0 = 4 * np.pi * 1e-7

# Dunavant order-4 rule (weight, l1, l2, l3) of surface_integral / compute_avg_power_flux
# (fem/mth/integrals.py:51-70): the full-precision table
_DUNAVANT4 = np.array([
    [0.223381589678011, 0.223381589678011, 0.223381589678011, 0.109951743655322, 0.109951743655322, 0.109951743655322],
    [0.108103018168070, 0.445948490915965, 0.445948490915965, 0.816847572980459, 0.091576213509771, 0.091576213509771],
    [0.445948490915965, 0.445948490915965, 0.108103018168070, 0.091576213509771, 0.091576213509771, 0.816847572980459],
    [0.445948490915965, 0.108103018168070, 0.445948490915965, 0.091576213509771, 0.816847572980459, 0.091576213509771]])


def _dev(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()


class PortFace:
    """A port face as a NedelecLegrange2 field (fem/elements/nedleg2.py:30-126).

    DOF order: [e1 of every edge | f1 of every triangle | e2 | f2 | Ez at nodes | Ez at
    edge midpoints]; ``tri_to_field`` (14, nt) and ``edge_to_field`` (5, ne: e1, e2, Ez at
    the midpoint, Ez at both vertices) follow nedleg2.py:54-68."""

    def __init__(self, mesh, tri_ids, cs):
        self.mesh = mesh
        self.cs = cs
        self.tri_ids = np.asarray(tri_ids, dtype=np.int64)
        # normals away from origin - zhat (assembler.py:258-260)
        from .farfield import SurfaceMesh
        self.surface = SurfaceMesh(mesh, self.tri_ids, cs.origin - cs.zax)
        s = self.surface
        nt, ne, nn = s.n_tris, s.n_edges, s.n_nodes
        self.n_tris, self.n_edges, self.n_nodes = nt, ne, nn
        self.n_xy = 2 * ne + 2 * nt
        self.n_field = self.n_xy + nn + ne
        t2f = np.empty((14, nt), dtype=np.int64)
        t2f[0:3] = s.tri_to_edge
        t2f[3] = np.arange(nt) + ne
        t2f[4:7] = s.tri_to_edge + ne + nt
        t2f[7] = np.arange(nt) + 2 * ne + nt
        t2f[8:11] = s.tris + 2 * ne + 2 * nt
        t2f[11:14] = s.tri_to_edge + 2 * ne + 2 * nt + nn
        self.tri_to_field = t2f
        e2f = np.empty((5, ne), dtype=np.int64)
        e2f[0] = np.arange(ne)
        e2f[1] = np.arange(ne) + ne + nt
        e2f[2] = np.arange(ne) + 2 * ne + 2 * nt + nn
        e2f[3:5] = s.edges + 2 * ne + 2 * nt
        self.edge_to_field = e2f
        # face edge / triangle number of each 3D-mesh edge / triangle on the face (-1 elsewhere)
        self.edge_of_src = np.full(mesh.n_edges, -1, dtype=np.int64)
        self.edge_of_src[mesh.tri_to_edge[:, self.tri_ids].ravel()] = s.tri_to_edge.ravel()
        self.tri_of_src = np.full(mesh.n_tris, -1, dtype=np.int64)
        self.tri_of_src[self.tri_ids] = np.arange(nt)
        B = cs._basis_inv
        self.frame_nodes = B @ s.nodes                                       # nedeleclegrange2.py:42 (no origin)
        self.local_nodes = B @ (s.nodes - cs.origin[:, None])                # nedleg2.py:82-105
        self._dev = None

    # ------------------------------------------------------------------ DOFs
    def pec_dofs(self, bcs) -> np.ndarray:
        """assembler.py:268-296: the 5 DOFs of every face edge that is an edge of a PEC
        triangle, and the 14 of every face triangle that is itself PEC."""
        out = []
        for bc in bcs:
            if not isinstance(bc, PEC):
                continue
            tris = self.mesh.get_triangles(bc.tags)
            fe = self.edge_of_src[self.mesh.tri_to_edge[:, tris].ravel()]
            out.append(self.edge_to_field[:, fe[fe >= 0]].ravel())
            ft = self.tri_of_src[tris]
            out.append(self.tri_to_field[:, ft[ft >= 0]].ravel())
        return np.unique(np.concatenate(out)) if out else np.zeros(0, dtype=np.int64)

    def device(self):
        """Face geometry and maps on the GPU (uploaded once)."""
        import torch
        if self._dev is None:
            s = self.surface
            self._dev = dict(
                tris=_dev(s.tris.T, torch.int64),
                t2f=_dev(self.tri_to_field.T, torch.int64),
                nodes=_dev(self.local_nodes[:2].T, torch.float64),
                basis=_dev(self.cs._basis, torch.float64))
        return self._dev

    # ---------------------------------------------------------------- fields
    def fields_local(self, field_dev, xl, yl, beta=0.0, diadic_dev=None, hconst=0j):
        """E (and H when diadic_dev is given) in global components at face-frame points (xl, yl):
        two (3, n) complex arrays on the GPU (H None without diadic_dev)."""
        import torch
        L = _lib.load()
        d = self.device()
        pts = _dev(np.stack([np.ravel(xl), np.ravel(yl)], axis=1), torch.float64)
        npts = pts.shape[0]
        E = torch.empty((npts, 3), dtype=torch.complex128, device=pts.device)
        H = torch.empty((npts, 3), dtype=torch.complex128, device=pts.device) if diadic_dev is not None else None
        if field_dev.numel() != self.n_field or field_dev.dtype != torch.complex128:
            raise ValueError(f"a face field is {self.n_field} complex128 values, not {tuple(field_dev.shape)}")
        _lib.check(L.emg_port_field(npts, _lib.ptr(pts), self.n_tris, _lib.ptr(d["tris"]), self.n_nodes,
                                    _lib.ptr(d["nodes"]), _lib.ptr(d["t2f"]), self.n_field, _lib.ptr(field_dev),
                                    float(beta),
                                    _lib.ptr(diadic_dev) if diadic_dev is not None else None, _lib.ptr(d["basis"]),
                                    float(np.real(hconst)), float(np.imag(hconst)), _lib.ptr(E),
                                    _lib.ptr(H) if H is not None else None, _lib.stream_ptr()), "emg_port_field")
        return E.T, (H.T if H is not None else None)

    def field_at(self, field, xg, yg, zg) -> np.ndarray:
        """The face field E at global points, global components (3, n) (nedleg2.py:82-90)."""
        import torch
        fd = field if isinstance(field, torch.Tensor) else _dev(field, torch.complex128)
        xl, yl, _ = self.cs.in_local_cs(np.ravel(xg), np.ravel(yg), np.ravel(zg))
        E, _ = self.fields_local(fd, xl, yl)
        return E.cpu().numpy()

    def quadrature_points(self):
        """Order-4 quadrature points (3, 6 nt, point-major per triangle) and triangle areas
        (fem/mth/optimized.py:140-171)."""
        s = self.surface
        P = np.einsum("cvt,vq->cqt", s.nodes[:, s.tris], _DUNAVANT4[1:])
        return P.reshape(3, -1), s.areas


class ModalSolve:
    """Dense pencil of one port face at one frequency on the GPU and its windowed eigenpairs."""

    def __init__(self, face: PortFace, er_tri, ur_tri, k0: float, solve_ids: np.ndarray, face_block: bool = False):
        import torch
        self.torch = torch
        self.face = face
        self.k0 = k0
        self.solve_ids = solve_ids
        n = solve_ids.shape[0]
        self.n = n
        red = np.full(face.n_field, -1, dtype=np.int64)
        red[solve_ids] = np.arange(n)
        s = face.surface
        fn = face.frame_nodes
        verts = np.concatenate([fn[0, s.tris].T, fn[1, s.tris].T], axis=1)          # (nt, 6)
        Ms = np.linalg.inv(np.moveaxis(ur_tri, 2, 0))                                 # ur^-1 per triangle
        Mm = np.moveaxis(er_tri, 2, 0)
        self.A = torch.zeros((n, n), dtype=torch.complex128, device="cuda")
        self.B = torch.zeros((n, n), dtype=torch.complex128, device="cuda")
        d_verts = _dev(verts, torch.float64)
        d_Ms, d_Mm = _dev(Ms, torch.complex128), _dev(Mm, torch.complex128)
        d_t2r = _dev(red[face.tri_to_field].T, torch.int64)
        L = _lib.load()
        self.L = L
        _lib.check(L.emg_port_assemble(face.n_tris, _lib.ptr(d_verts), _lib.ptr(d_Ms), _lib.ptr(d_Mm), float(k0),
                                       1 if face_block else 0, _lib.ptr(d_t2r), n, _lib.ptr(self.A), _lib.ptr(self.B),
                                       None, None, _lib.stream_ptr()), "emg_port_assemble")

    def _gemm(self, ta, tb, X, Y, out, alpha=1.0, beta=0.0):
        """out = alpha op(X) op(Y) + beta out on column-major views: a torch (r, c) row-major tensor
        is the column-major (c, r) matrix, so X, Y, out are passed as their transposes."""
        # all tensors here are stored column-major as torch tensors of shape (cols, rows)
        m = out.shape[1]
        n = out.shape[0]
        k = X.shape[0] if ta == 0 else X.shape[1]
        _lib.check(self.L.emg_zgemm(ta, tb, m, n, k, float(np.real(alpha)), float(np.imag(alpha)), _lib.ptr(X),
                                    X.shape[1], _lib.ptr(Y), Y.shape[1], float(np.real(beta)), float(np.imag(beta)),
                                    _lib.ptr(out), out.shape[1], _lib.stream_ptr()), "emg_zgemm")

    M_MAX = 768        # largest Krylov basis before the search gives up on unconverged window values

    def eigenpairs(self, lo: float, hi: float, tol: float = 1e-6, m0: int = 96, seed: int = 0):
        """Every eigenpair with lo <= Re lambda <= hi, by shift-invert Arnoldi about the window
        centre and Rayleigh-Ritz: the basis doubles (up to M_MAX) until every Ritz value in the
        window has a relative residual <= tol (the pairs are then refined by ``refine``).
        Returns (lambda (k,), X (k, n) device rows, residuals (k,)); ``self.unconverged``
        counts window Ritz values left above tol."""
        torch = self.torch
        n = self.n
        sigma = 0.5 * (lo + hi)
        M = torch.empty_like(self.A)
        ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
        _lib.check(self.L.emg_port_factor(n, _lib.ptr(self.A), _lib.ptr(self.B), sigma, 0.0, _lib.ptr(M),
                                          _lib.ptr(ipiv), _lib.stream_ptr()), "emg_port_factor")
        T = torch.empty_like(self.A)                                              # (A - sigma B)^{-1} B
        _lib.check(self.L.emg_port_operator(n, _lib.ptr(M), _lib.ptr(ipiv), _lib.ptr(self.B), _lib.ptr(T),
                                            _lib.stream_ptr()), "emg_port_operator")
        del M
        rng = np.random.default_rng(seed)
        m = min(m0, n)
        while True:
            # columns of the Krylov basis are the rows of V (column-major n x (m + 1))
            V = torch.zeros((m + 1, n), dtype=torch.complex128, device="cuda")
            V[0] = torch.from_numpy(rng.standard_normal(n) + 1j * rng.standard_normal(n)).cuda()
            H = torch.zeros((m, m + 1), dtype=torch.complex128, device="cuda")
            _lib.check(self.L.emg_port_krylov(n, _lib.ptr(T), m, _lib.ptr(V), _lib.ptr(H), _lib.stream_ptr()),
                       "emg_port_krylov")
            sub = torch.diagonal(H[:, 1:]).abs().cpu().numpy()                    # H[j + 1, j]
            brk = np.flatnonzero(sub <= 1e-14 * max(sub.max(), 1e-300))
            k = int(brk[0]) + 1 if brk.size else m
            Vk = V[:k]
            lam, X, res = self._ritz(Vk)
            inwin = (lam.real >= lo) & (lam.real <= hi)
            if np.all(res[inwin] <= tol) or m >= min(n, self.M_MAX) or brk.size:
                self.unconverged = int(np.count_nonzero(inwin & (res > tol)))
                self.krylov_size = k
                idx = np.flatnonzero(inwin & (res <= tol))
                return lam[idx], X[torch.from_numpy(idx).cuda()], res[idx]
            m = min(2 * m, n, self.M_MAX)

    def refine(self, lam: complex, x, steps: int = 2):
        """Inverse iteration at the Ritz value itself: x <- (A - lam B)^{-1} B x, normalised,
        `steps` times, then lam <- x^H A x / x^H B x.  Brings the Ritz pair (residual ~ the
        Rayleigh-Ritz tolerance) to working accuracy, so that small components (Ez of a TE
        mode) are resolved as a dense LAPACK eigenvector resolves them.  Returns (lam, x)."""
        torch = self.torch
        n = self.n
        M = torch.empty_like(self.A)
        ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
        _lib.check(self.L.emg_port_factor(n, _lib.ptr(self.A), _lib.ptr(self.B), float(np.real(lam)),
                                          float(np.imag(lam)), _lib.ptr(M), _lib.ptr(ipiv), _lib.stream_ptr()),
                   "emg_port_factor")
        x = x.contiguous().clone()
        for _ in range(steps):
            y = torch.empty_like(x)
            _lib.check(self.L.emg_port_inverse_step(n, _lib.ptr(M), _lib.ptr(ipiv), _lib.ptr(self.B), 1, _lib.ptr(x),
                                                    _lib.ptr(y), _lib.stream_ptr()), "emg_port_inverse_step")
            x = y / torch.linalg.vector_norm(y)
        X = x.reshape(1, -1)
        AX, BX = torch.empty_like(X), torch.empty_like(X)
        self._gemm(0, 0, self.A, X, AX)
        self._gemm(0, 0, self.B, X, BX)
        xbx = complex(torch.vdot(X[0], BX[0]).item())
        if abs(xbx) > 1e-12 * float(torch.linalg.vector_norm(BX).item()):
            lam = complex(torch.vdot(X[0], AX[0]).item()) / xbx
        return lam, x

    def residual(self, lam: complex, x) -> float:
        """||A x - lam B x|| / (||A x|| + |lam| ||B x||) for one vector x (device, n)."""
        torch = self.torch
        X = x.reshape(1, -1).contiguous()
        AX, BX = torch.empty_like(X), torch.empty_like(X)
        self._gemm(0, 0, self.A, X, AX)
        self._gemm(0, 0, self.B, X, BX)
        r = torch.linalg.vector_norm(AX - lam * BX)
        d = torch.linalg.vector_norm(AX) + abs(lam) * torch.linalg.vector_norm(BX)
        return float((r / d).item())

    def _ritz(self, Vk):
        """Rayleigh-Ritz of the pencil on the rows of Vk: Ritz values, Ritz vectors (rows, unit
        2-norm) and relative residuals ||A x - lambda B x|| / (||A x|| + |lambda| ||B x||)."""
        torch = self.torch
        k = Vk.shape[0]
        AV = torch.empty_like(Vk)
        BV = torch.empty_like(Vk)
        # column-major: AV (n x k) = A (n x n) V (n x k); torch stores A as (n, n) = A^T column-major,
        # the pencil is assembled column-major, i.e. self.A[c, r] = A[r, c]
        self._gemm(0, 0, self.A, Vk, AV)
        self._gemm(0, 0, self.B, Vk, BV)
        Ap = torch.empty((k, k), dtype=torch.complex128, device="cuda")
        Bp = torch.empty((k, k), dtype=torch.complex128, device="cuda")
        self._gemm(2, 0, Vk, AV, Ap)                                               # V^H A V
        self._gemm(2, 0, Vk, BV, Bp)
        lam, Y = sla.eig(Ap.T.cpu().numpy(), Bp.T.cpu().numpy())
        ok = np.isfinite(lam)
        lam, Y = lam[ok], Y[:, ok]
        Yd = torch.from_numpy(np.ascontiguousarray(Y.T)).cuda()                   # column-major k x r
        X = torch.empty((Y.shape[1], self.n), dtype=torch.complex128, device="cuda")
        self._gemm(0, 0, Vk, Yd, X)                                                # V Y
        X = X / torch.linalg.vector_norm(X, dim=1, keepdim=True)
        AX = torch.empty_like(X)
        BX = torch.empty_like(X)
        self._gemm(0, 0, self.A, X, AX)
        self._gemm(0, 0, self.B, X, BX)
        lt = torch.from_numpy(lam).cuda()[:, None]
        R = torch.linalg.vector_norm(AX - lt * BX, dim=1)
        S = torch.linalg.vector_norm(AX, dim=1) + lt.abs()[:, 0] * torch.linalg.vector_norm(BX, dim=1)
        res = (R / S).cpu().numpy()
        return lam, X, res


def _filter_order(lam, target):
    """filter_real_modes (fem/solver.py:34-68) with ermax = urmax = 2: the window mask with
    numpy's complex ordering, then ascending |sqrt(-lambda) - target|."""
    upper = -(target ** 2) * 2 * 2 * 2
    mask = (lam <= -1) & (lam >= upper)
    idx = np.flatnonzero(mask)
    order = np.argsort(np.abs(np.sqrt(-lam[idx]) - target))
    return idx[order]


def modal_analysis(phys, port: ModalPort, nmodes: int = 6, TEM: bool = False, freq: float | None = None,
                   target_kz=None, face_block: bool = False, tol: float = 1e-10):
    """Electrodynamics3D.modal_analysis (emfreq3d.py:201-364) on the GPU; fills port.modes
    and port.face and returns port.modes.  `nmodes` is accepted as the reference's direct
    path does and, like it, not used to truncate: every mode in the window is kept.
    face_block=True departs from the reference by filling Dtt's face-face entries."""
    import torch
    if not phys._bc_initialized:
        from .physics import SimulationError
        raise SimulationError("Cannot run a modal analysis because no boundary conditions have been assigned.")
    mesh = phys.mesh
    er_tet, ur_tet = phys.material_tensors()
    first = mesh.tri_to_tet[0]
    er = er_tet[:, :, first]                                                    # (3, 3, n_tris), emfreq3d.py:237-241
    ur = ur_tet[:, :, first]
    ermean = np.mean(er[er > 0].flatten())
    urmean = np.mean(ur[ur > 0].flatten())
    ermax, urmax = np.max(er.flatten()), np.max(ur.flatten())
    if freq is None:
        freq = phys.frequencies[0]
    k0 = 2 * np.pi * freq / C0
    kmax = k0 * np.sqrt(ermax * urmax)
    if target_kz is None:
        target_kz = ermean * urmean * (1.1 if TEM else 0.65) * k0
    face = PortFace(mesh, mesh.get_triangles(port.tags), port.cs)
    pec = face.pec_dofs(phys.boundary_conditions)
    solve_ids = np.setdiff1d(np.arange(face.n_field), pec)
    ms = ModalSolve(face, er[:, :, face.tri_ids], ur[:, :, face.tri_ids], k0, solve_ids, face_block)
    upper = float(np.real(-(target_kz ** 2) * 8))
    lam, X, res = ms.eigenpairs(upper, -1.0)
    order = _filter_order(lam, target_kz)
    # H of the power flux: diag(ur^-1) of the 3D triangle with the face triangle's number (nedleg2.py:92-105)
    urinv = np.linalg.inv(np.moveaxis(ur[:, :, :face.n_tris], 2, 0))
    diadic = _dev(urinv, torch.complex128)
    pts, areas = face.quadrature_points()
    xl, yl, _ = port.cs.in_local_cs(pts[0], pts[1], pts[2])
    nrm = face.surface.normals[:, 0]
    hconst = 1j / ((k0 * C0) * MU0)
    port.face = face
    port.modes = []
    port.modal_solve = dict(n=ms.n, krylov=ms.krylov_size, unconverged=ms.unconverged, refined_out=0)
    for i in order:
        Em = torch.zeros(face.n_field, dtype=torch.complex128, device="cuda")
        lam[i], x = ms.refine(lam[i], X[i])
        res[i] = ms.residual(lam[i], x)
        if res[i] > tol:                          # refinement did not reach working accuracy: not a mode
            port.modal_solve["refined_out"] += 1
            continue
        # a deterministic phase first (largest entry real positive), then the reference's rule
        x = x * torch.exp(-1j * torch.angle(x[torch.argmax(x.abs())]))
        Em[torch.from_numpy(solve_ids).cuda()] = x
        Eh = Em.cpu().numpy()
        Eh = Eh * np.exp(-1j * np.angle(np.max(Eh)))                          # emfreq3d.py:330
        Em = torch.from_numpy(Eh).cuda()
        beta = min(np.emath.sqrt(-lam[i]).real, kmax.real)
        E, H = face.fields_local(Em, xl, yl, beta, diadic, hconst)
        E, H = E.cpu().numpy(), H.cpu().numpy()
        S = 0.5 * np.real(np.cross(E.T, np.conj(H.T))) @ nrm                      # port_functions.py:28-43
        P = float(np.sum(areas * (_DUNAVANT4[0] @ S.reshape(6, -1))))
        if float(np.mean(np.abs(Eh) ** 2)) < 1e-4:                               # fem/bc.py:433-435
            continue
        Ez = np.max(np.abs(Eh[face.n_xy:]))
        Exy = np.max(np.abs(Eh[:face.n_xy]))
        modetype = "TEM" if TEM else ("TE" if Ez / Exy < 1e-5 else "TM")
        port.modes.append(PortMode(Eh, beta, k0, freq, P, modetype, TEM, eigenvalue=lam[i]))
        port.modes[-1].residual = float(res[i])
        port.modes[-1].ez_ratio = float(Ez / Exy)
    port._er, port._ur = er, ur
    return port.modes
