"""Device-resident frequency-domain problem: K(f), port sources, GPU solve, S-parameters.

The reference assembles K(f) = E - k0^2 B + sum_p B_p(f) per frequency on the
host (fem/physics/edm/assembler.py:312-388), removes PEC DOFs by fancy
indexing (fem/solver.py:430-435) and solves directly.  Here E and B live on the
GPU once per mesh; the Robin terms are affine in frequency,
B_p(f) = gamma_p(f) * P_p and b_p(f) = force_p(f) * q_p, so P_p and q_p are
built once by the boundary kernel and every frequency costs one affine combine
plus a scatter of the port blocks.  PEC DOFs are masked inside the solver.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .assembly import GPUPattern
from .bc import C0, PEC, PortBC, RobinBC
from .mesh import Mesh3D

# fem/mth/optimized.py:28-29 (order-4 Dunavant rule, 6 points) as gaus_quad_tri(4) lays it out
_DUNAVANT4 = [(3, 0.223381589678011, 0.108103018168070, 0.445948490915965, 0.445948490915965),
              (3, 0.109951743655322, 0.816847572980459, 0.091576213509771, 0.091576213509771)]


def dunavant4() -> np.ndarray:
    """(4, 6): weights then barycentric rows (fem/mth/optimized.py:77-108)."""
    pts = []
    for N, W, L1, L2, L3 in _DUNAVANT4:
        l1, l2, l3 = L1, L2, L3
        for n in range(N):
            if n == 3:
                l1, l2, l3 = L1, L3, L2
            pts.append([W, l1, l2, l3])
            l1, l2, l3 = l2, l3, l1
    return np.array(pts).T


def pec_dofs(mesh: Mesh3D, bcs) -> np.ndarray:
    """DOFs on PEC triangles: both modes of their edges and faces (assembler.py:348-359)."""
    ne, nent = mesh.n_edges, mesh.n_entities
    out = []
    for bc in bcs:
        if isinstance(bc, PEC):
            tri = mesh.get_triangles(bc.tags)
            e = np.unique(mesh.tri_to_edge[:, tri].ravel())
            f = ne + tri
            ent = np.concatenate([e, f])
            out.append(np.concatenate([ent, ent + nent]))
    return np.unique(np.concatenate(out)) if out else np.zeros(0, dtype=np.int64)


def tri_dofs(mesh: Mesh3D, tri: np.ndarray) -> np.ndarray:
    """(ntri, 8) DOFs in the reference's tri slot order (fem/elements/nedelec2.py:57-62)."""
    ne, nent = mesh.n_edges, mesh.n_entities
    e = mesh.tri_to_edge[:, tri].T
    f = (ne + tri)[:, None]
    return np.concatenate([e, f, e + nent, f + nent], axis=1)


def port_local_xy(bc, mesh: Mesh3D, tri: np.ndarray) -> np.ndarray:
    """(ntri, 6) local vertex coordinates in the frame the reference uses."""
    v = mesh.nodes[:, mesh.tris[:, tri]]                       # (3, 3 verts, ntri)
    if bc._include_force:
        # assembler.py:112: vertices_local = local_basis @ (nodes - origin)
        B = bc.get_inv_basis()
        loc = np.einsum("ij,jvt->ivt", B, v - bc.cs.origin[:, None, None])
        return np.concatenate([loc[0].T, loc[1].T], axis=1)
    # fem/mth/tri.py:708-721: per-triangle frame (ax1, -axn x ax1, axn) inverted by pinv
    ax1 = v[:, 1] - v[:, 0]
    ax2 = v[:, 2] - v[:, 0]
    ax1 = ax1 / np.linalg.norm(ax1, axis=0)
    ax2 = ax2 / np.linalg.norm(ax2, axis=0)
    axn = np.cross(ax1.T, ax2.T).T
    ax2b = -np.cross(axn.T, ax1.T).T
    n2 = np.sum(ax2b * ax2b, axis=0)
    lx = np.einsum("it,ivt->vt", ax1, v)
    ly = np.einsum("it,ivt->vt", ax2b, v) / n2
    return np.concatenate([lx.T, ly.T], axis=1)


@dataclass
class RobinTerm:
    bc: RobinBC
    tri: np.ndarray
    pos: object            # device int32 (ntri*64,) CSR positions
    vals: object           # device c128 (ntri*64,) block values for gamma = 1
    q: object | None       # device c128 (N,) source for force_scale = 1


class FrequencyProblem:
    """E, B, Robin terms and the PEC mask of one mesh on one GPU."""

    def __init__(self, mesh: Mesh3D, er: np.ndarray, ur: np.ndarray, bcs, exact_face_mass: bool = False,
                 device: str = "cuda", pattern: GPUPattern | None = None, reference_geometry: bool = False):
        import torch
        self.torch = torch
        self.mesh = mesh
        self.bcs = list(bcs)
        self.dev = torch.device(device)
        self.pat = GPUPattern(mesh, device=device) if pattern is None else pattern
        self.E, self.B = self.pat.assemble(er, ur, exact_face_mass=exact_face_mass,
                                           reference_geometry=reference_geometry)
        self.er, self.ur = er, ur
        self.N, self.nnz = self.pat.n, self.pat.nnz
        mask = np.ones(self.N, dtype=np.uint8)
        self.pec = pec_dofs(mesh, self.bcs)
        mask[self.pec] = 0
        self.mask = torch.from_numpy(mask).to(self.dev)
        self.ports = sorted([bc for bc in self.bcs if isinstance(bc, PortBC)], key=lambda b: b.port_number)
        self.dpts = dunavant4()
        self._dpts_d = torch.from_numpy(np.ascontiguousarray(self.dpts)).to(self.dev)
        self.robin = [self._robin_term(bc) for bc in self.bcs if isinstance(bc, RobinBC)]
        self.K = torch.empty(self.nnz, dtype=torch.complex128, device=self.dev)

    def _robin_term(self, bc) -> RobinTerm:
        torch, mesh, L = self.torch, self.mesh, _lib.load()
        tri = mesh.get_triangles(bc.tags)
        nt = tri.shape[0]
        lxy = torch.from_numpy(np.ascontiguousarray(port_local_xy(bc, mesh, tri))).to(self.dev)
        elen = None
        if not bc._include_force:
            elen = torch.from_numpy(np.ascontiguousarray(mesh.edge_lengths[mesh.tri_to_edge[:, tri]].T)).to(self.dev)
        Bblk = torch.empty((nt, 64), dtype=torch.float64, device=self.dev)
        bblk = None
        uinc = None
        if bc._include_force:
            l = port_local_xy(bc, mesh, tri)
            xq = l[:, 0:3] @ self.dpts[1:]                        # (nt, nq) local x of the points
            yq = l[:, 3:6] @ self.dpts[1:]
            U = bc.mode_shape(xq.ravel(), yq.ravel())               # (3, nt*nq)
            u = np.stack([U[0].reshape(nt, -1), U[1].reshape(nt, -1)], axis=-1).astype(np.complex128)
            uinc = torch.from_numpy(np.ascontiguousarray(u)).to(self.dev)
            bblk = torch.empty((nt, 8), dtype=torch.complex128, device=self.dev)
        _lib.check(L.emg_robin_blocks(nt, _lib.ptr(lxy), _lib.ptr(elen) if elen is not None else None,
                                      _lib.ptr(uinc) if uinc is not None else None, self.dpts.shape[1],
                                      _lib.ptr(self._dpts_d), _lib.ptr(Bblk),
                                      _lib.ptr(bblk) if bblk is not None else None, _lib.stream_ptr()),
                   "emg_robin_blocks")
        dofs = tri_dofs(mesh, tri)                                 # (nt, 8)
        rows = torch.from_numpy(np.repeat(dofs, 8, axis=1).ravel().astype(np.int32)).to(self.dev)
        cols = torch.from_numpy(np.tile(dofs, (1, 8)).ravel().astype(np.int32)).to(self.dev)
        pos = torch.empty(nt * 64, dtype=torch.int32, device=self.dev)
        _lib.check(L.emg_csr_find(nt * 64, _lib.ptr(self.pat.rowptr), _lib.ptr(self.pat.cols), _lib.ptr(rows),
                                  _lib.ptr(cols), _lib.ptr(pos), _lib.stream_ptr()), "emg_csr_find")
        vals = Bblk.reshape(-1).to(torch.complex128)
        q = None
        if bblk is not None:
            q = torch.zeros(self.N, dtype=torch.complex128, device=self.dev)
            idx = torch.from_numpy(dofs.ravel().astype(np.int32)).to(self.dev)
            _lib.check(L.emg_scatter_add(nt * 8, _lib.ptr(idx), _lib.ptr(bblk.reshape(-1)), 1.0, 0.0, _lib.ptr(q),
                                         _lib.stream_ptr()), "emg_scatter_add")
        return RobinTerm(bc, tri, pos, vals, q)

    def system(self, freq: float, out=None):
        """K(f) values on the device (complex128, nnz)."""
        L = _lib.load()
        k0 = 2 * np.pi * freq / C0
        K = self.K if out is None else out
        _lib.check(L.emg_affine_combine(self.nnz, _lib.ptr(self.E), _lib.ptr(self.B), k0 * k0, _lib.ptr(K),
                                        _lib.stream_ptr()), "emg_affine_combine")
        for rt in self.robin:
            if not rt.bc._include_stiff:
                continue
            g = complex(rt.bc.get_gamma(k0))
            _lib.check(L.emg_scatter_add(rt.pos.shape[0], _lib.ptr(rt.pos), _lib.ptr(rt.vals), g.real, g.imag,
                                         _lib.ptr(K), _lib.stream_ptr()), "emg_scatter_add")
        return K

    def rhs(self, port: PortBC, freq: float, masked: bool = True):
        """Port source vector b_p(f) on the device (masked to free DOFs unless masked=False,
        which gives the reference's port_vectors entry, assembler.py:378)."""
        k0 = 2 * np.pi * freq / C0
        for rt in self.robin:
            if rt.bc is port:
                q = rt.q * complex(port.force_scale(k0))
                return q * self.mask if masked else q
        raise KeyError(f"port {port.port_number} has no source term")

    def to_scipy(self, vals):
        return self.pat.to_scipy(vals)

    def gradient(self):
        """Device CSR of the P2 gradient on the free DOFs (built once per problem)."""
        if getattr(self, "_gradient", None) is None:
            from types import SimpleNamespace
            from .gradient import gradient_matrix
            torch = self.torch
            G, nodes = gradient_matrix(self.mesh, self.mask.cpu().numpy())
            Gt = G.T.tocsr()
            Gt.sort_indices()
            d = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(self.dev)
            self._gradient = SimpleNamespace(np2=G.shape[1], nodes=nodes, host=G,
                                             rowptr=d(G.indptr, np.int32), cols=d(G.indices, np.int32),
                                             vals=d(G.data, np.float64), t_rowptr=d(Gt.indptr, np.int32),
                                             t_cols=d(Gt.indices, np.int32), t_vals=d(Gt.data, np.float64))
        return self._gradient


def poly_coefficients(H: np.ndarray) -> np.ndarray | None:
    """Coefficients c_i = 1/theta_i of the GMRES polynomial preconditioner from
    the (d+1) x d Hessenberg matrix of d Arnoldi steps on K M: theta are the
    harmonic Ritz values (eigenvalues of H_d + |h_{d+1,d}|^2 H_d^{-H} e_d e_d^T,
    the roots of the GMRES residual polynomial of degree d), in Leja order
    (largest first, then each maximising the product of distances to those
    before) so the product form 1 - z p(z) = prod (1 - z/theta_i) is evaluated
    stably.  None if a root is zero or not finite."""
    H = np.asarray(H, dtype=np.complex128)
    d = H.shape[1]
    if H.shape[0] != d + 1 or d < 1:
        raise ValueError(f"H must be (d+1) x d, got {H.shape}")
    Hd = H[:d, :d]
    ed = np.zeros(d)
    ed[-1] = 1.0
    try:
        f = np.linalg.solve(Hd.conj().T, ed)
        theta = np.linalg.eigvals(Hd + abs(H[d, d - 1]) ** 2 * np.outer(f, ed))
    except np.linalg.LinAlgError:
        return None
    if not np.all(np.isfinite(theta)) or np.min(np.abs(theta)) == 0.0:
        return None
    order = [int(np.argmax(np.abs(theta)))]
    while len(order) < d:
        rest = [i for i in range(d) if i not in order]
        score = [np.sum(np.log(np.abs(theta[i] - theta[order]) + 1e-300)) for i in rest]
        order.append(rest[int(np.argmax(score))])
    return 1.0 / theta[order]


class GPUSolver:
    """Restarted GMRES on the GPU (emg_solver_* in the C ABI).

    preconditioner: "hybrid" (2x2 block Jacobi plus the gradient correction
    G diag(G^T K G)^{-1} G^T, emerge_amd.gradient) or "jacobi" (block Jacobi only).
    poly_degree: d > 1 wraps the smoother M in the GMRES polynomial of degree d,
    M p(K M), recomputed at every set_matrix from d Arnoldi steps on K M
    (poly_coefficients); each GMRES step then applies K and M d times but the
    Krylov space, and with it the orthogonalisation work, grows d times slower.
    """

    def __init__(self, problem: FrequencyProblem, restart: int = 150, precision: int = 64,
                 preconditioner: str = "hybrid", poly_degree: int = 1):
        if preconditioner not in ("hybrid", "jacobi"):
            raise ValueError(f"preconditioner must be 'hybrid' or 'jacobi', not {preconditioner!r}")
        if not 1 <= int(poly_degree) <= 16:
            raise ValueError(f"poly_degree must be in [1, 16], not {poly_degree!r}")
        self.poly_degree = int(poly_degree)
        self.poly_active = 1
        self.L = _lib.load()
        self.p = problem
        h = ctypes.c_void_p()
        _lib.check(self.L.emg_solver_create(problem.N, problem.nnz, _lib.ptr(problem.pat.rowptr),
                                            _lib.ptr(problem.pat.cols), restart, precision, ctypes.byref(h)),
                   "emg_solver_create")
        self.h = h
        self.precision = precision
        self.restart = restart
        self.preconditioner = preconditioner
        if preconditioner == "hybrid":
            G = problem.gradient()
            _lib.check(self.L.emg_solver_set_gradient(self.h, G.np2, _lib.ptr(G.rowptr), _lib.ptr(G.cols),
                                                      _lib.ptr(G.vals), _lib.ptr(G.t_rowptr), _lib.ptr(G.t_cols),
                                                      _lib.ptr(G.t_vals), _lib.stream_ptr()),
                       "emg_solver_set_gradient")

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                self.L.emg_solver_destroy(h)
            except Exception:
                pass
            self.h = None

    learned = None          # LearnedSmoother whose weights scale the smoother at each frequency
    POLY_STEP_CAP = 3000    # steps at the configured degree before it is halved for good
    POLY_CHUNK = 1000       # GMRES steps between progress checks while the polynomial is active
    POLY_MIN_GAIN = 2.0     # a chunk must cut the residual this much, else the degree halves

    def spmv_tuning(self) -> dict:
        """The autotuned SpMV shape of this solver (chosen on its first matrix): lanes per
        entity for complex128 and complex64 steps and the measured ms per launch of each
        candidate (4, 8, 16, 32 lanes); lanes 0 before the first set_matrix."""
        lanes = (ctypes.c_int * 2)()
        ms = (ctypes.c_float * 8)()
        _lib.check(self.L.emg_solver_spmv_tuning(self.h, lanes, ms), "emg_solver_spmv_tuning")
        cand = (4, 8, 16, 32)
        return {"lanes_c128": int(lanes[0]), "lanes_c64": int(lanes[1]),
                "ms_c128": dict(zip(cand, [float(v) for v in ms[0:4]])),
                "ms_c64": dict(zip(cand, [float(v) for v in ms[4:8]]))}

    def smoother_tuning(self) -> dict:
        """The autotuned lanes per row of the smoother's G^T product (chosen with the SpMV on
        the solver's first matrix when the gradient correction is set; 0 otherwise, 8 used) and
        the measured ms per launch of each candidate (4, 8, 16)."""
        lanes = ctypes.c_int()
        ms = (ctypes.c_float * 3)()
        _lib.check(self.L.emg_solver_smoother_tuning(self.h, ctypes.byref(lanes), ms), "emg_solver_smoother_tuning")
        return {"gt_lanes": int(lanes.value), "ms": dict(zip((4, 8, 16), [float(v) for v in ms]))}

    def set_weights(self, we, wp):
        """Learned smoother weights (complex128 device tensors or None) for the next set_matrix."""
        self._w = (we, wp)
        _lib.check(self.L.emg_solver_set_weights(self.h, _lib.ptr(we) if we is not None else None,
                                                 _lib.ptr(wp) if wp is not None else None), "emg_solver_set_weights")

    def set_matrix(self, K, freq: float | None = None):
        """K(f) values; with a learned smoother attached and `freq` given, its weights for f are applied."""
        self._K = K
        if self.learned is not None and freq is not None:
            self.set_weights(*self.learned.weights(freq))
        self._set_matrix_c()

    def _set_matrix_c(self):
        _lib.check(self.L.emg_solver_set_matrix(self.h, _lib.ptr(self._K), _lib.ptr(self.p.mask), _lib.stream_ptr()),
                   "emg_solver_set_matrix")
        if self.poly_degree > 1:
            self._update_poly()

    def _update_poly(self):
        """Polynomial of the preconditioner for the current matrix (plain smoother if it fails)."""
        torch = self.p.torch
        d = self.poly_degree
        if getattr(self, "_v0", None) is None:
            g = torch.Generator(device=self.p.dev)
            g.manual_seed(20251)
            v = torch.randn((self.p.N, 2), generator=g, dtype=torch.float64, device=self.p.dev)
            self._v0 = torch.view_as_complex(v).contiguous() * self.p.mask.to(torch.float64)
        _lib.check(self.L.emg_solver_set_poly(self.h, 1, None), "emg_solver_set_poly")
        self.poly_active = 1
        if (d + 1) * 16 > (self.restart + 1) * (16 if self.precision == 64 else 8):
            return                                          # the probe needs d + 1 complex128 basis slots
        H = np.zeros((d, d + 1), dtype=np.complex128)     # column-major (d+1) x d
        rc = self.L.emg_solver_arnoldi_probe(self.h, d, _lib.ptr(self._v0), H.ctypes.data, _lib.stream_ptr())
        if rc != 0:
            msg = self.L.emg_last_error().decode(errors="replace")
            if "breakdown" not in msg:
                _lib.check(rc, "emg_solver_arnoldi_probe")
            return                                          # Krylov space exhausted: the smoother alone
        c = poly_coefficients(H.T)
        if c is None:
            self.poly_active = 1
            return
        cc = np.ascontiguousarray(c, dtype=np.complex128)
        _lib.check(self.L.emg_solver_set_poly(self.h, d, cc.ctypes.data), "emg_solver_set_poly")
        self.poly_active = d

    def solve(self, b, x0=None, tol: float = 1e-8, maxiter: int = 20000, max_refine: int = 0):
        """GMRES from x0.  maxiter bounds the smoother applications (a step with the degree-d
        polynomial costs d).  While the polynomial is active the solve runs in chunks of
        POLY_CHUNK steps, each continuing from the last iterate.  The degree halves (6 -> 3
        -> 1) for this and every later solve of this solver when the configured degree has
        not converged in POLY_STEP_CAP steps, or when a chunk at any degree does not cut the
        FP64 residual POLY_MIN_GAIN-fold (complex64 steps can stall with a high-degree
        polynomial on hard matrices); info.reserved bit 2 is set and info.iterations then
        counts smoother applications.  With learned smoother weights active, a solve that
        does not converge is continued from its last iterate with the plain smoother (bit 1)."""
        torch = self.p.torch
        x = torch.zeros_like(b) if x0 is None else x0.clone()
        info = _lib.SolveInfo()
        apps = steps = steps_d = 0
        lowered = False
        res_prev = None
        while True:
            d = max(1, self.poly_active)
            budget = (maxiter - apps) // d
            if budget <= 0:
                break
            n = min(budget, self.POLY_CHUNK) if d > 1 else budget
            _lib.check(self.L.emg_solver_solve(self.h, _lib.ptr(b), _lib.ptr(x), tol, n, max_refine,
                                               ctypes.byref(info), _lib.stream_ptr()), "emg_solver_solve")
            apps += info.iterations * d
            steps += info.iterations
            steps_d += info.iterations
            if info.converged or d == 1:
                break
            stalled = res_prev is not None and info.residual * self.POLY_MIN_GAIN > res_prev
            capped = not getattr(self, "_poly_lowered", False) and steps_d >= self.POLY_STEP_CAP
            if stalled or capped:
                # the polynomial does not pay off on this matrix: halve its degree from here on
                self._poly_lowered = True
                steps_d = 0
                self.poly_degree = d // 2
                if self.poly_degree > 1:
                    self._update_poly()
                else:
                    _lib.check(self.L.emg_solver_set_poly(self.h, 1, None), "emg_solver_set_poly")
                    self.poly_active = 1
                lowered = True
                res_prev = None
                continue
            res_prev = info.residual
        info.iterations = apps if lowered else steps
        if lowered:
            info.reserved |= 2
        w = getattr(self, "_w", (None, None))
        if not info.converged and (w[0] is not None or w[1] is not None):
            used = info.iterations
            self.set_weights(None, None)
            self._set_matrix_c()
            _lib.check(self.L.emg_solver_solve(self.h, _lib.ptr(b), _lib.ptr(x), tol, maxiter, max_refine,
                                               ctypes.byref(info), _lib.stream_ptr()), "emg_solver_solve")
            info.iterations += used
            info.reserved |= 1
        return x, info
