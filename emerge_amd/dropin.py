"""Entry points in the reference's own call shapes, in the caller's DOF numbering.

* ``tet_mass_stiffness_matrices(field, er, ur)``
  (fem/physics/edm/optimized_assembly.py:43-64): E and B as scipy CSR in the
  caller's numbering.
* ``Assembler.assemble_freq_matrix(field, er, ur, bcs, frequency,
  cache_matrices)`` (fem/physics/edm/assembler.py:312-388): (K, b, solve_ids,
  port_vectors) in the caller's numbering.
* ``GPUSolveRoutine.solve(A, b, solve_ids, reuse)`` and ``.eig(A, B,
  solve_ids, nmodes, direct, target_k0, which)`` (fem/solver.py:405-505): the
  routine object a caller assigns to ``physics.solveroutine``.

``field`` is the reference's Nedelec-2 field or anything with ``.mesh``
holding ``nodes`` (3, nv), ``tets`` (4, nt), ``edges`` (2, ne), ``tris``
(3, nf) and, for boundary conditions, ``get_triangles(tags)`` (the reference's
gmsh-backed selection, fem/mesh3d.py:191-207).  Assembly, renumbering and
solves run on the GPU; only the returned scipy matrices and numpy vectors
live on the host.
"""
from __future__ import annotations

import zlib

import numpy as np
import scipy.sparse as sp

from .assembly import GPUPattern
from .bc import PortBC
from .numbering import CallerNumbering, _caller_mesh

_CONTEXTS: dict[int, tuple] = {}


def _context(field, device="cuda"):
    """(CallerNumbering, GPUPattern) of a caller field, built once per field object."""
    key = id(field)
    ctx = _CONTEXTS.get(key)
    if ctx is None or ctx[0] is not field:
        num = CallerNumbering(field, device=device)
        pat = GPUPattern(num.mesh, device=device)
        ctx = (field, num, pat)
        _CONTEXTS[key] = ctx
    return ctx[1], ctx[2]


def tet_mass_stiffness_matrices(field, er: np.ndarray, ur: np.ndarray) -> tuple[sp.csr_matrix, sp.csr_matrix]:
    """E (curl-curl, 1/mu_r) and B (mass, eps_r), N x N complex128 CSR in the caller's numbering
    (fem/physics/edm/optimized_assembly.py:43-64); assembled and renumbered on the GPU."""
    num, pat = _context(field)
    E, B = pat.assemble(np.asarray(er, dtype=np.complex128), np.asarray(ur, dtype=np.complex128))
    return num.to_caller_csr(pat, E), num.to_caller_csr(pat, B)


def _tag_triangles(field, num: CallerNumbering, tags) -> None:
    """Make num.mesh.face_tags hold the caller's triangles of every tag."""
    cm = _caller_mesh(field)
    mesh = num.mesh
    for t in tags:
        t = int(t)
        if t in mesh.face_tags:
            continue
        if hasattr(cm, "get_triangles"):
            ids = np.atleast_1d(np.asarray(cm.get_triangles([t]), dtype=np.int64))
            tri = np.asarray(cm.tris)[:, ids]
        elif hasattr(cm, "face_tags"):
            tri = np.asarray(cm.face_tags[t]).reshape(3, -1)
        else:
            raise ValueError(f"the caller's mesh cannot select the triangles of face tag {t}")
        ours = mesh.get_tri(tri[0], tri[1], tri[2]) if tri.size else np.zeros(0, dtype=np.int64)
        if np.any(ours < 0):
            raise ValueError(f"face tag {t} names triangles that are not in the mesh")
        mesh.face_tags[t] = np.unique(ours)


class Assembler:
    """fem/physics/edm/assembler.py:204-388 on the GPU."""

    def __init__(self):
        self.cached_matrices = None       # the device-resident problem of the cached (field, materials, bcs)
        self._key = None

    def assemble_freq_matrix(self, field, er: np.ndarray, ur: np.ndarray, bcs, frequency: float,
                             cache_matrices: bool = False):
        """K(f) = E - k0^2 B + sum of Robin blocks, b, solve_ids and the port source vectors,
        all in the caller's numbering (assembler.py:312-388).  With cache_matrices the E and B
        of the first call are reused, as the reference does (assembler.py:324-331)."""
        from .fdsolve import FrequencyProblem
        num, pat = _context(field)
        bcs = list(bcs)
        _tag_triangles(field, num, sorted({t for bc in bcs for t in bc.tags}))
        key = (id(field), tuple(id(b) for b in bcs))
        if not (cache_matrices and self.cached_matrices is not None and self._key == key):
            self.cached_matrices = FrequencyProblem(num.mesh, np.asarray(er, dtype=np.complex128),
                                                    np.asarray(ur, dtype=np.complex128), bcs, pattern=pat)
            self._key = key
        prob = self.cached_matrices
        K = num.to_caller_csr(pat, prob.system(frequency))
        b = np.zeros(num.n, dtype=np.complex128)
        port_vectors = {}
        for port in (bc for bc in bcs if isinstance(bc, PortBC)):
            port_vectors[port.port_number] = num.to_caller(prob.rhs(port, frequency, masked=False)).cpu().numpy()
        free = np.flatnonzero(prob.mask.cpu().numpy())
        solve_ids = np.sort(num.pinv[free])
        return K, b, solve_ids, port_vectors


class _MatrixProblem:
    """A caller's matrix (and mask) on the library's pattern, in the shape GPUSolver and
    ShiftInvertArnoldi read (N, nnz, pat, mask, dev, E/B, gradient())."""

    def __init__(self, num: CallerNumbering, pat: GPUPattern, mask):
        import torch
        self.torch = torch
        self.mesh = num.mesh
        self.pat = pat
        self.dev = pat.device
        self.N, self.nnz = pat.n, pat.nnz
        self.mask = mask
        self._gradient = None

    def gradient(self):
        if self._gradient is None:
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


def _structure_key(A: sp.csr_matrix):
    return (A.shape, A.nnz, zlib.crc32(np.ascontiguousarray(A.indptr).view(np.uint8)),
            zlib.crc32(np.ascontiguousarray(A.indices).view(np.uint8)))


class GPUSolveRoutine:
    """A ``SolveRoutine`` (fem/solver.py:361-505) whose solves run on the GPU.

    Assign it with ``physics.solveroutine = GPUSolveRoutine(field)``.  The
    field gives the DOF numbering and the mesh for the smoother's gradient
    correction.  Solver settings: precision 64 (complex128 GMRES) or 32
    (complex64 GMRES in FP64 refinement), tol (FP64 relative residual over the
    free DOFs), preconditioner "hybrid" or "jacobi", poly_degree (GMRES
    polynomial around the smoother, 1 = off).  Like the reference, a solver
    failure is not raised: ``last_info`` holds iterations and the FP64
    residual of the last solve (the reference logs its code, solver.py:467-468).
    """

    def __init__(self, field, precision: int = 32, tol: float = 1e-8, preconditioner: str = "hybrid",
                 poly_degree: int = 6, restart: int = 150, maxiter: int = 60000, device: str = "cuda"):
        self.field = field
        self.precision, self.tol = precision, tol
        self.preconditioner, self.poly_degree = preconditioner, poly_degree
        self.restart, self.maxiter = restart, maxiter
        self.device = device
        self.use_sorter = self.use_preconditioner = self.use_direct = True   # solver.py:383-385 flags
        self.last_info = None
        self._pos = {}
        self._mask_key = None
        self._problem = None
        self._solver = None
        self._K = None

    def __str__(self) -> str:
        return f"GPUSolveRoutine(precision={self.precision}, preconditioner={self.preconditioner})"

    def _setup(self, solve_ids):
        import torch
        num, pat = _context(self.field, self.device)
        sid = np.asarray(solve_ids, dtype=np.int64)
        mkey = (sid.shape[0], zlib.crc32(np.ascontiguousarray(sid).view(np.uint8)))
        if self._mask_key != mkey:
            m = np.zeros(num.n, dtype=np.uint8)
            m[num.p[sid]] = 1
            self._problem = _MatrixProblem(num, pat, torch.from_numpy(m).to(pat.device))
            self._solver = None
            self._mask_key = mkey
        return num, pat, self._problem

    def _values(self, num, pat, A, out=None):
        A = sp.csr_matrix(A)
        A.sort_indices()
        key = _structure_key(A)
        pos = self._pos.get(key)
        if pos is None:
            pos = num.positions(pat, A)
            self._pos = {key: pos}
        return num.from_caller_csr(pat, A, pos=pos, out=out)

    def solve(self, A, b: np.ndarray, solve_ids: np.ndarray, reuse: bool = False) -> np.ndarray:
        """x with A[solve_ids][:, solve_ids] x[solve_ids] = b[solve_ids] and x = 0 elsewhere
        (solver.py:405-469).  reuse: A is the matrix of the previous call (its
        preconditioner set-up is kept, like reuse_factorization)."""
        from .fdsolve import GPUSolver
        num, pat, prob = self._setup(solve_ids)
        if self._solver is None:
            self._solver = GPUSolver(prob, restart=self.restart, precision=self.precision,
                                     preconditioner=self.preconditioner, poly_degree=self.poly_degree)
            reuse = False
        if not reuse or self._K is None:
            self._K = self._values(num, pat, A, out=self._K)
            self._solver.set_matrix(self._K)
        bd = num.from_caller(b) * prob.mask
        x, info = self._solver.solve(bd, tol=self.tol, maxiter=self.maxiter)
        self.last_info = info
        out = num.to_caller(x).cpu().numpy()
        keep = np.zeros(num.n, dtype=bool)
        keep[np.asarray(solve_ids, dtype=np.int64)] = True
        out[~keep] = 0
        return out

    def eig(self, A, B, solve_ids: np.ndarray, nmodes: int = 6, direct: bool | None = None,
            target_k0: float | None = None, which: str = "LM"):
        """Modes of A x = lambda B x on the free DOFs nearest lambda = target_k0^2 (the 3D
        cavity convention of fem/physics/edm/emfreq3d.py:838-848): shift-invert Krylov-Schur
        Arnoldi with GPU inner solves and FP64 Rayleigh-Ritz, gradient modes removed.  Returns
        (eigvals (nmodes,), eigvecs (len(solve_ids), nmodes)) like eigs on the reduced pencil."""
        import torch
        from .eigen import ShiftInvertArnoldi
        if which != "LM":
            raise ValueError("only which='LM' (nearest the shift) is supported")
        if target_k0 is None:
            raise ValueError("eig needs target_k0")
        num, pat, prob = self._setup(solve_ids)
        prob.E = self._values(num, pat, A)
        prob.B = self._values(num, pat, B)
        arn = ShiftInvertArnoldi(prob, float(target_k0) ** 2, ncv=max(4 * nmodes + 20, 24),
                                 preconditioner=self.preconditioner, poly_degree=self.poly_degree)
        lam, X, resid = arn.run(nmodes)
        self.last_info = {"residuals": resid, "restarts": arn.restarts, "inner_iterations": list(arn.inner_iters)}
        sid = np.asarray(solve_ids, dtype=np.int64)
        vecs = np.stack([num.to_caller(X[i].contiguous()).cpu().numpy()[sid] for i in range(X.shape[0])], axis=1)
        del arn
        torch.cuda.empty_cache()
        return lam, vecs
