"""Shift-invert Arnoldi eigenmodes of the curl-curl pencil on the GPU.

Solves E x = lambda B x (lambda = k^2) on the PEC-reduced space for the modes
nearest sigma = (2 pi f_target / c)^2, as the reference's deprecated 3D call
does with ARPACK (fem/physics/edm/emfreq3d.py:838-848:
``eigs(A, k, M=B, sigma, which='LM')``) and its port-mode path with eigsh
(fem/solver.py:343-357).

* OP v = (E - sigma B)^{-1} B v, every application is a GPU GMRES solve
  (complex128, tight tolerance) with the shifted matrix set once;
* Arnoldi with classical Gram-Schmidt + one re-orthogonalisation, all vector
  work through the library's complex128 primitives;
* FP64 Rayleigh-Ritz: the pencil is projected onto the Krylov basis
  (V^H E V, V^H B V) and the small dense problem is solved in FP64, which
  gives eigenvalues far more accurate than the inner-solve tolerance;
* spurious modes: the gradient null space of E has lambda = 0 with huge
  multiplicity, but a Krylov space holds only one direction per distinct
  eigenvalue, so it surfaces as a few Ritz pairs with lambda ~ 0, removed by
  a threshold relative to sigma.
"""
from __future__ import annotations

import ctypes

import numpy as np
import scipy.linalg as sla

from . import _lib
from .bc import C0


class ShiftInvertArnoldi:
    def __init__(self, problem, sigma: float, ncv: int = 40, inner_tol: float = 1e-12, restart: int = 150,
                 inner_maxiter: int = 200000, seed: int = 0, preconditioner: str = "hybrid", verbose: bool = False,
                 poly_degree: int = 1):
        import torch
        from .fdsolve import GPUSolver
        self.torch = torch
        self.p = problem
        self.sigma = float(sigma)
        self.ncv = ncv
        self.inner_tol = inner_tol
        self.inner_maxiter = inner_maxiter
        self.L = _lib.load()
        N = problem.N
        dev = problem.dev
        self.Ks = torch.empty(problem.nnz, dtype=torch.complex128, device=dev)
        _lib.check(self.L.emg_affine_combine(problem.nnz, _lib.ptr(problem.E), _lib.ptr(problem.B), self.sigma,
                                             _lib.ptr(self.Ks), _lib.stream_ptr()), "emg_affine_combine")
        self.solver = GPUSolver(problem, restart=restart, precision=64, preconditioner=preconditioner,
                                poly_degree=poly_degree)
        self.solver.set_matrix(self.Ks)
        self.V = torch.zeros((ncv + 1, N), dtype=torch.complex128, device=dev)
        self.work = torch.empty(((ncv + 8) // 8) * 1024 * 16 + 1024, dtype=torch.float64, device=dev)
        self.h = torch.empty(2 * (ncv + 1), dtype=torch.float64, device=dev)
        self.nrm = torch.empty(4, dtype=torch.float64, device=dev)
        self.seed = seed
        self.verbose = verbose
        self.inner_iters = []
        self.inner_resid = []

    def _spmv(self, vals, x, y):
        p = self.p
        _lib.check(self.L.emg_spmv(p.N, _lib.ptr(p.pat.rowptr), _lib.ptr(p.pat.cols), _lib.ptr(vals), _lib.ptr(p.mask),
                                   _lib.ptr(x), _lib.ptr(y), _lib.stream_ptr()), "emg_spmv")

    def _norm(self, x) -> float:
        _lib.check(self.L.emg_maxpy(self.p.N, 0, _lib.ptr(self.V), _lib.ptr(self.h), _lib.ptr(x), _lib.ptr(self.nrm),
                                    _lib.ptr(self.work), _lib.stream_ptr()), "emg_maxpy")
        return float(self.nrm[0].item())

    def _comb(self, V, coef: np.ndarray, out):
        """out = sum_i coef[i] V[i] (library kernel)."""
        c = np.empty(2 * coef.shape[0])
        c[0::2], c[1::2] = coef.real, coef.imag
        self.h[: c.shape[0]].copy_(self.torch.from_numpy(c))
        _lib.check(self.L.emg_mcomb(self.p.N, coef.shape[0], _lib.ptr(V), _lib.ptr(self.h), _lib.ptr(out),
                                    _lib.stream_ptr()), "emg_mcomb")

    def _axpy_norm(self, w, x, a: complex) -> float:
        """w -= a x, returns ||w||."""
        self.h[:2].copy_(self.torch.tensor([a.real, a.imag], dtype=self.torch.float64))
        _lib.check(self.L.emg_maxpy(self.p.N, 1, _lib.ptr(x), _lib.ptr(self.h), _lib.ptr(w), _lib.ptr(self.nrm),
                                    _lib.ptr(self.work), _lib.stream_ptr()), "emg_maxpy")
        return float(self.nrm[0].item())

    def _orth(self, nv, w) -> float:
        """CGS2 of w against V[:nv]; returns ||w|| after projection, H column on host."""
        torch, L = self.torch, self.L
        hsum = np.zeros(nv, dtype=np.complex128)
        for _ in range(2):
            _lib.check(L.emg_mdotc(self.p.N, nv, _lib.ptr(self.V), _lib.ptr(w), _lib.ptr(self.h), _lib.ptr(self.work),
                                   _lib.stream_ptr()), "emg_mdotc")
            _lib.check(L.emg_maxpy(self.p.N, nv, _lib.ptr(self.V), _lib.ptr(self.h), _lib.ptr(w), _lib.ptr(self.nrm),
                                   _lib.ptr(self.work), _lib.stream_ptr()), "emg_maxpy")
            hv = self.h[: 2 * nv].cpu().numpy()
            hsum += hv[0::2] + 1j * hv[1::2]
        return float(self.nrm[0].item()), hsum

    def run(self, nev: int, spurious_rel: float = 1e-3):
        torch = self.torch
        p = self.p
        m = self.ncv
        rng = np.random.default_rng(self.seed)
        v0 = rng.standard_normal(p.N) * p.mask.cpu().numpy()
        v = torch.from_numpy(v0.astype(np.complex128)).to(p.dev)
        self._comb(v[None, :], np.array([1.0 / self._norm(v)], dtype=np.complex128), self.V[0])
        H = np.zeros((m + 1, m), dtype=np.complex128)
        y = torch.empty(p.N, dtype=torch.complex128, device=p.dev)
        k = m
        for j in range(m):
            self._spmv(p.B, self.V[j], y)        # y = B v; PEC rows stay 0
            z, info = self.solver.solve(y, tol=self.inner_tol, maxiter=self.inner_maxiter)
            self.inner_iters.append(info.iterations)
            self.inner_resid.append(info.residual)
            if self.verbose:
                import sys
                print(f"arnoldi: step {j + 1}/{m}, inner GMRES {info.iterations} iterations, residual {info.residual:.1e}",
                      file=sys.stderr, flush=True)
            hn, hcol = self._orth(j + 1, z)
            H[: j + 1, j] = hcol
            H[j + 1, j] = hn
            if hn < 1e-14:
                k = j + 1
                break
            self._comb(z[None, :], np.array([1.0 / hn], dtype=np.complex128), self.V[j + 1])
        Vk = self.V[:k]
        # FP64 Rayleigh-Ritz on span(V): (V^H E V) c = lambda (V^H B V) c
        EV = torch.empty_like(Vk)
        BV = torch.empty_like(Vk)
        for i in range(k):
            self._spmv(p.E, Vk[i], EV[i])
            self._spmv(p.B, Vk[i], BV[i])
        Ap = np.empty((k, k), dtype=np.complex128)
        Bp = np.empty((k, k), dtype=np.complex128)
        for i in range(k):
            for M, out in ((EV, Ap), (BV, Bp)):
                _lib.check(self.L.emg_mdotc(p.N, k, _lib.ptr(Vk), _lib.ptr(M[i]), _lib.ptr(self.h), _lib.ptr(self.work),
                                            _lib.stream_ptr()), "emg_mdotc")
                hv = self.h[: 2 * k].cpu().numpy()
                out[:, i] = hv[0::2] + 1j * hv[1::2]
        lam, C = sla.eig(Ap, Bp)
        keep = np.isfinite(lam) & (np.abs(lam) > spurious_rel * abs(self.sigma))
        lam, C = lam[keep], C[:, keep]
        order = np.argsort(np.abs(lam - self.sigma))[:nev]
        lam, C = lam[order], C[:, order]
        # Ritz vectors and their true residuals ||E x - lambda B x|| / ||E x||
        X = torch.empty((lam.shape[0], p.N), dtype=torch.complex128, device=p.dev)
        resid = []
        ex = torch.empty(p.N, dtype=torch.complex128, device=p.dev)
        bx = torch.empty_like(ex)
        for i in range(lam.shape[0]):
            self._comb(Vk, C[:, i], X[i])
            self._spmv(p.E, X[i], ex)
            self._spmv(p.B, X[i], bx)
            ne = self._norm(ex)
            resid.append(self._axpy_norm(ex, bx, complex(lam[i])) / ne)
        return lam, X, np.array(resid)


def eigen_frequencies(lam) -> np.ndarray:
    """k^2 -> frequency in Hz (emfreq3d.py:850)."""
    return np.sqrt(lam) * C0 / (2 * np.pi)
