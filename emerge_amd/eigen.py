"""Shift-invert Arnoldi eigenmodes of the curl-curl pencil on the GPU.

Solves E x = lambda B x (lambda = k^2) on the PEC-reduced space for the modes
nearest sigma = (2 pi f_target / c)^2, as the reference's deprecated 3D call
does with ARPACK (fem/physics/edm/emfreq3d.py:838-848:
``eigs(A, k, M=B, sigma, which='LM')``) and its port-mode path with eigsh
(fem/solver.py:343-357).

* OP v = (E - sigma B)^{-1} B v, every application is a GPU GMRES solve to the
  FP64 relative residual inner_tol (default 1e-10; complex128, or complex64 cycles
  in FP64 refinement) with the shifted matrix set once;
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

import numpy as np
import scipy.linalg as sla

from . import _lib
from .bc import C0


class ShiftInvertArnoldi:
    def __init__(self, problem, sigma: float, ncv: int = 40, inner_tol: float = 1e-10, restart: int = 150,
                 inner_maxiter: int = 200000, seed: int = 0, preconditioner: str = "hybrid", verbose: bool = False,
                 poly_degree: int = 1, precision: int = 64):
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
        # inner solves: complex128 GMRES (precision 64) or complex64 cycles in FP64 refinement (32),
        # both to the FP64 residual inner_tol
        self.solver = GPUSolver(problem, restart=restart, precision=precision, preconditioner=preconditioner,
                                poly_degree=poly_degree)
        self.solver.set_matrix(self.Ks)
        self.V = torch.zeros((ncv + 1, N), dtype=torch.complex128, device=dev)
        self.restarts = 0
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

    def _arnoldi_step(self, j, H) -> bool:
        """V[j+1], H[:j+2, j] from OP V[j]; False on breakdown (invariant subspace)."""
        p = self.p
        y = self._y
        self._spmv(p.B, self.V[j], y)        # y = B v; PEC rows stay 0
        z, info = self.solver.solve(y, tol=self.inner_tol, maxiter=self.inner_maxiter)
        self.inner_iters.append(info.iterations)
        self.inner_resid.append(info.residual)
        if not info.converged:
            raise EigenError(f"inner GMRES did not reach {self.inner_tol:g} (residual {info.residual:.2e} after "
                             f"{info.iterations} iterations)")
        if self.verbose:
            import sys
            print(f"arnoldi: step {j + 1}/{self.ncv}, inner GMRES {info.iterations} iterations, "
                  f"residual {info.residual:.1e}", file=sys.stderr, flush=True)
        hn, hcol = self._orth(j + 1, z)
        H[: j + 1, j] = hcol
        H[j + 1, j] = hn
        if hn < 1e-14 * max(1.0, float(np.abs(hcol).max())):
            return False
        self._comb(z[None, :], np.array([1.0 / hn], dtype=np.complex128), self.V[j + 1])
        return True

    def _wanted(self, theta, nkeep, spurious_rel):
        """Boolean mask of the nkeep Ritz values theta of OP with the largest |theta|
        whose eigenvalue sigma + 1/theta is not a gradient (lambda ~ 0) mode."""
        with np.errstate(divide="ignore", invalid="ignore"):
            lam = self.sigma + 1.0 / theta
        ok = np.isfinite(lam) & (np.abs(lam) > spurious_rel * abs(self.sigma))
        order = [i for i in np.argsort(-np.abs(theta)) if ok[i]][:nkeep]
        mask = np.zeros(theta.shape[0], dtype=bool)
        mask[order] = True
        return mask

    def run(self, nev: int, spurious_rel: float = 1e-3, tol: float = 1e-10, max_restarts: int = 30,
            resid_tol: float = 1e-8):
        """The nev modes nearest sigma.  Krylov-Schur restarted Arnoldi on OP =
        (E - sigma B)^{-1} B: after each pass over ncv vectors the Schur form of
        the projected OP is reordered so the wanted Ritz values (largest |theta|,
        gradient modes excluded) lead, and the Krylov relation is truncated to
        them; converged Ritz values stay in the kept block (locked in place) while
        the rest keep improving.  Stops when nev Ritz pairs have residual estimates
        |h_{m+1,m} e_m^T s| <= tol |theta|, then an FP64 Rayleigh-Ritz step on
        span(V) gives the eigenvalues; every returned pair must have a true
        relative residual ||E x - lambda B x|| / ||E x|| <= resid_tol, else
        EigenError (no silent partial answer)."""
        import scipy.linalg as sla_
        torch = self.torch
        p = self.p
        m = self.ncv
        if nev < 1 or 2 * nev + 4 > m:
            raise ValueError(f"need 1 <= nev and ncv >= 2 nev + 4, got nev={nev}, ncv={m}")
        rng = np.random.default_rng(self.seed)
        v0 = rng.standard_normal(p.N) * p.mask.cpu().numpy()
        v = torch.from_numpy(v0.astype(np.complex128)).to(p.dev)
        self._comb(v[None, :], np.array([1.0 / self._norm(v)], dtype=np.complex128), self.V[0])
        self._y = torch.empty(p.N, dtype=torch.complex128, device=p.dev)
        H = np.zeros((m + 1, m), dtype=np.complex128)
        k = 0                      # OP V[:k] = V[:k+1] H[:k+1, :k] holds
        nkeep = min(m - 2, max(nev + 2, (nev + m) // 2))
        self.restarts = 0
        size = m
        for it in range(max_restarts + 1):
            size = m
            for j in range(k, m):
                if not self._arnoldi_step(j, H):
                    size = j + 1                       # invariant subspace: Ritz values are exact
                    break
            T, Q = sla_.schur(H[:size, :size], output="complex")
            theta = np.diag(T).copy()
            # residual estimates of the Ritz pairs (eigenvectors s of T, Ritz vectors V Q s)
            w, S = np.linalg.eig(T)
            est = np.abs(H[size, size - 1]) * np.abs((Q[size - 1, :] @ S)) if size == m else np.zeros(size)
            want = self._wanted(w, nev, spurious_rel)
            nconv = int(np.sum(want & (est <= tol * np.abs(w))))
            if self.verbose:
                import sys
                print(f"arnoldi: pass {it + 1}: {nconv}/{nev} Ritz pairs converged", file=sys.stderr, flush=True)
            if nconv >= nev or size < m or it == max_restarts:
                break
            # Krylov-Schur restart: the nkeep wanted Schur vectors first, truncate to them
            sel = self._wanted(theta, nkeep, spurious_rel)
            keep_vals = theta[sel]
            T2, Q2, sdim = sla_.schur(H[:m, :m], output="complex",
                                      sort=lambda z: bool(np.any(np.isclose(z, keep_vals, rtol=1e-12, atol=0.0))))
            kk = int(sdim)
            if kk < 1:
                break
            Vnew = torch.empty((kk, p.N), dtype=torch.complex128, device=p.dev)
            for i in range(kk):
                self._comb(self.V[:m], Q2[:, i].copy(), Vnew[i])
            self.V[:kk].copy_(Vnew)
            self.V[kk].copy_(self.V[m])
            del Vnew
            Hn = np.zeros_like(H)
            Hn[:kk, :kk] = T2[:kk, :kk]
            Hn[kk, :kk] = H[m, m - 1] * Q2[m - 1, :kk]
            H = Hn
            k = kk
            self.restarts += 1
        # the wanted Schur vectors of the projected OP span the converged invariant subspace;
        # Rayleigh-Ritz on all of span(V) would admit spurious interior Ritz values near sigma
        T, Q = sla_.schur(H[:size, :size], output="complex")
        keep_vals = np.diag(T)[self._wanted(np.diag(T), nev, spurious_rel)]
        _, Qw, sdim = sla_.schur(H[:size, :size], output="complex",
                                 sort=lambda z: bool(np.any(np.isclose(z, keep_vals, rtol=1e-12, atol=0.0))))
        k = int(sdim)
        if k < nev:
            raise EigenError(f"only {k} non-gradient Ritz values for {nev} requested modes (Krylov space of {size})")
        Vk = torch.empty((k, p.N), dtype=torch.complex128, device=p.dev)
        for i in range(k):
            self._comb(self.V[:size], Qw[:, i].copy(), Vk[i])
        # FP64 Rayleigh-Ritz on that subspace: (W^H E W) c = lambda (W^H B W) c
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
        del EV, BV
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
        resid = np.array(resid)
        good = int(np.sum(resid <= resid_tol))
        if lam.shape[0] < nev or good < nev:
            raise EigenError(f"{good} of {nev} requested modes reached relative residual {resid_tol:g} after "
                             f"{self.restarts} restarts (residuals {np.array2string(resid, precision=2)}); "
                             "raise ncv or max_restarts")
        order = np.argsort(lam.real)
        return lam[order], X[torch.from_numpy(order).to(p.dev)], resid[order]


class EigenError(RuntimeError):
    """The eigen solve could not deliver the requested modes to the stated residual."""


def eigen_frequencies(lam) -> np.ndarray:
    """k^2 -> frequency in Hz (emfreq3d.py:850)."""
    return np.sqrt(lam) * C0 / (2 * np.pi)
