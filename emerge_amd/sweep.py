"""Batched frequency sweep with FP64 residual certificates.

The reference solves every (frequency, port) pair from scratch: assemble K(f),
factorise, solve, sample S (fem/physics/edm/emfreq3d.py:607-732).  Over a
sweep the solutions x_p(f) are smooth in f, and K(f) is affine in a few
frequency functions (E - k0^2 B + sum_t gamma_t(f) P_t, b = c_p(f) q_p).  This
driver exploits both, on the GPU:

1. every converged GMRES solution is orthonormalised (CGS2) into a basis V
   held in HBM; the columns E v, B v, P_t v of each basis vector are
   orthonormalised in turn into Q (K(f) V = Q R C(f) for every f, with R
   upper triangular and C(f) holding 1, -k0^2, gamma_t(f)), so the trial
   y_j = argmin ||b_p(f) - K(f) V y|| is a small least-squares problem
   solved in FP64 on the host;
2. a block of up to 16 (frequency, port) pairs gets its trial solutions
   x_j = V y_j from one kernel that reads V once for the whole block;
3. one affine SpMM pass over E, B and the CSR indices computes every FP64
   residual ||c_j q_p - K(f_j) x_j|| of the block (K(f) is never formed);
4. a pair whose trial residual is above `tol` is solved by the GPU GMRES
   (mixed or FP64) warm-started from its trial solution, to tol * basis_tol
   (a basis more accurate than the target keeps later trials certifiable),
   and the solution enlarges the basis, after which the block's remaining
   failures are re-projected and re-checked.

Every accepted solution, trial or GMRES, has a measured FP64 relative
residual <= tol; the certificate of each pair is kept in its SolveRecord.
S-parameters are linear in the solution, so the port fields of each basis
vector are sampled once and combined with the coefficients y_j.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import _lib
from .bc import C0
from .emdata import SolveRecord
from .fdsolve import tri_dofs

MAX_BLOCK = 16
MAX_BASIS = 128           # emg_basis_combine takes nvec * m <= 2048 with nvec = 16
MAX_MULTI = 512           # vectors per emg_mdotc / emg_maxpy call


class StoredSolution:
    """A GMRES solution kept outside the basis (the basis was full): its vector
    and the port-point fields sampled from it."""

    def __init__(self, x, fields):
        self.x = x
        self.fields = fields


def _c128_to_f64(a: np.ndarray) -> np.ndarray:
    out = np.empty(2 * a.size)
    out[0::2], out[1::2] = a.real.ravel(), a.imag.ravel()
    return out


class BasisSweep:
    def __init__(self, problem, solver, sampler, tol: float = 1e-8, max_basis: int = 128, maxiter: int = 60000,
                 max_refine: int = 0, block: int = MAX_BLOCK, basis_tol: float = 1e-1, verbose: bool = False):
        import torch
        self.torch = torch
        self.p = problem
        self.solver = solver
        self.sampler = sampler
        self.tol = tol
        self.basis_tol = basis_tol
        self.verbose = verbose
        self.maxiter = maxiter
        self.max_refine = max_refine
        self.L = _lib.load()
        if not 1 <= int(max_basis) <= MAX_BASIS:
            raise ValueError(f"max_basis must be in [1, {MAX_BASIS}], not {max_basis!r}")
        self.max_basis = max_basis
        self.block = block
        p, dev, N = problem, problem.dev, problem.N
        self.ports = p.ports
        self.V = torch.zeros((max_basis, N), dtype=torch.complex128, device=dev)
        self.m = 0
        self.Fb: list[np.ndarray] = []                     # port-point fields of each basis vector
        self.work = torch.empty(max(((MAX_MULTI + 8) // 8) * 1024 * 16 + 1024, 4096 * MAX_BLOCK * 2),
                                dtype=torch.float64, device=dev)
        self.h = torch.empty(2 * max(max_basis * MAX_BLOCK, MAX_MULTI), dtype=torch.float64, device=dev)
        self.nrm = torch.empty(2 * MAX_BLOCK, dtype=torch.float64, device=dev)
        self.x = torch.empty(N, dtype=torch.complex128, device=dev)
        # port sources q_p (force scale 1), masked
        qs = []
        for port in self.ports:
            rt = next(r for r in p.robin if r.bc is port)
            qs.append(rt.q * p.mask)
        self.qsrc = torch.stack(qs).contiguous()
        # Robin rows: CSR over the rows that carry port / ABC blocks, entries tagged by term
        self.terms = [rt for rt in p.robin if rt.bc._include_stiff]
        rows, cols, term, vals = [], [], [], []
        for ti, rt in enumerate(self.terms):
            d = tri_dofs(p.mesh, rt.tri)
            rows.append(np.repeat(d, 8, axis=1).ravel())
            cols.append(np.tile(d, (1, 8)).ravel())
            term.append(np.full(d.shape[0] * 64, ti))
            vals.append(rt.vals.cpu().numpy())
        q_host = self.qsrc.cpu().numpy()
        q_support = np.flatnonzero(np.any(q_host != 0, axis=0))
        if rows:
            rows, cols = np.concatenate(rows), np.concatenate(cols)
            term, vals = np.concatenate(term), np.concatenate(vals)
            keep = p.mask.cpu().numpy()[rows].astype(bool) & p.mask.cpu().numpy()[cols].astype(bool)
            rows, cols, term, vals = rows[keep], cols[keep], term[keep], vals[keep]
            prows, rinv = np.unique(rows, return_inverse=True)
            order = np.argsort(rinv, kind="stable")
            ptr = np.zeros(prows.shape[0] + 1, dtype=np.int32)
            ptr[1:] = np.cumsum(np.bincount(rinv, minlength=prows.shape[0]))
            prow_of_row = -np.ones(N, dtype=np.int32)
            prow_of_row[prows] = np.arange(prows.shape[0], dtype=np.int32)
            to = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(dev)
            self.nprow = prows.shape[0]
            self.p_ptr = to(ptr, np.int32)
            self.p_col = to(cols[order], np.int32)
            self.p_term = to(term[order], np.int32)
            self.p_val = to(vals[order], np.complex128)
            self.prow_of_row = to(prow_of_row, np.int32)
            self.Zp = torch.zeros(self.nprow * MAX_BLOCK, dtype=torch.complex128, device=dev)
            self.pdofs = np.unique(np.concatenate([rows, cols, q_support]))
        else:
            self.nprow = 0
            self.pdofs = q_support
        pos = {d: i for i, d in enumerate(self.pdofs.tolist())}
        npd = self.pdofs.shape[0]
        self.P_pd = []
        for ti in range(len(self.terms)):
            sel = term == ti
            r = np.fromiter((pos[v] for v in rows[sel]), dtype=np.int64, count=int(sel.sum()))
            c = np.fromiter((pos[v] for v in cols[sel]), dtype=np.int64, count=int(sel.sum()))
            self.P_pd.append(sp.csr_matrix((vals[sel], (r, c)), shape=(npd, npd)))
        self.q_pd = q_host[:, self.pdofs]                      # (nports, npd)
        self.pdofs_d = torch.from_numpy(self.pdofs.astype(np.int64)).to(dev)
        self.Vpd = np.zeros((max_basis, npd), dtype=np.complex128)
        self.log: list[tuple] = []          # (basis size, [(freq index, port index, trial residual)])
        self.learned_info = None            # learned-vs-plain check of the first GMRES solve
        # K(f) V = U C(f) with U = [E v_i, B v_i, P_t v_i]_i; U = Q R (Q orthonormal, in HBM)
        self.ncol = 2 + len(self.terms)
        qmax = max_basis * self.ncol
        self.Q = torch.zeros((qmax, N), dtype=torch.complex128, device=dev)
        self.R = np.zeros((qmax, qmax), dtype=np.complex128)
        self.nq = 0
        self.Qq = np.zeros((len(self.ports), qmax), dtype=np.complex128)
        self.u = torch.empty(N, dtype=torch.complex128, device=dev)

    # ------------------------------------------------------------ primitives
    def _spmv(self, vals, x, y):
        p = self.p
        _lib.check(self.L.emg_spmv(p.N, _lib.ptr(p.pat.rowptr), _lib.ptr(p.pat.cols), _lib.ptr(vals), _lib.ptr(p.mask),
                                   _lib.ptr(x), _lib.ptr(y), _lib.stream_ptr()), "emg_spmv")

    def _mdotc(self, V, nv, w) -> np.ndarray:
        """conj(V_i) . w for i < nv (in calls of at most MAX_MULTI vectors)."""
        out = np.zeros(nv, dtype=np.complex128)
        for i0 in range(0, nv, MAX_MULTI):
            k = min(MAX_MULTI, nv - i0)
            _lib.check(self.L.emg_mdotc(self.p.N, k, _lib.ptr(V[i0:i0 + k]), _lib.ptr(w), _lib.ptr(self.h),
                                        _lib.ptr(self.work), _lib.stream_ptr()), "emg_mdotc")
            hv = self.h[: 2 * k].cpu().numpy()
            out[i0:i0 + k] = hv[0::2] + 1j * hv[1::2]
        return out

    def _maxpy(self, V, nv, coef: np.ndarray, w) -> float:
        """w -= sum_i coef_i V_i; returns ||w|| (in calls of at most MAX_MULTI vectors)."""
        for i0 in range(0, max(nv, 1), MAX_MULTI):
            k = min(MAX_MULTI, nv - i0)
            if k > 0:
                self.h[: 2 * k].copy_(self.torch.from_numpy(_c128_to_f64(coef[i0:i0 + k])))
            _lib.check(self.L.emg_maxpy(self.p.N, k, _lib.ptr(V[i0:i0 + max(k, 1)]), _lib.ptr(self.h), _lib.ptr(w),
                                        _lib.ptr(self.nrm), _lib.ptr(self.work), _lib.stream_ptr()), "emg_maxpy")
        return float(self.nrm[0].item())

    def _comb(self, V, nv, coef: np.ndarray, out):
        self.h[: 2 * nv].copy_(self.torch.from_numpy(_c128_to_f64(coef)))
        _lib.check(self.L.emg_mcomb(self.p.N, nv, _lib.ptr(V), _lib.ptr(self.h), _lib.ptr(out), _lib.stream_ptr()),
                   "emg_mcomb")

    # ------------------------------------------------------------ basis
    def add(self, x):
        """Orthonormalise x into the basis; returns y with x = V[:m] y (m after the call), or a
        StoredSolution holding x itself when the basis is full and x has a component outside it."""
        m = self.m
        w = self.x
        w.copy_(x)
        xn = self._maxpy(self.V, 0, np.zeros(0), w)
        y = np.zeros(m, dtype=np.complex128)
        r = xn
        if m > 0:
            for _ in range(2):
                h = self._mdotc(self.V, m, w)
                r = self._maxpy(self.V, m, h, w)
                y += h
        if r <= 1e-13 * xn:
            return y
        if m >= self.max_basis:
            xs = x.clone()
            return StoredSolution(xs, self.sampler.fields(xs[None, :])[0])
        self._comb(w[None, :], 1, np.array([1.0 / r], dtype=np.complex128), self.V[m])
        v = self.V[m]
        # the columns of K(f) v for every frequency: E v, B v, P_t v (masked operator)
        self.Vpd[m] = v[self.pdofs_d].cpu().numpy()
        u = self.u
        for k in range(self.ncol):
            if k == 0:
                self._spmv(self.p.E, v, u)
            elif k == 1:
                self._spmv(self.p.B, v, u)
            else:
                u.zero_()
                u[self.pdofs_d] = self.torch.from_numpy(self.P_pd[k - 2] @ self.Vpd[m]).to(self.p.dev)
            self._add_column(u, m * self.ncol + k)
        self.Fb.append(self.sampler.fields(v[None, :])[0])
        self.m = m + 1
        return np.append(y, r)

    def _add_column(self, u, col: int):
        """CGS2 of u against Q; records u = Q R[:, col], extending Q when u adds a direction."""
        nq = self.nq
        un = self._maxpy(self.Q, 0, np.zeros(0), u)
        h = np.zeros(nq, dtype=np.complex128)
        rho = un
        if nq > 0:
            for _ in range(2):
                hh = self._mdotc(self.Q, nq, u)
                rho = self._maxpy(self.Q, nq, hh, u)
                h += hh
        self.R[:nq, col] = h
        if rho > 1e-14 * un and nq < self.Q.shape[0]:
            self._comb(u[None, :], 1, np.array([1.0 / rho], dtype=np.complex128), self.Q[nq])
            self.R[nq, col] = rho
            for pi in range(len(self.ports)):
                self.Qq[pi, nq] = self._mdotc(self.Q[nq:nq + 1], 1, self.qsrc[pi])[0]
            self.nq = nq + 1

    def trial(self, freq: float, port_index: int) -> np.ndarray:
        """y minimising ||b_p(f) - K(f) V y|| (least squares through K V = Q R C(f))."""
        m, nq, nc = self.m, self.nq, self.ncol
        k0 = 2 * np.pi * freq / C0
        R = self.R[:nq, : m * nc]
        RC = R[:, 0::nc] - k0 * k0 * R[:, 1::nc]
        for t, rt in enumerate(self.terms):
            RC = RC + complex(rt.bc.get_gamma(k0)) * R[:, 2 + t::nc]
        rhs = complex(self.ports[port_index].force_scale(k0)) * self.Qq[port_index, :nq]
        return np.linalg.lstsq(RC, rhs, rcond=None)[0]

    # ------------------------------------------------------------ certificates
    def residuals(self, pairs, Y: list[np.ndarray]) -> np.ndarray:
        """FP64 relative residuals of x_j = V y_j for the (freq, port index) pairs."""
        torch, L, p = self.torch, self.L, self.p
        n = len(pairs)
        nv = 1
        while nv < n:
            nv *= 2
        m = self.m
        Ym = np.zeros((nv, m), dtype=np.complex128)
        s = np.zeros(nv)
        c = np.zeros(nv, dtype=np.complex128)
        qidx = np.zeros(nv, dtype=np.int32)
        g = np.zeros((max(len(self.terms), 1), nv), dtype=np.complex128)
        for j in range(nv):
            f, pi = pairs[min(j, n - 1)]
            y = Y[min(j, n - 1)]
            Ym[j, : y.shape[0]] = y
            k0 = 2 * np.pi * f / C0
            s[j] = k0 * k0
            c[j] = complex(self.ports[pi].force_scale(k0))
            qidx[j] = pi
            for ti, rt in enumerate(self.terms):
                g[ti, j] = complex(rt.bc.get_gamma(k0))
        dev = p.dev
        X = torch.empty(p.N * nv, dtype=torch.complex128, device=dev)
        Yd = torch.from_numpy(Ym).to(dev)
        _lib.check(L.emg_basis_combine(p.N, m, _lib.ptr(self.V), nv, _lib.ptr(Yd), _lib.ptr(X), _lib.stream_ptr()),
                   "emg_basis_combine")
        sd = torch.from_numpy(s).to(dev)
        cd = torch.from_numpy(c).to(dev)
        qd = torch.from_numpy(qidx).to(dev)
        zp = prow = None
        if self.nprow:
            gd = torch.from_numpy(np.ascontiguousarray(g)).to(dev)
            _lib.check(L.emg_port_apply(self.nprow, nv, _lib.ptr(self.p_ptr), _lib.ptr(self.p_col),
                                        _lib.ptr(self.p_term), _lib.ptr(self.p_val), _lib.ptr(gd), _lib.ptr(X),
                                        _lib.ptr(self.Zp), _lib.stream_ptr()), "emg_port_apply")
            zp, prow = _lib.ptr(self.Zp), _lib.ptr(self.prow_of_row)
        _lib.check(L.emg_sweep_residual(p.N, _lib.ptr(p.pat.rowptr), _lib.ptr(p.pat.cols), _lib.ptr(p.E),
                                        _lib.ptr(p.B), _lib.ptr(p.mask), nv, _lib.ptr(sd), _lib.ptr(X), _lib.ptr(self.qsrc),
                                        _lib.ptr(qd), _lib.ptr(cd), prow, zp, None, _lib.ptr(self.nrm),
                                        _lib.ptr(self.work), _lib.stream_ptr()), "emg_sweep_residual")
        nr = self.nrm[: 2 * nv].cpu().numpy()
        return np.sqrt(nr[0::2][:n] / np.maximum(nr[1::2][:n], 1e-300))

    # ------------------------------------------------------------ full solve
    def full_solve(self, freq: float, port_index: int, y0: np.ndarray | None):
        p = self.p
        if getattr(self, "_matrix_freq", None) != freq:
            self.solver.set_matrix(p.system(freq), freq)
            self._matrix_freq = freq
        b = p.rhs(self.ports[port_index], freq)
        x0 = None
        if y0 is not None and self.m > 0:
            x0 = self.torch.empty(p.N, dtype=self.torch.complex128, device=p.dev)
            self._comb(self.V, self.m, y0, x0)
        if self.solver.learned is not None and self.learned_info is None:
            # first GMRES of the sweep with a learned smoother: solve with the plain smoother, then
            # with the learned weights capped at the plain solve's iteration count, and keep the
            # network only if it converged in fewer (the plain smoother is the fallback)
            learned = self.solver.learned
            self.solver.learned = None
            self.solver.set_weights(None, None)
            self.solver.set_matrix(p.K)
            x2, info2 = self.solver.solve(b, x0=x0, tol=self.tol * self.basis_tol, maxiter=self.maxiter,
                                          max_refine=self.max_refine)
            self.solver.learned = learned
            self.solver.set_matrix(p.K, freq)
            # the cap in smoother applications (what maxiter bounds): steps times the degree used
            apps2 = int(info2.iterations) if info2.reserved & 2 else int(info2.iterations) * max(1, self.solver.poly_active)
            cap = max(1, apps2 - 1)
            x, info = self.solver.solve(b, x0=x0, tol=self.tol * self.basis_tol, maxiter=cap,
                                        max_refine=self.max_refine)
            use = bool(info.converged) and not (info.reserved & 1) and info.iterations < info2.iterations
            self.learned_info = {"learned_iterations": int(info.iterations), "plain_iterations": int(info2.iterations),
                                 "used": use}
            if not use:
                self.solver.learned = None
                self.solver.set_weights(None, None)
                self.solver.set_matrix(p.K)
                x, info2.iterations = x2, info2.iterations + info.iterations
                info = info2
        else:
            x, info = self.solver.solve(b, x0=x0, tol=self.tol * self.basis_tol, maxiter=self.maxiter,
                                        max_refine=self.max_refine)
        if self.verbose:
            import sys
            print(f"sweep: GMRES f={freq:.6g} port {self.ports[port_index].port_number}: {info.iterations} iterations, "
                  f"residual {info.residual:.2e}{' (plain smoother fallback)' if info.reserved & 1 else ''}"
                  f"{f' (polynomial degree lowered to {self.solver.poly_degree})' if info.reserved & 2 else ''}",
                  file=sys.stderr,
                  flush=True)
        return x, info

    # ------------------------------------------------------------ driver
    def run(self, freqs, label: str):
        """Returns per-frequency lists of (y, record) per port, in port order."""
        nf, npt = len(freqs), len(self.ports)
        out = [[None] * npt for _ in range(nf)]
        # solve the two ends and the middle first, so later trials interpolate
        seeds = sorted({0, nf - 1, nf // 2}) if nf >= 3 else list(range(nf))
        for fi in seeds:
            for pi in range(npt):
                y0 = self.trial(freqs[fi], pi) if self.m else None
                x, info = self.full_solve(freqs[fi], pi, y0)
                y = self.add(x)
                out[fi][pi] = (y, SolveRecord(self.ports[pi].port_number, info.iterations, info.refinements,
                                              info.residual, bool(info.converged), label))
        rest = [(fi, pi) for fi in range(nf) for pi in range(npt) if out[fi][pi] is None]
        for b0 in range(0, len(rest), self.block):
            todo = rest[b0:b0 + self.block]
            for _ in range(len(todo) + 1):
                if not todo:
                    break
                Y = [self.trial(freqs[fi], pi) for fi, pi in todo]
                res = self.residuals([(freqs[fi], pi) for fi, pi in todo], Y)
                self.log.append((self.m, [(fi, pi, float(r)) for (fi, pi), r in zip(todo, res)]))
                still = []
                for (fi, pi), y, r in zip(todo, Y, res):
                    if r <= self.tol:
                        out[fi][pi] = (y, SolveRecord(self.ports[pi].port_number, 0, 0, float(r), True, "basis"))
                    else:
                        still.append(((fi, pi), y))
                if not still:
                    break
                # GMRES on the worst pair, warm-started from its trial; enlarge the basis
                (fi, pi), y = still[int(np.argmax([res[todo.index(t)] for t, _ in still]))]
                x, info = self.full_solve(freqs[fi], pi, y)
                y = self.add(x)
                out[fi][pi] = (y, SolveRecord(self.ports[pi].port_number, info.iterations, info.refinements,
                                              info.residual, bool(info.converged), label))
                todo = [t for t, _ in still if t != (fi, pi)]
            if self.verbose:
                import sys
                nfull = sum(1 for row in out for e in row if e is not None and e[1].iterations > 0)
                print(f"sweep: block {b0 // self.block + 1}/{(len(rest) + self.block - 1) // self.block} done, "
                      f"basis {self.m}, GMRES solves {nfull}", file=sys.stderr, flush=True)
        return out

    def fields_of(self, y) -> np.ndarray:
        if isinstance(y, StoredSolution):
            return y.fields
        Fb = np.stack(self.Fb[: y.shape[0]])                   # (m, npts, 3)
        return np.tensordot(y, Fb, axes=(0, 0))

    def vector_of(self, y):
        if isinstance(y, StoredSolution):
            return y.x
        x = self.torch.empty(self.p.N, dtype=self.torch.complex128, device=self.p.dev)
        self._comb(self.V, y.shape[0], y, x)
        return x
