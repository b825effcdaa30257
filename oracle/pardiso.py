"""CPU ORACLE -- test/baseline infrastructure only.

MKL PARDISO through its C ABI (the library pypardiso wraps), used the way the
reference's x86 default solver does (fem/solver.py:294-309 SolverPardiso):
the complex system is turned into the real 2N block system
[[Re A, -Im A], [Im A, Re A]] (fem/solver.py:70-90) and solved with a real
nonsymmetric factorisation (mtype 11, pypardiso's default).  Timed by
bench.py as the host-core reference solve.
"""
from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp

_MKL = None
MKL_PATHS = ["/opt/conda/lib/libmkl_rt.so.1", "libmkl_rt.so.1", "libmkl_rt.so"]


def mkl():
    global _MKL
    if _MKL is None:
        for p in MKL_PATHS:
            try:
                _MKL = ctypes.CDLL(p)
                break
            except OSError:
                continue
        if _MKL is None:
            raise OSError("MKL (libmkl_rt) not found")
    return _MKL


def available() -> bool:
    try:
        mkl()
        return True
    except OSError:
        return False


def complex_to_real_block(A, b):
    """fem/solver.py:70-83."""
    Ar = sp.csr_matrix(A.real)
    Ai = sp.csr_matrix(A.imag)
    upper = sp.hstack([Ar, -Ai])
    lower = sp.hstack([Ai, Ar])
    return sp.vstack([upper, lower]).tocsr(), np.hstack([b.real, b.imag])


class Pardiso:
    """Real nonsymmetric PARDISO factorisation of a CSR matrix (zero-based indices)."""

    def __init__(self, A: sp.csr_matrix, threads: int | None = None, mtype: int = 11, iparm: dict | None = None):
        L = mkl()
        if threads:
            L.MKL_Set_Num_Threads(ctypes.c_int(threads))
        A = sp.csr_matrix(A)
        A.sort_indices()
        self.n = A.shape[0]
        self.data = np.ascontiguousarray(A.data, dtype=np.float64)
        self.ia = np.ascontiguousarray(A.indptr, dtype=np.int32)
        self.ja = np.ascontiguousarray(A.indices, dtype=np.int32)
        self.pt = (ctypes.c_int64 * 64)()
        self.iparm = (ctypes.c_int32 * 64)()
        self.mtype = ctypes.c_int32(mtype)
        L.pardisoinit(self.pt, ctypes.byref(self.mtype), self.iparm)
        self.iparm[0] = 1          # use the values below
        self.iparm[1] = 2          # nested dissection (METIS)
        self.iparm[34] = 1         # zero-based indices
        for k, v in (iparm or {}).items():
            self.iparm[k] = v
        self._call(12, np.zeros(1), np.zeros(1), 1)

    def _call(self, phase, b, x, nrhs):
        L = mkl()
        i32 = ctypes.c_int32
        err = i32(0)
        perm = np.zeros(1, dtype=np.int32)
        P = ctypes.c_void_p
        L.pardiso(self.pt, ctypes.byref(i32(1)), ctypes.byref(i32(1)), ctypes.byref(self.mtype),
                  ctypes.byref(i32(phase)), ctypes.byref(i32(self.n)), self.data.ctypes.data_as(P),
                  self.ia.ctypes.data_as(P), self.ja.ctypes.data_as(P), perm.ctypes.data_as(P),
                  ctypes.byref(i32(nrhs)), self.iparm, ctypes.byref(i32(0)), b.ctypes.data_as(P), x.ctypes.data_as(P),
                  ctypes.byref(err))
        if err.value != 0:
            raise RuntimeError(f"PARDISO phase {phase} error {err.value}")

    @property
    def memory_kb(self) -> int:
        """Peak memory of analysis + factorisation in KB (iparm(15), iparm(16)+iparm(17))."""
        return int(max(self.iparm[14], self.iparm[15] + self.iparm[16]))

    def solve(self, b: np.ndarray) -> np.ndarray:
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        self._call(33, b, x, 1)
        return x

    def release(self):
        self._call(-1, np.zeros(1), np.zeros(1), 1)


def inertia(S: sp.spmatrix, threads: int | None = None) -> tuple[int, int]:
    """(positive, negative) eigenvalue counts of a real symmetric matrix, from PARDISO's
    symmetric indefinite factorisation (mtype -2, upper triangle; iparm(22), iparm(23)).
    With S = A - tau B (B positive definite) the negative count is the number of
    generalized eigenvalues below tau (Sylvester's law of inertia)."""
    U = sp.triu(sp.csr_matrix(S)).tocsr()
    # pivots perturbed by at most 1e-13 |S| (iparm(10)), Bunch-Kaufman pivoting (iparm(21)) with
    # weighted matching (iparm(13)); a count made with perturbed pivots is refused, since a
    # perturbation can flip the sign of an eigenvalue of S smaller than it
    P = Pardiso(U, threads=threads, mtype=-2, iparm={9: 13, 12: 1, 20: 1})
    out, perturbed = (int(P.iparm[21]), int(P.iparm[22])), int(P.iparm[13])
    P.release()
    if perturbed:
        raise RuntimeError(f"inertia: {perturbed} perturbed pivots; the eigenvalue count is not exact")
    return out


def solve_complex(Asel: sp.csr_matrix, bsel: np.ndarray, threads: int | None = None) -> np.ndarray:
    """SolverPardiso.solve on the real block system (fem/solver.py:304-309, 437-464)."""
    Ab, bb = complex_to_real_block(Asel, bsel)
    p = Pardiso(Ab, threads=threads)
    x = p.solve(bb)
    p.release()
    n = bsel.shape[0]
    return x[:n] + 1j * x[n:]
