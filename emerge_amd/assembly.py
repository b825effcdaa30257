"""GPU assembly of the Nedelec-2 curl-curl (E) and mass (B) matrices.

Drop-in for ``tet_mass_stiffness_matrices`` (fem/physics/edm/optimized_assembly.py:43-64):
same inputs (field/mesh, er and ur as (3,3,nt) complex tensors), same outputs
(E, B as N x N complex128 CSR with the reference's DOF numbering layout).
The heavy lifting is done by libemerge_amd.so; torch only owns device memory.
"""
from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp

from . import _lib
from .mesh import Mesh3D

EXACT_FACE_MASS = 1
FORCE_TWO_PHASE = 2
REFERENCE_GEOMETRY = 4


def is_isotropic(t: np.ndarray) -> bool:
    """True if every (3,3) tensor of a (3,3,nt) array is a scalar times I."""
    off = t.copy()
    for k in range(3):
        off[k, k] = 0
    d = np.stack([t[0, 0], t[1, 1], t[2, 2]])
    return not np.any(off) and np.array_equal(d[0], d[1]) and np.array_equal(d[0], d[2])


def matinv_reference(ur: np.ndarray) -> np.ndarray:
    """Per-tet inverse of mu_r as the reference computes it (fem/mth/optimized.py:440-461):
    diagonal tensors are inverted entry-wise, others get adjugate * det (sic)."""
    out = np.zeros_like(ur)
    offdiag = np.zeros(ur.shape[2], dtype=bool)
    for a in range(3):
        for b in range(3):
            if a != b:
                offdiag |= ur[a, b] != 0
    d = ~offdiag
    for k in range(3):
        out[k, k, d] = 1.0 / ur[k, k, d]
    if np.any(offdiag):
        s = ur[:, :, offdiag]
        det = (s[0, 0] * s[1, 1] * s[2, 2] - s[0, 0] * s[1, 2] * s[2, 1] - s[0, 1] * s[1, 0] * s[2, 2]
               + s[0, 1] * s[1, 2] * s[2, 0] + s[0, 2] * s[1, 0] * s[2, 1] - s[0, 2] * s[1, 1] * s[2, 0])
        adj = np.empty_like(s)
        adj[0, 0] = s[1, 1] * s[2, 2] - s[1, 2] * s[2, 1]
        adj[0, 1] = -s[0, 1] * s[2, 2] + s[0, 2] * s[2, 1]
        adj[0, 2] = s[0, 1] * s[1, 2] - s[0, 2] * s[1, 1]
        adj[1, 0] = -s[1, 0] * s[2, 2] + s[1, 2] * s[2, 0]
        adj[1, 1] = s[0, 0] * s[2, 2] - s[0, 2] * s[2, 0]
        adj[1, 2] = -s[0, 0] * s[1, 2] + s[0, 2] * s[1, 0]
        adj[2, 0] = s[1, 0] * s[2, 1] - s[1, 1] * s[2, 0]
        adj[2, 1] = -s[0, 0] * s[2, 1] + s[0, 1] * s[2, 0]
        adj[2, 2] = s[0, 0] * s[1, 1] - s[0, 1] * s[1, 0]
        out[:, :, offdiag] = adj * det
    return out


class GPUPattern:
    """Device-resident mesh + symbolic CSR pattern (built once per mesh)."""

    def __init__(self, mesh: Mesh3D, device: str = "cuda"):
        import torch
        L = _lib.load()
        self.mesh = mesh
        self.device = torch.device(device)
        if mesh.n_nodes >= 2_000_000:
            raise ValueError("meshes with >= 2e6 vertices are not supported by the face keys")
        self.nodes = torch.from_numpy(np.ascontiguousarray(mesh.nodes.T)).to(self.device)      # (nv,3) f64
        self.tets = torch.from_numpy(mesh.tets_canonical).to(self.device)                      # (nt,4) i32
        self.tet_ent = torch.from_numpy(mesh.tet_ent).to(self.device)                          # (nt,10) i32
        handle = ctypes.c_void_p()
        _lib.check(L.emg_pattern_build(_lib.ptr(self.tet_ent), mesh.n_tets, mesh.n_entities, _lib.stream_ptr(),
                                       ctypes.byref(handle)), "emg_pattern_build")
        self._handle = handle
        n_rows, nnz, nnz_ent, n_inc = (ctypes.c_int64() for _ in range(4))
        _lib.check(L.emg_pattern_info(handle, ctypes.byref(n_rows), ctypes.byref(nnz), ctypes.byref(nnz_ent),
                                      ctypes.byref(n_inc)), "emg_pattern_info")
        self.n, self.nnz, self.nnz_ent, self.n_inc = n_rows.value, nnz.value, nnz_ent.value, n_inc.value
        self.rowptr = torch.empty(self.n + 1, dtype=torch.int32, device=self.device)
        self.cols = torch.empty(self.nnz, dtype=torch.int32, device=self.device)
        _lib.check(L.emg_pattern_csr(handle, _lib.ptr(self.rowptr), _lib.ptr(self.cols), _lib.stream_ptr()),
                   "emg_pattern_csr")

    @property
    def handle(self):
        return self._handle

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                _lib.load().emg_pattern_destroy(h)
            except Exception:
                pass
            self._handle = None

    def assemble(self, er: np.ndarray, ur: np.ndarray, exact_face_mass: bool = False, out=None,
                 two_phase: bool = False, reference_geometry: bool = False):
        """E, B values (nnz,) complex128 on the device for tensors er, ur (3,3,nt).

        two_phase selects the per-tet-block isotropic kernels instead of the
        fused one (same results; used as a fallback and in tests).
        reference_geometry evaluates the vertex gradients and volumes with the
        reference's absolute-coordinate formulas (fem/mth/optimized.py:343-369)
        on the mesh's own tet vertex order, reproducing the reference's rounding;
        isotropic materials only."""
        import torch
        L = _lib.load()
        flags = ((EXACT_FACE_MASS if exact_face_mass else 0) | (FORCE_TWO_PHASE if two_phase else 0)
                 | (REFERENCE_GEOMETRY if reference_geometry else 0))
        if out is None:
            E = torch.empty(self.nnz, dtype=torch.complex128, device=self.device)
            B = torch.empty(self.nnz, dtype=torch.complex128, device=self.device)
        else:
            E, B = out
        nu = matinv_reference(np.asarray(ur, dtype=np.complex128))
        if is_isotropic(er) and is_isotropic(nu):
            eps = torch.from_numpy(np.ascontiguousarray(er[0, 0])).to(self.device)
            nus = torch.from_numpy(np.ascontiguousarray(nu[0, 0])).to(self.device)
            self.assemble_iso_device(eps, nus, flags, E, B)
        else:
            if reference_geometry:
                raise ValueError("reference_geometry needs isotropic materials")
            mm = torch.from_numpy(np.ascontiguousarray(np.transpose(er, (2, 0, 1)))).to(self.device)
            ms = torch.from_numpy(np.ascontiguousarray(np.transpose(nu, (2, 0, 1)))).to(self.device)
            _lib.check(L.emg_assemble_tensor(self._handle, _lib.ptr(self.nodes), _lib.ptr(self.tets), _lib.ptr(mm),
                                             _lib.ptr(ms), flags, _lib.ptr(E), _lib.ptr(B), _lib.stream_ptr()),
                       "emg_assemble_tensor")
        return E, B

    def assembly_tuning(self) -> dict:
        """The fused assembly's autotuned grid (chosen on the pattern's first fused assembly):
        workgroups per CU and the measured ms per assembly of each candidate (4, 8, 16, 32, 64);
        wg_per_cu 0 before the first fused assembly."""
        wg = ctypes.c_int()
        ms = (ctypes.c_float * 5)()
        _lib.check(_lib.load().emg_pattern_assembly_tuning(self._handle, ctypes.byref(wg), ms),
                   "emg_pattern_assembly_tuning")
        return {"wg_per_cu": int(wg.value), "ms": dict(zip((4, 8, 16, 32, 64), [float(v) for v in ms]))}

    def assemble_iso_device(self, eps, nu, flags, E, B):
        """Isotropic assembly from device-resident per-tet eps_r and 1/mu_r (complex128)."""
        L = _lib.load()
        tets = self.tets
        if flags & REFERENCE_GEOMETRY:
            if getattr(self, "tets_ref", None) is None:
                import torch
                self.tets_ref = torch.from_numpy(np.ascontiguousarray(self.mesh.tets.T.astype(np.int32))).to(self.device)
            tets = self.tets_ref
        _lib.check(L.emg_assemble_iso(self._handle, _lib.ptr(self.nodes), _lib.ptr(tets), _lib.ptr(eps),
                                      _lib.ptr(nu), flags, _lib.ptr(E), _lib.ptr(B), _lib.stream_ptr()),
                   "emg_assemble_iso")

    def assemble_iso_batched(self, nodes, eps, nu, flags: int = 0):
        """E and B of several meshes that share this pattern's tets, in one fused launch pair.

        nodes: (M, nv, 3) float64 coordinates per mesh (device tensor or numpy), eps, nu:
        (M, nt) complex128 eps_r and 1/mu_r per tet.  Returns device tensors E, B of shape
        (M, nnz); mesh b's values equal assemble_iso_device's on that mesh."""
        import torch
        L = _lib.load()
        dev = self.device
        nodes = torch.as_tensor(nodes, dtype=torch.float64, device=dev).contiguous()
        eps = torch.as_tensor(eps, dtype=torch.complex128, device=dev).contiguous()
        nu = torch.as_tensor(nu, dtype=torch.complex128, device=dev).contiguous()
        M, nv, nt = nodes.shape[0], self.mesh.n_nodes, self.mesh.n_tets
        if nodes.shape != (M, nv, 3) or eps.shape != (M, nt) or nu.shape != (M, nt):
            raise ValueError(f"batch shapes: nodes {tuple(nodes.shape)}, eps {tuple(eps.shape)}, nu {tuple(nu.shape)}; "
                             f"expected ({M}, {nv}, 3), ({M}, {nt}), ({M}, {nt})")
        tets = self.tets
        if flags & REFERENCE_GEOMETRY:
            if getattr(self, "tets_ref", None) is None:
                self.tets_ref = torch.from_numpy(np.ascontiguousarray(self.mesh.tets.T.astype(np.int32))).to(dev)
            tets = self.tets_ref
        E = torch.empty((M, self.nnz), dtype=torch.complex128, device=dev)
        B = torch.empty_like(E)
        _lib.check(L.emg_assemble_iso_batched(self._handle, M, nv, _lib.ptr(nodes), _lib.ptr(tets), _lib.ptr(eps),
                                              _lib.ptr(nu), flags, _lib.ptr(E), _lib.ptr(B), _lib.stream_ptr()),
                   "emg_assemble_iso_batched")
        return E, B

    def to_scipy(self, vals) -> sp.csr_matrix:
        """Device values -> host scipy CSR with this pattern."""
        return sp.csr_matrix((vals.cpu().numpy(), self.cols.cpu().numpy(), self.rowptr.cpu().numpy()),
                             shape=(self.n, self.n))


def tet_mass_stiffness_matrices(mesh: Mesh3D, er: np.ndarray, ur: np.ndarray, exact_face_mass: bool = False,
                                pattern: GPUPattern | None = None) -> tuple[sp.csr_matrix, sp.csr_matrix]:
    """E (curl-curl, with 1/mu_r) and B (mass, with eps_r) as host CSR matrices,
    assembled on the GPU (fem/physics/edm/optimized_assembly.py:43-64)."""
    pat = GPUPattern(mesh) if pattern is None else pattern
    E, B = pat.assemble(er, ur, exact_face_mass=exact_face_mass)
    return pat.to_scipy(E), pat.to_scipy(B)
