"""Write golden input/output vectors of the reference's closed-form element
routines, for the CPU tests that pin the oracle's restatements of them.

Run in the build container only (the reference source is not available
elsewhere):  python tests/golden/write_ref_vectors.py /root/reference

The three routines are the ones a restatement could get wrong in a way that
brute-force quadrature would flag as a reference quirk rather than a mistake
(the face-face mass coefficient of fem/mth/tet.py:1036 and the
non-orthonormal local frame of fem/mth/tri.py:708-721):

  * ned2_tet_stiff_mass   fem/mth/tet.py:854-1088
  * ned2_tri_stiff_force  fem/mth/tri.py:575-695
  * ned2_tri_stiff        fem/mth/tri.py:697-796

fem/mth/{optimized,tet,tri}.py import only numpy and numba; numba is absent
here, so a stub module whose ``njit`` returns the function unchanged is put on
sys.path (the same idea as the commented-out stub at fem/mth/tri.py:24-27) and
the three files are loaded as a synthetic package without running
fem/__init__.py.  The results are plain numpy arrays saved to
tests/golden/ned2_ref.npz; nothing of the reference is stored.
"""
from __future__ import annotations

import importlib
import os
import sys
import tempfile
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "ned2_ref.npz")

EPAIRS = [(0, 1), (0, 2), (0, 3), (1, 2), (3, 1), (2, 3)]      # fem/mesh3d.py:292 local edge pairs
FTRIP = [(0, 1, 2), (0, 2, 3), (0, 3, 1), (1, 2, 3)]          # fem/mesh3d.py:296 local faces

_NUMBA_STUB = '''
class _T:
    def __call__(self, *a, **k):
        return self
    def __getitem__(self, k):
        return self
    def __getattr__(self, k):
        return _T()

f8 = i8 = c16 = f4 = i4 = b1 = _T()
types = _T()
prange = range

def njit(*args, **kwargs):
    if len(args) == 1 and callable(args[0]) and not isinstance(args[0], _T) and not kwargs:
        return args[0]
    return lambda fn: fn

jit = njit
'''


def load_reference_mth(ref_root: str):
    stub_dir = tempfile.mkdtemp(prefix="numba_stub_")
    os.makedirs(os.path.join(stub_dir, "numba"))
    with open(os.path.join(stub_dir, "numba", "__init__.py"), "w") as f:
        f.write(_NUMBA_STUB)
    sys.path.insert(0, stub_dir)
    pkg = types.ModuleType("refmth")
    pkg.__path__ = [os.path.join(ref_root, "fem", "mth")]
    sys.modules["refmth"] = pkg
    opt = importlib.import_module("refmth.optimized")
    tet = importlib.import_module("refmth.tet")
    tri = importlib.import_module("refmth.tri")
    return opt, tet, tri


def local_mapping(vertex_ids, ids):
    out = np.zeros(ids.shape, dtype=np.int64)
    for t in range(ids.shape[1]):
        for j in range(ids.shape[0]):
            out[j, t] = int(np.flatnonzero(vertex_ids == ids[j, t])[0])
    return out


def tet_cases(rng, n):
    cases = []
    for i in range(n):
        scale = [1.0, 1e-3, 2e-2][i % 3]
        off = rng.normal(size=(3, 1)) * [0.0, 1.0, 5.0][i % 3] * scale
        verts = off + rng.normal(size=(3, 4)) * scale
        # keep a positive orientation (all meshes here are positively oriented)
        T = np.stack([verts[:, 1] - verts[:, 0], verts[:, 2] - verts[:, 0], verts[:, 3] - verts[:, 0]], 1)
        if np.linalg.det(T) < 0:
            verts[:, [2, 3]] = verts[:, [3, 2]]
        gids = rng.permutation(80)[:4]
        gedge = np.array([sorted([gids[a], gids[b]]) for a, b in EPAIRS]).T
        gtri = np.array([sorted([gids[a], gids[b], gids[c]]) for a, b, c in FTRIP]).T
        lem = local_mapping(gids, gedge)
        ltm = local_mapping(gids, gtri)
        el = np.array([np.linalg.norm(verts[:, a] - verts[:, b]) for a, b in EPAIRS])
        if i % 4 == 3:
            A = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
            Mm = A + A.T + 6 * np.eye(3)
            B = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
            Ms = 0.1 * (B + B.T) + np.eye(3)
        else:
            Mm = rng.uniform(1, 10) * (1 - 1j * rng.uniform(0, 0.02)) * np.eye(3)
            Ms = np.eye(3) / rng.uniform(1, 3)
        cases.append((verts, el, lem, ltm, Ms.astype(np.complex128), Mm.astype(np.complex128)))
    return cases


def main(ref_root: str) -> None:
    opt, tet, tri = load_reference_mth(ref_root)
    rng = np.random.default_rng(20260)
    out = {}
    # ---- tets
    tc = tet_cases(rng, 20)
    out["tet_verts"] = np.stack([c[0] for c in tc])
    out["tet_elen"] = np.stack([c[1] for c in tc])
    out["tet_lem"] = np.stack([c[2] for c in tc])
    out["tet_ltm"] = np.stack([c[3] for c in tc])
    out["tet_Ms"] = np.stack([c[4] for c in tc])
    out["tet_Mm"] = np.stack([c[5] for c in tc])
    D, F = [], []
    for verts, el, lem, ltm, Ms, Mm in tc:
        d, f = tet.ned2_tet_stiff_mass(verts, el, lem, ltm, Ms, Mm)
        D.append(np.array(d))
        F.append(np.array(f))
    out["tet_D"], out["tet_F"] = np.stack(D), np.stack(F)
    # ---- Robin triangles with a source (port faces, local 2D coordinates)
    dpts = np.array(opt.gaus_quad_tri(4))
    nq = dpts.shape[1]
    L, G, U, B, b = [], [], [], [], []
    for i in range(20):
        scale = [1.0, 1e-3, 5e-3][i % 3]
        lcs = np.zeros((3, 3))
        lcs[:2] = rng.normal(size=(2, 3)) * scale + rng.normal(size=(2, 1)) * scale
        gamma = complex(rng.normal() + 1j * rng.uniform(0.1, 3.0))
        Uinc = np.zeros((3, nq), dtype=np.complex128)
        Uinc[:2] = rng.normal(size=(2, nq)) + 1j * rng.normal(size=(2, nq))
        Bm, bv = tri.ned2_tri_stiff_force(lcs, gamma, Uinc, dpts)
        L.append(lcs); G.append(gamma); U.append(Uinc); B.append(np.array(Bm)); b.append(np.array(bv))
    out["trif_lcs"], out["trif_gamma"], out["trif_uinc"] = np.stack(L), np.array(G), np.stack(U)
    out["trif_B"], out["trif_b"], out["trif_dpts"] = np.stack(B), np.stack(b), dpts
    # ---- absorbing-boundary triangles (3D vertices, per-triangle frame of tri.py:708-721)
    V, E, G2, B2 = [], [], [], []
    for i in range(20):
        scale = [1.0, 1e-3, 2e-2][i % 3]
        verts = rng.normal(size=(3, 3)) * scale + rng.normal(size=(3, 1)) * scale * (i % 2)
        el = np.array([np.linalg.norm(verts[:, 0] - verts[:, 1]), np.linalg.norm(verts[:, 1] - verts[:, 2]),
                       np.linalg.norm(verts[:, 0] - verts[:, 2])])
        gamma = complex(1j * rng.uniform(10, 100))
        Bm = tri.ned2_tri_stiff(verts, el, gamma)
        V.append(verts); E.append(el); G2.append(gamma); B2.append(np.array(Bm))
    out["tris_verts"], out["tris_elen"], out["tris_gamma"], out["tris_B"] = (np.stack(V), np.stack(E), np.array(G2),
                                                                           np.stack(B2))
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: " + ", ".join(f"{k}{tuple(v.shape)}" for k, v in out.items()))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
