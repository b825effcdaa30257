"""Write golden input/output vectors of the reference's closed-form element
routines, for the CPU tests that pin the oracle's restatements of them.

Run in the build container only (the reference source is not available
elsewhere):  python tests/golden/write_ref_vectors.py /root/reference

The routines are the ones a restatement could get wrong in a way that
brute-force quadrature would flag as a reference quirk rather than a mistake
(the face-face mass coefficient of fem/mth/tet.py:1036, the non-orthonormal
local frame of fem/mth/tri.py:708-721, the 8-digit quadrature table of
fem/physics/edm/nedeleclegrange2.py:201-204, the longitudinal-field
expression of fem/mth/tri.py:303):

  * ned2_tet_stiff_mass    fem/mth/tet.py:854-1088
  * ned2_tri_stiff_force   fem/mth/tri.py:575-695
  * ned2_tri_stiff         fem/mth/tri.py:697-796
  * generalized_matrix_GQ  fem/physics/edm/nedeleclegrange2.py:206-374 (port modes)
  * ned2_tri_interp_full, ned2_tri_interp_curl  fem/mth/tri.py:228-401 (port-mode E, H)

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


_NUMBA_PROGRESS_STUB = '''
class ProgressBar:
    def __init__(self, *a, **k):
        pass
    def __enter__(self):
        return self
    def __exit__(self, *a):
        return False
    def update(self, n=1):
        pass
ProgressBarType = None
'''


def load_reference_port_elements(ref_root: str):
    """fem/physics/edm/nedeleclegrange2.py as a module: its generalized_matrix_GQ needs only
    numpy and fem/mth/optimized.py; the names it imports for its assembly driver (the
    NedelecLegrange2 field class, numba_progress's ProgressBar) are given empty stand-ins."""
    stub_dir = os.path.dirname(sys.modules["numba"].__file__)
    os.makedirs(os.path.join(os.path.dirname(stub_dir), "numba_progress"), exist_ok=True)
    with open(os.path.join(os.path.dirname(stub_dir), "numba_progress", "__init__.py"), "w") as f:
        f.write(_NUMBA_PROGRESS_STUB)
    for name in ("refpkg", "refpkg.physics", "refpkg.physics.edm", "refpkg.elements"):
        m = types.ModuleType(name)
        m.__path__ = []
        sys.modules[name] = m
    sys.modules["refpkg.physics.edm"].__path__ = [os.path.join(ref_root, "fem", "physics", "edm")]
    ned = types.ModuleType("refpkg.elements.nedleg2")
    ned.NedelecLegrange2 = object
    sys.modules["refpkg.elements.nedleg2"] = ned
    sys.modules["refpkg.mth"] = sys.modules["refmth"]
    sys.modules["refpkg.mth.optimized"] = sys.modules["refmth.optimized"]
    return importlib.import_module("refpkg.physics.edm.nedeleclegrange2")


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
    # ---- port-mode element matrices (generalized_matrix_GQ, nedeleclegrange2.py:223-374) and the
    # 2D field / curl interpolation of a port mode (tri.py:228-401)
    nl2 = load_reference_port_elements(ref_root)
    PV, PA, PB, PMs, PMm, PK = [], [], [], [], [], []
    for i in range(16):
        scale = [1.0, 1e-3, 5e-3, 2e-2][i % 4]
        v = rng.normal(size=(2, 3)) * scale + rng.normal(size=(2, 1)) * scale
        if (v[0, 1] - v[0, 0]) * (v[1, 2] - v[1, 0]) - (v[0, 2] - v[0, 0]) * (v[1, 1] - v[1, 0]) < 0:
            v[:, [1, 2]] = v[:, [2, 1]]
        if i % 5 == 4:
            v[:, [1, 2]] = v[:, [2, 1]]                     # a clockwise triangle too
        Mm = (rng.uniform(1, 5) * (1 - 0.02j)) * np.eye(3, dtype=np.complex128)
        Ms = np.eye(3, dtype=np.complex128) / rng.uniform(1, 2)
        if i % 3 == 2:
            Mm = np.diag(rng.uniform(1, 4, 3) + 1j * rng.uniform(-0.1, 0, 3)).astype(np.complex128)
            Ms = np.diag(1 / rng.uniform(1, 2, 3)).astype(np.complex128)
        k0 = rng.uniform(20, 80) / scale * 1e-3
        lem = np.array([[0, 1, 0], [1, 2, 2]], dtype=np.int64)
        A, B = nl2.generalized_matrix_GQ(np.ascontiguousarray(v), lem, Ms, Mm, float(k0))
        PV.append(v); PA.append(np.array(A)); PB.append(np.array(B)); PMs.append(Ms); PMm.append(Mm); PK.append(k0)
    out["gq_verts"], out["gq_A"], out["gq_B"] = np.stack(PV), np.stack(PA), np.stack(PB)
    out["gq_Ms"], out["gq_Mm"], out["gq_k0"] = np.stack(PMs), np.stack(PMm), np.array(PK)
    # a small planar triangle mesh (2 x 2 squares, 8 triangles) with a random 14-DOF-per-triangle field
    gx, gy = np.meshgrid(np.linspace(0, 2e-3, 3), np.linspace(0, 1.5e-3, 3))
    nodes2 = np.stack([gx.ravel(), gy.ravel()]) + rng.normal(size=(2, 9)) * 5e-5 * np.array([[0], [1]])
    tris = []
    for r in range(2):
        for c in range(2):
            a0, a1, a2, a3 = r * 3 + c, r * 3 + c + 1, (r + 1) * 3 + c, (r + 1) * 3 + c + 1
            tris += [sorted([a0, a1, a3]), sorted([a0, a3, a2])]
    tris = np.array(tris, dtype=np.int64).T
    nf = 40
    t2f = rng.integers(0, nf, size=(14, tris.shape[1])).astype(np.int64)
    field = rng.normal(size=nf) + 1j * rng.normal(size=nf)
    pts = []
    for t in range(tris.shape[1]):
        lam = rng.dirichlet(np.ones(3), size=3).T
        pts.append(nodes2[:, tris[:, t]] @ lam)
    pts = np.concatenate(pts, axis=1)
    Ef = tri.ned2_tri_interp_full(pts, field, tris, nodes2, t2f)
    dia = np.zeros((3, 3, tris.shape[1]), dtype=np.complex128)
    for t in range(tris.shape[1]):
        dia[:, :, t] = np.diag(rng.uniform(0.5, 2, 3))
    beta = 123.4
    Hf = tri.ned2_tri_interp_curl(pts, field, tris, nodes2, t2f, dia, beta)
    out["it_nodes"], out["it_tris"], out["it_t2f"], out["it_field"], out["it_pts"] = nodes2, tris, t2f, field, pts
    out["it_E"], out["it_diadic"], out["it_beta"], out["it_H"] = np.array(Ef), dia, np.array(beta), np.array(Hf)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: " + ", ".join(f"{k}{tuple(v.shape)}" for k, v in out.items()))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
