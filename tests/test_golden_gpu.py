"""The GPU boundary kernel (emg_robin_blocks, with the host-side local frames of
emerge_amd.fdsolve.port_local_xy) against golden vectors written by the
reference's own ned2_tri_stiff_force / ned2_tri_stiff (fem/mth/tri.py:575-796)."""
import os
from types import SimpleNamespace

import numpy as np
import pytest

from emerge_amd import _lib
from emerge_amd.bc import AbsorbingBoundary
from emerge_amd.fdsolve import port_local_xy

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ned2_ref.npz")


def _robin(cuda, lxy, elen, uinc, dpts):
    import torch
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(cuda)
    nt = lxy.shape[0]
    B = torch.empty((nt, 64), dtype=torch.float64, device=cuda)
    b = torch.empty((nt, 8), dtype=torch.complex128, device=cuda) if uinc is not None else None
    dl, de = t(lxy, np.float64), (t(elen, np.float64) if elen is not None else None)
    du = t(uinc, np.complex128) if uinc is not None else None
    dd = t(dpts, np.float64)
    L = _lib.load()
    _lib.check(L.emg_robin_blocks(nt, _lib.ptr(dl), _lib.ptr(de) if de is not None else None,
                                  _lib.ptr(du) if du is not None else None, dpts.shape[1], _lib.ptr(dd), _lib.ptr(B),
                                  _lib.ptr(b) if b is not None else None, _lib.stream_ptr()), "emg_robin_blocks")
    return B.cpu().numpy().reshape(nt, 8, 8), (b.cpu().numpy() if b is not None else None)


def test_abc_blocks_match_reference(cuda):
    g = np.load(GOLD)
    V = g["tris_verts"]                                   # (n, 3, 3): xyz rows, vertex columns
    n = V.shape[0]
    mesh = SimpleNamespace(nodes=np.concatenate(list(V), axis=1), tris=np.arange(3 * n).reshape(n, 3).T)
    lxy = port_local_xy(AbsorbingBoundary([1]), mesh, np.arange(n))
    B, _ = _robin(cuda, lxy, g["tris_elen"], None, g["trif_dpts"])
    for i in range(n):
        ref = g["tris_B"][i] / g["tris_gamma"][i]
        err = np.abs(B[i] - ref).max() / np.abs(ref).max()
        assert err <= 1e-12, (i, err)


def test_port_blocks_and_sources_match_reference(cuda):
    g = np.load(GOLD)
    lcs = g["trif_lcs"]
    lxy = np.concatenate([lcs[:, 0, :], lcs[:, 1, :]], axis=1)
    uinc = np.transpose(g["trif_uinc"][:, :2, :], (0, 2, 1))     # (n, nq, 2)
    B, b = _robin(cuda, lxy, None, uinc, g["trif_dpts"])
    for i in range(lcs.shape[0]):
        ref = g["trif_B"][i] / g["trif_gamma"][i]
        assert np.abs(B[i] - ref).max() / np.abs(ref).max() <= 1e-12, i
        assert np.abs(b[i] - g["trif_b"][i]).max() / np.abs(g["trif_b"][i]).max() <= 1e-12, i
