"""Mesh topology on the GPU (emg_topology_build): the arrays Mesh3D.update builds
(fem/mesh3d.py:224-355) and the entity table of the Nedelec-2 DOF map
(fem/elements/nedelec2.py:32-70), from radix sorts of 64-bit edge and face keys.
Used by ``Mesh3D(..., topology="device")``; the arrays equal the host build's."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def device_topology(nodes: np.ndarray, tets: np.ndarray, device: str = "cuda") -> dict:
    """Topology arrays of a mesh (nodes (3, nv) float64, tets (4, nt) int) built on the GPU
    and returned as host numpy arrays under Mesh3D's attribute names.  Raises ValueError for
    a non-manifold mesh or invalid vertex ids, as the host build does."""
    import torch
    L = _lib.load()
    nv, nt = int(nodes.shape[1]), int(tets.shape[1])
    dev = torch.device(device)
    d_tets = torch.from_numpy(np.ascontiguousarray(tets, dtype=np.int64)).to(dev)
    d_nodes = torch.from_numpy(np.ascontiguousarray(nodes, dtype=np.float64)).to(dev)
    h = ctypes.c_void_p()
    rc = L.emg_topology_build(nv, nt, _lib.ptr(d_tets), _lib.ptr(d_nodes), ctypes.byref(h), _lib.stream_ptr())
    if rc == -1:
        raise ValueError(L.emg_last_error().decode(errors="replace"))
    _lib.check(rc, "emg_topology_build")
    try:
        ne, nf = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(L.emg_topology_sizes(h, ctypes.byref(ne), ctypes.byref(nf)), "emg_topology_sizes")
        ne, nf = ne.value, nf.value
        i64 = lambda *sh: torch.empty(sh, dtype=torch.int64, device=dev)
        i32 = lambda *sh: torch.empty(sh, dtype=torch.int32, device=dev)
        out = {"edges": i64(2, ne), "tris": i64(3, nf), "tets_canonical": i32(nt, 4), "tet_ent": i32(nt, 10),
               "tet_to_edge": i64(6, nt), "tet_to_edge_sign": i64(6, nt), "tet_to_tri": i64(4, nt),
               "tet_to_tri_sign": i64(4, nt), "tri_to_tet": i64(2, nf), "tri_to_edge": i64(3, nf),
               "edge_lengths": torch.empty(ne, dtype=torch.float64, device=dev)}
        P = lambda k: _lib.ptr(out[k])
        _lib.check(L.emg_topology_get(h, P("edges"), P("tris"), P("tets_canonical"), P("tet_ent"), P("tet_to_edge"),
                                      P("tet_to_edge_sign"), P("tet_to_tri"), P("tet_to_tri_sign"), P("tri_to_tet"),
                                      P("tri_to_edge"), P("edge_lengths"), _lib.stream_ptr()), "emg_topology_get")
        torch.cuda.synchronize(dev)
    finally:
        L.emg_topology_destroy(h)
    return {k: np.ascontiguousarray(v.cpu().numpy()) for k, v in out.items()}
