"""Shared test helpers: DOF-key permutations between numberings."""
import numpy as np

from oracle import oracle as O


def mesh_dof_keys(mesh) -> np.ndarray:
    """(N, 4) DOF keys of emerge_amd's numbering (sorted vertex ids, -1 pad, mode)."""
    ne, nent = mesh.n_edges, mesh.n_entities
    ent = np.concatenate([np.stack([mesh.edges[0], mesh.edges[1], -np.ones(ne, np.int64)], 1), mesh.tris.T])
    return np.concatenate([np.concatenate([ent, np.zeros((nent, 1), np.int64)], 1),
                           np.concatenate([ent, np.ones((nent, 1), np.int64)], 1)])


def oracle_in_mesh_order(mesh, Eo, Bo, ofield):
    """Permute oracle CSR matrices into emerge_amd's DOF numbering."""
    p = O.permutation_to(mesh_dof_keys(mesh), ofield.dof_keys())
    return Eo[p][:, p].tocsr(), Bo[p][:, p].tocsr()


def rel_fro(a, b) -> float:
    import scipy.sparse as sp
    if sp.issparse(a) or sp.issparse(b):
        from scipy.sparse.linalg import norm
        return float(norm(sp.csr_matrix(a - b)) / norm(sp.csr_matrix(b)))
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rel_max(a, b) -> float:
    d = abs(a - b)
    d = d.max() if hasattr(d, "max") else np.max(d)
    return float(d / abs(b).max())
