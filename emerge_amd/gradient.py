"""Discrete gradient of the second-order Lagrange space into the Nedelec-2 space.

The curl-curl matrix E of the reference's element (fem/mth/tet.py:854-1088)
has the gradients of P2 functions (vertices and edge midpoints) as its null
space.  Writing grad(phi) in the edge/face basis of fem/mth/tet.py:451-492
gives exact integer coefficients that depend only on the entity a DOF sits on,
divided by that basis function's length factor:

  edge (a<b), length L:           mode 0: ( 3 v_a + 1 v_b - 4 e_ab) / L
                                  mode 1: (-1 v_a - 3 v_b + 4 e_ab) / L
  face (a<b<c):                   mode 0: ( v_a + v_b - 2 v_c - 8 e_ab + 4 e_ac + 4 e_bc) / L_ac
                                  mode 1: (-v_a + 2 v_b - v_c - 4 e_ab + 8 e_ac - 4 e_bc) / L_ab

(v_x: vertex function lambda_x (2 lambda_x - 1); e_xy: edge function
4 lambda_x lambda_y).  G is used by the GPU solver's hybrid smoother, which
adds the correction G diag(G^T K G)^{-1} G^T to the 2x2 block Jacobi step, so
the gradient components of the field -- scaled by -k0^2 instead of the curl-curl
stiffness, the low-frequency trouble of E - k0^2 B -- are preconditioned by their
own diagonal.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

EDGE_ROWS = ((3.0, 1.0, -4.0), (-1.0, -3.0, 4.0))                      # (v_a, v_b, e_ab)
FACE_ROWS = ((1.0, 1.0, -2.0, -8.0, 4.0, 4.0), (-1.0, 2.0, -1.0, -4.0, 8.0, -4.0))   # (v_a, v_b, v_c, e_ab, e_ac, e_bc)


def gradient_matrix(mesh, mask: np.ndarray | None = None):
    """(G, nodes): G (N x n_nodes) CSR float64 with rows = DOFs (mode * n_ent + entity)
    and columns = P2 nodes (vertex v -> v, edge e -> n_vertices + e).

    With a DOF mask (1 = free), the P2 nodes whose gradient touches a masked DOF
    (nodes on PEC surfaces) are dropped; `nodes` lists the kept node ids.
    """
    nv, ne, nf = mesh.n_nodes, mesh.n_edges, mesh.n_tris
    nent = ne + nf
    a, b = mesh.edges
    L = mesh.edge_lengths
    e = np.arange(ne)
    rows, cols, vals = [], [], []
    for mode, coef in enumerate(EDGE_ROWS):
        r = mode * nent + e
        for node, c in zip((a, b, nv + e), coef):
            rows.append(r)
            cols.append(node)
            vals.append(c / L)
    ta, tb, tc = mesh.tris
    eab, ebc, eac = mesh.tri_to_edge                      # (a,b), (b,c), (a,c)
    f = np.arange(nf)
    scale = (L[eac], L[eab])
    for mode, coef in enumerate(FACE_ROWS):
        r = mode * nent + ne + f
        for node, c in zip((ta, tb, tc, nv + eab, nv + eac, nv + ebc), coef):
            rows.append(r)
            cols.append(node)
            vals.append(c / scale[mode])
    rows, cols, vals = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    nodes = np.arange(nv + ne)
    if mask is not None:
        bad = np.zeros(nv + ne, dtype=bool)
        bad[cols[~np.asarray(mask, dtype=bool)[rows]]] = True
        nodes = np.flatnonzero(~bad)
        remap = -np.ones(nv + ne, dtype=np.int64)
        remap[nodes] = np.arange(nodes.shape[0])
        keep = remap[cols] >= 0
        rows, cols, vals = rows[keep], remap[cols[keep]], vals[keep]
    G = sp.csr_matrix((vals, (rows, cols)), shape=(2 * nent, nodes.shape[0]))
    G.sort_indices()
    return G, nodes
