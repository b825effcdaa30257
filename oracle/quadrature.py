"""CPU ORACLE -- test infrastructure only.

Brute-force pins for the closed-form element matrices: the reference's own
basis definitions evaluated pointwise and integrated by a high-order
collapsed-coordinate Gauss rule.

* ``ned2_tet_basis`` restates the basis of ``ned2_tet_interp``
  (fem/mth/tet.py:451-492): for each edge (a<b, global order) of length L,
  N_e1 = L*la*(lb*grad(la) - la*grad(lb)), N_e2 = L*lb*(same); for each face
  (a<b<c) N_f1 = -L_ac*lb*(lc*grad(la) - la*grad(lc)),
  N_f2 = L_ab*lc*(lb*grad(la) - la*grad(lb)).
* curls are taken by central differences, which are exact for these
  quadratic polynomials up to rounding.
* ``tet_rule`` is a Duffy-collapsed tensor Gauss-Legendre rule, exact for
  polynomials of total degree <= 2n-3 on the tetrahedron.
"""
from __future__ import annotations

import numpy as np


def tet_rule(n: int = 6):
    """Points (3, P) in the unit tetrahedron and weights (P,) summing to 1/6."""
    g, w = np.polynomial.legendre.leggauss(n)
    g = 0.5 * (g + 1)
    w = 0.5 * w
    u, v, t = np.meshgrid(g, g, g, indexing="ij")
    wu, wv, wt = np.meshgrid(w, w, w, indexing="ij")
    x = u
    y = v * (1 - u)
    z = t * (1 - u) * (1 - v)
    jac = (1 - u) ** 2 * (1 - v)
    return np.stack([x.ravel(), y.ravel(), z.ravel()]), (wu * wv * wt * jac).ravel()


def tri_rule(n: int = 6):
    """Points (2, P) in the unit triangle and weights (P,) summing to 1/2."""
    g, w = np.polynomial.legendre.leggauss(n)
    g = 0.5 * (g + 1)
    w = 0.5 * w
    u, v = np.meshgrid(g, g, indexing="ij")
    wu, wv = np.meshgrid(w, w, indexing="ij")
    return np.stack([u.ravel(), (v * (1 - u)).ravel()]), (wu * wv * (1 - u)).ravel()


def barycentric(verts: np.ndarray, pts: np.ndarray) -> np.ndarray:
    """lambda_k(x) (4, P) for tet vertices verts (3,4) at physical points (3,P)."""
    T = np.stack([verts[:, 1] - verts[:, 0], verts[:, 2] - verts[:, 0], verts[:, 3] - verts[:, 0]], axis=1)
    loc = np.linalg.solve(T, pts - verts[:, :1])
    return np.vstack([1 - loc.sum(axis=0), loc])


def grad_barycentric(verts: np.ndarray) -> np.ndarray:
    """grad(lambda_k) (4, 3)."""
    T = np.stack([verts[:, 1] - verts[:, 0], verts[:, 2] - verts[:, 0], verts[:, 3] - verts[:, 0]], axis=1)
    Ti = np.linalg.inv(T)  # rows: grad of local coords
    return np.vstack([-Ti.sum(axis=0), Ti])


def ned2_tet_basis(verts: np.ndarray, gids: np.ndarray, pts: np.ndarray) -> np.ndarray:
    """(20, 3, P) basis values in the reference's local slot order.

    verts (3,4) in the tet's own vertex order with global ids ``gids`` (4,);
    slots follow fem/mesh3d.py:292,296 and fem/elements/nedelec2.py:46-50.
    """
    lam = barycentric(verts, pts)
    G = grad_barycentric(verts)
    out = np.zeros((20, 3, pts.shape[1]))
    epairs = [(0, 1), (0, 2), (0, 3), (1, 2), (3, 1), (2, 3)]
    ftrip = [(0, 1, 2), (0, 2, 3), (0, 3, 1), (1, 2, 3)]
    for s, (i, j) in enumerate(epairs):
        a, b = sorted((i, j), key=lambda k: gids[k])
        L = np.linalg.norm(verts[:, a] - verts[:, b])
        w = lam[b][None] * G[a][:, None] - lam[a][None] * G[b][:, None]
        out[s] = L * lam[a][None] * w
        out[10 + s] = L * lam[b][None] * w
    for s, tri in enumerate(ftrip):
        a, b, c = sorted(tri, key=lambda k: gids[k])
        Lac = np.linalg.norm(verts[:, a] - verts[:, c])
        Lab = np.linalg.norm(verts[:, a] - verts[:, b])
        out[6 + s] = -Lac * lam[b][None] * (lam[c][None] * G[a][:, None] - lam[a][None] * G[c][:, None])
        out[16 + s] = Lab * lam[c][None] * (lam[b][None] * G[a][:, None] - lam[a][None] * G[b][:, None])
    return out


def ned2_tet_curl(verts: np.ndarray, gids: np.ndarray, pts: np.ndarray, h: float | None = None) -> np.ndarray:
    """(20, 3, P) curl of the basis by central differences (exact for quadratics)."""
    if h is None:
        h = 1e-3 * np.max(np.ptp(verts, axis=1))
    d = []
    for k in range(3):
        e = np.zeros((3, 1))
        e[k] = h
        d.append((ned2_tet_basis(verts, gids, pts + e) - ned2_tet_basis(verts, gids, pts - e)) / (2 * h))
    # d[k][:, c] = dN_c / dx_k
    cx = d[1][:, 2] - d[2][:, 1]
    cy = d[2][:, 0] - d[0][:, 2]
    cz = d[0][:, 1] - d[1][:, 0]
    return np.stack([cx, cy, cz], axis=1)


def brute_force_element(verts: np.ndarray, gids: np.ndarray, Ms: np.ndarray, Mm: np.ndarray, n: int = 6):
    """D_ij = int curl N_i . Ms curl N_j, F_ij = int N_i . Mm N_j by quadrature."""
    ref, w = tet_rule(n)
    T = np.stack([verts[:, 1] - verts[:, 0], verts[:, 2] - verts[:, 0], verts[:, 3] - verts[:, 0]], axis=1)
    pts = verts[:, :1] + T @ ref
    wj = w * abs(np.linalg.det(T))
    N = ned2_tet_basis(verts, gids, pts)
    C = ned2_tet_curl(verts, gids, pts)
    F = np.einsum("ikp,kl,jlp,p->ij", N, Mm, N, wj)
    D = np.einsum("ikp,kl,jlp,p->ij", C, Ms, C, wj)
    return D, F
