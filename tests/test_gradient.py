"""Discrete P2 gradient G used by the hybrid smoother: pinned by the curl-curl null space and by
brute-force least squares on the oracle's own basis functions."""
import numpy as np
import pytest

from emerge_amd import meshgen
from emerge_amd.gradient import gradient_matrix
from oracle import oracle as O
from oracle.quadrature import barycentric, grad_barycentric, ned2_tet_basis
from tests.helpers import mesh_dof_keys, oracle_in_mesh_order


@pytest.fixture(scope="module")
def box():
    mesh = meshgen.structured_box(3, 2, 2, size=(3.0, 2.0, 1.5), jitter=0.12, seed=5)
    om = O.OracleMesh(mesh.nodes, mesh.tets, order_seed=1)
    field = O.OracleNedelec2(om)
    er, ur = O.material_tensors(np.full(mesh.n_tets, 2.0 - 0.1j))
    Eo, Bo = O.tet_mass_stiffness_matrices(field, er, ur)
    E, B = oracle_in_mesh_order(mesh, Eo, Bo, field)
    return mesh, E, B


def test_gradient_is_curl_free(box):
    """E G = 0: every column is the gradient of a P2 function (tet.py curl-curl null space)."""
    mesh, E, B = box
    G, nodes = gradient_matrix(mesh)
    assert G.shape == (2 * mesh.n_entities, mesh.n_nodes + mesh.n_edges)
    EG = (E @ G).toarray()
    BG = (B @ G).toarray()
    assert np.abs(EG).max() < 1e-13 * np.abs(BG).max() * np.abs(E).max() / np.abs(B).max()
    # and G has full column rank up to the constant: P2 has nv + ne functions, one constant
    s = np.linalg.svd(G.toarray(), compute_uv=False)
    assert np.sum(s > 1e-10 * s[0]) == G.shape[1] - 1


def test_gradient_matches_least_squares_on_oracle_basis(box):
    """grad(phi) = sum_i G_i,phi N_i at random points of every tet, N from the oracle's basis."""
    mesh, _, _ = box
    G, _ = gradient_matrix(mesh)
    Gd = G.toarray()
    om = O.OracleMesh(mesh.nodes, mesh.tets, order_seed=1)
    field = O.OracleNedelec2(om)
    perm = O.permutation_to(field.dof_keys(), mesh_dof_keys(mesh))    # mesh DOF of each oracle DOF
    rng = np.random.default_rng(3)
    nv = mesh.n_nodes
    for t in range(mesh.n_tets):
        vid = om.tets[:, t]
        v = om.nodes[:, vid]
        pts = v @ rng.dirichlet(np.ones(4), 12).T
        Nb = ned2_tet_basis(v, vid, pts)                                 # (20, 3, P) oracle slots
        dofs = perm[field.tet_to_field[:, t]]
        lam = barycentric(v, pts)
        Gl = grad_barycentric(v)
        for a in range(4):                                               # vertex functions
            grad = (4 * lam[a] - 1)[None] * Gl[a][:, None]
            rec = np.einsum("i,icp->cp", Gd[dofs, vid[a]], Nb)
            assert np.abs(rec - grad).max() < 1e-10 * np.abs(grad).max()
        for s_, (a, b) in enumerate([(0, 1), (0, 2), (0, 3), (1, 2), (3, 1), (2, 3)]):
            e = mesh.get_edge(min(vid[a], vid[b]), max(vid[a], vid[b]))
            grad = 4 * (lam[b][None] * Gl[a][:, None] + lam[a][None] * Gl[b][:, None])
            rec = np.einsum("i,icp->cp", Gd[dofs, nv + int(e)], Nb)
            assert np.abs(rec - grad).max() < 1e-10 * np.abs(grad).max()


def test_masked_columns_dropped(box):
    mesh, _, _ = box
    rng = np.random.default_rng(0)
    mask = np.ones(2 * mesh.n_entities, dtype=np.uint8)
    mask[rng.choice(mask.shape[0], 20, replace=False)] = 0
    Gf, _ = gradient_matrix(mesh)
    G, nodes = gradient_matrix(mesh, mask)
    assert np.all(np.asarray(abs(G[mask == 0]).sum(1)).ravel() == 0)
    kept = Gf[:, nodes]
    assert abs(kept - G).max() == 0
    dropped = np.setdiff1d(np.arange(Gf.shape[1]), nodes)
    assert np.all(np.asarray(abs(Gf[mask == 0][:, dropped]).sum(0)).ravel() > 0)
