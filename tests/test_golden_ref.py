"""The oracle's element routines against golden vectors written by the
reference's own code (tests/golden/write_ref_vectors.py): 20 random elements
each of ned2_tet_stiff_mass (fem/mth/tet.py:854-1088), ned2_tri_stiff_force
(fem/mth/tri.py:575-695) and ned2_tri_stiff (fem/mth/tri.py:697-796), the
last with the reference's non-orthonormal per-triangle frame."""
import os

import numpy as np
import pytest

from oracle import fd
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ned2_ref.npz")
TOL = 1e-13


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_tet_stiff_mass_matches_reference(gold):
    for i in range(gold["tet_D"].shape[0]):
        D, F = O.element_matrices(gold["tet_verts"][i], gold["tet_elen"][i], gold["tet_lem"][i], gold["tet_ltm"][i],
                                  gold["tet_Ms"][i], gold["tet_Mm"][i])
        assert _rel(D, gold["tet_D"][i]) <= TOL, (i, _rel(D, gold["tet_D"][i]))
        assert _rel(F, gold["tet_F"][i]) <= TOL, (i, _rel(F, gold["tet_F"][i]))


def test_tri_stiff_force_matches_reference(gold):
    dpts = gold["trif_dpts"]
    assert np.abs(dpts - fd.gaus_quad_tri4()).max() < 1e-15      # the oracle's Dunavant table is the reference's
    for i in range(gold["trif_B"].shape[0]):
        B, b = fd._tri_stiff_force(gold["trif_lcs"][i], gold["trif_gamma"][i], gold["trif_uinc"][i], dpts)
        assert _rel(B, gold["trif_B"][i]) <= TOL, (i, _rel(B, gold["trif_B"][i]))
        assert _rel(b, gold["trif_b"][i]) <= TOL, (i, _rel(b, gold["trif_b"][i]))


def test_tri_stiff_abc_frame_matches_reference(gold):
    for i in range(gold["tris_B"].shape[0]):
        B = fd._tri_stiff(gold["tris_verts"][i], gold["tris_elen"][i], gold["tris_gamma"][i])
        assert _rel(B, gold["tris_B"][i]) <= TOL, (i, _rel(B, gold["tris_B"][i]))


def test_planted_frame_mistake_is_caught(gold):
    """An orthonormal frame (the 'obvious' reading of tri.py:708-721) differs from the reference's."""
    i = 1
    v = gold["tris_verts"][i]
    ax1 = (v[:, 1] - v[:, 0]) / np.linalg.norm(v[:, 1] - v[:, 0])
    ax2 = (v[:, 2] - v[:, 0]) / np.linalg.norm(v[:, 2] - v[:, 0])
    n = np.cross(ax1, ax2)
    n /= np.linalg.norm(n)
    e2 = np.cross(n, ax1)
    loc = np.stack([ax1 @ v, e2 @ v, np.zeros(3)])
    # the force variant on orthonormal local coordinates gives the exact 2D block
    B_ortho, _ = fd._tri_stiff_force(loc, gold["tris_gamma"][i], np.zeros((3, 6), complex), fd.gaus_quad_tri4())
    assert _rel(B_ortho, gold["tris_B"][i]) > 1e-6


def test_port_element_matrices_match_reference(gold):
    """oracle.modal.generalized_matrix against generalized_matrix_GQ (nedeleclegrange2.py:206-374),
    including its 8-digit quadrature table and clockwise triangles."""
    from oracle import modal as OM
    for i in range(gold["gq_A"].shape[0]):
        A, B = OM.generalized_matrix(gold["gq_verts"][i], gold["gq_Ms"][i], gold["gq_Mm"][i], float(gold["gq_k0"][i]))
        assert _rel(A, gold["gq_A"][i]) <= 1e-12, (i, _rel(A, gold["gq_A"][i]))
        assert _rel(B, gold["gq_B"][i]) <= 1e-12, (i, _rel(B, gold["gq_B"][i]))


def test_port_field_interpolation_matches_reference(gold):
    """oracle.modal.interp_full / interp_curl against ned2_tri_interp_full / _curl (tri.py:228-401)."""
    from oracle import modal as OM
    args = (gold["it_pts"], gold["it_field"], gold["it_tris"], gold["it_nodes"], gold["it_t2f"])
    E = OM.interp_full(*args)
    H = OM.interp_curl(*args, gold["it_diadic"], float(gold["it_beta"]))
    assert _rel(E, gold["it_E"]) <= 1e-12, _rel(E, gold["it_E"])
    assert _rel(H, gold["it_H"]) <= 1e-12, _rel(H, gold["it_H"])
