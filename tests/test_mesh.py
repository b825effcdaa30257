"""Topology and DOF numbering follow the reference's conventions
(fem/mesh3d.py:224-355, fem/elements/nedelec2.py:32-70)."""
import numpy as np
import pytest

from emerge_amd import meshgen
from emerge_amd.mesh import CANON_EDGES, CANON_FACES
from oracle import oracle as O


@pytest.fixture(scope="module")
def small():
    return meshgen.structured_box(3, 2, 2, size=(1.0, 0.8, 0.6), jitter=0.1, seed=4)


def test_topology_matches_oracle(small):
    om = O.OracleMesh(small.nodes, small.tets)
    for name in ["edges", "tris", "tet_to_edge", "tet_to_tri", "tri_to_edge", "tri_to_tet"]:
        assert np.array_equal(getattr(small, name), getattr(om, name)), name
    assert np.allclose(small.edge_lengths, om.edge_lengths, rtol=0, atol=0)
    assert np.allclose(small.areas, om.areas)


def test_canonical_entity_table(small):
    c = small.tets_canonical
    for s, (a, b) in enumerate(CANON_EDGES):
        e = small.tet_ent[:, s]
        assert np.array_equal(small.edges[0, e], c[:, a]) and np.array_equal(small.edges[1, e], c[:, b])
    ne = small.n_edges
    for s, (a, b, cc) in enumerate(CANON_FACES):
        f = small.tet_ent[:, 6 + s] - ne
        assert np.array_equal(small.tris[:, f], c[:, [a, b, cc]].T)


def test_euler_and_manifold(small):
    V, E, F, T = small.n_nodes, small.n_edges, small.n_tris, small.n_tets
    assert V - E + F - T == 1            # a ball
    inner = small.tri_to_tet[1] >= 0
    assert np.all(small.tri_to_tet[0] >= 0)
    nb = (~inner).sum()
    assert nb == sum(v.shape[0] for v in small.face_tags.values())   # every boundary face is tagged


def test_positive_orientation(small):
    v = small.nodes[:, small.tets]
    det = np.einsum("in,in->n", v[:, 1] - v[:, 0], np.cross((v[:, 2] - v[:, 0]).T, (v[:, 3] - v[:, 0]).T).T)
    assert np.all(det > 0)


def test_assembly_box_sizes():
    """Sizes of the 50x50x67 Kuhn assembly workload (1,005,000 tets)."""
    mesh, eps = meshgen.assembly_box()
    assert mesh.n_tets == 1_005_000 and mesh.n_nodes == 176_868
    assert mesh.n_edges == 1_200_267 and mesh.n_tris == 2_028_400
    assert 2 * mesh.n_entities == 6_457_334
    assert eps.shape == (1_005_000,) and np.all(eps.real >= 1) and np.all(eps.real <= 10)
