"""Mesh topology built on the GPU (emg_topology_build) against the host build of
Mesh3D (fem/mesh3d.py:224-355 conventions): every array equal, integer arrays bit
for bit and the edge lengths to the last bit, on the WR-90, posts and 1,005,000-tet
assembly meshes; invalid and non-manifold meshes refused as the host build refuses them."""
import numpy as np
import pytest

from emerge_amd import meshgen
from emerge_amd.mesh import Mesh3D

pytestmark = pytest.mark.gpu

ARRAYS = ("edges", "tris", "tets_canonical", "tet_ent", "tet_to_edge", "tet_to_edge_sign", "tet_to_tri",
          "tet_to_tri_sign", "tri_to_tet", "tri_to_edge", "tri_to_edge_sign", "edge_lengths", "areas",
          "edge_centers", "tri_centers", "centers")


@pytest.mark.parametrize("name", ["wr90", "posts", "assembly_box"])
def test_device_topology_equals_host(cuda, name):
    if name == "wr90":
        host = meshgen.waveguide_wr90()
    elif name == "posts":
        host = meshgen.posts_box(80, 40, 16, seed=2)
    else:
        host, _ = meshgen.assembly_box()
    print(f"{name}: {host.n_tets} tets", flush=True)
    dev = Mesh3D(host.nodes, host.tets, topology="device")
    dev = Mesh3D(host.nodes, host.tets, topology="device")          # second build: no first-call costs
    print(f"{name}: host {host.topology_seconds:.3f} s, device {dev.topology_seconds:.3f} s "
          f"(incl. upload and copy back)", flush=True)
    for k in ARRAYS:
        a, b = getattr(host, k), getattr(dev, k)
        assert a.shape == b.shape, (k, a.shape, b.shape)
        assert a.dtype == b.dtype or (a.dtype.kind == b.dtype.kind == "i"), (k, a.dtype, b.dtype)
        assert np.array_equal(a, b), (k, np.abs(a - b).max())
    # lookups go through the same sorted keys
    t = host.tets[:, :100]
    assert np.array_equal(host.get_tri(t[0], t[1], t[2]), dev.get_tri(t[0], t[1], t[2]))
    assert host.n_entities == dev.n_entities


def test_device_topology_refuses_bad_meshes(cuda):
    from emerge_amd.topology import device_topology
    nodes = np.array([[0, 1, 0, 0, 1], [0, 0, 1, 0, 1], [0, 0, 0, 1, 1]], dtype=np.float64)
    # a repeated vertex in a tet
    with pytest.raises(ValueError):
        device_topology(nodes, np.array([[0], [1], [1], [3]]))
    # an id beyond the node count (Mesh3D checks this before building)
    with pytest.raises(ValueError):
        device_topology(nodes, np.array([[0], [1], [2], [5]]))
    # three tets on one triangle (0, 1, 2)
    nodes6 = np.concatenate([nodes, np.array([[0.3], [0.3], [-1.0]])], axis=1)
    tets = np.array([[0, 0, 0], [1, 1, 1], [2, 2, 2], [3, 4, 5]])
    with pytest.raises(ValueError):
        Mesh3D(nodes6, tets, topology="host")
    with pytest.raises(ValueError):
        Mesh3D(nodes6, tets, topology="device")
