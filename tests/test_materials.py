"""Materials through the physics API (fem/material.py:20-106, fem/mesh3d.py:358-378,
fem/modeling/pmlbox.py:25-92): scalar, tensor, callable and PML tensors per tet."""
import numpy as np

from emerge_amd import meshgen
from emerge_amd.bc import Material, pml_material
from emerge_amd.physics import Electrodynamics3D


def _box():
    return meshgen.structured_box(4, 3, 2, size=(4.0, 3.0, 2.0), jitter=0.1, seed=1,
                                  tag_fn=lambda c: np.where(c[0] > 3.0, 2, 1))


def test_scalar_and_tensor_materials():
    assert np.allclose(Material(er=4.0, tand=0.01).ermat, 4.0 * (1 - 0.01j) * np.eye(3))
    T = np.array([[2.0, 0.1, 0], [0.1, 3.0, 0], [0, 0, 4.0]])
    m = Material(er=T, ur=np.diag([1.0, 2.0, 3.0]), tand=0.02)
    assert np.allclose(m.ermat, T * (1 - 0.02j))
    assert np.allclose(m.urmat, np.diag([1.0, 2.0, 3.0]))
    mesh = _box()
    phys = Electrodynamics3D(mesh)
    phys.set_material(2, m)
    er, ur = phys.material_tensors()
    ids = np.flatnonzero(mesh.tet_tags == 2)
    assert ids.size > 0
    assert np.allclose(er[:, :, ids], (T * (1 - 0.02j))[:, :, None])
    assert np.allclose(er[:, :, mesh.tet_tags == 1], np.eye(3)[:, :, None])


def test_callable_material_at_centroids():
    mesh = _box()
    f = lambda x, y, z: np.stack([np.stack([1 + x, 0 * x, 0 * x]), np.stack([0 * x, 2 + y, 0 * x]),
                                  np.stack([0 * x, 0 * x, 3 + z])]).astype(np.complex128)
    phys = Electrodynamics3D(mesh)
    phys.set_material([1, 2], Material(_fer=f))
    er, _ = phys.material_tensors()
    c = mesh.centers
    assert np.allclose(er[0, 0], 1 + c[0]) and np.allclose(er[1, 1], 2 + c[1]) and np.allclose(er[2, 2], 3 + c[2])
    assert not np.any(er[0, 1])


def test_pml_tensors_closed_form():
    """A layer on +x of a box centred at the origin with width 2: s_x = 1 - j ((x - 1) / t)^n delta,
    er = eps diag(1 / s_x, s_x, s_x), ur = diag(1 / s_x, s_x, s_x)."""
    base = Material(er=2.0, tand=0.01)
    t, n, dmax = 0.5, 2.0, 4.0
    pml = pml_material(base, (0.0, 0.0, 0.0), (2.0, 1.0, 1.0), (1, 0, 0), t, exponent=n, deltamax=dmax)
    x = np.array([1.1, 1.3, 1.5])
    sx = 1 - 1j * ((x - 1.0) / t) ** n * dmax
    er = pml.fer3d_mat(x, 0 * x, 0 * x)
    ur = pml.fur3d_mat(x, 0 * x, 0 * x)
    eps = 2.0 * (1 - 0.01j)
    assert np.allclose(er[0, 0], eps / sx) and np.allclose(er[1, 1], eps * sx) and np.allclose(er[2, 2], eps * sx)
    assert np.allclose(ur[0, 0], 1 / sx) and np.allclose(ur[1, 1], sx) and np.allclose(ur[2, 2], sx)
    assert not np.any(er[0, 1]) and not np.any(ur[1, 2])
