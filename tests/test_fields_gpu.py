"""Field reconstruction at arbitrary points (EMDataSet.interpolate, fem/physics/edm/emdata.py:181-199):
GPU point location, E and curl E against the oracle's brute-force search and interpolation
(fem/mth/tet.py:371-626)."""
import numpy as np
import pytest

from emerge_amd import meshgen
from emerge_amd.fields import PointLocator
from oracle import fd
from oracle import oracle as O
from tests.helpers import mesh_dof_keys

pytestmark = pytest.mark.gpu


def test_locate_and_interpolate_match_oracle(cuda):
    mesh = meshgen.structured_box(4, 3, 3, size=(1.0, 0.8, 0.6), jitter=0.1, seed=9)
    er, ur = O.material_tensors(np.ones(mesh.n_tets, complex))
    orc = fd.OracleProblem(mesh, er, ur, [], order_seed=4)
    rng = np.random.default_rng(2)
    xo = rng.standard_normal(orc.N) + 1j * rng.standard_normal(orc.N)
    # the same field in this library's numbering
    p = O.permutation_to(mesh_dof_keys(mesh), orc.field.dof_keys())
    x = xo[p]
    pts = rng.uniform([0, 0, 0], [1.0, 0.8, 0.6], size=(200, 3)).T
    pts = np.concatenate([pts, np.array([[1.5, 0.1, 0.1], [-0.2, 0.4, 0.3]]).T], axis=1)   # two outside
    loc = PointLocator(mesh)
    tet, bary = loc.locate(pts)
    tet = tet.cpu().numpy()
    assert np.all(tet[:200] >= 0) and np.all(tet[200:] == -1)
    c = rng.standard_normal(mesh.n_tets) + 1j * rng.standard_normal(mesh.n_tets)
    E, H = loc.fields(x, pts, curl_scale=c)
    Eo = orc.interpolate(xo, pts)
    Ho = orc.interpolate_curl(xo, pts, c=c)
    assert np.abs(E - Eo).max() <= 1e-12 * np.abs(Eo).max()
    assert np.abs(H - Ho).max() <= 1e-11 * np.abs(Ho).max()
    assert not np.any(E[:, 200:]) and not np.any(H[:, 200:])


def test_emdataset_interpolate_waveguide(cuda):
    """A solved WR-90 section: EMDataSet.interpolate gives E and H = curl E / (-j w mu0) at points;
    in the middle of the guide the TE10 wave has |Ey| dominant and H tangential to the walls."""
    from emerge_amd.physics import Electrodynamics3D
    from tests.test_freq_gpu import waveguide
    mesh, ports = waveguide()
    phys = Electrodynamics3D(mesh)
    phys.assign(*ports)
    phys.set_frequency([10e9])
    phys.solveroutine.tol = 1e-10
    data = phys.frequency_domain()
    ds = data.item(0)
    xs = np.linspace(2e-3, 20e-3, 7)
    ys = np.full_like(xs, 5.08e-3)
    zs = np.full_like(xs, 15e-3)
    ds.interpolate(xs, ys, zs)
    assert ds.Ex.shape == xs.shape
    assert np.all(np.abs(ds.Ey) > 10 * np.abs(ds.Ex)) and np.all(np.abs(ds.Ey) > 10 * np.abs(ds.Ez))
    # against the oracle on the same solution vector
    er, ur = phys.material_tensors()
    orc = fd.OracleProblem(mesh, er, ur, phys.boundary_conditions, order_seed=6)
    p = O.permutation_to(orc.field.dof_keys(), mesh_dof_keys(mesh))
    xo = np.asarray(ds._field)[p]
    pts = np.stack([xs, ys, zs])
    const = 1.0 / (-1j * 2 * np.pi * 10e9 * 4 * np.pi * 1e-7)
    Eo = orc.interpolate(xo, pts)
    Ho = orc.interpolate_curl(xo, pts) * const
    assert np.abs(np.stack(ds.E) - Eo).max() <= 1e-11 * np.abs(Eo).max()
    assert np.abs(np.stack(ds.H) - Ho).max() <= 1e-10 * np.abs(Ho).max()
