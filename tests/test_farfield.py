"""Stratton-Chu far field: the oracle (oracle/farfield.py, sc.py:27-180) pinned by the closed
form of a Hertzian dipole, the surface sampling of the product against the oracle's."""
import numpy as np

from emerge_amd import meshgen
from oracle import farfield as OF

Z0 = 376.73031366857


def dipole_fields(pts, k, Il=1.0):
    """Exact E, H of a z-directed Hertzian dipole at the origin (exp(+j w t), outgoing exp(-j k r))."""
    x, y, z = pts
    r = np.sqrt(x * x + y * y + z * z)
    ct, st = z / r, np.sqrt(x * x + y * y) / r
    cp, sp = x / np.hypot(x, y), y / np.hypot(x, y)
    e = np.exp(-1j * k * r)
    Er = Z0 * Il / (2 * np.pi * r ** 2) * (1 + 1 / (1j * k * r)) * ct * e
    Et = 1j * k * Z0 * Il / (4 * np.pi * r) * (1 + 1 / (1j * k * r) - 1 / (k * r) ** 2) * st * e
    Hp = 1j * k * Il / (4 * np.pi * r) * (1 + 1 / (1j * k * r)) * st * e
    rhat = np.stack([st * cp, st * sp, ct])
    that = np.stack([ct * cp, ct * sp, -st])
    phat = np.stack([-sp, cp, 0 * cp])
    return Er * rhat + Et * that, Hp * phat


def _box_surface(n=14, a=0.6):
    mesh = meshgen.structured_box(n, n, n, size=(a, a, a), origin=(-a / 2, -a / 2, -a / 2))
    tags = sorted(mesh.face_tags)
    tri = mesh.get_triangles(tags)
    return mesh, tags, tri


def test_oracle_dipole_far_field():
    """|E_theta| = k Z0 I l |cos(el)| / (4 pi) in the elevation convention of sc.py (theta = elevation),
    E along the elevation unit vector only; the magnetic-current and electric-current halves add."""
    k = 2 * np.pi / 1.0
    mesh, tags, tri = _box_surface()
    c, w = OF.surface(mesh.nodes, mesh.tris[:, tri], (0.0, 0.0, 0.0))
    E, H = dipole_fields(c, k)
    el = np.linspace(-1.2, 1.2, 9)
    ph = np.full_like(el, 0.7)
    Eo, Ho = OF.stratton_chu(E, H, c, w, el, ph, k)
    # far-field amplitude F (E ~ F exp(-j k r) / r): F = j k Z0 I l sin(polar) along the polar unit vector
    tp = np.pi / 2 - el
    that = np.stack([np.cos(tp) * np.cos(ph), np.cos(tp) * np.sin(ph), -np.sin(tp)])
    ref = 1j * k * Z0 / (4 * np.pi) * np.sin(tp)
    assert np.abs(np.sum(Eo * that, axis=0) - ref).max() <= 1e-4 * np.abs(ref).max()
    rhat = np.stack([np.cos(el) * np.cos(ph), np.cos(el) * np.sin(ph), np.sin(el)])
    assert np.abs(np.sum(Eo * rhat, axis=0)).max() <= 1e-4 * np.abs(ref).max()     # transverse
    assert np.allclose(Ho, np.cross(rhat.T, Eo.T).T / Z0)


def test_surface_sampling_matches_oracle():
    mesh, tags, tri = _box_surface(n=4)
    surf = mesh.boundary_surface(tags, (0.0, 0.0, 0.0))
    c, w = OF.surface(mesh.nodes, mesh.tris[:, tri], (0.0, 0.0, 0.0))
    key = lambda P: np.lexsort(np.round(P, 12)[::-1])
    ours = surf.edge_centers[:, key(surf.edge_centers)]
    theirs = c[:, key(c)]
    assert np.allclose(ours, theirs, atol=1e-15)
    assert np.allclose(surf.weighted_normals()[:, key(surf.edge_centers)], w[:, key(c)], atol=1e-15)
    assert np.isclose(surf.areas.sum(), 6 * 0.6 ** 2)
