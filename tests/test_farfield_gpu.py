"""GPU Stratton-Chu (emg_farfield) against the oracle restatement of sc.py:27-180."""
import numpy as np
import pytest

from emerge_amd import meshgen
from emerge_amd.farfield import stratton_chu
from oracle import farfield as OF
from tests.test_farfield import Z0, dipole_fields

pytestmark = pytest.mark.gpu


def test_farfield_matches_oracle(cuda):
    a = 0.6
    mesh = meshgen.structured_box(10, 10, 10, size=(a, a, a), origin=(-a / 2, -a / 2, -a / 2), jitter=0.05, seed=2)
    tags = sorted(mesh.face_tags)
    surf = mesh.boundary_surface(tags, (0.0, 0.0, 0.0))
    rng = np.random.default_rng(3)
    E = rng.standard_normal((3, surf.n_edges)) + 1j * rng.standard_normal((3, surf.n_edges))
    H = (rng.standard_normal((3, surf.n_edges)) + 1j * rng.standard_normal((3, surf.n_edges))) / Z0
    E[:, :5] *= 1e-6                                               # below the 1e-3 level: skipped by both
    k = 2 * np.pi / 0.8
    th = np.linspace(-np.pi / 2, 1.5 * np.pi, 41)
    ph = np.linspace(0, 1.0, 41)
    Eg, Hg = stratton_chu(E, H, surf, th, ph, k)
    c = surf.edge_centers
    Eo, Ho = OF.stratton_chu(E, H, c, surf.weighted_normals(), th, ph, k)
    assert np.abs(Eg - Eo).max() <= 1e-12 * np.abs(Eo).max()
    assert np.abs(Hg - Ho).max() <= 1e-12 * np.abs(Ho).max()


def test_dipole_far_field_on_gpu(cuda):
    a, k = 0.6, 2 * np.pi
    mesh = meshgen.structured_box(14, 14, 14, size=(a, a, a), origin=(-a / 2, -a / 2, -a / 2))
    surf = mesh.boundary_surface(sorted(mesh.face_tags), (0.0, 0.0, 0.0))
    E, H = dipole_fields(surf.edge_centers, k)
    el = np.linspace(-1.2, 1.2, 9)
    ph = np.full_like(el, 0.7)
    Eg, _ = stratton_chu(E, H, surf, el, ph, k)
    tp = np.pi / 2 - el
    that = np.stack([np.cos(tp) * np.cos(ph), np.cos(tp) * np.sin(ph), -np.sin(tp)])
    ref = 1j * k * Z0 / (4 * np.pi) * np.sin(tp)
    assert np.abs(np.sum(Eg * that, axis=0) - ref).max() <= 1e-4 * np.abs(ref).max()
