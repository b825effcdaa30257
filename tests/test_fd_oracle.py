"""Pins of the frequency-domain oracle: closed forms, invariants, quadrature."""
import numpy as np
import pytest

from emerge_amd import meshgen
from emerge_amd.bc import C0, Z_VAC, PEC, CoordinateSystem, LumpedPort, RectangularWaveguide
from oracle import fd
from oracle import oracle as O
from oracle import quadrature as Q

A_WG, B_WG = 22.86e-3, 10.16e-3


def test_port_closed_forms():
    cs = CoordinateSystem((1, 0, 0), (0, 1, 0), (0, 0, 1), (A_WG / 2, B_WG / 2, 0))
    wg = RectangularWaveguide([5], 1, cs=cs, dims=(A_WG, B_WG))
    k0 = 2 * np.pi * 10e9 / C0
    assert abs(wg.get_beta(k0) - 158.2382563) < 1e-6                   # sqrt(k0^2 - (pi/a)^2)
    assert wg.Zmode(k0) == Z_VAC                                         # reference: modetype 'TEM'
    assert np.isclose(wg.get_amplitude(k0), np.sqrt(4 * 377 / (A_WG * B_WG)))
    lp = LumpedPort([1], 1, width=2e-3, height=1e-3, direction=(0, 0, 1), cs=cs, Z0=50)
    assert np.isclose(lp.surfZ, 100) and np.isclose(lp.voltage, 10)
    assert np.isclose(lp.get_gamma(k0), 1j * k0 * 376.7303 / 100)
    assert np.allclose(lp.port_mode_3d(np.zeros(3), np.zeros(3), k0), np.array([[0] * 3, [0] * 3, [1] * 3]))


def _tri_basis(xs, ys, pts, Le):
    """2D NED2 basis of fem/mth/tri.py:670-687 at points (2,P): 8 x 2 x P."""
    det = (xs[1] - xs[0]) * (ys[2] - ys[0]) - (xs[2] - xs[0]) * (ys[1] - ys[0])
    b = np.array([ys[1] - ys[2], ys[2] - ys[0], ys[0] - ys[1]])
    c = np.array([xs[2] - xs[1], xs[0] - xs[2], xs[1] - xs[0]])
    a = np.array([xs[1] * ys[2] - ys[1] * xs[2], xs[2] * ys[0] - ys[2] * xs[0], xs[0] * ys[1] - ys[0] * xs[1]])
    lam = (a[:, None] + b[:, None] * pts[0] + c[:, None] * pts[1]) / det
    G = np.stack([b, c], 1) / det
    out = np.zeros((8, 2, pts.shape[1]))
    for k, (u, v) in enumerate([(0, 1), (1, 2), (0, 2)]):
        w = lam[v] * G[u][:, None] - lam[u] * G[v][:, None]
        out[k] = Le[k] * lam[u] * w
        out[4 + k] = Le[k] * lam[v] * w
    L02 = np.hypot(xs[2] - xs[0], ys[2] - ys[0])
    L01 = np.hypot(xs[1] - xs[0], ys[1] - ys[0])
    out[3] = -L02 * lam[1] * (lam[2] * G[0][:, None] - lam[0] * G[2][:, None])
    out[7] = L01 * lam[2] * (lam[1] * G[0][:, None] - lam[0] * G[1][:, None])
    return out


def test_tri_robin_vs_quadrature():
    rng = np.random.default_rng(4)
    dpts = fd.gaus_quad_tri4()
    for _ in range(3):
        xs, ys = rng.normal(size=3), rng.normal(size=3)
        gamma = 0.3 + 2.0j
        U = np.zeros((3, 6), dtype=complex)
        U[0], U[1] = 1.5 - 0.5j, -0.7 + 0.2j                        # constant field: the 6-point rule is exact
        Bo, bo = fd._tri_stiff_force(np.stack([xs, ys, np.zeros(3)]), gamma, U, dpts)
        ref, w = Q.tri_rule(8)
        T = np.array([[xs[1] - xs[0], xs[2] - xs[0]], [ys[1] - ys[0], ys[2] - ys[0]]])
        pts = np.array([[xs[0]], [ys[0]]]) + T @ ref
        wj = w * abs(np.linalg.det(T))
        Le = [np.hypot(xs[0] - xs[1], ys[0] - ys[1]), np.hypot(xs[1] - xs[2], ys[1] - ys[2]), np.hypot(xs[0] - xs[2], ys[0] - ys[2])]
        N = _tri_basis(xs, ys, pts, Le)
        Bq = gamma * np.einsum("ikp,jkp,p->ij", N, N, wj)
        assert np.abs(Bo - Bq).max() / np.abs(Bq).max() < 1e-12
        bq = np.einsum("ikp,k,p->i", N, U[:2, 0], wj)
        assert np.abs(bo + bq).max() / np.abs(bq).max() < 1e-12        # the reference's source is -int N.U


@pytest.fixture(scope="module")
def wg_small():
    length = 30e-3
    mesh = meshgen.waveguide_wr90(8, 4, 16, seed=0, length=length)
    cs1 = CoordinateSystem((1, 0, 0), (0, 1, 0), (0, 0, 1), (A_WG / 2, B_WG / 2, 0))
    cs2 = CoordinateSystem((-1, 0, 0), (0, 1, 0), (0, 0, -1), (A_WG / 2, B_WG / 2, length))
    p1 = RectangularWaveguide([meshgen.ZMIN], 1, cs=cs1, dims=(A_WG, B_WG))
    p2 = RectangularWaveguide([meshgen.ZMAX], 2, cs=cs2, dims=(A_WG, B_WG))
    pec = PEC([meshgen.XMIN, meshgen.XMAX, meshgen.YMIN, meshgen.YMAX])
    er, ur = O.material_tensors(np.ones(mesh.n_tets, complex))
    prob = fd.OracleProblem(mesh, er, ur, [pec, p1, p2], order_seed=9)
    return prob, length


def test_oracle_waveguide_closed_form(wg_small):
    """TE10 section: |S21| ~ 1, phase -beta L, S11 ~ 0, reciprocity (coarse mesh: 2% / 0.05 rad)."""
    prob, length = wg_small
    f = 10e9
    S, _ = prob.frequency_domain([f])
    S = S[0]
    k0 = 2 * np.pi * f / C0
    beta = np.sqrt(k0 ** 2 - (np.pi / A_WG) ** 2)
    assert abs(abs(S[1, 0]) - 1) < 0.02
    assert abs(np.angle(S[1, 0] * np.exp(1j * beta * length))) < 0.05
    assert abs(S[0, 0]) < 0.05 and abs(S[1, 1]) < 0.05
    assert abs(S[1, 0] - S[0, 1]) < 0.02


def test_oracle_numbering_invariance(wg_small):
    """S is independent of the arbitrary edge/face numbering (set order in the reference)."""
    prob, _ = wg_small
    S1, _ = prob.frequency_domain([9e9])
    other = fd.OracleProblem(prob.src, prob.er, prob.ur, prob.bcs, order_seed=123)
    S2, _ = other.frequency_domain([9e9])
    assert np.abs(S1 - S2).max() < 1e-9
