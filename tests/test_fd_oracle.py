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


# ---------------------------------------------------------------- port physics of the oracle (oracle/ports.py)
from oracle import ports as OP  # noqa: E402


def test_oracle_port_constants():
    """Literal values of the reference's port formulas (fem/bc.py:549-556, 571-575, 684-713)."""
    fr = OP.Frame((1, 0, 0), (0, 1, 0), (0, 0, 1), (A_WG / 2, B_WG / 2, 0))
    wg = OP.OraclePort("waveguide", [5], 1, fr, dims=(A_WG, B_WG))
    k0 = 2 * np.pi * 10e9 / C0
    assert abs(wg.beta(k0) - 158.2382563) < 1e-6
    assert wg.gamma(k0) == 1j * wg.beta(k0)
    assert wg.zmode(k0) == 376.730313412
    e = wg.mode_local(np.array([0.0, A_WG / 2]), np.zeros(2))
    assert np.isclose(e[1, 0], np.sqrt(4 * 377 / (A_WG * B_WG))) and abs(e[1, 1]) < 1e-12 * e[1, 0]
    lp = OP.OraclePort("lumped", [1], 1, fr, width=2e-3, height=1e-3, direction=(0, 0, 1), Z0=50)
    assert np.isclose(lp.voltage(), 10.0)                                     # sqrt(2 P Z0)
    assert np.isclose(lp.gamma(k0), 1j * k0 * 376.7303 / 100.0)               # sheet impedance 50 * 2 / 1
    assert np.allclose(lp.uinc_local(np.zeros(2), np.zeros(2), k0)[2], -2j * k0 * 10 / 1e-3 * 376.7303 / 100)


def _uniform_field_dofs(prob, E):
    """DOFs of a uniform field by per-tet least squares on the oracle's basis
    (fem/mth/tet.py:451-492); NED2 holds constants exactly."""
    f, om = prob.field, prob.om
    x = np.zeros(prob.N, dtype=np.complex128)
    cnt = np.zeros(prob.N)
    seen = {}
    rng = np.random.default_rng(0)
    for t in range(om.n_tets):
        v = om.nodes[:, om.tets[:, t]]
        pts = v @ rng.dirichlet(np.ones(4), size=12).T
        dofs = f.tet_to_field[:, t]
        Phi = np.zeros((36, 20))
        for k, d in enumerate(dofs):
            e = np.zeros(prob.N)
            e[d] = 1.0
            Phi[:, k] = prob.interpolate(e, pts, [t]).real.ravel()
        rhs = np.tile(np.asarray(E, dtype=np.float64)[:, None], (1, 12)).ravel()
        c = np.linalg.lstsq(Phi, rhs, rcond=None)[0]
        assert np.abs(Phi @ c - rhs).max() < 1e-9 * np.abs(rhs).max()      # the constant is in the space
        for d, cd in zip(dofs, c):
            if d in seen:
                assert abs(seen[d] - cd) <= 1e-8 * max(1.0, abs(cd))          # shared DOFs agree between tets
            seen[d] = cd
        x[dofs] = c
        cnt[dofs] += 1
    return x


def test_oracle_lumped_voltage_of_uniform_field():
    """V = int E.dl over the port's voltage line (emfreq3d.py:366-389, 753-768) of a uniform
    field E0 z equals E0 h; the S-data field term is |V| / sqrt(2 Z0)."""
    size = (6e-3, 4e-3, 3e-3)
    mesh = meshgen.structured_box(3, 2, 2, size=size, jitter=0.1, seed=5)
    cs = CoordinateSystem((0, 1, 0), (0, 0, 1), (1, 0, 0), (0.0, size[1] / 2, size[2] / 2))
    lp = LumpedPort([meshgen.XMIN], 1, width=size[1], height=size[2], direction=(0, 0, 1), cs=cs, Z0=50)
    er, ur = O.material_tensors(np.ones(mesh.n_tets, complex))
    prob = fd.OracleProblem(mesh, er, ur, [lp], order_seed=4)
    E0 = 3.7
    x = _uniform_field_dofs(prob, [0.0, 0.0, E0])
    port = prob.bcs[0]
    mid, dl = port.line
    assert np.allclose(dl.sum(axis=1), [0, 0, size[2]]) and np.allclose(mid[:2], [[0], [size[1] / 2]])
    fp, mp = prob.s_data(port, x, 2 * np.pi * 1e9 / C0)
    assert abs(abs(fp) - E0 * size[2] / np.sqrt(2 * 50)) <= 1e-12 * E0 * size[2]
    assert mp == 0


def test_oracle_waveguide_mode_power():
    """sparam_mode_power (fem/mth/sparam.py:72-139) of the TE10 mode over the port face:
    amp^2 a b / 4 / Z_vac = 377 / Z_vac for unit power (4th-order rule on a fine face mesh)."""
    length = 10e-3
    mesh = meshgen.waveguide_wr90(16, 8, 4, seed=0, length=length)
    cs1 = CoordinateSystem((1, 0, 0), (0, 1, 0), (0, 0, 1), (A_WG / 2, B_WG / 2, 0))
    p1 = RectangularWaveguide([meshgen.ZMIN], 1, cs=cs1, dims=(A_WG, B_WG))
    er, ur = O.material_tensors(np.ones(mesh.n_tets, complex))
    prob = fd.OracleProblem(mesh, er, ur, [p1])
    _, mp = prob.s_data(prob.bcs[0], np.zeros(prob.N, dtype=np.complex128), 2 * np.pi * 10e9 / C0)
    assert abs(mp - 377 / 376.730313412) < 2e-5


def test_oracle_curl_interpolation_by_finite_differences():
    """The oracle's analytic curl (tet.py:499-626) equals central differences of its own
    interpolated E (tet.py:371-497) inside a tet, for a random field."""
    mesh = meshgen.structured_box(2, 2, 2, size=(1.0, 0.8, 0.6), jitter=0.1, seed=3)
    er, ur = O.material_tensors(np.ones(mesh.n_tets, complex))
    prob = fd.OracleProblem(mesh, er, ur, [], order_seed=5)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(prob.N) + 1j * rng.standard_normal(prob.N)
    t = 7
    v = prob.om.nodes[:, prob.om.tets[:, t]]
    p0 = v @ np.array([0.3, 0.25, 0.2, 0.25])
    h = 1e-5
    J = np.zeros((3, 3), dtype=complex)                       # J[i, j] = dE_i / dx_j
    for j in range(3):
        d = np.zeros(3)
        d[j] = h
        pts = np.stack([p0 + d, p0 - d], axis=1)
        E = prob.interpolate(x, pts, [t])
        J[:, j] = (E[:, 0] - E[:, 1]) / (2 * h)
    fdcurl = np.array([J[2, 1] - J[1, 2], J[0, 2] - J[2, 0], J[1, 0] - J[0, 1]])
    an = prob.interpolate_curl(x, p0[:, None], [t])[:, 0]
    assert np.abs(an - fdcurl).max() <= 1e-6 * np.abs(an).max()


def test_pardiso_inertia_counts_eigenvalues_below_the_shift():
    """oracle.pardiso.inertia (the completeness check of the cavity fixture) against dense
    generalized eigenvalues: the negative count of A - tau B is the number below tau."""
    import scipy.linalg as sla
    import scipy.sparse as sp
    from oracle import pardiso
    if not pardiso.available():
        pytest.skip("MKL PARDISO is not available")
    rng = np.random.default_rng(5)
    n = 60
    X = rng.standard_normal((n, n))
    A = X @ X.T
    A[:, :8] = 0.0
    A[:8, :] = 0.0                                  # an 8-dimensional null space, as the curl-curl has
    Y = rng.standard_normal((n, n))
    B = Y @ Y.T + n * np.eye(n)
    lam = sla.eigh(A, B, eigvals_only=True)
    for tau in (0.5 * lam[8] if lam[8] > 0 else 1e-3, 0.5 * (lam[20] + lam[21]), lam[-1] * 1.01):
        npos, nneg = pardiso.inertia(sp.csr_matrix(A - tau * B), threads=1)
        assert nneg == int(np.sum(lam < tau)) and npos + nneg == n, (tau, npos, nneg)
