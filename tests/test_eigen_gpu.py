"""Shift-invert Arnoldi eigenmodes on the GPU vs ARPACK on the oracle's matrices."""
import numpy as np
import pytest

from emerge_amd import meshgen
from emerge_amd.bc import C0, Material
from emerge_amd.physics import Electrodynamics3D
from oracle import fd

pytestmark = pytest.mark.gpu


def _cavity(block=False):
    size = (30e-3, 20e-3, 15e-3)
    tag = None
    if block:   # eps_r = 6 block 10 x 10 x 5 mm, 3 mm from the corner
        tag = lambda c: np.where((c[0] > 3e-3) & (c[0] < 13e-3) & (c[1] > 3e-3) & (c[1] < 13e-3)
                                 & (c[2] > 3e-3) & (c[2] < 8e-3), 2, 1)
    mesh = meshgen.structured_box(6, 4, 3, size=size, jitter=0.05, seed=3, tag_fn=tag)
    phys = Electrodynamics3D(mesh)
    if block:
        phys.set_material(2, Material(er=6.0))
    phys._bc_initialized = True
    return mesh, phys, size


@pytest.mark.parametrize("block", [False, True])
def test_eigenmodes_match_arpack(cuda, block):
    print(f"eigen block={block}", flush=True)
    mesh, phys, size = _cavity(block)
    ftarget = 10e9
    data = phys.eigenmode(num_sols=4, target_frequency=ftarget)
    lam_g = np.array([ds.eigenvalue for ds in data.datasets])
    res = np.array([ds.residual for ds in data.datasets])
    print("inner iterations", data.info["inner_iterations"][:5], "residuals", res, flush=True)
    er, ur = phys.material_tensors()
    orc = fd.OracleProblem(mesh, er, ur, phys.boundary_conditions, order_seed=5)
    lam_o = fd.oracle_eigenmodes(orc, (2 * np.pi * ftarget / C0) ** 2, 4)
    lam_g = lam_g[np.argsort(lam_g.real)]
    lam_o = lam_o[np.argsort(lam_o.real)]
    print("gpu", lam_g, "oracle", lam_o, flush=True)
    assert np.all(np.abs(lam_g - lam_o) / np.abs(lam_o) < 1e-8)
    if not block:   # empty PEC box: k^2 = (m pi/a)^2 + (n pi/b)^2 + (p pi/d)^2 (coarse mesh: 2 %)
        a, b, d = size
        exact = []
        for m in range(4):
            for n in range(4):
                for p in range(4):
                    nz = (m > 0) + (n > 0) + (p > 0)
                    if nz >= 2:   # TE and TM are degenerate when all three indices are nonzero
                        exact += [(m * np.pi / a) ** 2 + (n * np.pi / b) ** 2 + (p * np.pi / d) ** 2] * (2 if nz == 3 else 1)
        near = sorted(exact, key=lambda v: abs(v - (2 * np.pi * ftarget / C0) ** 2))[:4]
        assert np.allclose(np.sort(lam_g.real), np.sort(near), rtol=0.02)


def test_restarted_arnoldi_finds_the_oracle_set(cuda):
    """Krylov-Schur restarts with a short basis (ncv = 2 nev + 4) return the same sorted
    set of eigenvalues as ARPACK (a skipped or spurious mode fails the set comparison)."""
    from emerge_amd.eigen import ShiftInvertArnoldi
    mesh, phys, size = _cavity(True)
    ftarget, nev = 12e9, 6
    sigma = (2 * np.pi * ftarget / C0) ** 2
    arn = ShiftInvertArnoldi(phys.problem(), sigma, ncv=2 * nev + 4)
    lam_g, X, res = arn.run(nev)
    print("restarts", arn.restarts, "residuals", res, flush=True)
    er, ur = phys.material_tensors()
    lam_o = fd.oracle_eigenmodes(fd.OracleProblem(mesh, er, ur, phys.boundary_conditions, order_seed=2), sigma, nev)
    lam_o = np.sort(lam_o.real)
    assert lam_g.shape[0] == nev and np.all(res <= 1e-8)
    assert np.all(np.abs(np.sort(lam_g.real) - lam_o) / lam_o < 1e-8), (lam_g, lam_o)


def test_eigen_refuses_unmet_residual(cuda):
    """A residual target the pairs cannot meet raises instead of returning them."""
    from emerge_amd.eigen import EigenError, ShiftInvertArnoldi
    mesh, phys, size = _cavity(False)
    arn = ShiftInvertArnoldi(phys.problem(), (2 * np.pi * 10e9 / C0) ** 2, ncv=12)
    with pytest.raises(EigenError):
        arn.run(4, resid_tol=1e-30, max_restarts=0)


def test_cavity_workload_modes_match_oracle_fixture(cuda):
    """Full cavity workload (221,184 tets, eps_r 6 block): the 10 modes nearest 5 GHz equal the
    oracle's (ARPACK shift-invert with PARDISO, tests/golden/oracle_fixtures.npz) as a sorted set;
    the fixture's inertia counts (no mode below 5 GHz, exactly 10 up to the 10th) make that set
    the complete one."""
    import os
    from emerge_amd import workloads
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "oracle_fixtures.npz"))
    lam_o = np.sort(gold["c4_lambda"].real)
    assert list(gold["c4_modes_below"]) == [0, 10]
    phys = workloads.cavity()
    print(f"cavity: {phys.mesh.n_tets} tets", flush=True)
    data = phys.eigenmode(num_sols=10, target_frequency=workloads.CAVITY_TARGET_HZ)
    lam_g = np.sort(np.array([ds.eigenvalue for ds in data.datasets]).real)
    res = np.array([ds.residual for ds in data.datasets])
    rel = np.abs(lam_g - lam_o) / lam_o
    print("gpu", lam_g, "max rel", rel.max(), "residuals", res.max(), flush=True)
    assert float(gold["c4_sigma"]) == pytest.approx((2 * np.pi * workloads.CAVITY_TARGET_HZ / C0) ** 2, rel=1e-15)
    assert lam_g.shape == lam_o.shape and np.all(rel < 1e-8)
