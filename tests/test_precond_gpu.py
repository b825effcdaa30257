"""Hybrid smoother (block Jacobi + gradient correction) on the GPU."""
import numpy as np
import pytest

from emerge_amd import meshgen
from emerge_amd.bc import CoordinateSystem, LumpedPort
from emerge_amd.physics import Electrodynamics3D

pytestmark = pytest.mark.gpu


def _posts(nx=20, ny=10, nz=4):
    mesh = meshgen.posts_box(nx, ny, nz)
    H = 3.75e-3
    cs1 = CoordinateSystem((0, 1, 0), (0, 0, 1), (1, 0, 0), (0.0, 10e-3, H / 2))
    cs2 = CoordinateSystem((0, -1, 0), (0, 0, 1), (-1, 0, 0), (40e-3, 10e-3, H / 2))
    phys = Electrodynamics3D(mesh)
    phys.assign(LumpedPort([meshgen.PORT1_TAG], 1, width=2e-3, height=H, direction=(0, 0, 1), cs=cs1, Z0=50),
                LumpedPort([meshgen.PORT2_TAG], 2, width=2e-3, height=H, direction=(0, 0, 1), cs=cs2, Z0=50))
    return mesh, phys


def test_gradient_in_gpu_null_space(cuda):
    """The GPU-assembled curl-curl matrix annihilates every column of G."""
    mesh, phys = _posts()
    prob = phys.problem()
    G = prob.gradient().host
    E = prob.to_scipy(prob.E).tocsr()
    B = prob.to_scipy(prob.B).tocsr()
    EG, BG = abs(E @ G).max(), abs(B @ G).max()
    assert EG < 1e-12 * BG * abs(E).max() / abs(B).max()
    # kept P2 nodes touch no PEC DOF
    mask = prob.mask.cpu().numpy().astype(bool)
    assert abs(G[~mask]).sum() == 0


@pytest.mark.parametrize("precision", [64, 32])
def test_hybrid_cuts_iterations_same_solution(cuda, precision):
    """Low-frequency posts box: the gradient correction reduces GMRES iterations several-fold
    and both preconditioners reach the same FP64-certified solution."""
    print(f"hybrid vs jacobi precision={precision}", flush=True)
    mesh, phys = _posts()
    phys.set_frequency([2e9])
    out = {}
    for pc in ("jacobi", "hybrid"):
        phys.solveroutine.precision = precision
        phys.solveroutine.preconditioner = pc
        phys.solveroutine.poly_degree = 1
        phys.solveroutine.sweep = "plain"
        phys.solveroutine.tol = 1e-9
        data = phys.frequency_domain()
        recs = data.item(0).solves
        for r in recs:
            assert r.converged and r.residual <= 1e-9, (pc, r)
        out[pc] = (sum(r.iterations for r in recs), data.S_array()[0])
        print(pc, "iterations", out[pc][0], flush=True)
    assert out["hybrid"][0] * 3 < out["jacobi"][0]
    assert np.abs(out["hybrid"][1] - out["jacobi"][1]).max() < 1e-6


@pytest.mark.parametrize("precision", [64, 32])
def test_polynomial_cuts_steps_same_solution(cuda, precision):
    """Degree-6 GMRES polynomial around the hybrid smoother: several times fewer GMRES steps
    (each applying K and the smoother 6 times), the same FP64-certified S-parameters."""
    print(f"polynomial vs plain hybrid precision={precision}", flush=True)
    mesh, phys = _posts()
    phys.set_frequency([3e9])
    out = {}
    for deg in (1, 6):
        phys.solveroutine.precision = precision
        phys.solveroutine.sweep = "plain"
        phys.solveroutine.tol = 1e-9
        phys.solveroutine.poly_degree = deg
        data = phys.frequency_domain()
        recs = data.item(0).solves
        for r in recs:
            assert r.converged and r.residual <= 1e-9, (deg, r)
        out[deg] = (sum(r.iterations for r in recs), data.S_array()[0])
        print("degree", deg, "GMRES steps", out[deg][0], flush=True)
    assert out[6][0] * 2 < out[1][0]
    assert np.abs(out[6][1] - out[1][1]).max() < 1e-6


def test_blocked_spmv_matches_csr_spmv(cuda):
    """Entity-blocked SpMV (the solver's) against the scalar CSR SpMV, with the PEC mask,
    on a posts-box system matrix and random complex vectors."""
    import ctypes
    import torch
    from emerge_amd import _lib
    mesh, phys = _posts(10, 5, 2)
    prob = phys.problem()
    K = prob.system(3e9)
    L = _lib.load()
    g = torch.Generator(device=prob.dev)
    g.manual_seed(5)
    x = torch.view_as_complex(torch.randn((prob.N, 2), generator=g, dtype=torch.float64, device=prob.dev)).contiguous()
    y1 = torch.empty_like(x)
    y2 = torch.empty_like(x)
    for mask in (prob.mask, None):
        mp = _lib.ptr(mask) if mask is not None else None
        _lib.check(L.emg_spmv(prob.N, _lib.ptr(prob.pat.rowptr), _lib.ptr(prob.pat.cols), _lib.ptr(K), mp,
                              _lib.ptr(x), _lib.ptr(y1), _lib.stream_ptr()), "emg_spmv")
        _lib.check(L.emg_spmv_blocked(prob.N, prob.nnz, _lib.ptr(prob.pat.rowptr), _lib.ptr(prob.pat.cols),
                                      _lib.ptr(K), mp, _lib.ptr(x), _lib.ptr(y2), _lib.stream_ptr()),
                   "emg_spmv_blocked")
        torch.cuda.synchronize()
        err = float((y1 - y2).abs().max() / y1.abs().max())
        print("blocked vs csr", "masked" if mask is not None else "plain", err, flush=True)
        assert err < 1e-14


@pytest.mark.parametrize("precision", [64, 32])
def test_spmv_autotuner_picks_the_fastest_shape(cuda, precision):
    """The solver's SpMV autotuner: nothing tuned before the first matrix; then every
    candidate (4, 8, 16, 32 lanes per entity) has a measured time, the kept shape is the
    fastest measured, and the tuned product equals the scalar CSR SpMV."""
    import torch
    from emerge_amd import _lib
    from emerge_amd.fdsolve import GPUSolver
    mesh, phys = _posts(10, 5, 2)
    prob = phys.problem()
    s = GPUSolver(prob, precision=precision)
    assert s.spmv_tuning()["lanes_c128"] == 0
    K = prob.system(3e9)
    s.set_matrix(K)
    t = s.spmv_tuning()
    print("spmv tuning", t, flush=True)
    for key in ("c128",) + (("c64",) if precision == 32 else ()):
        ms = t["ms_" + key]
        assert all(v > 0 for v in ms.values())
        assert t["lanes_" + key] == min(ms, key=ms.get)
    L = _lib.load()
    g = torch.Generator(device=prob.dev)
    g.manual_seed(7)
    x = torch.view_as_complex(torch.randn((prob.N, 2), generator=g, dtype=torch.float64, device=prob.dev)).contiguous()
    y1 = torch.empty_like(x)
    _lib.check(L.emg_spmv(prob.N, _lib.ptr(prob.pat.rowptr), _lib.ptr(prob.pat.cols), _lib.ptr(K), _lib.ptr(prob.mask),
                          _lib.ptr(x), _lib.ptr(y1), _lib.stream_ptr()), "emg_spmv")
    if precision == 64:
        y2 = torch.empty_like(x)
        _lib.check(L.emg_solver_apply(s.h, 64, _lib.ptr(x), _lib.ptr(y2), _lib.stream_ptr()), "emg_solver_apply")
        tol = 1e-14
    else:
        x32 = x.to(torch.complex64).contiguous()
        y32 = torch.empty_like(x32)
        _lib.check(L.emg_solver_apply(s.h, 32, _lib.ptr(x32), _lib.ptr(y32), _lib.stream_ptr()), "emg_solver_apply")
        y2 = y32.to(torch.complex128)
        tol = 1e-5
    torch.cuda.synchronize()
    err = float((y1 - y2).abs().max() / y1.abs().max())
    print("tuned vs csr", precision, err, flush=True)
    assert err < tol


def test_solver_rejects_csr_off_the_entity_pattern(cuda):
    """A PEC-reduced or re-sorted CSR breaks the blocked SpMV's layout: emg_solver_create refuses it."""
    import ctypes
    import torch
    import scipy.sparse as sp
    from emerge_amd import _lib
    A = sp.random(40, 40, density=0.2, random_state=1, format="csr") + sp.eye(40, format="csr")
    A.sort_indices()
    d = lambda a: torch.from_numpy(a.astype(np.int32)).to(cuda)
    rp, ci = d(A.indptr), d(A.indices)
    L = _lib.load()
    h = ctypes.c_void_p()
    rc = L.emg_solver_create(40, A.nnz, _lib.ptr(rp), _lib.ptr(ci), 20, 64, ctypes.byref(h))
    assert rc != 0
    assert b"entity pattern" in L.emg_last_error()


@pytest.mark.parametrize("precision", [64, 32])
def test_smoother_autotuner_same_result(cuda, precision, monkeypatch):
    """The smoother's G^T product shape (4, 8 or 16 lanes per row) is tuned with the SpMV on
    the solver's first matrix: every candidate has a measured time and the kept one is the
    fastest; solvers forced to each shape (EMG_GT_LANES) give the same smoother parts up to
    the order of the row sums."""
    import torch
    from emerge_amd import _lib
    from emerge_amd.fdsolve import GPUSolver
    mesh, phys = _posts(10, 5, 2)
    prob = phys.problem()
    K = prob.system(3e9)
    s = GPUSolver(prob, precision=precision)
    assert s.smoother_tuning()["gt_lanes"] == 0
    s.set_matrix(K)
    t = s.smoother_tuning()
    print("smoother tuning", t, flush=True)
    assert all(v > 0 for v in t["ms"].values())
    assert t["gt_lanes"] == min(t["ms"], key=t["ms"].get)
    L = _lib.load()
    g = torch.Generator(device=prob.dev)
    g.manual_seed(11)
    r = torch.view_as_complex(torch.randn((prob.N, 2), generator=g, dtype=torch.float64, device=prob.dev)).contiguous()
    np2 = prob.gradient().np2
    parts = []
    for lanes in (4, 8, 16):
        monkeypatch.setenv("EMG_GT_LANES", str(lanes))
        s2 = GPUSolver(prob, precision=64)
        s2.set_matrix(K)
        assert s2.smoother_tuning()["gt_lanes"] == lanes
        u = torch.empty_like(r)
        tg = torch.empty(np2, dtype=torch.complex128, device=prob.dev)
        _lib.check(L.emg_solver_smoother_parts(s2.h, _lib.ptr(r), _lib.ptr(u), _lib.ptr(tg), _lib.stream_ptr()),
                   "emg_solver_smoother_parts")
        parts.append(tg)
    for tg in parts[1:]:
        assert float((tg - parts[0]).abs().max() / parts[0].abs().max()) < 1e-13
