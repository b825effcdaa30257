"""The configured workloads at their full sizes on the GPU, against the oracle
(run here, or its answers in tests/golden/oracle_fixtures.npz written by
tests/golden/write_oracle_fixtures.py) and against FP64 algebra that does not
share the sweep's residual kernel."""
import os

import numpy as np
import pytest

from emerge_amd import _lib, workloads
from emerge_amd.bc import C0
from oracle import oracle as O
from tests.helpers import oracle_in_mesh_order, rel_max

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_fixtures.npz")


def _gold():
    return np.load(GOLD)


def test_wr90_workload_matrices_and_sparams(cuda):
    """WR-90 section at 20,160 tets, 10 GHz: E and B by DOF key against the FP64 oracle
    (<= 1e-12 relative, reference-geometry mode: the reference's own absolute-coordinate
    cofactors) and against the long-double oracle (default mode), and S11/S21 against the
    oracle's SuperLU solve (<= 1e-6)."""
    phys = workloads.wr90()
    mesh = phys.mesh
    print(f"wr90: {mesh.n_tets} tets, {2 * mesh.n_entities} DOFs", flush=True)
    assert mesh.n_tets == 20160
    prob = phys.problem()
    er, ur = phys.material_tensors()
    of = O.OracleNedelec2(O.OracleMesh(mesh.nodes, mesh.tets, order_seed=1))
    Eo, Bo = oracle_in_mesh_order(mesh, *O.tet_mass_stiffness_matrices(of, er, ur), of)
    Eld, Bld = oracle_in_mesh_order(mesh, *O.tet_mass_stiffness_matrices(of, er, ur, extended=True), of)
    Er, Br = prob.pat.assemble(er, ur, reference_geometry=True)
    for name, g, gr, o, old in (("E", prob.E, Er, Eo, Eld), ("B", prob.B, Br, Bo, Bld)):
        e_ref = rel_max(prob.to_scipy(gr), o)
        e_def = rel_max(prob.to_scipy(g), old)
        print(f"wr90 {name}: reference geometry vs fp64 oracle {e_ref:.2e}; default vs long-double oracle "
              f"{e_def:.2e}; default vs fp64 oracle {rel_max(prob.to_scipy(g), o):.2e}", flush=True)
        assert e_ref <= 1e-12, (name, e_ref)
        assert e_def <= 1e-12, (name, e_def)
    phys.solveroutine.tol = 1e-10
    data = phys.frequency_domain()
    S = data.S_array()[0]
    So = _gold()["c1_S"][0]
    print("wr90 S gpu", S.ravel(), "oracle", So.ravel(), flush=True)
    assert np.abs(S - So).max() <= 1e-6
    assert max(r.residual for ds in data.datasets for r in ds.solves) <= 1e-10


def _run_sweep(phys, tol=1e-8):
    """The basis sweep as frequency_domain runs it, keeping the driver for checks."""
    from emerge_amd.fdsolve import GPUSolver
    from emerge_amd.sparam import PortSampler
    from emerge_amd.sweep import BasisSweep
    st = phys.solveroutine
    prob = phys.problem()
    sampler = PortSampler(prob)
    solver = GPUSolver(prob, restart=st.restart, precision=32, preconditioner="hybrid", poly_degree=st.poly_degree)
    bs = BasisSweep(prob, solver, sampler, tol=tol, max_basis=st.max_basis, maxiter=st.maxiter)
    out = bs.run(list(phys.frequencies), "gmres-c64")
    return prob, sampler, bs, out


def _s_matrix(phys, sampler, bs, out, fi):
    k0 = 2 * np.pi * phys.frequencies[fi] / C0
    ports = phys.problem().ports
    return sampler.s_matrix({p.port_number: bs.fields_of(out[fi][i][0]) for i, p in enumerate(ports)}, k0)


def test_posts_workload_full_sweep_certified(cuda):
    """Posts box at full size (298,447 tets, 201 points over 2-4 GHz, two lumped ports):
    every one of the 402 (frequency, port) results has an FP64 residual <= 1e-8 recomputed
    through K(f) = E - k0^2 B + gamma P (emg_affine_combine + emg_scatter_add) and emg_spmv,
    and S is reciprocal."""
    import torch
    phys = workloads.posts()
    print(f"posts: {phys.mesh.n_tets} tets", flush=True)
    assert phys.mesh.n_tets == 298447
    prob, sampler, bs, out = _run_sweep(phys)
    print(f"posts: sweep done, basis {bs.m}", flush=True)
    L = _lib.load()
    Kv = torch.empty(prob.nnz, dtype=torch.complex128, device=prob.dev)
    r = torch.empty(prob.N, dtype=torch.complex128, device=prob.dev)
    worst = 0.0
    recip = 0.0
    for fi, f in enumerate(phys.frequencies):
        prob.system(f, out=Kv)
        for pi, port in enumerate(prob.ports):
            x = bs.vector_of(out[fi][pi][0])
            b = prob.rhs(port, f)
            _lib.check(L.emg_spmv(prob.N, _lib.ptr(prob.pat.rowptr), _lib.ptr(prob.pat.cols), _lib.ptr(Kv),
                                  _lib.ptr(prob.mask), _lib.ptr(x), _lib.ptr(r), _lib.stream_ptr()), "emg_spmv")
            m = prob.mask.to(torch.bool)
            rel = float(torch.linalg.vector_norm((b - r)[m]) / torch.linalg.vector_norm(b[m]))
            worst = max(worst, rel)
        S = _s_matrix(phys, sampler, bs, out, fi)
        recip = max(recip, abs(S[0, 1] - S[1, 0]))
    print(f"posts: worst recomputed residual {worst:.2e}, reciprocity {recip:.2e}", flush=True)
    assert worst <= 1e-8
    assert recip <= 1e-6


def test_posts_geometry_sweep_matches_direct_solves(cuda):
    """The posts geometry at 40 x 20 x 8 hexes, all 201 points through the basis sweep:
    S at indices 0, 20, ..., 200 against the oracle's PARDISO direct solves (<= 1e-6)."""
    g = _gold()
    phys = workloads.posts(*[int(v) for v in g["c3_hexes"]])
    print(f"posts small: {phys.mesh.n_tets} tets", flush=True)
    prob, sampler, bs, out = _run_sweep(phys)
    err = 0.0
    for k, fi in enumerate(g["c3_indices"]):
        assert phys.frequencies[fi] == pytest.approx(float(g["c3_freqs"][k]), rel=1e-15)
        S = _s_matrix(phys, sampler, bs, out, int(fi))
        err = max(err, float(np.abs(S - g["c3_S"][k]).max()))
    print(f"posts small: max |S - S_direct| {err:.2e}, basis {bs.m}", flush=True)
    assert err <= 1e-6
