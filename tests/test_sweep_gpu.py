"""Batched sweep kernels and the basis sweep driver against explicit FP64 algebra and the oracle."""
import numpy as np
import pytest

from emerge_amd import meshgen
from emerge_amd.bc import C0, AbsorbingBoundary, CoordinateSystem, LumpedPort, Material
from emerge_amd.physics import Electrodynamics3D
from oracle import fd
from tests.test_freq_gpu import waveguide

pytestmark = pytest.mark.gpu


def _lumped_box():
    size = (20e-3, 8e-3, 6e-3)
    mesh = meshgen.structured_box(10, 4, 3, size=size, jitter=0.05, seed=2,
                                  tag_fn=lambda c: np.where(c[0] > 10e-3, 2, 1))
    cs1 = CoordinateSystem((0, 1, 0), (0, 0, 1), (1, 0, 0), (0.0, size[1] / 2, size[2] / 2))
    cs2 = CoordinateSystem((0, -1, 0), (0, 0, 1), (-1, 0, 0), (size[0], size[1] / 2, size[2] / 2))
    p1 = LumpedPort([meshgen.XMIN], 1, width=size[1], height=size[2], direction=(0, 0, 1), cs=cs1, Z0=50)
    p2 = LumpedPort([meshgen.XMAX], 2, width=size[1], height=size[2], direction=(0, 0, 1), cs=cs2, Z0=50)
    abc = AbsorbingBoundary([meshgen.YMAX])
    phys = Electrodynamics3D(mesh)
    phys.set_material(2, Material(er=3.0, tand=0.02))
    phys.assign(p1, p2, abc)
    return mesh, phys


@pytest.mark.parametrize("nvec", [1, 2, 8, 16])
def test_sweep_residual_matches_explicit(cuda, nvec):
    """r_j = mask.(c_j q_p - K(f_j) x_j) for a block of pairs, in one pass (assembler.py:312-388)."""
    import torch
    from emerge_amd.sweep import BasisSweep
    from emerge_amd.fdsolve import GPUSolver
    from emerge_amd.sparam import PortSampler
    print(f"sweep residual nvec={nvec}", flush=True)
    mesh, phys = _lumped_box()
    prob = phys.problem()
    bs = BasisSweep(prob, GPUSolver(prob, precision=64), PortSampler(prob), max_basis=16)
    rng = np.random.default_rng(nvec)
    mask = prob.mask.cpu().numpy().astype(bool)
    m = 5
    for i in range(m):
        v = (rng.standard_normal(prob.N) + 1j * rng.standard_normal(prob.N)) * mask
        bs.add(torch.from_numpy(v).to(prob.dev))
    assert bs.m == m
    pairs = [(float(f), int(i % 2)) for i, f in enumerate(np.linspace(2e9, 6e9, nvec))]
    Y = [rng.standard_normal(m) + 1j * rng.standard_normal(m) for _ in pairs]
    res = bs.residuals(pairs, Y)
    V = bs.V[:m].cpu().numpy()
    for (f, pi), y, r in zip(pairs, Y, res):
        K = prob.to_scipy(prob.system(f)).tocsr()
        x = y @ V
        b = prob.rhs(prob.ports[pi], f).cpu().numpy()
        rr = b - mask * (K @ (mask * x))
        ref = np.linalg.norm(rr[mask]) / np.linalg.norm(b[mask])
        assert abs(r - ref) <= 1e-12 * ref, (f, pi, r, ref)


def test_basis_qr_and_least_squares_trial(cuda):
    """V orthonormal, K(f) V = Q R C(f), and the trial is the least-squares solution over span(V)."""
    import torch
    from emerge_amd.sweep import BasisSweep
    from emerge_amd.fdsolve import GPUSolver
    from emerge_amd.sparam import PortSampler
    mesh, phys = _lumped_box()
    prob = phys.problem()
    bs = BasisSweep(prob, GPUSolver(prob, precision=64), PortSampler(prob), max_basis=16)
    rng = np.random.default_rng(0)
    mask = prob.mask.cpu().numpy().astype(bool)
    xs = []
    for i in range(6):
        x = (rng.standard_normal(prob.N) + 1j * rng.standard_normal(prob.N)) * mask
        xs.append(x)
        y = bs.add(torch.from_numpy(x).to(prob.dev))
        V = bs.V[: bs.m].cpu().numpy()
        assert np.abs(y @ V - x).max() <= 1e-12 * np.abs(x).max()
    V = bs.V[: bs.m].cpu().numpy()
    assert np.abs(np.conj(V) @ V.T - np.eye(bs.m)).max() < 1e-13
    Q = bs.Q[: bs.nq].cpu().numpy()
    assert np.abs(np.conj(Q) @ Q.T - np.eye(bs.nq)).max() < 1e-12
    for f, pi in ((3e9, 0), (5.5e9, 1)):
        K = prob.to_scipy(prob.system(f)).tocsr()
        KV = np.stack([mask * (K @ (mask * v)) for v in V], axis=1)
        k0 = 2 * np.pi * f / C0
        RC = bs.R[: bs.nq, : bs.m * bs.ncol]
        M = RC[:, 0::bs.ncol] - k0 * k0 * RC[:, 1::bs.ncol]
        for t, rt in enumerate(bs.terms):
            M = M + complex(rt.bc.get_gamma(k0)) * RC[:, 2 + t::bs.ncol]
        assert np.abs(Q.T @ M - KV).max() <= 1e-11 * np.abs(KV).max()
        b = prob.rhs(prob.ports[pi], f).cpu().numpy()
        y_ref = np.linalg.lstsq(KV, b, rcond=None)[0]
        y = bs.trial(f, pi)
        r, r_ref = np.linalg.norm(b - KV @ y), np.linalg.norm(b - KV @ y_ref)
        assert r <= r_ref * (1 + 1e-8)
    # a dependent vector adds nothing
    y = bs.add(torch.from_numpy(xs[0] * 2.0 - 3j * xs[1]).to(prob.dev))
    assert bs.m == 6 and y.shape[0] == 6


@pytest.mark.parametrize("precision", [64, 32])
def test_basis_sweep_waveguide_matches_plain_and_direct(cuda, precision):
    print(f"basis sweep waveguide precision={precision}", flush=True)
    mesh, ports = waveguide()
    freqs = list(np.linspace(8.5e9, 11.5e9, 13))
    S = {}
    recs = {}
    for mode in ("plain", "basis"):
        phys = Electrodynamics3D(mesh)
        phys.assign(*ports)
        phys.set_frequency(freqs)
        st = phys.solveroutine
        st.precision, st.tol, st.sweep = precision, 1e-10, mode
        data = phys.frequency_domain()
        S[mode] = data.S_array()
        recs[mode] = [r for ds in data.datasets for r in ds.solves]
        print(mode, "seconds", round(data.info["sweep_seconds"], 3), "iterations",
              sum(r.iterations for r in recs[mode]), flush=True)
        for r in recs[mode]:
            assert r.converged and r.residual <= 1e-10, (mode, r)
    full = sum(r.iterations > 0 for r in recs["basis"])
    print("full solves", full, "of", len(recs["basis"]), "basis size", data.info["basis_size"],
          "max |dS| basis-plain", np.abs(S["basis"] - S["plain"]).max(), flush=True)
    assert np.abs(S["basis"] - S["plain"]).max() < 1e-7
    assert full < len(recs["basis"])
    er, ur = phys.material_tensors()
    orc = fd.OracleProblem(mesh, er, ur, phys.boundary_conditions, order_seed=3)
    pick = [0, 6, 12]
    So, _ = orc.frequency_domain([freqs[i] for i in pick])
    print("max |dS| basis-direct", np.abs(S["basis"][pick] - So).max(), flush=True)
    assert np.abs(S["basis"][pick] - So).max() < 1e-6


def test_basis_sweep_lumped_abc_matches_direct(cuda):
    print("basis sweep lumped + abc", flush=True)
    mesh, phys = _lumped_box()
    freqs = list(np.linspace(3e9, 6e9, 9))
    phys.set_frequency(freqs)
    phys.solveroutine.tol = 1e-10
    data = phys.frequency_domain()
    for ds in data.datasets:
        for r in ds.solves:
            assert r.converged and r.residual <= 1e-10, r
    Sg = data.S_array()
    er, ur = phys.material_tensors()
    orc = fd.OracleProblem(mesh, er, ur, phys.boundary_conditions, order_seed=3)
    pick = [0, 4, 8]
    So, _ = orc.frequency_domain([freqs[i] for i in pick])
    print("max |dS|", np.abs(Sg[pick] - So).max(), "basis size", data.info["basis_size"], flush=True)
    assert np.abs(Sg[pick] - So).max() < 1e-6
    f, s21 = data.ax("freq").S(2, 1)
    assert np.allclose(s21, Sg[:, 1, 0])
    # stored fields are full-length with zeros on PEC DOFs
    x = data.item(2)._fields[1]
    assert x.shape == (phys.problem().N,)
    assert np.all(x[~phys.problem().mask.cpu().numpy().astype(bool)] == 0)


def test_full_basis_keeps_certified_solutions(cuda):
    """With max_basis=2 the sweep must still certify every stored result: trial results through
    residuals(), GMRES results kept outside the full basis through their own FP64 residual."""
    from emerge_amd.fdsolve import GPUSolver
    from emerge_amd.sparam import PortSampler
    from emerge_amd.sweep import BasisSweep, StoredSolution
    mesh, phys = _lumped_box()
    prob = phys.problem()
    freqs = list(np.linspace(2e9, 6e9, 9))
    tol = 1e-8
    bs = BasisSweep(prob, GPUSolver(prob, precision=64), PortSampler(prob), tol=tol, max_basis=2)
    out = bs.run(freqs, "gmres")
    assert bs.m == 2
    mask = prob.mask.cpu().numpy().astype(bool)
    stored = 0
    for fi, f in enumerate(freqs):
        for pi in range(len(prob.ports)):
            y, rec = out[fi][pi]
            assert rec.converged
            if isinstance(y, StoredSolution):
                stored += 1
                K = prob.to_scipy(prob.system(f)).tocsr()
                x = y.x.cpu().numpy()
                b = prob.rhs(prob.ports[pi], f).cpu().numpy()
                r = b - mask * (K @ (mask * x))
                rel = np.linalg.norm(r[mask]) / np.linalg.norm(b[mask])
                assert rel <= tol, (f, pi, rel)
            else:
                rel = bs.residuals([(f, pi)], [y])[0]
                assert rel <= tol * 1.0001, (f, pi, rel)
    assert stored > 0
