"""Frequency-domain path on the GPU against the CPU oracle (direct solves)."""
import numpy as np
import pytest

from emerge_amd import meshgen
from emerge_amd.bc import AbsorbingBoundary, CoordinateSystem, LumpedPort, Material, RectangularWaveguide
from emerge_amd.physics import Electrodynamics3D
from oracle import fd
from oracle import oracle as O
from tests.helpers import mesh_dof_keys

pytestmark = pytest.mark.gpu

A_WG, B_WG = 22.86e-3, 10.16e-3


def waveguide(nx=8, ny=4, nz=16, length=30e-3, seed=0):
    mesh = meshgen.waveguide_wr90(nx, ny, nz, seed=seed, length=length)
    cs1 = CoordinateSystem((1, 0, 0), (0, 1, 0), (0, 0, 1), (A_WG / 2, B_WG / 2, 0))
    cs2 = CoordinateSystem((-1, 0, 0), (0, 1, 0), (0, 0, -1), (A_WG / 2, B_WG / 2, length))
    p1 = RectangularWaveguide([meshgen.ZMIN], 1, cs=cs1, dims=(A_WG, B_WG))
    p2 = RectangularWaveguide([meshgen.ZMAX], 2, cs=cs2, dims=(A_WG, B_WG))
    return mesh, [p1, p2]


def _oracle(mesh, phys, order_seed=3):
    er, ur = phys.material_tensors()
    return fd.OracleProblem(mesh, er, ur, phys.boundary_conditions, order_seed=order_seed)


def test_system_matrix_and_sources_match_oracle(cuda):
    """K(f) over free DOFs and the port source vectors (assembler.py:312-388)."""
    print("system parity", flush=True)
    mesh, ports = waveguide(4, 2, 6, length=12e-3)
    phys = Electrodynamics3D(mesh)
    phys.assign(*ports)
    prob = phys.problem()
    f = 9.5e9
    Kg = prob.to_scipy(prob.system(f)).tocsr()
    orc = _oracle(mesh, phys)
    Ko, _, sid, pv = orc.assemble_freq_matrix(f)
    p = O.permutation_to(mesh_dof_keys(mesh), orc.field.dof_keys())     # oracle index of each mesh DOF
    free = np.flatnonzero(prob.mask.cpu().numpy())
    assert np.array_equal(np.sort(p[free]), np.sort(sid))
    Kg_f = Kg[free][:, free].toarray()
    Ko_f = Ko[p[free]][:, p[free]].toarray()
    assert np.abs(Kg_f - Ko_f).max() / np.abs(Ko_f).max() < 1e-12
    for port in ports:
        bg = prob.rhs(port, f).cpu().numpy()[free]
        bo = pv[port.port_number][p[free]]
        assert np.abs(bg - bo).max() / np.abs(bo).max() < 1e-12


@pytest.mark.parametrize("precision", [64, 32])
def test_waveguide_sparams_match_direct(cuda, precision):
    print(f"waveguide S precision={precision}", flush=True)
    mesh, ports = waveguide()
    phys = Electrodynamics3D(mesh)
    phys.assign(*ports)
    freqs = [8.5e9, 10e9, 11.5e9]
    phys.set_frequency(freqs)
    phys.solveroutine.precision = precision
    phys.solveroutine.tol = 1e-10
    phys.solveroutine.sweep = "plain"
    data = phys.frequency_domain()
    for ds in data.datasets:
        for rec in ds.solves:
            print(rec, flush=True)
            assert rec.converged and rec.residual <= 1e-10, rec
    Sg = data.S_array()
    So, _ = _oracle(mesh, phys).frequency_domain(freqs)
    print("max |dS|", np.abs(Sg - So).max(), flush=True)
    assert np.abs(Sg - So).max() < 1e-6
    # physics sanity: near-lossless, reciprocal
    for k in range(len(freqs)):
        S = Sg[k]
        assert abs(abs(S[0, 0]) ** 2 + abs(S[1, 0]) ** 2 - 1) < 0.02
        assert abs(S[1, 0] - S[0, 1]) < 0.02
    freqs_ax, s21 = data.ax("freq").S(2, 1)
    assert np.allclose(freqs_ax, freqs) and np.allclose(s21, Sg[:, 1, 0])


def test_lumped_ports_and_absorbing_boundary(cuda):
    """Coax-like lumped ports (x faces) in a box with a lossy block and an ABC side."""
    print("lumped + abc", flush=True)
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
    freqs = [3e9, 6e9]
    phys.set_frequency(freqs)
    phys.solveroutine.tol = 1e-10
    data = phys.frequency_domain()
    Sg = data.S_array()
    So, _ = _oracle(mesh, phys).frequency_domain(freqs)
    print("max |dS|", np.abs(Sg - So).max(), flush=True)
    assert np.abs(Sg - So).max() < 1e-6


def test_pml_terminated_waveguide(cuda):
    """WR-90 section ended by a PML layer (pml_material, callable diagonal tensors per tet,
    fem/modeling/pmlbox.py:25-92): K(f) and S11 equal the oracle's direct solve, and the
    layer absorbs the TE10 wave (|S11| well below the PEC-backed 1)."""
    from emerge_amd.bc import Material, pml_material
    a, b, L, t = 22.86e-3, 10.16e-3, 30e-3, 8e-3
    mesh = meshgen.structured_box(8, 4, 24, size=(a, b, L), jitter=0.05, seed=5,
                                  tag_fn=lambda c: np.where(c[2] > L - t, 2, 1))
    cs1 = CoordinateSystem((1, 0, 0), (0, 1, 0), (0, 0, 1), (a / 2, b / 2, 0))
    phys = Electrodynamics3D(mesh)
    phys.set_material(2, pml_material(Material(), (a / 2, b / 2, (L - t) / 2), (a, b, L - t), (0, 0, 1), t,
                                      exponent=2.0, deltamax=4.0))
    phys.assign(RectangularWaveguide([meshgen.ZMIN], 1, cs=cs1, dims=(a, b)))
    phys.set_frequency([10e9])
    phys.solveroutine.tol = 1e-10
    data = phys.frequency_domain()
    er, ur = phys.material_tensors()
    assert np.any(er[0, 0].imag != 0) and np.any(er[2, 2] != er[0, 0])
    orc = fd.OracleProblem(mesh, er, ur, phys.boundary_conditions, order_seed=2)
    So, _ = orc.frequency_domain([10e9])
    S = data.S_array()
    print("pml S11 gpu", S[0, 0, 0], "oracle", So[0, 0, 0], flush=True)
    assert abs(S[0, 0, 0] - So[0, 0, 0]) <= 1e-6
    assert abs(S[0, 0, 0]) < 0.3
