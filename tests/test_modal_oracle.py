"""Pins of the oracle's port modal analysis (oracle/modal.py) beyond the golden element
vectors: with the face-face block of Dtt filled, the TE10 mode of a WR-90 face against its
closed form, unit power after normalisation and a two-port sweep with modal ports against the
analytic rectangular-waveguide ports; with the reference's unfilled block, the size of the
propagation-constant error that omission causes."""
import numpy as np
import pytest

from emerge_amd import meshgen
from emerge_amd.bc import C0, CoordinateSystem, ModalPort, RectangularWaveguide
from emerge_amd.physics import Electrodynamics3D
from oracle import fd
from oracle import modal as OM
from oracle import oracle as O

A_WG, B_WG = 22.86e-3, 10.16e-3


def _guide(length=12e-3, nz=6, port_cls="modal"):
    mesh = meshgen.waveguide_wr90(8, 4, nz, seed=0, length=length)
    cs1 = CoordinateSystem((1, 0, 0), (0, 1, 0), (0, 0, 1), (A_WG / 2, B_WG / 2, 0))
    cs2 = CoordinateSystem((-1, 0, 0), (0, 1, 0), (0, 0, -1), (A_WG / 2, B_WG / 2, length))
    phys = Electrodynamics3D(mesh)
    if port_cls == "modal":
        phys.assign(ModalPort([meshgen.ZMIN], 1, cs=cs1), ModalPort([meshgen.ZMAX], 2, cs=cs2))
    else:
        phys.assign(RectangularWaveguide([meshgen.ZMIN], 1, cs=cs1, dims=(A_WG, B_WG)),
                    RectangularWaveguide([meshgen.ZMAX], 2, cs=cs2, dims=(A_WG, B_WG)))
    er, ur = O.material_tensors(np.ones(mesh.n_tets, complex))
    return mesh, phys, er, ur


def _modal(face_block):
    mesh, phys, er, ur = _guide()
    orc = fd.OracleProblem(mesh, er, ur, phys.boundary_conditions, order_seed=3)
    f = 10e9
    for bc in orc.bcs:
        if bc.kind == "modal":
            OM.modal_analysis(orc, bc, f, face_block=face_block)
    return mesh, phys, orc, f


@pytest.fixture(scope="module")
def guide_modes():
    return _modal(True)


def test_reference_face_block_shifts_beta():
    """The reference's port element (nedeleclegrange2.py:350-358) leaves Dtt's face-face
    entries zero; on this face that raises beta of TE10 by about 4.9 % and couples Ez
    (Ez/Et ~ 1.3e-5, enough for the reference's own TE test to call the mode 'TM')."""
    mesh, phys, orc, f = _modal(False)
    k0 = 2 * np.pi * f / C0
    beta = np.sqrt(k0 ** 2 - (np.pi / A_WG) ** 2)
    m = [bc for bc in orc.bcs if bc.kind == "modal"][0].modes[0]
    assert 0.03 < m["beta"] / beta - 1 < 0.07
    assert m["ez_ratio"] > 1e-6


def test_te10_mode_of_the_face(guide_modes):
    mesh, phys, orc, f = guide_modes
    k0 = 2 * np.pi * f / C0
    beta = np.sqrt(k0 ** 2 - (np.pi / A_WG) ** 2)
    for bc in orc.bcs:
        if bc.kind != "modal":
            continue
        m = bc.modes[0]
        assert abs(m["beta"] - beta) < 1e-5 * beta                       # 8 x 4 quadratic face cells
        assert m["modetype"] == "TE" and m["ez_ratio"] < 1e-12
        P = OM.avg_power_flux(bc.port_field, m["field"] * m["norm"], k0, orc.ur[:, :, orc.om.tri_to_tet[0]],
                              m["beta"], fd.gaus_quad_tri4())
        assert abs(abs(P) - 1.0) < 1e-12
        # the field is the TE10 sine across the broad wall: |E| at the centre line above that near the walls
        pts = np.array([[A_WG / 2, A_WG / 8], [B_WG / 2, B_WG / 2], [bc.frame.origin[2]] * 2])
        E = bc.mode_global(*pts)
        assert np.linalg.norm(E[:, 0]) > 2 * np.linalg.norm(E[:, 1])


def test_modal_ports_sweep_matches_analytic_ports(guide_modes):
    mesh, phys, orc, f = guide_modes
    S, _ = orc.frequency_domain([f])
    mesh2, phys2, er, ur = _guide(port_cls="wg")
    S2, _ = fd.OracleProblem(mesh2, er, ur, phys2.boundary_conditions, order_seed=3).frequency_domain([f])
    assert abs(abs(S[0, 1, 0]) - 1) < 0.02 and abs(S[0, 0, 0]) < 0.05
    assert abs(abs(S[0, 1, 0]) - abs(S2[0, 1, 0])) < 0.02
    assert abs(S[0, 1, 0] - S[0, 0, 1]) < 2e-3        # reciprocity, up to the two faces' own meshes
