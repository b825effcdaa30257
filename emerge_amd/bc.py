"""Boundary conditions, coordinate systems and materials of the frequency-domain path.

Mirrors the reference's classes and their physics (fem/bc.py:135-754,
fem/cs.py:221-461, fem/material.py:20-106, fem/coord.py:21-50) with the same
names, arguments and formulas, but selections are plain face-tag lists of
``emerge_amd.mesh.Mesh3D`` instead of gmsh selections.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Literal

import numpy as np

C0 = 299792458.0
Z_VAC = 376.730313412


# ------------------------------------------------------------------ geometry
class CoordinateSystem:
    """Local frame (fem/cs.py:221-461): axes xax, yax, zax (3-vectors) and origin."""

    def __init__(self, xax, yax, zax, origin=(0.0, 0.0, 0.0)):
        self.xax = np.asarray(xax, dtype=np.float64)
        self.yax = np.asarray(yax, dtype=np.float64)
        self.zax = np.asarray(zax, dtype=np.float64)
        self.origin = np.asarray(origin, dtype=np.float64)
        self._basis = np.array([self.xax, self.yax, self.zax]).T          # fem/cs.py:244
        self._basis_inv = np.linalg.pinv(self._basis)                      # fem/cs.py:246

    def in_global_cs(self, x, y, z):                                        # fem/cs.py:354-368
        return (self.xax[0] * x + self.yax[0] * y + self.zax[0] * z + self.origin[0],
                self.xax[1] * x + self.yax[1] * y + self.zax[1] * z + self.origin[1],
                self.xax[2] * x + self.yax[2] * y + self.zax[2] * z + self.origin[2])

    def in_local_cs(self, x, y, z):                                         # fem/cs.py:370-390
        B = self._basis_inv
        xg, yg, zg = x - self.origin[0], y - self.origin[1], z - self.origin[2]
        return (B[0, 0] * xg + B[0, 1] * yg + B[0, 2] * zg,
                B[1, 0] * xg + B[1, 1] * yg + B[1, 2] * zg,
                B[2, 0] * xg + B[2, 1] * yg + B[2, 2] * zg)

    def in_global_basis(self, x, y, z):                                     # fem/cs.py:392-408
        return (self.xax[0] * x + self.yax[0] * y + self.zax[0] * z,
                self.xax[1] * x + self.yax[1] * y + self.zax[1] * z,
                self.xax[2] * x + self.yax[2] * y + self.zax[2] * z)

    def in_local_basis(self, x, y, z):                                      # fem/cs.py:410-431
        B = self._basis_inv
        return (B[0, 0] * x + B[0, 1] * y + B[0, 2] * z,
                B[1, 0] * x + B[1, 1] * y + B[1, 2] * z,
                B[2, 0] * x + B[2, 1] * y + B[2, 2] * z)


GCS = CoordinateSystem((1, 0, 0), (0, 1, 0), (0, 0, 1))


class Line:
    """Polyline with segment midpoints (fem/coord.py:21-50)."""

    def __init__(self, xs, ys, zs):
        self.xs, self.ys, self.zs = xs, ys, zs
        self.dxs, self.dys, self.dzs = xs[1:] - xs[:-1], ys[1:] - ys[:-1], zs[1:] - zs[:-1]
        self.xmid, self.ymid, self.zmid = 0.5 * (xs[:-1] + xs[1:]), 0.5 * (ys[:-1] + ys[1:]), 0.5 * (zs[:-1] + zs[1:])

    @property
    def cmid(self):
        return self.xmid, self.ymid, self.zmid

    @staticmethod
    def from_points(start, end, npts: int) -> "Line":
        return Line(np.linspace(start[0], end[0], npts), np.linspace(start[1], end[1], npts),
                    np.linspace(start[2], end[2], npts))


# ------------------------------------------------------------------ materials
def _is_scalar(v) -> bool:
    return isinstance(v, (float, complex, int, np.floating, np.complexfloating, np.integer))


@dataclass
class Material:
    """fem/material.py:20-106.

    er, ur: scalars (times the identity) or (3, 3) tensors; tand multiplies
    er by (1 - j tand).  _fer, _fur: optional callables (x, y, z) -> (3, 3, N)
    complex tensors at N points (tet centroids), e.g. PML stretches; they take
    precedence over er / ur, as in the reference."""
    er: float | complex | np.ndarray = 1.0
    ur: float | complex | np.ndarray = 1.0
    tand: float = 0.0
    _fer: object = None
    _fur: object = None

    @property
    def ermat(self) -> np.ndarray:                                          # fem/material.py:44-49
        if _is_scalar(self.er):
            return self.er * (1 - 1j * self.tand) * np.eye(3)
        return np.asarray(self.er) * (1 - 1j * self.tand)

    @property
    def urmat(self) -> np.ndarray:                                          # fem/material.py:51-56
        if _is_scalar(self.ur):
            return self.ur * np.eye(3)
        return np.asarray(self.ur)

    @property
    def fer3d_mat(self):                                                    # fem/material.py:91-96
        if self._fer is None:
            return lambda x, y, z: np.repeat(self.ermat[:, :, None], np.shape(x)[0], axis=2)
        return self._fer

    @property
    def fur3d_mat(self):                                                    # fem/material.py:98-103
        if self._fur is None:
            return lambda x, y, z: np.repeat(self.urmat[:, :, None], np.shape(x)[0], axis=2)
        return self._fur

    def key(self):
        """Identity for caching: value fields, or the callables / tensors by object identity."""
        v = lambda a: a if _is_scalar(a) else id(a)
        return (v(self.er), v(self.ur), self.tand, id(self._fer) if self._fer is not None else None,
                id(self._fur) if self._fur is not None else None)


AIR = Material()


def pml_material(material: Material, center, dims, direction, thickness: float, exponent: float = 3.0,
                 deltamax: float = 8.0) -> Material:
    """A perfectly matched layer of `material` on the side `direction` (one of +-1/0 per
    axis) of a box with centre `center` and size `dims`: diagonal er and ur tensors
    er * diag(sy sz / sx, sz sx / sy, sx sy / sz) with the stretches
    s_i = 1 - j (d_i (x_i - c_i - d_i D_i / 2) / thickness)^exponent deltamax on the
    layered axes and 1 elsewhere (fem/modeling/pmlbox.py:25-92)."""
    (px, py, pz), (W, D, H), (dx, dy, dz) = center, dims, direction

    def stretch(d, c, size):
        if d == 0:
            return lambda x: np.ones(np.shape(x)[0], dtype=np.complex128)
        return lambda x: 1 - 1j * (d * (x - c - d * size / 2) / thickness) ** exponent * deltamax

    sxf, syf, szf = stretch(dx, px, W), stretch(dy, py, D), stretch(dz, pz, H)

    def tensor(base):
        def f(x, y, z):
            sx, sy, sz = sxf(x), syf(y), szf(z)
            out = np.zeros((3, 3, np.shape(x)[0]), dtype=np.complex128)
            out[0, 0] = base * sy * sz / sx
            out[1, 1] = base * sz * sx / sy
            out[2, 2] = base * sx * sy / sz
            return out
        return f

    er = material.er * (1 - 1j * material.tand)
    return Material(er=material.er, ur=material.ur, tand=material.tand, _fer=tensor(er), _fur=tensor(material.ur))


# ------------------------------------------------------------------ boundary conditions
class BoundaryCondition:
    """A face-tag based boundary condition (fem/bc.py:43-133)."""

    def __init__(self, tags):
        self.tags: list[int] = sorted(set([int(tags)] if np.isscalar(tags) else [int(t) for t in tags]))

    def exclude_bc(self, other: "BoundaryCondition") -> list[int]:
        removed = [t for t in self.tags if t in other.tags]
        self.tags = [t for t in self.tags if t not in other.tags]
        return removed

    def __repr__(self):
        return f"{type(self).__name__}{self.tags}"


class PEC(BoundaryCondition):
    """Perfect electric conductor (fem/bc.py:135-147)."""


class RobinBC(BoundaryCondition):
    """Boundary condition of the third kind (fem/bc.py:149-178)."""
    _include_stiff = False
    _include_mass = False
    _include_force = False

    def __init__(self, tags):
        super().__init__(tags)
        self.v_integration = False
        self.vintline: Line | None = None

    def get_basis(self):
        return np.eye(3)

    def get_inv_basis(self):
        return np.eye(3)


class PortBC(RobinBC):
    """Port base class (fem/bc.py:180-252)."""
    Zvac = Z_VAC

    def __init__(self, tags):
        super().__init__(tags)
        self.port_number: int | None = None
        self.cs: CoordinateSystem | None = None
        self.selected_mode = 0
        self.Z0 = None
        self.active = False
        self.power = 1.0

    def get_basis(self):
        return self.cs._basis

    def get_inv_basis(self):
        return self.cs._basis_inv

    @property
    def modetype(self) -> Literal["TEM", "TE", "TM"]:
        # fem/bc.py:203-205: every port subclass here inherits 'TEM'
        return "TEM"

    def Zmode(self, k0: float) -> float:                                     # fem/bc.py:207-215
        if self.modetype == "TEM":
            return self.Zvac
        if self.modetype == "TE":
            return k0 * C0 / self.get_beta(k0) * 4 * np.pi * 1e-7
        return self.get_beta(k0) / (k0 * C0 * 8.854187818814 * 1e-12)

    @property
    def mode_number(self) -> int:
        return self.selected_mode + 1

    def get_beta(self, k0):
        return k0

    def get_gamma(self, k0):
        return 1j * self.get_beta(k0)

    def port_mode_3d(self, x_local, y_local, k0):
        raise NotImplementedError

    def port_mode_3d_global(self, xg, yg, zg, k0):                           # fem/bc.py:244-252
        xl, yl, _ = self.cs.in_local_cs(xg, yg, zg)
        Ex, Ey, Ez = self.port_mode_3d(xl, yl, k0)
        return np.array(self.cs.in_global_basis(Ex, Ey, Ez))

    # affine-in-frequency decomposition used by the batched sweep:
    #   gamma(k0) = gamma_scale(k0), U_inc(x, k0) = force_scale(k0) * mode_shape(x)
    def force_scale(self, k0):
        raise NotImplementedError

    def mode_shape(self, x_local, y_local):
        raise NotImplementedError


class AbsorbingBoundary(RobinBC):
    """First-order absorbing boundary: gamma = j k0 (fem/bc.py:254-300)."""
    _include_stiff = True
    _include_mass = True
    _include_force = False

    def __init__(self, tags, order: int = 1, origin=None):
        super().__init__(tags)
        self.order = order
        self.origin = origin
        self.cs = GCS

    def get_beta(self, k0):
        return k0

    def get_gamma(self, k0):
        return 1j * self.get_beta(k0)

    def get_Uinc(self, x_local, y_local, k0):
        return np.zeros((3, len(x_local)), dtype=np.complex128)


class RectangularWaveguide(PortBC):
    """TE_mn rectangular waveguide port with analytic mode (fem/bc.py:496-618)."""
    _include_stiff = True
    _include_mass = False
    _include_force = True

    def __init__(self, tags, port_number: int, active: bool = False, cs: CoordinateSystem | None = None,
                 dims: tuple[float, float] | None = None, power: float = 1.0):
        super().__init__(tags)
        if cs is None or dims is None:
            raise ValueError("RectangularWaveguide needs an explicit cs and dims here (no gmsh face fitting)")
        self.port_number = port_number
        self.active = active
        self.power = power
        self.type = "TE"
        self.mode = (1, 0)
        self.cs = cs
        self.dims = dims

    def get_amplitude(self, k0: float) -> float:                             # fem/bc.py:549-556
        Zte = 377
        return np.sqrt(self.power * 4 * Zte / (self.dims[0] * self.dims[1]))

    def get_beta(self, k0: float):                                          # fem/bc.py:571-575
        w, h = self.dims
        return np.sqrt(k0 ** 2 - (np.pi * self.mode[0] / w) ** 2 - (np.pi * self.mode[1] / h) ** 2)

    def get_gamma(self, k0: float):
        return 1j * self.get_beta(k0)

    def get_Uinc(self, x_local, y_local, k0):                                # fem/bc.py:588-589
        return -2 * 1j * self.get_beta(k0) * self.port_mode_3d(x_local, y_local, k0)

    def port_mode_3d(self, x_local, y_local, k0):                            # fem/bc.py:591-606
        w, h = self.dims
        E = self.get_amplitude(k0) * np.cos(np.pi * self.mode[0] * x_local / w) * np.cos(np.pi * self.mode[1] * y_local / h)
        return np.array([0 * E, E, 0 * E])

    def force_scale(self, k0):
        return -2 * 1j * self.get_beta(k0)

    def mode_shape(self, x_local, y_local):
        return self.port_mode_3d(x_local, y_local, 0.0)


class LumpedPort(PortBC):
    """Lumped port with uniform field along `direction` (fem/bc.py:620-754)."""
    _include_stiff = True
    _include_mass = False
    _include_force = True

    def __init__(self, tags, port_number: int, width: float, height: float, direction, active: bool = False,
                 power: float = 1.0, Z0: float = 50.0, cs: CoordinateSystem | None = None):
        super().__init__(tags)
        self.port_number = port_number
        self.active = active
        self.power = power
        self.Z0 = Z0
        self.width = width
        self.height = height
        self.direction = np.asarray(direction, dtype=np.float64)
        self.type = "TEM"
        if cs is None:
            raise ValueError("LumpedPort needs the face coordinate system (its zax = face normal)")
        self.cs = cs
        self.vintline: Line | None = None
        self.v_integration = True

    @property
    def surfZ(self) -> float:                                                # fem/bc.py:678-682
        return self.Z0 * self.width / self.height

    @property
    def voltage(self) -> float:                                              # fem/bc.py:684-690
        return np.sqrt(2 * self.power * self.Z0)

    def get_gamma(self, k0):                                                 # fem/bc.py:700-709
        return 1j * k0 * 376.7303 / self.surfZ

    def get_Uinc(self, x_local, y_local, k0):                                # fem/bc.py:711-713
        Emag = -1j * 2 * k0 * self.voltage / self.height * (376.7303 / self.surfZ)
        return Emag * self.port_mode_3d(x_local, y_local, k0)

    def port_mode_3d(self, x_local, y_local, k0):                            # fem/bc.py:715-728
        px, py, pz = self.cs.in_local_basis(*self.direction)
        one = np.ones_like(x_local)
        return np.array([px * one, py * one, pz * one])

    def force_scale(self, k0):
        return -1j * 2 * k0 * self.voltage / self.height * (376.7303 / self.surfZ)

    def mode_shape(self, x_local, y_local):
        return self.port_mode_3d(x_local, y_local, 0.0)


class PortMode:
    """One eigenmode of a modal port (fem/bc.py:302-327): the 14-DOF face field, its
    propagation constant, the normalisation to unit power and the mode type."""

    def __init__(self, field, beta, k0, freq, power, modetype: str, TEM: bool, eigenvalue=None):
        self.modefield = field
        self.beta = beta
        self.k0 = k0
        self.freq = freq
        self.TEM = TEM
        self.modetype = modetype
        self.eigenvalue = eigenvalue
        self.neff = beta / k0
        self.energy = float(np.mean(np.abs(field) ** 2))
        self.norm_factor = float(np.sqrt(1 / np.abs(power)))                     # set_power, fem/bc.py:325-327
        self.power = power


class ModalPort(PortBC):
    """Port whose mode comes from an eigenmode analysis of its face (fem/bc.py:329-500):
    run Electrodynamics3D.modal_analysis(port) before the sweep.  `cs` is the face frame
    (zax = the face normal); the reference derives it from the face when omitted."""
    _include_stiff = True
    _include_mass = False
    _include_force = True

    def __init__(self, tags, port_number: int, active: bool = False, cs: CoordinateSystem | None = None,
                 power: float = 1.0):
        super().__init__(tags)
        if cs is None:
            raise ValueError("ModalPort needs the face coordinate system here (zax = face normal)")
        self.port_number = port_number
        self.active = active
        self.power = power
        self.cs = cs
        self.selected_mode = 0
        self.modes: list[PortMode] = []
        self.face = None                   # emerge_amd.modal.PortFace after modal_analysis

    @property
    def nmodes(self) -> int:
        return len(self.modes)

    def get_mode(self, i=None) -> PortMode:
        return self.modes[self.selected_mode if i is None else i]

    @property
    def modetype(self):                                                      # fem/bc.py:376-379
        return self.get_mode().modetype

    def get_beta(self, k0):                                                  # fem/bc.py:444-453
        m = self.get_mode()
        if m.TEM:
            return m.beta / m.k0 * k0
        freq = k0 * C0 / (2 * np.pi)
        return np.sqrt(m.beta ** 2 + k0 ** 2 * (1 - (m.freq / freq) ** 2))

    def get_gamma(self, k0):
        return 1j * self.get_beta(k0)

    def port_mode_3d(self, x_local, y_local, k0=None):
        """Mode E in local components at local face points (fem/bc.py:469-484)."""
        xg, yg, zg = self.cs.in_global_cs(np.asarray(x_local), np.asarray(y_local), 0 * np.asarray(x_local))
        E = self.port_mode_3d_global(xg, yg, zg, k0)
        return np.array(self.cs.in_local_basis(E[0], E[1], E[2]))

    def port_mode_3d_global(self, xg, yg, zg, k0=None):                      # fem/bc.py:486-500
        m = self.get_mode()
        return np.sqrt(self.power) * m.norm_factor * self.face.field_at(m.modefield, xg, yg, zg)

    def get_Uinc(self, x_local, y_local, k0):
        return -2j * self.get_beta(k0) * self.port_mode_3d(x_local, y_local, k0)

    def force_scale(self, k0):
        return -2j * self.get_beta(k0)

    def mode_shape(self, x_local, y_local):
        return self.port_mode_3d(x_local, y_local)
