"""Result containers with the reference's access pattern
(fem/physics/edm/emdata.py:34-333, fem/dataset.py:61-83):

    freqs, S21 = data.ax('freq').S(2, 1)
    S21_at_3 = data.item(3).S(2, 1)
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


class Sparam:
    """Port-labelled square S matrix (emdata.py:34-88)."""

    def __init__(self, port_nrs):
        self.map = {label: i for i, label in enumerate(port_nrs)}
        self.arry = np.zeros((len(port_nrs), len(port_nrs)), dtype=np.complex128)

    def get(self, p1, p2):
        try:
            return self.arry[self.map[p1], self.map[p2]]
        except KeyError as e:
            raise KeyError(f"Port/mode {e.args[0]!r} not found in mapping") from None

    def set(self, p1, p2, value):
        self.arry[self.map[p1], self.map[p2]] = value

    __call__ = get

    def __getitem__(self, key):
        return self.get(*key)

    def __setitem__(self, key, value):
        self.set(key[0], key[1], value)


@dataclass
class PortProperties:
    """emdata.py:90-97."""
    port_number: int | None = None
    k0: float | None = None
    beta: complex | None = None
    Z0: float | None = None
    Pout: float | None = None
    mode_number: int = 1


@dataclass
class SolveRecord:
    """Per-solve diagnostics (no counterpart in the reference, which discards them)."""
    port: int
    iterations: int
    refinements: int
    residual: float
    converged: bool
    solver: str


class EMDataSet:
    """One frequency point (emdata.py:99-203)."""

    def __init__(self, **kw):
        self.freq: float | None = None
        self.k0: float | None = None
        self.er = None
        self.ur = None
        self.Sp: Sparam | None = None
        self._fields: dict[int, np.ndarray] = {}
        self.port_modes: list[PortProperties] = []
        self.excitation: dict[int, complex] = {}
        self.solves: list[SolveRecord] = []
        self.mode = None
        self.beta = None
        self._mode_field = None
        self._basis = None                  # emerge_amd.fields.PointLocator of the mesh
        self.Ex = self.Ey = self.Ez = self.Hx = self.Hy = self.Hz = None
        for k, v in kw.items():
            setattr(self, k, v)

    def init_sp(self, port_numbers):
        self.Sp = Sparam(port_numbers)

    def add_port_properties(self, port_number, mode_number, k0, beta, Z0, Pout):
        self.port_modes.append(PortProperties(port_number, k0, beta, Z0, Pout, mode_number))

    def write_S(self, i1, i2, value):
        self.Sp[i1, i2] = value

    def S(self, i1, i2):
        return self.Sp(i1, i2)

    def set_field_vector(self):
        self.excitation = {k: 0.0 for k in self._fields}
        if self.port_modes:
            self.excitation[self.port_modes[0].port_number] = 1.0 + 0j

    def interpolate(self, xs, ys, zs, freq=None) -> "EMDataSet":
        """E and H of this dataset's field at the points (fem/physics/edm/emdata.py:181-199):
        sets Ex, Ey, Ez and Hx, Hy, Hz (shaped like xs) with H = curl E / (-j 2 pi f mu0 mu_r)
        per tet, and returns self.  The field is the excitation-weighted sum of the stored port
        solutions (needs store_fields) or the mode field of an eigen dataset."""
        if self._basis is None:
            raise ValueError("this dataset has no mesh attached to interpolate on")
        xs, ys, zs = (np.asarray(v, dtype=np.float64) for v in (xs, ys, zs))
        shp = xs.shape
        pts = np.stack([xs.ravel(), ys.ravel(), zs.ravel()])
        ur = np.broadcast_to(np.asarray(self.ur if self.ur is not None else 1.0), (self._basis.mesh.n_tets,))
        const = 1.0 / (-1j * 2 * np.pi * self.freq * (ur * 4 * np.pi * 1e-7))
        E, H = self._basis.fields(np.asarray(self._field, dtype=np.complex128), pts, curl_scale=const)
        self.Ex, self.Ey, self.Ez = (E[i].reshape(shp) for i in range(3))
        self.Hx, self.Hy, self.Hz = (H[i].reshape(shp) for i in range(3))
        return self

    @property
    def E(self):
        return self.Ex, self.Ey, self.Ez

    @property
    def H(self):
        return self.Hx, self.Hy, self.Hz

    @property
    def _field(self):
        if self._mode_field is not None:
            return self._mode_field
        return sum(self.excitation[m.port_number] * self._fields[m.port_number] for m in self.port_modes
                   if m.port_number in self._fields)


class _Proxy:
    """Collects a property or method result of every dataset along an axis (emdata.py:206-243)."""

    def __init__(self, axis: str, dss):
        self._axis = axis
        self._dss = dss

    def __getattr__(self, name):
        dss, axis = self._dss, self._axis
        if callable(getattr(dss[0], name)):
            def wrapped(*a, **k):
                return (np.array([getattr(d, axis) for d in dss]),
                        np.array([getattr(d, name)(*a, **k) for d in dss]))
            return wrapped
        return np.array([getattr(d, axis) for d in dss]), np.array([getattr(d, name) for d in dss])


class EMSimData:
    """All frequency points of a sweep (emdata.py:246-333)."""

    def __init__(self, basis=None):
        self._basis = basis
        self.datasets: list[EMDataSet] = []
        self._axis = "freq"
        self.info: dict = {}

    def new(self, **vars) -> EMDataSet:
        ds = EMDataSet(**vars)
        ds._basis = self._basis
        self.datasets.append(ds)
        return ds

    def item(self, k: int) -> EMDataSet:
        return self.datasets[k]

    def ax(self, field: str) -> _Proxy:
        return _Proxy(field, self.datasets)

    def __len__(self):
        return len(self.datasets)

    def S_array(self) -> np.ndarray:
        """(nfreq, P, P) S matrices."""
        return np.stack([d.Sp.arry for d in self.datasets])

    def export_touchstone(self, filename: str, format: str = "RI") -> str:
        """Touchstone export of all frequency points (emdata.py:306-333); ports are
        assumed numbered 1..N, as the reference assumes."""
        from .touchstone import generate_touchstone
        nports = len(self.datasets[0].Sp.map)
        freqs = np.array([d.freq for d in self.datasets])
        Smat = np.zeros((len(freqs), nports, nports), dtype=np.complex128)
        for i in range(1, nports + 1):
            for j in range(1, nports + 1):
                _, S = self.ax("freq").S(i, j)
                Smat[:, i - 1, j - 1] = S
        return generate_touchstone(filename, freqs, Smat, data_format=format)
