"""emerge_amd: MI355X-native Nedelec-2 frequency-domain / eigenmode hot path.

Drop-in for the physics/solver layer of EMerge (fem/physics/edm, fem/solver.py):
assembly, frequency sweeps and eigenmodes run as HIP kernels on gfx950 behind
the reference's Python API.
"""
from .mesh import Mesh3D  # noqa: F401

__version__ = "0.1.0"
