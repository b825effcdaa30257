"""The configured workloads as ready-to-run physics objects.

Each function returns an ``Electrodynamics3D`` with mesh, materials and
boundary conditions assigned (meshes from ``meshgen``: jittered Kuhn boxes
standing in for gmsh), plus the workload's frequencies or eigen target.

* ``wr90``:   air-filled WR-90 section 22.86 x 10.16 x 50 mm, PEC walls, two
              rectangular waveguide ports on the z faces, 15 x 7 x 32 hexes
              (20,160 tets), interior nodes jittered 5 % (seed 0); 10 GHz.
* ``posts``:  box 40 x 20 x 10 mm with 4 PEC posts (r 1.5 mm, h 8 mm) on the
              centre line and two 50-ohm lumped ports (2 x 3.75 mm strips on the
              x faces), 80 x 40 x 16 hexes carved to 298,447 tets (seed 2);
              201 points over 2-4 GHz.
* ``cavity``: PEC box 30 x 20 x 15 mm with a lossless eps_r = 6 block
              10 x 10 x 5 mm 3 mm from a corner, 48 x 32 x 24 hexes
              (221,184 tets, seed 3); modes nearest 5 GHz.
* ``patch``:  microstrip patch 38 x 29 mm (inset feed, 2.5 mm line) on a 60 x 60 mm
              ground with an eps_r 4.4, tan d 0.02, 1.6 mm substrate, in a 90 x 90 x 39 mm
              air box with a first-order absorbing boundary, a 50-ohm lumped port across
              the substrate at the line's end; 995,328 tets (seed 4); 101 points over 2-3 GHz.
"""
from __future__ import annotations

import numpy as np

from . import meshgen
from .bc import CoordinateSystem, LumpedPort, Material, RectangularWaveguide
from .physics import Electrodynamics3D

WR90_A, WR90_B = 22.86e-3, 10.16e-3


def wr90(nx=15, ny=7, nz=32, length=50e-3, seed=0, freqs=(10e9,)):
    mesh = meshgen.waveguide_wr90(nx, ny, nz, seed=seed, length=length)
    a, b = WR90_A, WR90_B
    cs1 = CoordinateSystem((1, 0, 0), (0, 1, 0), (0, 0, 1), (a / 2, b / 2, 0))
    cs2 = CoordinateSystem((-1, 0, 0), (0, 1, 0), (0, 0, -1), (a / 2, b / 2, length))
    phys = Electrodynamics3D(mesh)
    phys.assign(RectangularWaveguide([meshgen.ZMIN], 1, cs=cs1, dims=(a, b)),
                RectangularWaveguide([meshgen.ZMAX], 2, cs=cs2, dims=(a, b)))
    phys.set_frequency(list(freqs))
    return phys


def posts_ports(size_x=40e-3, yc=10e-3):
    W, H = meshgen.PORT_WIDTH, meshgen.PORT_HEIGHT
    cs1 = CoordinateSystem((0, 1, 0), (0, 0, 1), (1, 0, 0), (0.0, yc, H / 2))
    cs2 = CoordinateSystem((0, -1, 0), (0, 0, 1), (-1, 0, 0), (size_x, yc, H / 2))
    return [LumpedPort([meshgen.PORT1_TAG], 1, width=W, height=H, direction=(0, 0, 1), cs=cs1, Z0=50),
            LumpedPort([meshgen.PORT2_TAG], 2, width=W, height=H, direction=(0, 0, 1), cs=cs2, Z0=50)]


POSTS_FREQS = tuple(np.linspace(2e9, 4e9, 201))
POSTS_CHECK_INDICES = tuple(range(0, 201, 20))          # direct-solve comparison points


def posts(nx=80, ny=40, nz=16, seed=2, freqs=POSTS_FREQS):
    mesh = meshgen.posts_box(nx, ny, nz, seed=seed)
    phys = Electrodynamics3D(mesh)
    phys.assign(*posts_ports())
    phys.set_frequency(list(freqs))
    return phys


CAVITY_TARGET_HZ = 5e9


def cavity(nx=48, ny=32, nz=24, seed=3):
    mesh = meshgen.cavity_block(nx, ny, nz, seed=seed)
    phys = Electrodynamics3D(mesh)
    phys.set_material(meshgen.BLOCK_TAG, Material(er=6.0))
    phys._bc_initialized = True
    return phys


PATCH_FREQS = tuple(np.linspace(2e9, 3e9, 101))


def patch_boundary_conditions(h_sub=1.6e-3, port_y=-27.5e-3, line_w=2.5e-3):
    """The patch antenna's 50-ohm lumped port across the substrate at the line's end and the
    first-order absorbing boundary on the six faces of the air box (new objects per call)."""
    from .bc import AbsorbingBoundary
    cs = CoordinateSystem((1, 0, 0), (0, 0, 1), (0, -1, 0), (0.0, port_y, -h_sub / 2))
    port = LumpedPort([meshgen.FEED_PORT_TAG], 1, width=line_w, height=h_sub, direction=(0, 0, 1), cs=cs, Z0=50)
    abc = AbsorbingBoundary([meshgen.XMIN, meshgen.XMAX, meshgen.YMIN, meshgen.YMAX, meshgen.ZMIN, meshgen.ZMAX])
    return [port, abc]


def patch_materials(er_sub=4.4, tand=0.02) -> dict:
    return {meshgen.SUBSTRATE_TAG: Material(er=er_sub, tand=tand)}


def patch(seed=4, freqs=PATCH_FREQS, er_sub=4.4, tand=0.02, scale_w=1.0, scale_l=1.0, **mesh_kw):
    kw = dict(mesh_kw)
    mesh = meshgen.patch_antenna(seed=seed, scale_w=scale_w, scale_l=scale_l, **kw)
    phys = Electrodynamics3D(mesh)
    for tag, mat in patch_materials(er_sub, tand).items():
        phys.set_material(tag, mat)
    phys.assign(*patch_boundary_conditions(kw.get("h_sub", 1.6e-3), kw.get("port_y", -27.5e-3),
                                           kw.get("line_w", 2.5e-3)))
    phys.set_frequency(list(freqs))
    return phys
