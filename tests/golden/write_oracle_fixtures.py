"""Write the oracle's answers for the configured workloads at full size, for the
GPU parity tests that cannot run the oracle inside their time budget.

    python tests/golden/write_oracle_fixtures.py c1 c3 c4

* c1: WR-90 section (20,160 tets) at 10 GHz: S (2 x 2) from the oracle's
      SuperLU direct solve with the reference's settings (fem/solver.py:243-275).
* c3: posts box at 40 x 20 x 8 hexes (the full-size geometry at a size the
      oracle solves in seconds per point): S at the 11 check frequencies
      (indices 0, 20, ..., 200 of 201 points over 2-4 GHz) from MKL PARDISO on
      the real block system (fem/solver.py:70-90, 294-309).
* c4: cavity with the eps_r = 6 block (221,184 tets): the 10 eigenvalues
      k^2 nearest (2 pi 5 GHz / c)^2 from ARPACK shift-invert with a PARDISO
      factorisation (fem/physics/edm/emfreq3d.py:838-848), gradient modes removed.

Results go to tests/golden/oracle_fixtures.npz (keys are replaced one
workload at a time).  Every number comes from the CPU oracle in oracle/.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
OUT = os.path.join(HERE, "oracle_fixtures.npz")

POSTS_SMALL = (40, 20, 8)


def _save(**kw):
    old = dict(np.load(OUT)) if os.path.exists(OUT) else {}
    old.update(kw)
    np.savez_compressed(OUT, **old)


def c1():
    from emerge_amd import workloads
    from oracle import fd
    phys = workloads.wr90()
    er, ur = phys.material_tensors()
    orc = fd.OracleProblem(phys.mesh, er, ur, phys.boundary_conditions)
    S, _ = orc.frequency_domain(phys.frequencies)
    _save(c1_freqs=np.array(phys.frequencies), c1_S=S)
    print("c1 S", S[0])


def c3():
    from emerge_amd import workloads
    from oracle import fd
    phys = workloads.posts(*POSTS_SMALL)
    er, ur = phys.material_tensors()
    idx = np.array(workloads.POSTS_CHECK_INDICES)
    freqs = np.array(workloads.POSTS_FREQS)[idx]
    threads = len(os.sched_getaffinity(0))
    orc = fd.OracleProblem(phys.mesh, er, ur, phys.boundary_conditions, solver="pardiso", threads=threads)
    S = []
    for f in freqs:
        t0 = time.perf_counter()
        s, _ = orc.frequency_domain([f])
        S.append(s[0])
        print(f"c3 f={f:.4g} S21={s[0][1, 0]:.6f} ({time.perf_counter() - t0:.1f} s)", flush=True)
    orc.release()
    _save(c3_hexes=np.array(POSTS_SMALL), c3_indices=idx, c3_freqs=freqs, c3_S=np.array(S))


def c4():
    from emerge_amd import workloads
    from emerge_amd.bc import C0
    from oracle import fd
    phys = workloads.cavity()
    er, ur = phys.material_tensors()
    orc = fd.OracleProblem(phys.mesh, er, ur, phys.boundary_conditions)
    sigma = (2 * np.pi * workloads.CAVITY_TARGET_HZ / C0) ** 2
    t0 = time.perf_counter()
    print(f"c4: {phys.mesh.n_tets} tets", flush=True)
    w = fd.oracle_eigenmodes(orc, sigma, 10, solver="pardiso", threads=len(os.sched_getaffinity(0)), extra=6,
                             verbose=True)
    print(f"c4 eigenvalues {np.sort(w.real)} ({time.perf_counter() - t0:.1f} s)", flush=True)
    _save(c4_sigma=np.array(sigma), c4_lambda=w)


if __name__ == "__main__":
    for name in sys.argv[1:]:
        {"c1": c1, "c3": c3, "c4": c4}[name]()
