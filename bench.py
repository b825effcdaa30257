"""Benchmark of the MI355X Nedelec-2 hot path (one GPU).

Default workload: the assembly configuration -- unit box, 50x50x67 hexes x 6
tets = 1,005,000 tets, interior vertices jittered +-10% (seed 1), per-tet
complex eps_r, mu_r = 1; second-order Nedelec curl-curl and mass matrices in
FP64 scattered into complex128 CSR (6,457,334 rows, 279,117,868 nnz each).
A step is one numeric assembly of both matrices into the pattern, which is
built once per mesh before timing (symbolic time reported separately).

Prints one JSON line (see the contract in the repository docs).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sweep freq-pts/s at ≤1e-8 FP64 residual, learned vs plain GPU; assembly Mtet/s"
HBM_PEAK_GBS = 8000.0
# HBM bytes per numeric assembly (k_tet_gram + k_assemble_staged) on the
# 1,005,000-tet workload, from rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE
# runs of this benchmark (--sweep none --no-eigen --copy-calibration): FETCH (53,747 +
# 1,301,548) KiB x 2 + WRITE (430,714 + 8,764,685) KiB x 1, the factors a 2 GiB
# device-to-device copy in the same runs gives.
ASSEMBLY_TRAFFIC_BYTES = ((53_747 + 1_301_548) * 2 + 430_714 + 8_764_685) * 1024
# HBM bytes per launch of the sweep's complex64 SpMV (k_spmv_blk<float,16,false>) on the
# posts box, from rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of this benchmark
# (--spmv-kernel-only --copy-calibration): FETCH 602,184 KiB x 2 + WRITE 17,228 KiB x 1,
# with the factors the 2 GiB copy in the same runs gives (FETCH 1,048,612 KiB, WRITE 2,097,152 KiB).
SPMV_TRAFFIC_BYTES = (602_184 * 2 + 17_228) * 1024
FP64_PEAK_TFLOPS = 78.6
# FP64 flop convention per tet for the dense contraction of the closed form
FLOP_PER_TET = 15449


def cpu_baseline(mesh, eps, sample_tets: int):
    """Oracle (C/OpenMP restatement of the reference element loop) + scipy
    COO->CSR, on the first `sample_tets` tets, host cores."""
    import scipy.sparse as sp
    from oracle import oracle as O
    O.build()
    field = O.OracleNedelec2(mesh)
    er, ur = O.material_tensors(eps)
    n = min(sample_tets, mesh.n_tets)
    O.matrix_builder(field, er, ur, 0, min(n, 2000))          # warm-up
    t0 = time.perf_counter()
    dE, dB, r, c = O.matrix_builder(field, er, ur, 0, n)
    N = field.n_field
    sp.coo_matrix((dE, (r, c)), shape=(N, N)).tocsr()
    sp.coo_matrix((dB, (r, c)), shape=(N, N)).tocsr()
    dt = time.perf_counter() - t0
    cores = len(os.sched_getaffinity(0))
    return {"value": n / dt / 1e6, "unit": "Mtet/s", "cores": cores, "kind": "port",
            "sample": f"first {n} of {mesh.n_tets} tets: C/OpenMP element loop ({cores} threads) + "
                      f"scipy coo->csr of E and B (single thread, as the reference), {dt:.2f} s"}


def sweep_workload(name: str):
    """(mesh, port factory, frequencies, description) of a sweep workload."""
    from emerge_amd import meshgen
    from emerge_amd.bc import CoordinateSystem, RectangularWaveguide
    if name == "wr90":
        a, b, L = 22.86e-3, 10.16e-3, 50e-3
        mesh = meshgen.waveguide_wr90()
        cs1 = CoordinateSystem((1, 0, 0), (0, 1, 0), (0, 0, 1), (a / 2, b / 2, 0))
        cs2 = CoordinateSystem((-1, 0, 0), (0, 1, 0), (0, 0, -1), (a / 2, b / 2, L))
        ports = lambda: [RectangularWaveguide([meshgen.ZMIN], 1, cs=cs1, dims=(a, b)),
                         RectangularWaveguide([meshgen.ZMAX], 2, cs=cs2, dims=(a, b))]
        return mesh, ports, (8e9, 12e9), ("WR-90 section 22.86x10.16x50 mm, 15x7x32 Kuhn hexes (20,160 tets), "
                                          "2 rectangular waveguide ports, air")
    if name == "posts":
        mesh = meshgen.posts_box()
        return mesh, lambda: sweep_ports(mesh), (2e9, 4e9), (
            "posts box 40x20x10 mm, 4 PEC posts r 1.5 mm h 8 mm, 80x40x16 Kuhn hexes "
            f"carved to {mesh.n_tets:,} tets, 2 lumped ports 2x3.75 mm (50 ohm) on the x faces, air")
    if name == "patch":
        from emerge_amd import workloads
        mesh = meshgen.patch_antenna(seed=4)
        bcs = lambda: workloads.patch_boundary_conditions()
        bcs.materials = workloads.patch_materials()
        return mesh, bcs, (2e9, 3e9), (
            "microstrip patch 38x29 mm with inset feed on a 60x60 mm ground, eps_r 4.4 tan d 0.02 1.6 mm "
            f"substrate, 90x90x39 mm air box with a first-order absorbing boundary, {mesh.n_tets:,} tets (seed 4), "
            "one 50-ohm lumped port")
    raise ValueError(name)


def make_phys(mesh, ports):
    """Electrodynamics3D of a sweep workload: its boundary conditions and, where the
    workload has them, its materials."""
    from emerge_amd.physics import Electrodynamics3D
    phys = Electrodynamics3D(mesh)
    for tag, mat in getattr(ports, "materials", {}).items():
        phys.set_material(tag, mat)
    phys.assign(*ports())
    return phys


def sweep_ports(mesh):
    """The two lumped ports of the posts box (strips from the floor on x = 0 and x = 40 mm)."""
    from emerge_amd import meshgen
    from emerge_amd.bc import CoordinateSystem, LumpedPort
    W, H, X = meshgen.PORT_WIDTH, meshgen.PORT_HEIGHT, 40e-3
    cs1 = CoordinateSystem((0, 1, 0), (0, 0, 1), (1, 0, 0), (0.0, 10e-3, H / 2))
    cs2 = CoordinateSystem((0, -1, 0), (0, 0, 1), (-1, 0, 0), (X, 10e-3, H / 2))
    return [LumpedPort([meshgen.PORT1_TAG], 1, width=W, height=H, direction=(0, 0, 1), cs=cs1, Z0=50),
            LumpedPort([meshgen.PORT2_TAG], 2, width=W, height=H, direction=(0, 0, 1), cs=cs2, Z0=50)]


def bench_sweep(args, dev):
    """Frequency sweep, every solve to FP64 relative residual <= tol.

    * basis: mixed-precision GMRES (complex64 cycles in FP64 refinement) at the
      frequencies the basis driver cannot certify, batched FP64 residual
      certificates for the rest -- all `sweep_freqs` points;
    * plain: complex128 GMRES for every (frequency, port), warm-started from the
      previous frequency -- the first `plain_sample` points, rate extrapolated;
    * learned: the basis sweep with the learned smoother weights, the network
      trained first on plain-solver residuals at the band's ends and middle
      (training time reported separately);
    * cpu_baseline: the host-core reference path on one frequency."""
    import torch
    from emerge_amd.physics import Electrodynamics3D
    mesh, ports, (f0, f1), desc = sweep_workload(args.sweep)
    # the workload's own point count unless given: 101 over 2-3 GHz for the patch antenna, 201 otherwise
    nf = args.sweep_freqs if args.sweep_freqs else (101 if args.sweep == "patch" else 201)
    freqs = list(np.linspace(f0, f1, nf))
    out = {"workload": desc, "tets": mesh.n_tets, "dofs": 2 * mesh.n_entities, "freqs": len(freqs),
           "band_hz": [f0, f1], "tol": args.sweep_tol}
    runs = {}
    if args.sweep_kernel_only:
        out["residual_kernel"] = residual_kernel_roofline(mesh, ports, freqs, args.sweep_tol)
        return out
    if args.spmv_kernel_only:
        out["spmv_kernel"] = spmv_kernel_roofline(mesh, ports, freqs[len(freqs) // 2])
        return out
    runs_wanted = set(args.sweep_runs.split(","))
    # plain FP64 GMRES on a sample across the band: both ends and the middle (then evenly spaced)
    plain_idx = sorted({int(round(i * (len(freqs) - 1) / max(args.plain_sample - 1, 1)))
                        for i in range(args.plain_sample)}) if args.plain_sample > 0 else []
    plain_f = [freqs[i] for i in plain_idx]
    for key, mode, prec, fl in (("basis_mixed", "basis", 32, freqs), ("plain_fp64", "plain", 64, plain_f)):
        if not fl or mode not in runs_wanted:
            continue
        phys = make_phys(mesh, ports)
        phys.set_frequency(fl)
        st = phys.solveroutine
        st.precision, st.tol, st.store_fields, st.sweep, st.verbose = prec, args.sweep_tol, False, mode, True
        st.poly_degree = args.poly_degree if mode == "basis" else 1
        phys.problem()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        data = phys.frequency_domain()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        recs = [r for ds in data.datasets for r in ds.solves]
        runs[key] = data.S_array()
        out[key] = {"freq_pts_per_s": len(fl) / dt, "seconds": dt, "freqs": len(fl),
                    "gmres_solves": int(sum(r.iterations > 0 for r in recs)),
                    "iterations_total": int(sum(r.iterations for r in recs)),
                    "max_fp64_residual": float(max(r.residual for r in recs)),
                    "all_converged": bool(all(r.converged and r.residual <= args.sweep_tol for r in recs))}
        if mode == "basis":
            out[key]["basis_size"] = int(data.info.get("basis_size", 0))
            out[key]["poly_degree"] = args.poly_degree
        else:
            out[key]["freq_indices"] = plain_idx
            out[key]["rate_note"] = (f"{len(fl)} of {len(freqs)} points timed (indices {plain_idx}); "
                                     "freq_pts_per_s is their rate")
        print(json.dumps({key: out[key]}), file=sys.stderr, flush=True)
    if args.learned and "learned" in runs_wanted:
        from emerge_amd.fdsolve import GPUSolver
        from emerge_amd.learned import LearnedSmoother
        phys = make_phys(mesh, ports)
        prob = phys.problem()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        net = LearnedSmoother(prob)
        s64 = GPUSolver(prob, precision=64)
        train_f = [f0, 0.5 * (f0 + f1), f1]
        log = lambda m: print(m, file=sys.stderr, flush=True)
        samples = net.collect(s64, train_f)
        log(f"learned: collected {sum(r.shape[0] for _, r, _ in samples)} residual samples at {len(train_f)} frequencies")
        prep = net.prepare(s64, samples)
        del samples
        hist = net.train(prep, steps=args.learned_steps, lr=3e-3, log=log)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        del prep, s64
        torch.cuda.empty_cache()
        phys.set_frequency(freqs)
        st = phys.solveroutine
        st.precision, st.tol, st.store_fields, st.sweep, st.learned = 32, args.sweep_tol, False, "basis", net
        st.verbose, st.maxiter, st.poly_degree = True, 20000, args.poly_degree
        t0 = time.perf_counter()
        data = phys.frequency_domain()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        recs = [r for ds in data.datasets for r in ds.solves]
        runs["learned"] = data.S_array()
        out["learned_basis_mixed"] = {
            "freq_pts_per_s": len(freqs) / dt, "seconds": dt, "freqs": len(freqs),
            "gmres_solves": int(sum(r.iterations > 0 for r in recs)),
            "iterations_total": int(sum(r.iterations for r in recs)),
            "max_fp64_residual": float(max(r.residual for r in recs)),
            "all_converged": bool(all(r.converged and r.residual <= args.sweep_tol for r in recs)),
            "basis_size": int(data.info.get("basis_size", 0)),
            "poly_degree": args.poly_degree,
            "learned_check": data.info.get("learned"),
            "training": {"seconds": t_train, "frequencies_hz": train_f, "steps": args.learned_steps,
                         "loss_first": hist[0], "loss_last": hist[-1]},
            "max_abs_dS_vs_basis": (float(np.abs(runs["learned"] - runs["basis_mixed"]).max())
                                    if "basis_mixed" in runs else None)}
        print(json.dumps({"learned_basis_mixed": out["learned_basis_mixed"]}), file=sys.stderr, flush=True)
    out["residual_kernel"] = residual_kernel_roofline(mesh, ports, freqs, args.sweep_tol)
    out["spmv_kernel"] = spmv_kernel_roofline(mesh, ports, freqs[len(freqs) // 2])
    if "plain_fp64" in runs and "basis_mixed" in runs:
        out["max_abs_dS_basis_vs_plain"] = float(np.abs(runs["basis_mixed"][plain_idx] - runs["plain_fp64"]).max())
        out["speedup_basis_vs_plain"] = out["basis_mixed"]["freq_pts_per_s"] / out["plain_fp64"]["freq_pts_per_s"]
        if "learned_basis_mixed" in out:
            out["speedup_learned_vs_plain"] = (out["learned_basis_mixed"]["freq_pts_per_s"]
                                               / out["plain_fp64"]["freq_pts_per_s"])
    if args.sweep_cpu:
        out["cpu_baseline"] = cpu_sweep_child(args.sweep, freqs[len(freqs) // 2])
        if "basis_mixed" in out:
            out["speedup_basis_vs_cpu"] = out["basis_mixed"]["freq_pts_per_s"] / out["cpu_baseline"]["value"]
        if "learned_basis_mixed" in out:
            out["speedup_learned_vs_cpu"] = out["learned_basis_mixed"]["freq_pts_per_s"] / out["cpu_baseline"]["value"]
    return out


def residual_kernel_roofline(mesh, ports, freqs, tol, nvec: int = 16, reps: int = 20):
    """The sweep's batched FP64 certificate kernel (k_sweep_residual, one affine SpMM
    pass over E, B and the CSR indices for 16 (frequency, port) pairs) timed with HIP
    events on the stream it runs on, against HBM peak.

    Algorithmic bytes per launch: nnz (E 16 + B 16 + column 4) + rows (row pointer 4 +
    mask 1 + row map 4) + the trial block read once (16 nvec per row) + the port
    sources read by each pair (16 nvec per row) -- the residual itself is not written."""
    import torch
    from emerge_amd import _lib
    from emerge_amd.fdsolve import GPUSolver
    from emerge_amd.physics import Electrodynamics3D
    from emerge_amd.sparam import PortSampler
    from emerge_amd.sweep import BasisSweep
    phys = make_phys(mesh, ports)
    prob = phys.problem()
    bs = BasisSweep(prob, GPUSolver(prob, precision=64, preconditioner="jacobi"), PortSampler(prob), max_basis=8)
    rng = np.random.default_rng(0)
    mask = prob.mask.cpu().numpy()
    for _ in range(4):
        v = (rng.standard_normal(prob.N) + 1j * rng.standard_normal(prob.N)) * mask
        bs.add(torch.from_numpy(v).to(prob.dev))
    pairs = [(float(freqs[(7 * j) % len(freqs)]), j % len(bs.ports)) for j in range(nvec)]
    Y = [rng.standard_normal(bs.m) + 1j * rng.standard_normal(bs.m) for _ in pairs]
    bs.residuals(pairs, Y)                                   # warm-up (also builds the block)
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    # time the kernel alone: the block X and coefficients are prepared by the first call's code path
    L = _lib.load()
    n, nv = prob.N, nvec
    X = torch.randn(n * nv, dtype=torch.complex128, device=prob.dev)
    X = X.view(n, nv).mul_(prob.mask[:, None]).reshape(-1).contiguous()
    sd = torch.tensor([(2 * np.pi * f / 299792458.0) ** 2 for f, _ in pairs], dtype=torch.float64, device=prob.dev)
    cd = torch.ones(nv, dtype=torch.complex128, device=prob.dev)
    qd = torch.tensor([pi for _, pi in pairs], dtype=torch.int32, device=prob.dev)
    zp = torch.zeros(max(bs.nprow, 1) * nv, dtype=torch.complex128, device=prob.dev)
    for i in range(reps):
        ev[i][0].record(stream)
        _lib.check(L.emg_sweep_residual(n, _lib.ptr(prob.pat.rowptr), _lib.ptr(prob.pat.cols), _lib.ptr(prob.E),
                                        _lib.ptr(prob.B), _lib.ptr(prob.mask), nv, _lib.ptr(sd), _lib.ptr(X),
                                        _lib.ptr(bs.qsrc), _lib.ptr(qd), _lib.ptr(cd), _lib.ptr(bs.prow_of_row),
                                        _lib.ptr(zp), None, _lib.ptr(bs.nrm), _lib.ptr(bs.work), _lib.stream_ptr()),
                   "emg_sweep_residual")
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    nnz = prob.nnz
    alg = nnz * 36 + n * (4 + 1 + 4) + n * nv * 16 * 2
    ach = alg / (ms * 1e-3) / 1e9
    return {"kernel": "k_sweep_residual<16>", "pairs_per_launch": nv, "kernel_ms": ms, "alg_bytes": alg,
            "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "pairs_per_s": nv / (ms * 1e-3)}


def bench_eigen(args):
    """Eigenmodes of the PEC cavity with an eps_r = 6 block (221,184 tets): the modes
    nearest 5 GHz by shift-invert Arnoldi with GPU GMRES inner solves and FP64
    Rayleigh-Ritz; spurious (gradient) modes filtered."""
    import torch
    from emerge_amd import meshgen
    from emerge_amd.bc import Material
    from emerge_amd.physics import Electrodynamics3D
    mesh = meshgen.cavity_block()
    phys = Electrodynamics3D(mesh)
    phys.set_material(meshgen.BLOCK_TAG, Material(er=6.0))
    phys._bc_initialized = True
    phys.solveroutine.verbose = True
    phys.problem()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    data = phys.eigenmode(num_sols=args.eigen_modes, target_frequency=5e9)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    its = data.info["inner_iterations"]
    lam = np.array([ds.eigenvalue for ds in data.datasets])
    parity = None
    if args.eigen_oracle:
        # the oracle (C assembly + ARPACK shift-invert with PARDISO) in a child process; its shift sits
        # amid the GPU's modes, above the curl-curl null space, so ARPACK returns the same physical modes
        lam_o = eigen_oracle_child(float(np.mean(lam.real)), len(lam) + 6)
        rel = [float(np.min(np.abs(lam_o - l)) / abs(l)) for l in lam]
        parity = {"max_rel_diff": max(rel), "rel_diff": rel, "oracle": "C assembly + ARPACK shift-invert (PARDISO)"}
    return {"workload": "PEC cavity 30x20x15 mm, eps_r 6 block 10x10x5 mm 3 mm from a corner, 48x32x24 Kuhn hexes "
                        f"({mesh.n_tets:,} tets, {2 * mesh.n_entities:,} DOFs)", "modes": len(data),
            "seconds": dt, "modes_per_s": len(data) / dt, "target_hz": 5e9,
            "freqs_hz": [float(ds.freq) for ds in data.datasets],
            "relative_residuals": [float(ds.residual) for ds in data.datasets],
            "inner_solves": len(its), "inner_iterations_mean": float(np.mean(its)) if its else 0.0,
            "inner_poly_degree": phys.solveroutine.poly_degree,
            "parity_vs_oracle": parity}


def eigen_oracle_child(sigma: float, k: int) -> np.ndarray:
    import subprocess
    env = dict(os.environ)
    cores = str(min(len(os.sched_getaffinity(0)), 16))
    env.update(OMP_NUM_THREADS=cores, MKL_NUM_THREADS=cores)
    cp = subprocess.run([sys.executable, os.path.abspath(__file__), "--eigen-oracle-child", "--eigen-sigma", str(sigma),
                         "--eigen-k", str(k)], env=env, stdout=subprocess.PIPE, text=True)
    if cp.returncode != 0:
        raise RuntimeError(f"eigen oracle failed ({cp.returncode})")
    v = json.loads(cp.stdout.strip().splitlines()[-1])
    return np.array(v["re"]) + 1j * np.array(v["im"])


def eigen_oracle(sigma: float, k: int) -> dict:
    from emerge_amd import meshgen
    from emerge_amd.bc import Material
    from emerge_amd.physics import Electrodynamics3D
    from oracle import fd
    mesh = meshgen.cavity_block()
    phys = Electrodynamics3D(mesh)
    phys.set_material(meshgen.BLOCK_TAG, Material(er=6.0))
    phys._bc_initialized = True
    er, ur = phys.material_tensors()
    orc = fd.OracleProblem(mesh, er, ur, phys.boundary_conditions)
    w = fd.oracle_eigenmodes(orc, sigma, k, solver="pardiso", threads=min(len(os.sched_getaffinity(0)), 16), extra=0)
    return {"re": w.real.tolist(), "im": w.imag.tolist()}


def spmv_kernel_roofline(mesh, ports, freq, reps: int = 50):
    """The sweep's dominant kernel, the complex64 entity-blocked SpMV of the
    mixed-precision GMRES steps (k_spmv_blk<float,L,false> with the lanes L the
    solver's autotuner picked, emg_solver_apply precision 32), timed with HIP
    events on its stream.  Algorithmic bytes per
    launch (complex64 CSR, int32 indices, one read of x and one write of y):
    12 nnz + 20 N; the blocked format itself moves about 9 nnz + 20 N (one
    column index per four entries)."""
    import torch
    from emerge_amd import _lib
    from emerge_amd.fdsolve import GPUSolver
    from emerge_amd.physics import Electrodynamics3D
    phys = make_phys(mesh, ports)
    prob = phys.problem()
    # the sweep's smoother (hybrid: the gradient correction set) so that the tuner also picks
    # the shape of its G^T product; the timed product is the SpMV alone
    solver = GPUSolver(prob, precision=32, preconditioner="hybrid", poly_degree=1)
    solver.set_matrix(prob.system(freq))
    n, nnz = prob.N, prob.nnz
    x = torch.randn(n, dtype=torch.complex64, device=prob.dev) * prob.mask
    y = torch.empty_like(x)
    L = _lib.load()
    call = lambda: _lib.check(L.emg_solver_apply(solver.h, 32, _lib.ptr(x), _lib.ptr(y), _lib.stream_ptr()),
                              "emg_solver_apply")
    for _ in range(5):
        call()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    torch.cuda.synchronize()
    for i in range(reps):
        ev[i][0].record(stream)
        call()
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    alg = 12 * nnz + 20 * n
    fmt = 9 * nnz + 20 * n
    ach = alg / (ms * 1e-3) / 1e9
    tuning = solver.spmv_tuning()
    return {"kernel": f"k_spmv_blk<float,{tuning['lanes_c64']},false>", "autotune": tuning,
            "smoother_autotune": solver.smoother_tuning(), "rows": n, "nnz": nnz,
            "kernel_ms": ms, "alg_bytes": alg,
            "format_bytes": fmt, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "format_frac": fmt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": SPMV_TRAFFIC_BYTES}


def learned_bound(name: str, log) -> dict:
    """What the learned smoother can at best gain under its training loss: the smoother
    weights optimised directly on the residuals of the very frequency then solved
    (LearnedSmoother.optimal_weights, every entity and node weight free, no network), then
    GMRES steps per port with and without them (emg_solver_solve, no fallback), in the
    sweep's precision (complex64 steps in FP64 refinement), with and without the degree-6
    polynomial.  posts: the posts box at 2.5 GHz, bounded (mode 2) and unbounded complex
    (mode 0) weights, at most 8,000 steps; patch: the patch antenna at 2.45 GHz, bounded
    weights, degree 1, at most 40,000 steps."""
    import ctypes
    import torch
    from emerge_amd import _lib, workloads
    from emerge_amd.fdsolve import GPUSolver
    from emerge_amd.learned import LearnedSmoother
    if name == "posts":
        phys, freqs, modes, degrees, maxit = workloads.posts(), (2.5e9,), (2, 0), (1, 6), 8000
    else:
        phys, freqs, modes, degrees, maxit = workloads.patch(), (2.45e9,), (2,), (1,), 40000
    prob = phys.problem()
    L = _lib.load()
    out = {"workload": name, "rows": prob.N, "max_steps": maxit, "cases": []}
    for f in freqs:
        K = prob.system(f)
        for mode in modes:
            t0 = time.perf_counter()
            sc = GPUSolver(prob, precision=64, poly_degree=1)
            net = LearnedSmoother(prob, mode=mode)
            prep = net.prepare(sc, net.collect(sc, [f], tol=1e-8))
            del sc
            we, wp, hist = net.optimal_weights(prep, steps=200, lr=0.02, log=log)
            log(f"learned bound {name} f={f / 1e9:.2f} GHz mode {mode}: loss {hist[0]:.4g} -> {hist[-1]:.4g} "
                f"({time.perf_counter() - t0:.0f} s)")
            rows = []
            for d in degrees:
                for wname, w in (("plain", (None, None)), ("optimal", (we, wp))):
                    s = GPUSolver(prob, precision=32, poly_degree=d)
                    s.POLY_STEP_CAP = s.POLY_CHUNK = maxit
                    its, conv, res = [], [], []
                    for port in prob.ports:
                        s.set_weights(*w)
                        s.set_matrix(K)
                        b = prob.rhs(port, f)
                        x = torch.zeros_like(b)
                        info = _lib.SolveInfo()
                        _lib.check(L.emg_solver_solve(s.h, _lib.ptr(b), _lib.ptr(x), 1e-8, maxit, 0,
                                                      ctypes.byref(info), _lib.stream_ptr()), "emg_solver_solve")
                        its.append(int(info.iterations))
                        conv.append(bool(info.converged))
                        res.append(float(info.residual))
                    del s
                    rows.append({"poly_degree": d, "weights": wname, "steps": its, "converged": conv, "residual": res})
                    log(f"  degree {d} {wname}: steps {its} converged {conv} residual {[f'{r:.2e}' for r in res]}")
            out["cases"].append({"freq_hz": f, "weight_mode": "bounded [0.61, 1.65]" if mode == 2 else "complex 1 + o",
                                 "loss_first": hist[0], "loss_last": hist[-1], "solves": rows})
    return out


def cpu_sweep_child(name: str, freq: float) -> dict:
    """Run cpu_sweep_baseline in a child process that never touches the GPU (MKL's
    threads and allocator then run as they do in a plain host program)."""
    import subprocess
    env = dict(os.environ)
    cores = str(min(len(os.sched_getaffinity(0)), 16))
    env.update(OMP_NUM_THREADS=cores, MKL_NUM_THREADS=cores)
    cp = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-sweep-child", name, "--cpu-sweep-freq", str(float(freq))],
                        env=env, stdout=subprocess.PIPE, text=True)
    if cp.returncode != 0:
        raise RuntimeError(f"host-core sweep baseline failed ({cp.returncode})")
    return json.loads(cp.stdout.strip().splitlines()[-1])


def cpu_sweep_baseline(name, mesh, phys, freq, desc):
    """The reference CPU path for one frequency: oracle (C/OpenMP) assembly,
    MKL PARDISO on the real block system (the reference's x86 default solver),
    both port solves and S-parameters -- host cores.

    WR-90: the same problem.  Posts box: its PARDISO factorisation needs
    ~120 GB and minutes, so the same geometry is timed at 40x20x8 and 60x30x12
    hexes and the time per frequency is extrapolated to 80x40x16 along the
    power law through the two points (in the number of real unknowns)."""
    from oracle import fd
    from emerge_amd.fdsolve import pec_dofs
    cores = min(len(os.sched_getaffinity(0)), 16)

    def one(m, ph):
        from emerge_amd.bc import LumpedPort
        from emerge_amd.sparam import lumped_port_line
        for p in ph.ports():
            if isinstance(p, LumpedPort) and p.vintline is None:
                lumped_port_line(m, p)
        er, ur = ph.material_tensors()
        t0 = time.perf_counter()
        prob = fd.OracleProblem(m, er, ur, ph.boundary_conditions, solver="pardiso", threads=cores)
        prob.frequency_domain([freq])
        prob.release()
        return time.perf_counter() - t0

    def n_real(m, ph):
        return 2 * (2 * m.n_entities - pec_dofs(m, ph.boundary_conditions).shape[0])

    if name != "posts":
        dt = one(mesh, phys)
        return {"value": 1.0 / dt, "unit": "freq-pts/s", "cores": cores, "kind": "port",
                "sample": f"one frequency ({freq / 1e9:.3f} GHz) of the same problem ({desc}): oracle C/OpenMP "
                          f"assembly, MKL PARDISO factorisation of the real 2N block system ({cores} threads), "
                          f"2 port solves, S-parameters: {dt:.1f} s"}
    from emerge_amd import meshgen
    from emerge_amd.physics import Electrodynamics3D
    pts = []
    for hx in ((40, 20, 8), (60, 30, 12)):
        m = meshgen.posts_box(*hx)
        ph = Electrodynamics3D(m)
        ph.assign(*sweep_ports(m))
        pts.append((n_real(m, ph), one(m, ph), hx))
        print(json.dumps({"cpu_sweep_point": {"hexes": hx, "real_unknowns": pts[-1][0], "seconds": pts[-1][1]}}),
              file=sys.stderr, flush=True)
    n_full = n_real(mesh, phys)
    (n1, t1, h1), (n2, t2, h2) = pts
    expo = np.log(t2 / t1) / np.log(n2 / n1)
    dt = t2 * (n_full / n2) ** expo
    return {"value": 1.0 / dt, "unit": "freq-pts/s", "cores": cores, "kind": "port",
            "sample": f"PARDISO path ({cores} threads) timed on the same geometry at {h1} hexes "
                      f"({n1:,} real unknowns, {t1:.1f} s/frequency) and {h2} ({n2:,}, {t2:.1f} s); "
                      f"extrapolated to {n_full:,} real unknowns with exponent {expo:.2f}: {dt:.0f} s/frequency",
            "extrapolated": True}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=200_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--exact-face-mass", action="store_true")
    ap.add_argument("--two-phase", action="store_true", help="time the per-tet-block fallback kernels")
    ap.add_argument("--sweep", default="posts", choices=["none", "wr90", "posts", "patch"],
                    help="frequency-sweep workload reported beside the assembly figure")
    ap.add_argument("--sweep-freqs", type=int, default=0, help="points of the sweep (0: 201, or 101 for patch)")
    ap.add_argument("--plain-sample", type=int, default=3,
                    help="frequencies solved by plain FP64 GMRES (band ends and middle first)")
    ap.add_argument("--sweep-tol", type=float, default=1e-8)
    ap.add_argument("--no-sweep-cpu", dest="sweep_cpu", action="store_false",
                    help="skip timing the host-core reference on one sweep frequency")
    ap.add_argument("--no-learned", dest="learned", action="store_false",
                    help="skip training / using the learned smoother in the sweep")
    ap.add_argument("--learned-steps", type=int, default=60)
    ap.add_argument("--poly-degree", type=int, default=6,
                    help="degree of the GMRES polynomial around the smoother in the basis sweeps (1 = off)")
    ap.add_argument("--sweep-runs", default="basis,plain,learned",
                    help="comma-separated sweep variants to run: basis, plain, learned")
    ap.add_argument("--no-eigen", dest="eigen", action="store_false", help="skip the cavity eigenmode workload")
    ap.add_argument("--eigen-modes", type=int, default=10)
    ap.add_argument("--eigen-oracle", action="store_true", help="compare the eigenvalues with the CPU oracle (slow)")
    ap.add_argument("--eigen-oracle-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--eigen-sigma", type=float, default=0.0, help=argparse.SUPPRESS)
    ap.add_argument("--eigen-k", type=int, default=10, help=argparse.SUPPRESS)
    ap.add_argument("--sweep-kernel-only", action="store_true",
                    help="measure only the batched residual kernel of the sweep (for counter runs)")
    ap.add_argument("--spmv-kernel-only", action="store_true",
                    help="measure only the sweep's complex64 SpMV kernel (for counter runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 2 GiB device-to-device copy (reference for PMC byte counters)")
    ap.add_argument("--learned-bound", default=None, choices=["posts", "patch"],
                    help="only measure what the learned smoother's loss-optimal weights do to GMRES")
    ap.add_argument("--cpu-sweep-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-sweep-freq", type=float, default=3e9, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.eigen_oracle_child:
        print(json.dumps(eigen_oracle(args.eigen_sigma, args.eigen_k)), flush=True)
        return
    if args.cpu_sweep_child:
        from emerge_amd.physics import Electrodynamics3D
        mesh, ports, _, desc = sweep_workload(args.cpu_sweep_child)
        phys = make_phys(mesh, ports)
        print(json.dumps(cpu_sweep_baseline(args.cpu_sweep_child, mesh, phys, args.cpu_sweep_freq, desc)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the frequency-sharded "
              "multi-GPU design is described in the docs but not run here", file=sys.stderr)
        sys.exit(2)

    if args.learned_bound:
        res = learned_bound(args.learned_bound, lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps({"learned_bound": res}), flush=True)
        return

    import torch
    from emerge_amd import meshgen
    from emerge_amd.assembly import GPUPattern, EXACT_FACE_MASS

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    mesh, eps = meshgen.assembly_box()
    t_mesh = time.perf_counter() - t0
    # the same topology built on the GPU (emg_topology_build), twice: the second without first-call costs
    from emerge_amd.mesh import Mesh3D
    for _ in range(2):
        t_topo_dev = Mesh3D(mesh.nodes, mesh.tets, topology="device").topology_seconds
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pat = GPUPattern(mesh, device="cuda:0")
    torch.cuda.synchronize()
    t_sym = time.perf_counter() - t0
    eps_d = torch.from_numpy(eps.astype(np.complex128)).to(dev)
    nu_d = torch.ones(mesh.n_tets, dtype=torch.complex128, device=dev)
    E = torch.empty(pat.nnz, dtype=torch.complex128, device=dev)
    B = torch.empty(pat.nnz, dtype=torch.complex128, device=dev)
    from emerge_amd.assembly import FORCE_TWO_PHASE
    flags = (EXACT_FACE_MASS if args.exact_face_mass else 0) | (FORCE_TWO_PHASE if args.two_phase else 0)

    # the pattern's first fused assembly tunes its launch grid (emg_pattern_assembly_tuning):
    # done here, before the warmup, so it never falls in the timed steps
    pat.assemble_iso_device(eps_d, nu_d, flags, E, B)
    for _ in range(args.warmup):
        pat.assemble_iso_device(eps_d, nu_d, flags, E, B)
    torch.cuda.synchronize()
    # kernel time with HIP events on the stream the kernel is launched on
    # (torch's current stream, which the C ABI receives)
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record(stream)
        pat.assemble_iso_device(eps_d, nu_d, flags, E, B)
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    ms_per_step = wall / args.steps * 1e3

    if args.copy_calibration:
        src = torch.empty(2 << 30, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    nt, nv = mesh.n_tets, mesh.n_nodes
    alg_bytes = 2 * pat.nnz * 16 + nt * (16 + 32) + nv * 24
    achieved_gbs = alg_bytes / (kern_ms * 1e-3) / 1e9
    flops = FLOP_PER_TET * nt
    out = {
        "metric": METRIC,
        "value": nt / (ms_per_step * 1e-3) / 1e6,
        "unit": "Mtet/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (structured Kuhn mesh, seeded jitter and materials)",
        "config": {"workload": "assembly: unit box 50x50x67 hex x 6 tets = 1,005,000 tets, jitter 0.1 seed 1, "
                               "NED2 curl-curl + mass, complex128 CSR",
                   "tets": nt, "dofs": pat.n, "nnz_per_matrix": pat.nnz,
                   "symbolic_pattern_s": round(t_sym, 3), "host_mesh_s": round(t_mesh, 3),
                   "topology_host_s": round(mesh.topology_seconds, 3), "topology_device_s": round(t_topo_dev, 3),
                   "face_mass": "exact" if args.exact_face_mass else "reference",
                   "assembly_tuning": pat.assembly_tuning()},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS,
                     "traffic": ASSEMBLY_TRAFFIC_BYTES if not (args.two_phase or args.exact_face_mass) else None,
                     "kernel": "k_tet_blocks + k_gather_iso" if args.two_phase else "k_tet_gram + k_assemble_staged",
                     "kernel_ms": kern_ms, "alg_bytes": alg_bytes,
                     "fp64_tflops": flops / (kern_ms * 1e-3) / 1e12,
                     "fp64_frac": flops / (kern_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(mesh, eps, args.cpu_sample)
    if args.sweep != "none":
        del E, B, pat
        torch.cuda.empty_cache()
        out["sweep"] = bench_sweep(args, dev)
    if args.eigen:
        torch.cuda.empty_cache()
        out["eigen"] = bench_eigen(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
