"""Electrodynamics3D: the frequency-domain / eigenmode driver (fem/physics/edm/emfreq3d.py).

Same call pattern as the reference -- set_frequency(), assign(*bcs),
frequency_domain() -> EMSimData, plus the 3D eigenmode() the reference has
commented out (emfreq3d.py:815-858) -- on an emerge_amd Mesh3D with materials
per volume tag instead of a gmsh session.
"""
from __future__ import annotations

import time

import numpy as np

from .bc import AIR, C0, PEC, LumpedPort, Material, PortBC
from .emdata import EMSimData, SolveRecord
from .mesh import Mesh3D


class SimulationError(Exception):
    """emfreq3d.py:37-38."""


class GPUSolveSettings:
    """Solver knobs (the reference's SolveRoutine flags, fem/solver.py:361-402, for the GPU path).

    precision: 64 (complex128 GMRES) or 32 (complex64 GMRES in FP64 refinement)
    tol: FP64 relative residual target ||b - K x|| / ||b|| over free DOFs
    max_refine: cap on FP64 refinement cycles for precision 32 (0 = no cap besides maxiter)
    sweep: "basis" (trial solutions from the span of earlier solutions, certified
           by batched FP64 residuals, GMRES where a trial fails) or "plain"
           (GMRES for every frequency and port, warm-started from the previous
           frequency)
    max_basis: cap on the number of stored basis vectors for sweep="basis"
    preconditioner: "hybrid" (2x2 block Jacobi + gradient-space correction) or "jacobi"
    basis_tol: GMRES solves that enlarge the basis go to tol * basis_tol
    verbose: one progress line per block of the basis sweep on stderr
    learned: a trained emerge_amd.learned.LearnedSmoother (None = plain hybrid smoother)
    poly_degree: d > 1 wraps the smoother in the degree-d GMRES polynomial of K M,
                 recomputed per frequency (emerge_amd.fdsolve.poly_coefficients); 1 = off
    """

    def __init__(self, precision: int = 32, tol: float = 1e-8, restart: int = 150, maxiter: int = 60000,
                 max_refine: int = 0, warm_start: bool = True, store_fields: bool = True, sweep: str = "basis",
                 max_basis: int = 128, preconditioner: str = "hybrid", basis_tol: float = 1e-1,
                 verbose: bool = False, learned=None, poly_degree: int = 6):
        if sweep not in ("basis", "plain"):
            raise ValueError(f"sweep must be 'basis' or 'plain', not {sweep!r}")
        self.sweep = sweep
        self.max_basis = max_basis
        self.preconditioner = preconditioner
        self.basis_tol = basis_tol
        self.verbose = verbose
        self.learned = learned
        self.poly_degree = poly_degree
        self.precision = precision
        self.tol = tol
        self.restart = restart
        self.maxiter = maxiter
        self.max_refine = max_refine
        self.warm_start = warm_start
        self.store_fields = store_fields


class Electrodynamics3D:
    """Time-harmonic Nedelec-2 physics on one GPU (emfreq3d.py:74-807)."""

    def __init__(self, mesh: Mesh3D, order: int = 2, materials: dict[int, Material] | None = None):
        self.frequencies: list[float] = []
        self.mesh = mesh
        self.set_order(order)
        self.materials: dict[int, Material] = dict(materials or {})
        self.solveroutine = GPUSolveSettings()
        self.exact_face_mass = False
        self.reference_geometry = False     # reference's absolute-coordinate element geometry (its rounding)
        self.data: EMSimData | None = None
        self._problem = None
        self._problem_key = None
        self._bc_initialized = False
        self.boundary_conditions = []
        self._initialize_bcs()

    def set_order(self, order: int) -> None:                                      # emfreq3d.py:100-114
        if order not in (2,):
            raise ValueError(f"Order {order} not supported. Only order-2 allowed.")
        self.order = order
        self.resolution = 0.3

    def _initialize_bcs(self) -> None:                                           # emfreq3d.py:137-147
        self.boundary_conditions = [PEC(sorted(self.mesh.face_tags.keys()))]

    def set_frequency(self, frequency) -> None:                                  # emfreq3d.py:149-164
        if isinstance(frequency, (tuple, list, np.ndarray)):
            self.frequencies = [float(f) for f in frequency]
        else:
            self.frequencies = [float(frequency)]

    def set_material(self, vol_tags, material: Material) -> None:
        for t in ([vol_tags] if np.isscalar(vol_tags) else vol_tags):
            self.materials[int(t)] = material

    @property
    def nports(self) -> int:
        return len([bc for bc in self.boundary_conditions if isinstance(bc, PortBC)])

    def ports(self) -> list[PortBC]:
        return sorted([bc for bc in self.boundary_conditions if isinstance(bc, PortBC)], key=lambda b: b.port_number)

    def assign(self, *bcs) -> None:                                              # emfreq3d.py:782-807
        self._bc_initialized = True
        for bc in bcs:
            for existing in self.boundary_conditions:
                existing.exclude_bc(bc)
            self.boundary_conditions.append(bc)
        self._problem = None

    def material_tensors(self):
        """(3,3,nt) complex eps_r and mu_r per tet: each volume tag's material evaluated at its
        tet centroids through fer3d_mat / fur3d_mat (fem/mesh3d.py:358-378,
        fem/material.py:91-103), so callable and tensor materials (PML) work as in the reference."""
        nt = self.mesh.n_tets
        er = np.zeros((3, 3, nt), dtype=np.complex128)
        ur = np.zeros((3, 3, nt), dtype=np.complex128)
        tags = self.mesh.tet_tags
        xs, ys, zs = self.mesh.centers
        for tag in np.unique(tags):
            mat = self.materials.get(int(tag), AIR)
            ids = np.flatnonzero(tags == tag)
            er[:, :, ids] = mat.fer3d_mat(xs[ids], ys[ids], zs[ids])
            ur[:, :, ids] = mat.fur3d_mat(xs[ids], ys[ids], zs[ids])
        return er, ur

    def locator(self):
        """Point locator of the mesh for EMDataSet.interpolate (built on first use)."""
        from .fields import LazyLocator
        if getattr(self, "_locator", None) is None or self._locator.mesh is not self.mesh:
            self._locator = LazyLocator(self.mesh)
        return self._locator

    def problem(self):
        """The device-resident frequency problem (built once per mesh/material/BC set)."""
        from .fdsolve import FrequencyProblem
        key = (id(self.mesh), tuple(sorted((k, v.key()) for k, v in self.materials.items())),
               tuple(id(b) for b in self.boundary_conditions), self.exact_face_mass, self.reference_geometry)
        if self._problem is None or self._problem_key != key:
            er, ur = self.material_tensors()
            for p in self.ports():
                if isinstance(p, LumpedPort) and p.vintline is None:
                    from .sparam import lumped_port_line
                    lumped_port_line(self.mesh, p)
            self._problem = FrequencyProblem(self.mesh, er, ur, self.boundary_conditions,
                                             exact_face_mass=self.exact_face_mass,
                                             reference_geometry=self.reference_geometry)
            self._problem_key = key
        return self._problem

    def frequency_domain(self) -> EMSimData:
        """Sweep all frequencies and ports on the GPU (emfreq3d.py:607-732)."""
        if self._bc_initialized is False:
            raise SimulationError("Cannot run a modal analysis because no boundary conditions have been assigned.")
        if not self.frequencies:
            raise SimulationError("No frequencies set (set_frequency).")
        import torch
        from .fdsolve import GPUSolver
        from .sparam import PortSampler
        st = self.solveroutine
        prob = self.problem()
        sampler = PortSampler(prob)
        solver = GPUSolver(prob, restart=st.restart, precision=st.precision, preconditioner=st.preconditioner,
                           poly_degree=st.poly_degree)
        if st.learned is not None:
            if st.preconditioner != "hybrid":
                raise SimulationError("a learned smoother needs preconditioner='hybrid'")
            solver.learned = st.learned
        ports = prob.ports
        nums = [p.port_number for p in ports]
        data = EMSimData(basis=self.locator())
        er, ur = prob.er, prob.ur
        prev: dict[int, object] = {}
        t0 = time.perf_counter()
        label = f"gmres-c{64 if st.precision == 32 else 128}"
        if st.sweep == "basis":
            from .sweep import BasisSweep
            bs = BasisSweep(prob, solver, sampler, tol=st.tol, max_basis=st.max_basis, maxiter=st.maxiter,
                            max_refine=st.max_refine, basis_tol=st.basis_tol, verbose=st.verbose)
            res = bs.run(list(self.frequencies), label)
            for fi, freq in enumerate(self.frequencies):
                k0 = 2 * np.pi * freq / C0
                ds = data.new(freq=freq, k0=k0)
                ds.init_sp(nums)
                ds.er = np.squeeze(er[0, 0, :])
                ds.ur = np.squeeze(ur[0, 0, :])
                for p in ports:
                    ds.add_port_properties(p.port_number, p.mode_number, k0, p.get_beta(k0), p.Z0, p.power)
                ys = [res[fi][i][0] for i in range(len(ports))]
                ds.solves.extend(res[fi][i][1] for i in range(len(ports)))
                S = sampler.s_matrix({p.port_number: bs.fields_of(ys[i]) for i, p in enumerate(ports)}, k0)
                ds.Sp.arry[:, :] = S
                if st.store_fields:
                    for i, p in enumerate(ports):
                        ds._fields[p.port_number] = bs.vector_of(ys[i]).cpu().numpy()
                ds.set_field_vector()
            torch.cuda.synchronize()
            data.info["sweep_seconds"] = time.perf_counter() - t0
            data.info["basis_size"] = bs.m
            data.info["trial_log"] = bs.log
            data.info["learned"] = bs.learned_info
            self.data = data
            return data
        for freq in self.frequencies:
            k0 = 2 * np.pi * freq / C0
            ds = data.new(freq=freq, k0=k0)
            ds.init_sp(nums)
            ds.er = np.squeeze(er[0, 0, :])
            ds.ur = np.squeeze(ur[0, 0, :])
            for p in ports:
                ds.add_port_properties(p.port_number, p.mode_number, k0, p.get_beta(k0), p.Z0, p.power)
            K = prob.system(freq)
            solver.set_matrix(K, freq)
            X = []
            for ap in ports:
                b = prob.rhs(ap, freq)
                x0 = prev.get(ap.port_number) if st.warm_start else None
                x, info = solver.solve(b, x0=x0, tol=st.tol, maxiter=st.maxiter, max_refine=st.max_refine)
                ds.solves.append(SolveRecord(ap.port_number, info.iterations, info.refinements, info.residual,
                                             bool(info.converged), label))
                prev[ap.port_number] = x
                X.append(x)
            Xs = torch.stack(X)
            E = sampler.fields(Xs)
            S = sampler.s_matrix({p.port_number: E[i] for i, p in enumerate(ports)}, k0)
            ds.Sp.arry[:, :] = S
            if st.store_fields:
                for i, p in enumerate(ports):
                    ds._fields[p.port_number] = X[i].cpu().numpy()
            ds.set_field_vector()
        torch.cuda.synchronize()
        data.info["sweep_seconds"] = time.perf_counter() - t0
        self.data = data
        return data

    def modal_analysis(self, port, nmodes: int = 6, direct: bool = True, TEM: bool = False, target_kz=None,
                       freq: float | None = None, face_block: bool = False):
        """The mode(s) of a ModalPort from its face, on the GPU (emfreq3d.py:201-364); fills
        port.modes and returns them.  `direct` is accepted for the reference's signature; the
        GPU search always returns every mode in the reference's window."""
        from .modal import modal_analysis
        modes = modal_analysis(self, port, nmodes=nmodes, TEM=TEM, freq=freq, target_kz=target_kz,
                               face_block=face_block)
        self._problem = None
        return modes

    def eigenmode(self, mesh: Mesh3D | None = None, solver=None, num_sols: int = 6,
                  target_frequency: float | None = None, ncv: int | None = None, inner_tol: float = 1e-10,
                  spurious_rel: float = 1e-3, inner_precision: int = 32) -> EMSimData:
        """Modes of E x = k^2 B x nearest k(target)^2 with PEC walls (emfreq3d.py:815-858).

        Same leading arguments as the reference's (mesh, solver, num_sols); the
        target is the first frequency of set_frequency (emfreq3d.py:837) unless
        target_frequency is given.  mesh: a mesh to solve on instead of the
        physics' own; solver: a GPUSolveRoutine whose preconditioner and
        polynomial settings the inner solves take (None: solveroutine's);
        inner_precision: 64 (complex128 inner GMRES) or 32 (complex64 cycles in FP64
        refinement), both to the FP64 residual inner_tol (1e-10: the Ritz residuals land near
        1e-10 and the eigenvalues, quadratic in the residual, within ~2e-14 of an inner
        tolerance of 1e-12 on the 221,184-tet cavity, in 21 % less time).
        Returns an EMSimData with one dataset per mode: freq (Hz), k0, mode,
        residual, and the mode field as _mode_field; raises EigenError if fewer
        than num_sols modes reach the residual target."""
        from .eigen import ShiftInvertArnoldi, eigen_frequencies
        if mesh is not None and mesh is not self.mesh:
            self.mesh = mesh
            self._problem = None
            if not self._bc_initialized:
                self._initialize_bcs()
        st = solver if solver is not None else self.solveroutine
        f_t = target_frequency if target_frequency is not None else (self.frequencies[0] if self.frequencies else None)
        if f_t is None:
            raise SimulationError("eigenmode needs a target frequency")
        sigma = (2 * np.pi * f_t / C0) ** 2                                      # emfreq3d.py:838-839
        prob = self.problem()
        arn = ShiftInvertArnoldi(prob, sigma, ncv=ncv or max(4 * num_sols + 20, 24), inner_tol=inner_tol,
                                 preconditioner=st.preconditioner, verbose=getattr(st, "verbose", False),
                                 poly_degree=st.poly_degree, precision=inner_precision)
        lam, X, resid = arn.run(num_sols, spurious_rel=spurious_rel)
        data = EMSimData(basis=self.locator())
        for i in range(lam.shape[0]):
            f = float(eigen_frequencies(lam[i].real))
            ds = data.new(freq=f, k0=float(np.sqrt(lam[i].real)), mode=i + 1)
            ds.eigenvalue = complex(lam[i])
            ds.residual = float(resid[i])
            ds._mode_field = X[i].cpu().numpy()
        data.info["inner_iterations"] = list(arn.inner_iters)
        data.info["sigma"] = sigma
        data.info["restarts"] = arn.restarts
        self.data = data
        return data
