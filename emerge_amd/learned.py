"""Learned smoother weights: a graph network over the mesh, trained on the GPU.

The hybrid smoother of the GPU GMRES is
    z = BJ(r) + G diag(G^T K G)^{-1} G^T r
(2x2 block Jacobi plus the gradient correction).  This module learns local
complex weights for both parts,
    z = w_e . BJ(r) + G (w_p . diag(G^T K G)^{-1} G^T r),
one w_e per edge/face entity and one w_p per second-order Lagrange node, with a
graph network over the mesh vertices whose messages run along the mesh edges:

    H0 = X W0 + b0                                   X: 8 features per vertex
    H_{l+1} = relu(mean_{u~v} H_l[u] W1_l + H_l W2_l + b_l) + H_l     (4 layers, width 64)
    Hp = H_4 Wp,  ordered readouts per vertex / edge (a<b) / face (a<b<c)

Features are local and frequency-aware (k0 times the mean incident edge
length), so one trained network serves every frequency of a sweep.  Layer
weights are float16 on the device with float32 accumulation; float32 master
copies are updated by Adam.  Forward and backward run in this library's own
kernels (emg_gnn_*); the loss is the preconditioned residual after the best
global rescaling (GMRES is blind to the smoother's overall scale, so the loss is
too):
    J = sum_s min_alpha ||r_s - alpha K z_s||^2 / ||r_s||^2
over residuals r_s that the plain solver produces at a few GMRES iteration
counts (each paired with the converged solution's error x* - x_k), so training
targets the components that plain GMRES has not yet removed.  The weights
fold into the smoother's tables once per frequency (emg_solver_set_weights):
a GMRES iteration costs the same with or without the network, and a network
that does not help is simply not used (the plain smoother is the fallback).
"""
from __future__ import annotations


import numpy as np

from . import _lib
from .bc import C0, PEC, RobinBC

FEATURES = 8
PROJ = 16


class LearnedSmoother:
    def __init__(self, problem, hidden: int = 64, layers: int = 4, seed: int = 0, mode: int = 2):
        """mode 2: real weights exp(0.5 tanh(o)) in [0.61, 1.65] (default); mode 1: exp(o);
        mode 0: complex weights 1 + o."""
        import torch
        self.mode = mode
        self.torch = torch
        self.p = problem
        self.L = _lib.load()
        self.hidden, self.layers = hidden, layers
        mesh, dev = problem.mesh, problem.dev
        self.dev = dev
        nv = mesh.n_nodes
        self.nv, self.ne, self.nf = nv, mesh.n_edges, mesh.n_tris
        # vertex adjacency along mesh edges (both directions), mean aggregation
        a, b = mesh.edges
        src = np.concatenate([a, b])
        dst = np.concatenate([b, a])
        order = np.lexsort((dst, src))
        src, dst = src[order], dst[order]
        deg = np.bincount(src, minlength=nv)
        ptr = np.zeros(nv + 1, dtype=np.int64)
        ptr[1:] = np.cumsum(deg)
        d = lambda x, dt: torch.from_numpy(np.ascontiguousarray(x).astype(dt)).to(dev)
        self.adj_ptr, self.adj_idx = d(ptr, np.int32), d(dst, np.int32)
        self.inv_deg = d(1.0 / np.maximum(deg, 1), np.float32)
        self.ev_edge = d(mesh.edges.T, np.int32)
        self.ev_face = d(mesh.tris.T, np.int32)
        self.ev_vert = d(np.arange(nv), np.int32)
        self.deg = deg
        # static per-vertex geometry: mean incident edge length, boundary flags, material
        L = mesh.edge_lengths
        self.hbar = (np.bincount(a, L, nv) + np.bincount(b, L, nv)) / np.maximum(deg, 1)
        self.on_pec = np.zeros(nv)
        self.on_port = np.zeros(nv)
        for bc in problem.bcs:
            tri = mesh.get_triangles(bc.tags) if bc.tags else np.zeros(0, np.int64)
            if isinstance(bc, PEC):
                self.on_pec[np.unique(mesh.tris[:, tri])] = 1
            elif isinstance(bc, RobinBC):
                self.on_port[np.unique(mesh.tris[:, tri])] = 1
        er = problem.er[0, 0]
        cnt = np.bincount(mesh.tets.ravel(), minlength=nv)
        self.eps_re = np.bincount(mesh.tets.ravel(), np.repeat(er.real[None], 4, 0).ravel(), nv) / np.maximum(cnt, 1)
        self.eps_im = np.bincount(mesh.tets.ravel(), np.repeat(-er.imag[None], 4, 0).ravel(), nv) / np.maximum(cnt, 1)
        # parameters (float32 masters; float16 device copies for the kernels)
        rng = np.random.default_rng(seed)
        H = hidden
        shapes = {"W0": (FEATURES, H), "b0": (H,), "Wp": (H, PROJ), "c_v": (2,), "c_e": (4,), "c_f": (2,)}
        for l in range(layers):
            shapes.update({f"W1_{l}": (H, H), f"W2_{l}": (H, H), f"b_{l}": (H,)})
        self.params = {}
        for k, s in shapes.items():
            if k.startswith("W"):
                fan_in = s[0]
                scale = 0.0 if k == "Wp" else np.sqrt(1.0 / fan_in) * (0.5 if k.startswith("W1") or k.startswith("W2") else 1.0)
                v = rng.standard_normal(s) * scale
            else:
                v = np.zeros(s)
            self.params[k] = torch.tensor(v, dtype=torch.float32, device=dev)
        self.grads = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self._half = {}
        self.refresh()
        G = problem.gradient()
        self.G = G
        nodes = G.nodes
        self.node_is_vertex = torch.from_numpy(nodes < nv).to(dev)
        self.node_vertex = d(np.where(nodes < nv, nodes, 0), np.int64)
        self.node_edge = d(np.where(nodes >= nv, nodes - nv, 0), np.int64)
        self.trained = False

    # ------------------------------------------------------------ plumbing
    def refresh(self):
        """float16 device copies of the layer weights."""
        for k, v in self.params.items():
            if k.startswith("W"):
                self._half[k] = v.to(self.torch.float16).contiguous()

    def state(self) -> dict:
        return {k: v.detach().cpu().numpy().copy() for k, v in self.params.items()}

    def features(self, freq: float):
        k0 = 2 * np.pi * freq / C0
        kh = 10.0 * k0 * self.hbar                         # ~0.1-1 for meshes at 10-60 points per wavelength
        X = np.stack([np.ones(self.nv), self.on_pec, self.on_port, kh, kh * kh, self.eps_re - 1.0, self.eps_im,
                      self.deg / 16.0], axis=1)
        return self.torch.from_numpy(X.astype(np.float32)).to(self.dev)

    def _buf(self, *shape):
        return self.torch.empty(shape, dtype=self.torch.float32, device=self.dev)

    # ------------------------------------------------------------ network
    def forward(self, X):
        L, P, st, H = self.L, _lib.ptr, _lib.stream_ptr(), self.hidden
        nv = self.nv
        c = {"X": X, "H": [], "Y": [], "P": []}
        H0 = self._buf(nv, H)
        _lib.check(L.emg_gnn_gemm(nv, FEATURES, H, P(X), P(self._half["W0"]), P(self.params["b0"]), 0.0, P(H0), st),
                   "emg_gnn_gemm")
        c["H"].append(H0)
        for l in range(self.layers):
            Hl = c["H"][-1]
            Y = self._buf(nv, H)
            _lib.check(L.emg_gnn_agg(nv, H, P(self.adj_ptr), P(self.adj_idx), P(self.inv_deg), None, P(Hl), P(Y), st),
                       "emg_gnn_agg")
            Pl = self._buf(nv, H)
            _lib.check(L.emg_gnn_gemm(nv, H, H, P(Y), P(self._half[f"W1_{l}"]), P(self.params[f"b_{l}"]), 0.0, P(Pl),
                                      st), "emg_gnn_gemm")
            _lib.check(L.emg_gnn_gemm(nv, H, H, P(Hl), P(self._half[f"W2_{l}"]), None, 1.0, P(Pl), st),
                       "emg_gnn_gemm")
            Hn = self._buf(nv, H)
            _lib.check(L.emg_gnn_relu(nv * H, P(Pl), P(Hl), P(Hn), None, None, st), "emg_gnn_relu")
            c["Y"].append(Y)
            c["P"].append(Pl)
            c["H"].append(Hn)
        Hp = self._buf(nv, PROJ)
        _lib.check(L.emg_gnn_gemm(nv, H, PROJ, P(c["H"][-1]), P(self._half["Wp"]), None, 0.0, P(Hp), st),
                   "emg_gnn_gemm")
        c["Hp"] = Hp
        Ov, Oe, Of = self._buf(nv, 2), self._buf(self.ne, 4), self._buf(self.nf, 2)
        for (n, K, C, ev, off, bias, O) in ((nv, 1, 2, self.ev_vert, 0, "c_v", Ov), (self.ne, 2, 4, self.ev_edge, 2, "c_e", Oe),
                                            (self.nf, 3, 2, self.ev_face, 10, "c_f", Of)):
            _lib.check(L.emg_gnn_readout(n, K, C, P(ev), P(Hp), PROJ, off, P(self.params[bias]), P(O), None, None,
                                         None, st), "emg_gnn_readout")
        self.cache = c
        return Ov, Oe, Of

    def backward(self, dOv, dOe, dOf):
        """Accumulates parameter gradients into self.grads (float32)."""
        L, P, st, H = self.L, _lib.ptr, _lib.stream_ptr(), self.hidden
        c, g = self.cache, self.grads
        nv = self.nv
        dHp = self.torch.zeros((nv, PROJ), dtype=self.torch.float32, device=self.dev)
        for (n, K, C, ev, off, bias, dO) in ((nv, 1, 2, self.ev_vert, 0, "c_v", dOv), (self.ne, 2, 4, self.ev_edge, 2, "c_e", dOe),
                                             (self.nf, 3, 2, self.ev_face, 10, "c_f", dOf)):
            _lib.check(L.emg_gnn_readout(n, K, C, P(ev), None, PROJ, off, None, None, P(dO.contiguous()), P(dHp),
                                         P(g[bias]), st), "emg_gnn_readout")
        _lib.check(L.emg_gnn_gemm_tn(nv, H, PROJ, P(c["H"][-1]), P(dHp), P(g["Wp"]), None, st), "emg_gnn_gemm_tn")
        dH = self._buf(nv, H)
        _lib.check(L.emg_gnn_gemm_nt(nv, PROJ, H, P(dHp), P(self._half["Wp"]), 0.0, P(dH), st), "emg_gnn_gemm_nt")
        for l in reversed(range(self.layers)):
            Hl, Y, Pl = c["H"][l], c["Y"][l], c["P"][l]
            dP = self._buf(nv, H)
            _lib.check(L.emg_gnn_relu(nv * H, P(Pl), None, None, P(dH), P(dP), st), "emg_gnn_relu")
            _lib.check(L.emg_gnn_gemm_tn(nv, H, H, P(Y), P(dP), P(g[f"W1_{l}"]), P(g[f"b_{l}"]), st), "emg_gnn_gemm_tn")
            _lib.check(L.emg_gnn_gemm_tn(nv, H, H, P(Hl), P(dP), P(g[f"W2_{l}"]), None, st), "emg_gnn_gemm_tn")
            # dH_prev = dH + dP W2^T + Agg^T (dP W1^T)
            _lib.check(L.emg_gnn_gemm_nt(nv, H, H, P(dP), P(self._half[f"W2_{l}"]), 1.0, P(dH), st), "emg_gnn_gemm_nt")
            Q = self._buf(nv, H)
            _lib.check(L.emg_gnn_gemm_nt(nv, H, H, P(dP), P(self._half[f"W1_{l}"]), 0.0, P(Q), st), "emg_gnn_gemm_nt")
            AQ = self._buf(nv, H)
            _lib.check(L.emg_gnn_agg(nv, H, P(self.adj_ptr), P(self.adj_idx), None, P(self.inv_deg), P(Q), P(AQ), st),
                       "emg_gnn_agg")
            dH += AQ
        _lib.check(L.emg_gnn_gemm_tn(nv, FEATURES, H, P(c["X"]), P(dH), P(g["W0"]), P(g["b0"]), st), "emg_gnn_gemm_tn")

    # ------------------------------------------------------------ weights
    def readouts_to_weights(self, Ov, Oe, Of):
        """Complex128 (w_e per entity, w_p per kept P2 node) and the float32 readouts they came from."""
        torch, L, P, st = self.torch, self.L, _lib.ptr, _lib.stream_ptr()
        Oent = torch.cat([Oe[:, 0:2], Of]).contiguous()
        Onode = torch.where(self.node_is_vertex[:, None], Ov[self.node_vertex], Oe[self.node_edge, 2:4]).contiguous()
        we = torch.empty(Oent.shape[0], dtype=torch.complex128, device=self.dev)
        wp = torch.empty(Onode.shape[0], dtype=torch.complex128, device=self.dev)
        _lib.check(L.emg_gnn_weights(Oent.shape[0], self.mode, P(Oent), P(we), None, None, st), "emg_gnn_weights")
        _lib.check(L.emg_gnn_weights(Onode.shape[0], self.mode, P(Onode), P(wp), None, None, st), "emg_gnn_weights")
        self._O = (Oent, Onode)
        return we, wp

    def weights_grad_to_readouts(self, dent, dnode):
        """Scatter d(entity weights), d(node weights) (float32, 2 per row) back to dOv, dOe, dOf."""
        torch = self.torch
        dOv = torch.zeros((self.nv, 2), dtype=torch.float32, device=self.dev)
        dOe = torch.zeros((self.ne, 4), dtype=torch.float32, device=self.dev)
        dOe[:, 0:2] = dent[: self.ne]
        dOf = dent[self.ne:].contiguous()
        iv = self.node_is_vertex
        dOv.index_add_(0, self.node_vertex[iv], dnode[iv])
        dOe[:, 2:4].index_add_(0, self.node_edge[~iv], dnode[~iv])
        return dOv, dOe, dOf

    def weights(self, freq: float):
        Ov, Oe, Of = self.forward(self.features(freq))
        return self.readouts_to_weights(Ov, Oe, Of)

    # ------------------------------------------------------------ training
    def collect(self, solver, freqs, iters=(0, 2, 8, 32), tol: float = 1e-9, probes: int = 2, seed: int = 0):
        """Residual samples of the plain smoother's GMRES at each frequency: r_k = b - K x_k
        after k iterations from zero, for every port, with the pairs' solution errors x* - x_k;
        plus `probes` masked white-noise vectors (their errors K^{-1} r are not needed by the
        loss and are left zero), which keep the whole spectrum of K M in the objective."""
        torch, p = self.torch, self.p
        data = []
        solver.set_weights(None, None)
        gen = torch.Generator(device=self.dev)
        gen.manual_seed(seed)
        for f in freqs:
            solver.set_matrix(p.system(f))
            R, E = [], []
            for port in p.ports:
                b = p.rhs(port, f)
                xs, _ = solver.solve(b, tol=tol, maxiter=20000)
                for k in iters:
                    xk, _ = solver.solve(b, tol=1e-30, maxiter=k) if k > 0 else (torch.zeros_like(b), None)
                    R.append(b - self._apply_K(p.K, xk))
                    E.append(xs - xk)
            for _ in range(probes):
                v = torch.randn(p.N, dtype=torch.complex128, device=self.dev, generator=gen) * p.mask
                R.append(v * (R[0].abs().norm() / v.abs().norm()))
                E.append(torch.zeros_like(v))
            data.append((f, torch.stack(R), torch.stack(E)))
        return data

    def _apply_K(self, vals, x, conj: bool = False):
        p = self.p
        y = self.torch.empty_like(x)
        xin = x.conj().resolve_conj().contiguous() if conj else x
        _lib.check(self.L.emg_spmv(p.N, _lib.ptr(p.pat.rowptr), _lib.ptr(p.pat.cols), _lib.ptr(vals), _lib.ptr(p.mask),
                                   _lib.ptr(xin), _lib.ptr(y), _lib.stream_ptr()), "emg_spmv")
        return y.conj().resolve_conj() if conj else y

    def prepare(self, solver, data):
        """Per frequency: smoother parts u = BJ(r), t = diag(G^T K G)^{-1} G^T r of every sample."""
        torch, p = self.torch, self.p
        out = []
        solver.set_weights(None, None)
        for f, R, E in data:
            K = p.system(f).clone()
            solver.set_matrix(K)
            U = torch.empty_like(R)
            T = torch.empty((R.shape[0], self.G.np2), dtype=torch.complex128, device=self.dev)
            for s in range(R.shape[0]):
                _lib.check(self.L.emg_solver_smoother_parts(solver.h, _lib.ptr(R[s]), _lib.ptr(U[s]), _lib.ptr(T[s]),
                                                            _lib.stream_ptr()), "emg_solver_smoother_parts")
            scale = (1.0 / (R.abs() ** 2).sum(1)).to(torch.float64).contiguous()
            out.append((f, K, R, U, T, scale))
        return out

    def loss_and_grad(self, f, K, R, U, T, scale, accumulate: bool = True, readouts=None):
        """J at frequency f and its gradient into self.grads (network forward + backward).
        With readouts = (Ov, Oe, Of) given, the network is bypassed: J of those readouts
        and their gradient (dOv, dOe, dOf) are returned instead."""
        torch, L, P, st = self.torch, self.L, _lib.ptr, _lib.stream_ptr()
        G = self.G
        Ov, Oe, Of = self.forward(self.features(f)) if readouts is None else readouts
        we, wp = self.readouts_to_weights(Ov, Oe, Of)
        n_ent = we.shape[0]
        W = torch.cat([we, we])
        ns = R.shape[0]
        Gv = torch.empty((ns, self.p.N), dtype=torch.complex128, device=self.dev)
        Gtg = torch.empty((ns, G.np2), dtype=torch.complex128, device=self.dev)
        J = 0.0
        ones = torch.ones(ns, dtype=torch.float64, device=self.dev)
        for s in range(ns):
            wt = (wp * T[s]).contiguous()
            z = W * U[s]
            _lib.check(L.emg_csr_real_apply(self.p.N, P(G.rowptr), P(G.cols), P(G.vals), P(wt), P(z), 1, st),
                       "emg_csr_real_apply")
            v = self._apply_K(K, z)
            a = complex(torch.vdot(v, R[s]))                    # v^H r
            nv = float((v.abs() ** 2).sum())
            rr = 1.0 / float(scale[s])
            res = R[s] - (a / nv) * v
            J += float((res.abs() ** 2).sum()) / rr
            # dJ/d conj(z) = -(conj(a) / (nv rr)) K^H (r - alpha v)   (K complex symmetric)
            Gv[s] = (-(a.conjugate() / (nv * rr))) * self._apply_K(K, res, conj=True)
            _lib.check(L.emg_csr_real_apply(G.np2, P(G.t_rowptr), P(G.t_cols), P(G.t_vals), P(Gv[s]), P(Gtg[s]), 0, st),
                       "emg_csr_real_apply")
        scale = ones
        dent = torch.zeros((n_ent, 2), dtype=torch.float32, device=self.dev)
        dnode = torch.zeros((G.np2, 2), dtype=torch.float32, device=self.dev)
        _lib.check(L.emg_gnn_weight_grad(n_ent, ns, self.p.N, n_ent, P(U), P(Gv), P(scale), P(dent), st),
                   "emg_gnn_weight_grad")
        _lib.check(L.emg_gnn_weight_grad(G.np2, ns, G.np2, 0, P(T), P(Gtg), P(scale), P(dnode), st),
                   "emg_gnn_weight_grad")
        Oent, Onode = self._O
        dOent, dOnode = torch.empty_like(dent), torch.empty_like(dnode)
        _lib.check(L.emg_gnn_weights(n_ent, self.mode, P(Oent), None, P(dent), P(dOent), st), "emg_gnn_weights")
        _lib.check(L.emg_gnn_weights(G.np2, self.mode, P(Onode), None, P(dnode), P(dOnode), st), "emg_gnn_weights")
        if readouts is not None:
            return J, self.weights_grad_to_readouts(dOent, dOnode)
        if not accumulate:
            for v in self.grads.values():
                v.zero_()
        self.backward(*self.weights_grad_to_readouts(dOent, dOnode))
        return J

    def optimal_weights(self, prepared, steps: int = 200, lr: float = 0.02, log=None):
        """The weights that minimise the training loss directly: every entity and node
        readout a free parameter (no network, no features), Adam on the same J and
        gradient.  What any network trained on this loss can at best produce, so its
        GMRES iterations bound what the network can gain.  Returns (we, wp, history)."""
        torch = self.torch
        O = [torch.zeros((self.nv, 2), dtype=torch.float32, device=self.dev),
             torch.zeros((self.ne, 4), dtype=torch.float32, device=self.dev),
             torch.zeros((self.nf, 2), dtype=torch.float32, device=self.dev)]
        prm = [torch.nn.Parameter(o) for o in O]
        opt = torch.optim.Adam(prm, lr=lr)
        hist = []
        for it in range(steps):
            J = 0.0
            gsum = [torch.zeros_like(o) for o in O]
            for item in prepared:
                j, g = self.loss_and_grad(*item, readouts=tuple(p.data for p in prm))
                J += j
                for a, b in zip(gsum, g):
                    a += b
            for p_, g in zip(prm, gsum):
                p_.grad = g
            opt.step()
            hist.append(J)
            if log is not None and (it % 50 == 0 or it == steps - 1):
                log(f"optimal weights step {it}: loss {J:.6g}")
        we, wp = self.readouts_to_weights(*(p.data for p in prm))
        return we, wp, hist

    def train(self, prepared, steps: int = 150, lr: float = 2e-3, log=None):
        torch = self.torch
        masters = {k: torch.nn.Parameter(v) for k, v in self.params.items()}
        opt = torch.optim.Adam(list(masters.values()), lr=lr)
        hist = []
        for it in range(steps):
            for v in self.grads.values():
                v.zero_()
            J = 0.0
            for item in prepared:
                J += self.loss_and_grad(*item)
            for k, prm in masters.items():
                prm.grad = self.grads[k].clone()
            opt.step()
            for k, prm in masters.items():
                self.params[k] = prm.data
            self.refresh()
            hist.append(J)
            if log is not None and (it % 25 == 0 or it == steps - 1):
                log(f"train step {it}: loss {J:.6g}")
        self.trained = True
        return hist
