"""CPU ORACLE -- test infrastructure only (imported by tests/ and nothing else).

Float64 numpy restatement of the learned-smoother network of
emerge_amd.learned and of its training loss, used to check the GPU forward
and backward kernels.  The network has no counterpart in the reference (which
factorises directly, fem/solver.py:294-309); it is pinned here by central
finite differences of this restatement (tests/test_gnn_oracle.py).

Network (weights rounded to float16 exactly as the device keeps them):
  H0 = X W0 + b0
  H_{l+1} = relu(Agg(H_l) W1_l + H_l W2_l + b_l) + H_l,   Agg = mean over mesh-edge neighbours
  Hp = H_L Wp                                   (16 columns: vertex 2 | edge slots 2x4 | face slots 3x2)
  O_v = Hp[v, 0:2] + c_v
  O_e = Hp[a, 2:6] + Hp[b, 6:10] + c_e          (edge a < b)
  O_f = Hp[a, 10:12] + Hp[b, 12:14] + Hp[c, 14:16] + c_f   (face a < b < c)
Smoother weights: w = exp(0.5 tanh(O[:, 0])) (mode 2, default), exp(O[:, 0]) (mode 1) or
1 + O[:, 0] + j O[:, 1] (mode 0); entity weight w_e from O_e[:, 0:2] (edges) and
O_f (faces); P2-node weight w_p from O_v (vertices) and O_e[:, 2:4] (edge midpoints).
Loss over residual samples r_s (invariant to a global scale of the smoother, as GMRES is):
  J = sum_s min_alpha ||r_s - alpha K z_s||^2 / ||r_s||^2 = sum_s (1 - |<K z_s, r_s>|^2 / (||K z_s||^2 ||r_s||^2)),
z_s = w_e . BJ(r_s) + G (w_p . t_s),  t_s = diag(G^T K G)^{-1} G^T r_s.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

PROJ = 16


def fp16(a):
    return np.asarray(a, dtype=np.float16).astype(np.float64)


class OracleNet:
    def __init__(self, params: dict):
        self.p = {k: (fp16(v) if k.startswith("W") else np.asarray(v, dtype=np.float64)) for k, v in params.items()}
        self.L = sum(1 for k in params if k.startswith("W1_"))

    @staticmethod
    def adjacency(nv, edges):
        a, b = edges
        A = sp.csr_matrix((np.ones(2 * a.shape[0]), (np.concatenate([a, b]), np.concatenate([b, a]))), shape=(nv, nv))
        deg = np.asarray(A.sum(1)).ravel()
        return sp.diags(1.0 / np.maximum(deg, 1)) @ A

    def forward(self, X, Agg, edges, faces):
        p = self.p
        c = {"X": X, "H": [X @ p["W0"] + p["b0"]], "Y": [], "P": []}
        for l in range(self.L):
            H = c["H"][-1]
            Y = Agg @ H
            P = Y @ p[f"W1_{l}"] + H @ p[f"W2_{l}"] + p[f"b_{l}"]
            c["Y"].append(Y)
            c["P"].append(P)
            c["H"].append(np.maximum(P, 0) + H)
        Hp = c["H"][-1] @ p["Wp"]
        c["Hp"] = Hp
        Ov = Hp[:, 0:2] + p["c_v"]
        Oe = Hp[edges[0], 2:6] + Hp[edges[1], 6:10] + p["c_e"]
        Of = Hp[faces[0], 10:12] + Hp[faces[1], 12:14] + Hp[faces[2], 14:16] + p["c_f"]
        self.cache = c
        return Ov, Oe, Of

    def backward(self, dOv, dOe, dOf, Agg, edges, faces):
        p, c = self.p, self.cache
        g = {}
        nv = c["Hp"].shape[0]
        dHp = np.zeros((nv, PROJ))
        dHp[:, 0:2] += dOv
        np.add.at(dHp, (edges[0], slice(2, 6)), dOe)
        np.add.at(dHp, (edges[1], slice(6, 10)), dOe)
        for q, f in enumerate(faces):
            np.add.at(dHp, (f, slice(10 + 2 * q, 12 + 2 * q)), dOf)
        g["c_v"], g["c_e"], g["c_f"] = dOv.sum(0), dOe.sum(0), dOf.sum(0)
        HL = c["H"][-1]
        g["Wp"] = HL.T @ dHp
        dH = dHp @ p["Wp"].T
        for l in reversed(range(self.L)):
            H, Y, P = c["H"][l], c["Y"][l], c["P"][l]
            dP = dH * (P > 0)
            g[f"W1_{l}"] = Y.T @ dP
            g[f"W2_{l}"] = H.T @ dP
            g[f"b_{l}"] = dP.sum(0)
            dH = dH + dP @ p[f"W2_{l}"].T + Agg.T @ (dP @ p[f"W1_{l}"].T)
        g["W0"] = c["X"].T @ dH
        g["b0"] = dH.sum(0)
        return g


def entity_node_weights(Ov, Oe, Of, nodes, nv, mode: int = 2):
    """(w_e per entity, w_p per kept P2 node) from the readouts: mode 1 exp(o_0), mode 0 1 + o_0 + j o_1."""
    Oent = np.concatenate([Oe[:, 0:2], Of])
    Onode = np.concatenate([Ov, Oe[:, 2:4]])[nodes]
    if mode == 1:
        return np.exp(Oent[:, 0]) + 0j, np.exp(Onode[:, 0]) + 0j
    if mode == 2:
        return np.exp(0.5 * np.tanh(Oent[:, 0])) + 0j, np.exp(0.5 * np.tanh(Onode[:, 0])) + 0j
    return 1 + Oent[:, 0] + 1j * Oent[:, 1], 1 + Onode[:, 0] + 1j * Onode[:, 1]


def weights_backward(O, dw, mode: int = 2):
    """d(readout) from d(Re w, Im w)."""
    if mode == 1:
        return np.stack([dw[:, 0] * np.exp(O[:, 0]), np.zeros(O.shape[0])], 1)
    if mode == 2:
        th = np.tanh(O[:, 0])
        return np.stack([dw[:, 0] * np.exp(0.5 * th) * 0.5 * (1 - th * th), np.zeros(O.shape[0])], 1)
    return dw.copy()


def smoother_loss(w_e, w_p, R, U, T, K, G):
    """J and dJ/d(Re, Im) of w_e and w_p (Wirtinger: 2 (Re, Im) conj(a) b)."""
    n_ent = w_e.shape[0]
    J = 0.0
    ge = np.zeros((n_ent, 2))
    gp = np.zeros((w_p.shape[0], 2))
    W = np.concatenate([w_e, w_e])
    for r, u, t in zip(R, U, T):
        rr = np.vdot(r, r).real
        z = W * u + G @ (w_p * t)
        v = K @ z
        a = np.vdot(v, r)                     # v^H r
        nv = np.vdot(v, v).real
        alpha = a / nv
        res = r - alpha * v
        J += np.vdot(res, res).real / rr
        # dJ/d conj(v) = -(conj(a) / (nv rr)) (r - alpha v)
        gvec = -(np.conj(a) / (nv * rr)) * (K.conj().T @ res)
        a2 = np.conj(u) * gvec
        de = a2[:n_ent] + a2[n_ent:]
        dp = np.conj(t) * (G.T @ gvec)
        ge += 2 * np.stack([de.real, de.imag], 1)
        gp += 2 * np.stack([dp.real, dp.imag], 1)
    return J, ge, gp
