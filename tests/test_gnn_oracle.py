"""CPU oracle of the learned smoother: network backward and loss gradient against central
finite differences (float64), so the GPU kernels can be checked against it."""
import numpy as np
import scipy.sparse as sp

from emerge_amd import meshgen
from oracle.gnn import OracleNet, entity_node_weights, smoother_loss, weights_backward


def _net(nv, F=8, H=16, L=2, seed=0):
    rng = np.random.default_rng(seed)
    p = {"W0": rng.standard_normal((F, H)) * 0.3, "b0": rng.standard_normal(H) * 0.1,
         "Wp": rng.standard_normal((H, 16)) * 0.2, "c_v": rng.standard_normal(2) * 0.1,
         "c_e": rng.standard_normal(4) * 0.1, "c_f": rng.standard_normal(2) * 0.1}
    for l in range(L):
        p[f"W1_{l}"] = rng.standard_normal((H, H)) * 0.2
        p[f"W2_{l}"] = rng.standard_normal((H, H)) * 0.2
        p[f"b_{l}"] = rng.standard_normal(H) * 0.1
    return p


def _mesh():
    m = meshgen.structured_box(2, 2, 1, size=(1.0, 1.0, 0.5), jitter=0.1, seed=1)
    return m, m.edges, m.tris


def test_network_backward_matches_finite_differences():
    m, edges, faces = _mesh()
    nv = m.n_nodes
    rng = np.random.default_rng(1)
    X = rng.standard_normal((nv, 8))
    Agg = OracleNet.adjacency(nv, edges)
    params = _net(nv)
    cv, ce, cf = (rng.standard_normal((nv, 2)), rng.standard_normal((edges.shape[1], 4)),
                  rng.standard_normal((faces.shape[1], 2)))

    def f(p):
        net = OracleNet(p)
        net.p = {k: np.asarray(v, dtype=np.float64) for k, v in p.items()}   # no fp16 rounding for FD
        Ov, Oe, Of = net.forward(X, Agg, edges, faces)
        return np.sum(cv * Ov) + np.sum(ce * Oe) + np.sum(cf * Of), net

    _, net = f(params)
    g = net.backward(cv, ce, cf, Agg, edges, faces)
    h = 1e-6
    for name in ("W0", "b0", "W1_0", "W2_1", "b_1", "Wp", "c_e"):
        idx = tuple(rng.integers(0, s) for s in params[name].shape)
        pp = {k: v.copy() for k, v in params.items()}
        pm = {k: v.copy() for k, v in params.items()}
        pp[name][idx] += h
        pm[name][idx] -= h
        fd = (f(pp)[0] - f(pm)[0]) / (2 * h)
        assert abs(fd - g[name][idx]) <= 1e-6 * max(1.0, abs(fd)), (name, fd, g[name][idx])


def test_smoother_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    n_ent, np2 = 30, 12
    N = 2 * n_ent
    K = sp.random(N, N, density=0.2, random_state=3) + 1j * sp.random(N, N, density=0.2, random_state=4)
    K = (K + K.T + 5 * sp.eye(N)).tocsr()
    G = sp.random(N, np2, density=0.2, random_state=5).tocsr()
    R = [rng.standard_normal(N) + 1j * rng.standard_normal(N) for _ in range(3)]
    U = [rng.standard_normal(N) + 1j * rng.standard_normal(N) for _ in range(3)]
    T = [rng.standard_normal(np2) + 1j * rng.standard_normal(np2) for _ in range(3)]
    we = 1 + 0.1 * (rng.standard_normal(n_ent) + 1j * rng.standard_normal(n_ent))
    wp = 1 + 0.1 * (rng.standard_normal(np2) + 1j * rng.standard_normal(np2))
    J, ge, gp = smoother_loss(we, wp, R, U, T, K, G)
    # invariant to a global scale of the smoother
    assert abs(smoother_loss(3.0 * we, 3.0 * wp, R, U, T, K, G)[0] - J) < 1e-12 * J
    h = 1e-7
    for i in (0, 7, 29):
        for c, d in ((0, 1.0), (1, 1j)):
            wpl, wmi = we.copy(), we.copy()
            wpl[i] += h * d
            wmi[i] -= h * d
            fd = (smoother_loss(wpl, wp, R, U, T, K, G)[0] - smoother_loss(wmi, wp, R, U, T, K, G)[0]) / (2 * h)
            assert abs(fd - ge[i, c]) <= 1e-5 * max(1.0, abs(fd))
    for i in (0, 11):
        for c, d in ((0, 1.0), (1, 1j)):
            wpl, wmi = wp.copy(), wp.copy()
            wpl[i] += h * d
            wmi[i] -= h * d
            fd = (smoother_loss(we, wpl, R, U, T, K, G)[0] - smoother_loss(we, wmi, R, U, T, K, G)[0]) / (2 * h)
            assert abs(fd - gp[i, c]) <= 1e-5 * max(1.0, abs(fd))


def test_zero_readout_gives_unit_weights():
    m, edges, faces = _mesh()
    nodes = np.arange(m.n_nodes + m.n_edges)
    Ov, Oe, Of = np.zeros((m.n_nodes, 2)), np.zeros((m.n_edges, 4)), np.zeros((m.n_tris, 2))
    we, wp = entity_node_weights(Ov, Oe, Of, nodes, m.n_nodes)
    assert np.all(we == 1) and np.all(wp == 1) and we.shape[0] == m.n_entities


def test_exponential_weight_backward():
    """d/d(o) of a function of w = exp(o) through weights_backward."""
    rng = np.random.default_rng(4)
    O = rng.standard_normal((5, 2)) * 0.3
    c = rng.standard_normal(5)
    f = lambda o: np.sum(c * np.exp(o[:, 0]) ** 2)
    dw = np.stack([2 * c * np.exp(O[:, 0]), np.zeros(5)], 1)          # d f / d Re(w)
    g = weights_backward(O, dw, mode=1)
    f2 = lambda o: np.sum(c * np.exp(0.5 * np.tanh(o[:, 0])) ** 2)
    th = np.tanh(O[:, 0])
    g2 = weights_backward(O, np.stack([2 * c * np.exp(0.5 * th), np.zeros(5)], 1), mode=2)
    h = 1e-6
    for i in range(5):
        Op, Om = O.copy(), O.copy()
        Op[i, 0] += h
        Om[i, 0] -= h
        assert abs((f(Op) - f(Om)) / (2 * h) - g[i, 0]) < 1e-6 * max(1, abs(g[i, 0]))
        assert abs((f2(Op) - f2(Om)) / (2 * h) - g2[i, 0]) < 1e-6 * max(1, abs(g2[i, 0]))
    assert np.all(g[:, 1] == 0)
