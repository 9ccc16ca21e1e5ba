"""The FP64 network restatement (oracle/gnn_oracle.py) pinned by the
mathematics: a hand-computed two-node forward pass, and central finite
differences of the loss for every parameter block."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import gnn_oracle as N


def test_forward_two_nodes_by_hand():
    p = N.init_params(hidden=2, layers=1, seed=0)
    p["W_in"] = np.array([[1.0, 0.0], [0.0, 1.0], [0.0, 0.0], [0.0, 0.0]])
    p["b_in"] = np.array([0.0, 0.5])
    p["Ws0"] = np.eye(2)
    p["Wm0"] = np.array([[0.0, 1.0], [1.0, 0.0]])
    p["Wg0"] = np.zeros((2, 2))
    p["b0"] = np.array([-1.0, 0.0])
    p["w_out"] = np.array([1.0, -1.0])
    p["b_out"] = np.array([0.25])
    Ah = sp.csr_matrix(np.array([[0.0, -0.5], [-0.5, 0.0]]))
    X = np.array([[2.0, 1.0, 0.0, 1.0], [-1.0, 3.0, 0.0, 1.0]])
    out, c = N.forward(Ah, X, p)
    h0 = np.array([[2.0, 1.5], [0.0, 3.5]])
    agg = np.array([[0.0, -1.75], [-1.0, -0.75]])
    pre = h0 @ np.eye(2) + agg @ p["Wm0"] + np.array([-1.0, 0.0])
    # pre = [[2-1.75-1, 1.5+0], [0-0.75-1, 3.5-1]] = [[-0.75, 1.5], [-1.75, 2.5]]
    assert np.allclose(pre, [[-0.75, 1.5], [-1.75, 2.5]])
    h1 = h0 + np.maximum(pre, 0)
    assert np.allclose(out, h1 @ p["w_out"] + 0.25)
    assert np.allclose(out, [2.0 - 3.0 + 0.25, 0.0 - 6.0 + 0.25])


@pytest.mark.parametrize("objective", ["error", "residual"])
def test_backward_matches_finite_differences(objective):
    rng = np.random.default_rng(3)
    n = 40
    M = sp.random(n, n, density=0.1, random_state=4)
    M = M + M.T
    A = sp.diags(np.asarray(abs(M).sum(axis=1)).ravel() + 1.0) - M
    Ah, sd, rowsum = N.scaled_graph(A)
    r = rng.standard_normal(n)
    z = rng.standard_normal(n)
    X, s = N.features(r, z, sd, rowsum)
    p = N.init_params(hidden=8, layers=3, seed=5)
    for k in p:
        if k.startswith("b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.1
    target = rng.standard_normal(n)
    res = (A.tocsr(), sd, r) if objective == "residual" else None
    loss, g, _ = N.loss_and_grad(Ah, X, p, target, res)
    eps = 1e-6
    for k in p:
        flat = p[k].reshape(-1)
        idx = rng.choice(flat.size, size=min(6, flat.size), replace=False)
        for i in idx:
            old = flat[i]
            flat[i] = old + eps
            lp = N.loss_and_grad(Ah, X, p, target, res)[0]
            flat[i] = old - eps
            lm = N.loss_and_grad(Ah, X, p, target, res)[0]
            flat[i] = old
            fd = (lp - lm) / (2 * eps)
            an = g[k].reshape(-1)[i]
            assert abs(fd - an) <= 1e-6 * max(1e-3, abs(fd)), (k, i, fd, an)
