"""TEST INFRASTRUCTURE ONLY: FP64 numpy restatement of the learned
preconditioner network, used to check the device forward/backward kernels.

The network has no counterpart in the reference (the learned part is new in
this build); this file defines its mathematics, and the finite-difference
test in tests/test_gnn_oracle.py pins the backward pass against it.

Graph = the matrix's off-diagonal pattern.  With d_i = |a_ii| and
sd_i = sqrt(d_i):
    A_hat_ij = a_ij / (sd_i sd_j)            (i != j; zero diagonal)
    r~_i = r_i / sd_i,  s = rms(r~)
    X_i = [r~_i / s,  z_i sd_i / s,  (sum_j a_ij) / a_ii,  1]
          (z = ILU(0)^{-1} r; the row-sum ratio marks Dirichlet rows)
    h0  = relu(X W_in + b_in)
    layer l = 1..L:
        agg = A_hat h ;  g = mean over rows of h
        pre = h W_s + agg W_m + g W_g + b
        h   = h + relu(pre)
    out = h w_out + b_out ;  correction e_i = out_i s / sd_i
Loss on a captured pair (r, e), with t_i = e_i sd_i / s:
    loss = 1 - (out . t)^2 / (|out|^2 |t|^2)
i.e. one minus the squared cosine between output and target.  The solver
applies the correction through a minimal-residual multiple alpha * e, so only
its direction matters; the loss is invariant to the output's scale and sign.
Residual objective (setup "learnedObjective" RESIDUAL): with the correction's
direction c_i = out_i / sd_i and q = A c,
    loss = 1 - (q . r)^2 / (|q|^2 |r|^2)
the part of ||r||^2 that the minimal-residual step alpha c leaves; for a
symmetric A, dloss/dout_i = (A dloss/dq)_i / sd_i.
"""
import numpy as np

D_IN = 4


def init_params(hidden=64, layers=6, seed=0):
    rng = np.random.default_rng(seed)
    H = hidden
    p = {"W_in": rng.standard_normal((D_IN, H)) * np.sqrt(2.0 / D_IN), "b_in": np.zeros(H)}
    for l in range(layers):
        p["Ws%d" % l] = rng.standard_normal((H, H)) * (0.5 / np.sqrt(H))
        p["Wm%d" % l] = rng.standard_normal((H, H)) * (0.5 / np.sqrt(H))
        p["Wg%d" % l] = rng.standard_normal((H, H)) * (0.1 / np.sqrt(H))
        p["b%d" % l] = np.zeros(H)
    p["w_out"] = rng.standard_normal(H) * (1.0 / np.sqrt(H))
    p["b_out"] = np.zeros(1)
    return p


def n_layers(p):
    return sum(1 for k in p if k.startswith("Ws"))


def scaled_graph(A):
    """A (scipy CSR) -> (A_hat off-diagonal CSR, sd, row-sum ratio)."""
    import scipy.sparse as sp

    A = sp.csr_matrix(A)
    diag = A.diagonal()
    sd = np.sqrt(np.abs(diag))
    off = A - sp.diags(diag)
    off.eliminate_zeros()
    Ah = sp.diags(1.0 / sd) @ off @ sp.diags(1.0 / sd)
    rowsum = np.asarray(A.sum(axis=1)).ravel() / diag
    return sp.csr_matrix(Ah), sd, rowsum


def features(r, z, sd, rowsum):
    rt = r / sd
    s = np.sqrt(np.mean(rt * rt))
    X = np.stack([rt / s, z * sd / s, rowsum, np.ones_like(r)], axis=1)
    return X, s


def forward(Ah, X, p):
    relu = lambda v: np.maximum(v, 0.0)  # noqa: E731
    cache = {"X": X, "pre0": X @ p["W_in"] + p["b_in"]}
    h = relu(cache["pre0"])
    hs, aggs, gs, pres = [h], [], [], []
    for l in range(n_layers(p)):
        agg = Ah @ h
        g = h.mean(axis=0)
        pre = h @ p["Ws%d" % l] + agg @ p["Wm%d" % l] + g @ p["Wg%d" % l] + p["b%d" % l]
        h = h + relu(pre)
        aggs.append(agg)
        gs.append(g)
        pres.append(pre)
        hs.append(h)
    out = h @ p["w_out"] + p["b_out"][0]
    cache.update(hs=hs, aggs=aggs, gs=gs, pres=pres)
    return out, cache


def loss_and_grad(Ah, X, p, target, residual=None):
    """Error objective against `target`; with residual = (A, sd, r) the
    residual objective instead (A symmetric)."""
    out, c = forward(Ah, X, p)
    N = len(out)
    if residual is None:
        a = out @ target
        pp = out @ out
        q = target @ target
        loss = 1.0 - a * a / (pp * q)
        dout = (2.0 * a / (pp * q)) * (a * out / pp - target)
    else:
        A, sd, r = residual
        qv = A @ (out / sd)
        a = qv @ r
        pp = qv @ qv
        rr = r @ r
        loss = 1.0 - a * a / (pp * rr)
        dq = (2.0 * a / (pp * rr)) * (a * qv / pp - r)
        dout = (A.T @ dq) / sd
    g = {}
    hL = c["hs"][-1]
    g["w_out"] = hL.T @ dout
    g["b_out"] = np.array([dout.sum()])
    dh = np.outer(dout, p["w_out"])
    for l in reversed(range(n_layers(p))):
        h_prev = c["hs"][l]
        dpre = dh * (c["pres"][l] > 0)
        g["Ws%d" % l] = h_prev.T @ dpre
        g["Wm%d" % l] = c["aggs"][l].T @ dpre
        col = dpre.sum(axis=0)
        g["Wg%d" % l] = np.outer(c["gs"][l], col)
        g["b%d" % l] = col
        dh = dh + dpre @ p["Ws%d" % l].T + Ah.T @ (dpre @ p["Wm%d" % l].T) + (col @ p["Wg%d" % l].T)[None, :] / N
    dpre0 = dh * (c["pre0"] > 0)
    g["W_in"] = X.T @ dpre0
    g["b_in"] = dpre0.sum(axis=0)
    return loss, g, out


def apply(Ah, sd, rowsum, p, r, z):
    """Correction e for residual r (z = ILU^{-1} r)."""
    X, s = features(r, z, sd, rowsum)
    out, _ = forward(Ah, X, p)
    return out * s / sd
