"""FP64 numpy restatement of the learned-proposal network (test infrastructure).

TEST INFRASTRUCTURE ONLY: the checker for the device MLP kernels.  The
reference has no learned component; this follows the engine's own definition:
n_in -> H -> H -> n_out with ReLU, parameters packed column-major as
W1 [H x n_in], b1, W2 [H x H], b2, W3 [n_out x H], b3; inputs standardised as
(x - x_mean)/x_std, outputs as y' = (y - y_mean)/y_std; loss = mean over rows
and outputs of (net(x') - y')^2.  The backward pass is written out by hand and
pinned against central finite differences in the tests.
"""
import numpy as np


def unpack(p, nin, H, nout):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    out = []
    for shape in [(H, nin), (H,), (H, H), (H,), (nout, H), (nout,)]:
        size = int(np.prod(shape))
        a = p[o:o + size]
        out.append(a.reshape(shape[::-1]).T if len(shape) == 2 else a.copy())
        o += size
    if o != p.size:
        raise ValueError("parameter count does not match the sizes")
    return out


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.T.reshape(-1), b1, W2.T.reshape(-1), b2, W3.T.reshape(-1), b3])


def num_params(nin, H, nout):
    return H * nin + H + H * H + H + nout * H + nout


def forward(p, X, nin, H, nout, scale=None):
    """Network outputs (de-standardised) for row-major X [n x nin]."""
    W1, b1, W2, b2, W3, b3 = unpack(p, nin, H, nout)
    xm, xs, ym, ys = _scale(scale, nin, nout)
    Xs = (np.asarray(X, dtype=np.float64) - xm) / xs
    h1 = np.maximum(Xs @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    o = h2 @ W3.T + b3
    return o * ys + ym


def loss_grad(p, X, Y, nin, H, nout, scale=None):
    """Loss and its gradient (packed like p)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, nin, H, nout)
    xm, xs, ym, ys = _scale(scale, nin, nout)
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    n = X.shape[0]
    Xs = (X - xm) / xs
    Ys = (Y - ym) / ys
    z1 = Xs @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    o = h2 @ W3.T + b3
    d = o - Ys
    loss = float(np.mean(d * d))
    do = 2.0 * d / (n * nout)
    gW3 = do.T @ h2
    gb3 = do.sum(0)
    dh2 = (do @ W3) * (z2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (z1 > 0)
    gW1 = dh1.T @ Xs
    gb1 = dh1.sum(0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def loss_only(p, X, Y, nin, H, nout, scale=None):
    xm, xs, ym, ys = _scale(scale, nin, nout)
    o = (forward(p, X, nin, H, nout, scale) - ym) / ys
    d = o - (np.asarray(Y, dtype=np.float64) - ym) / ys
    return float(np.mean(d * d))


def _scale(scale, nin, nout):
    if scale is None:
        return np.zeros(nin), np.ones(nin), np.zeros(nout), np.ones(nout)
    xm, xs, ym, ys = (np.asarray(a, dtype=np.float64) for a in scale)
    return xm, xs, ym, ys


def proposal(pred, grid, mix):
    """The learned proposal over the grid for one prediction of 16 labels:
    q = (1-mix) qhat/sum qhat + mix/N, qhat = max(0, x^T P x) with
    P[l, r] = pred[l + 4 r]; uniform when every qhat is 0."""
    P = np.asarray(pred, dtype=np.float64).reshape(4, 4).T
    g = np.asarray(grid, dtype=np.float64).reshape(-1, 4)
    qh = np.maximum(np.einsum("gl,lr,gr->g", g, P, g), 0.0)
    s = qh.sum()
    if s > 0:
        return (1.0 - mix) * qh / s + mix / g.shape[0]
    return np.full(g.shape[0], 1.0 / g.shape[0])
