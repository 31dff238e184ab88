"""f64 NumPy restatement of the learned-boundary network (TEST INFRASTRUCTURE
ONLY): forward, mean-squared-error gradient and Adam step, written from the
mathematics of the layout in include/seiswave.h (sw_mlp_*)."""
import numpy as np

NFEAT, NH, NOUT = 72, 64, 36


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    out = {}
    for name, size in (("W1", NFEAT * NH), ("b1", NH), ("W2", NH * NH), ("b2", NH), ("W3", NH * NOUT),
                       ("b3", NOUT), ("in_scale", NFEAT), ("out_scale", NOUT)):
        out[name] = p[o:o + size]
        o += size
    out["W1"] = out["W1"].reshape(NFEAT, NH)
    out["W2"] = out["W2"].reshape(NH, NH)
    out["W3"] = out["W3"].reshape(NH, NOUT)
    return out


def forward(p, X):
    q = unpack(p)
    A = np.asarray(X, np.float64) / q["in_scale"]
    H1 = np.tanh(A @ q["W1"] + q["b1"])
    H2 = np.tanh(H1 @ q["W2"] + q["b2"])
    Y = H2 @ q["W3"] + q["b3"]
    return H1, H2, Y


def band_features(cur, prev):
    """Features of band columns: [cur, cur - prev] (72 per column)."""
    cur = np.asarray(cur, np.float64)
    return np.concatenate([cur, cur - np.asarray(prev, np.float64)], 1)


def column_scale(p, X):
    """RMS over the 72 channel-normalised features of each column."""
    u = np.asarray(X, np.float64) / unpack(p)["in_scale"]
    return np.sqrt(np.mean(u * u, 1))


def correction(p, X):
    """The fused kernel's correction: s * out_scale * net(X / s), 0 where s = 0."""
    s = column_scale(p, X)
    safe = np.where(s > 0, s, 1.0)
    y = forward(p, np.asarray(X, np.float64) / safe[:, None])[2] * unpack(p)["out_scale"] * safe[:, None]
    y[s == 0] = 0.0
    return y


def loss_and_grad(p, X, T):
    """L = mean((Y - T/out_scale)^2); gradient w.r.t. W1 b1 W2 b2 W3 b3 (flat)."""
    q = unpack(p)
    A = np.asarray(X, np.float64) / q["in_scale"]
    H1, H2, Y = forward(p, X)
    E = Y - np.asarray(T, np.float64) / q["out_scale"]
    B = X.shape[0]
    L = np.mean(E ** 2)
    dY = 2 * E / (B * NOUT)
    gW3 = H2.T @ dY
    gb3 = dY.sum(0)
    dZ2 = (dY @ q["W3"].T) * (1 - H2 ** 2)
    gW2 = H1.T @ dZ2
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ q["W2"].T) * (1 - H1 ** 2)
    gW1 = A.T @ dZ1
    gb1 = dZ1.sum(0)
    return L, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])


def adam(p, g, m, v, lr, b1, b2, eps, t):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
