"""CPU oracle of the learned warm start (test infrastructure only, like the rest of oracle/).

A per-vertex MLP 7 -> 64 -> 64 -> 1 with tanh activations; parameters packed as W1[64][7],
b1[64], W2[64][64], b2[64], w3[64], b3 (4737 values).  No counterpart in the reference, whose
solves start from the previous solution (`numerics/solvers.py:322-372`).

Pinning: `grad` is checked against central finite differences of `loss` (float64) and on
special cases with closed forms; `forward` on networks whose output is known in closed form
(zero weights -> b3; a single active hidden path -> nested tanh); `adam_step` against the
published Adam update (Kingma & Ba 2015, bias-corrected); `features` against direct
evaluation of its definition.
"""
import numpy as np

NIN, NH = 7, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH
NPARAM = OFF_B3 + 1


def unpack(p):
    p = np.asarray(p)
    return (p[OFF_W1:OFF_B1].reshape(NH, NIN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH, NH),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, [b3]])


def init(seed=0, dtype=np.float32):
    """Glorot-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)

    def glorot(n_out, n_in):
        a = np.sqrt(6.0 / (n_in + n_out))
        return rng.uniform(-a, a, (n_out, n_in))

    return pack(glorot(NH, NIN), np.zeros(NH), glorot(NH, NH), np.zeros(NH),
                glorot(1, NH)[0], 0.0).astype(dtype)


def forward(p, X, keep=False):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    out = h2 @ w3 + b3
    return (out, h1, h2) if keep else out


def loss(p, X, y):
    r = forward(p, X) - y
    return float(np.mean(r * r))


def grad(p, X, y):
    """dL/dp of L = mean (f(X) - y)^2, by back-propagation."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    out, h1, h2 = forward(p, X, keep=True)
    g = 2.0 * (out - y) / len(y)            # dL/dout
    gw3 = h2.T @ g
    gb3 = g.sum()
    d2 = np.outer(g, w3) * (1.0 - h2 * h2)  # dL/da2
    gW2 = d2.T @ h1
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ W2) * (1.0 - h1 * h1)         # dL/da1
    gW1 = d1.T @ X
    gb1 = d1.sum(axis=0)
    return pack(gW1, gb1, gW2, gb2, gw3, gb3)


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mhat = m / (1 - b1 ** step)
    vhat = v / (1 - b2 ** step)
    return p - lr * mhat / (np.sqrt(vhat) + eps), m, v


def features(u, u_prev, phi, grad_phi, alpha, D, dt, eps, bc_mask=None):
    """Rows of the 7 features per vertex and sigma = max |u - u_prev|."""
    u = np.asarray(u, float)
    du = u - np.asarray(u_prev, float)
    sigma = max(float(np.abs(du).max()), 1e-30)
    n = len(u)
    al = np.broadcast_to(np.asarray(alpha, float), (n,))
    X = np.empty((n, NIN))
    X[:, 0] = u
    X[:, 1] = du / sigma
    X[:, 2] = np.clip(np.asarray(phi) / eps, -4, 4) / 4
    X[:, 3] = np.minimum(np.asarray(grad_phi), 4) / 4
    X[:, 4] = 0.5 * np.log(D)
    X[:, 5] = al * dt
    X[:, 6] = 0.0 if bc_mask is None else np.asarray(bc_mask, float)
    return X, sigma
