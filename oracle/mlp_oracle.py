"""FP64 CPU restatement of the learned cost model's network and loss, the
checker for libfot's MLP kernels (fot_mlp_*).  Test infrastructure only.

Network: h1 = relu(x W1^T + b1), h2 = relu(h1 W2^T + b2), y = h2 W3^T + b3;
loss = sum(mask (y - t)^2) / sum(mask).  Parameters flat in the order
W1[h][in], b1, W2[h][h], b2, W3[out][h], b3.  The backward pass is the chain
rule written out; `numeric_grad` (central differences) pins it."""
import numpy as np


def unpack(p, n_in, hidden, n_out):
    sizes = [(hidden, n_in), (hidden,), (hidden, hidden), (hidden,), (n_out, hidden), (n_out,)]
    out, k = [], 0
    for s in sizes:
        n = int(np.prod(s))
        out.append(np.asarray(p[k:k + n], dtype=np.float64).reshape(s))
        k += n
    return out


def forward(p, x, n_in, hidden, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, hidden, n_out)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3.T + b3, (x, h1, h2)


def loss(p, x, t, mask, n_in, hidden, n_out):
    y, _ = forward(p, x, n_in, hidden, n_out)
    m = np.ones_like(y) if mask is None else np.asarray(mask, np.float64)
    return float((m * (y - t) ** 2).sum() / max(m.sum(), 1.0))


def loss_grad(p, x, t, mask, n_in, hidden, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, hidden, n_out)
    y, (x, h1, h2) = forward(p, x, n_in, hidden, n_out)
    m = np.ones_like(y) if mask is None else np.asarray(mask, np.float64)
    cnt = max(m.sum(), 1.0)
    dy = 2.0 * m * (y - t) / cnt
    dW3, db3 = dy.T @ h2, dy.sum(0)
    dz2 = (dy @ W3) * (h2 > 0)
    dW2, db2 = dz2.T @ h1, dz2.sum(0)
    dz1 = (dz2 @ W2) * (h1 > 0)
    dW1, db1 = dz1.T @ x, dz1.sum(0)
    g = np.concatenate([a.ravel() for a in (dW1, db1, dW2, db2, dW3, db3)])
    return float((m * (y - t) ** 2).sum() / cnt), g


def numeric_grad(p, x, t, mask, n_in, hidden, n_out, idx, h=1e-6):
    """Central differences of the FP64 loss at the parameters `idx`."""
    p = np.asarray(p, np.float64).copy()
    out = []
    for i in idx:
        old = p[i]
        p[i] = old + h
        lp = loss(p, x, t, mask, n_in, hidden, n_out)
        p[i] = old - h
        lm = loss(p, x, t, mask, n_in, hidden, n_out)
        p[i] = old
        out.append((lp - lm) / (2 * h))
    return np.array(out)
