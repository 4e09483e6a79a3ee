"""CPU reference (numpy, FP64) of the learned preconditioner's network and loss, the
checker of the device kernels in crtomo_stable_amd/csrc/crt_learned_pc.hip:
8 -> 64 -> 64 -> 64 -> 5, tanh hidden layers, loss = sum of masked squared errors
divided by the number of cells.  Parameters are one flat vector in the order
W1 (64x8), b1, W2 (64x64), b2, W3 (64x64), b3, W4 (5x64), b4; the model vector of
the library prepends 16 standardisation values (mean, inverse spread of the
8 features)."""
import numpy as np

NI, NH, NO = 8, 64, 5
SHAPES = [(NH, NI), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NO, NH), (NO,)]
NPAR = sum(int(np.prod(s)) for s in SHAPES)


def unpack(params):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(params[o:o + n], np.float64).reshape(s))
        o += n
    return out


def loss_grad(params, norm, feat, tgt, mask):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(params)
    norm = np.asarray(norm, np.float64)
    x = (np.asarray(feat, np.float64) - norm[:NI]) * norm[NI:]
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    h3 = np.tanh(h2 @ W3.T + b3)
    y = h3 @ W4.T + b4
    m = x.shape[0]
    e = (y - tgt) * mask
    loss = float(np.sum(e * e) / m)
    d4 = 2.0 * e / m
    d3 = (d4 @ W4) * (1.0 - h3 * h3)
    d2 = (d3 @ W3) * (1.0 - h2 * h2)
    d1 = (d2 @ W2) * (1.0 - h1 * h1)
    grads = [d1.T @ x, d1.sum(0), d2.T @ h1, d2.sum(0), d3.T @ h2, d3.sum(0), d4.T @ h3, d4.sum(0)]
    return loss, np.concatenate([g.ravel() for g in grads])
