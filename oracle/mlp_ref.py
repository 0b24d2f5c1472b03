"""FP64 numpy restatement of the warm-start MLP (test infrastructure: only
tests/ use it, as the checker of the HIP forward and backward kernels).

Network 5 -> 64 (tanh) -> 64 (tanh) -> 5 on normalised inputs; parameter
block W1[64][5] b1[64] W2[64][64] b2[64] W3[5][64] b3[5] xm[5] xs[5] ym[5]
ys[5] (4889 floats, the first 4869 trainable). Loss over n rows with
weights w: L = sum_b w_b |y_b - (t_b - ym) / ys|^2 / (5 n), y the network's
normalised output. The backward pass is the textbook one (tanh' = 1 - h^2);
tests pin it against central finite differences of this forward.

Not in the reference: the MLP warm-starts the discrete-Maxwellian Newton
solve, a component this implementation adds.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 5, 64, 5
W1, B1 = 0, NH * NIN
W2 = B1 + NH
B2 = W2 + NH * NH
W3 = B2 + NH
B3 = W3 + NOUT * NH
TRAIN = B3 + NOUT
XM, XS = TRAIN, TRAIN + NIN
YM, YS = XS + NIN, XS + NIN + NOUT
PARAMS = YS + NOUT


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return dict(W1=p[W1:B1].reshape(NH, NIN), b1=p[B1:W2], W2=p[W2:B2].reshape(NH, NH),
                b2=p[B2:W3], W3=p[W3:B3].reshape(NOUT, NH), b3=p[B3:TRAIN],
                xm=p[XM:XS], xs=p[XS:YM], ym=p[YM:YS], ys=p[YS:PARAMS])


def _fwd(q, x):
    xn = (np.asarray(x, np.float64) - q["xm"]) / q["xs"]
    h1 = np.tanh(xn @ q["W1"].T + q["b1"])
    h2 = np.tanh(h1 @ q["W2"].T + q["b2"])
    y = h2 @ q["W3"].T + q["b3"]
    return xn, h1, h2, y


def forward(p, x):
    """De-normalised outputs, n x 5."""
    q = unpack(p)
    return _fwd(q, x)[3] * q["ys"] + q["ym"]


def bf16(a):
    """Round FP32 values to BF16 (round to nearest, ties to even), as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(p, x):
    """The warm start's inference as the solver runs it: W1, W2, W3 rounded to
    BF16 (from the FP32 parameters), the layer-1 activations rounded to BF16
    before layer 2, everything else in FP64 here (FP32 on the device)."""
    q = unpack(np.asarray(p, np.float32))
    for k in ("W1", "W2", "W3"):
        q[k] = bf16(q[k])
    xn = (np.asarray(x, np.float32).astype(np.float64) - q["xm"]) / q["xs"]
    h1 = bf16(np.tanh(xn @ q["W1"].T + q["b1"]))
    h2 = np.tanh(h1 @ q["W2"].T + q["b2"])
    y = h2 @ q["W3"].T + q["b3"]
    return y * q["ys"] + q["ym"]


def loss(p, x, t, w):
    q = unpack(p)
    y = _fwd(q, x)[3]
    tn = (np.asarray(t, np.float64) - q["ym"]) / q["ys"]
    n = y.shape[0]
    return float(np.sum(np.asarray(w, np.float64)[:, None] * (y - tn) ** 2) / (NOUT * n))


def loss_grad(p, x, t, w):
    """(loss, gradient of the 4869 trainable parameters)."""
    q = unpack(p)
    xn, h1, h2, y = _fwd(q, x)
    tn = (np.asarray(t, np.float64) - q["ym"]) / q["ys"]
    w = np.asarray(w, np.float64)
    n = y.shape[0]
    r = y - tn
    L = float(np.sum(w[:, None] * r ** 2) / (NOUT * n))
    dy = 2.0 * w[:, None] * r / (NOUT * n)
    dz2 = (dy @ q["W3"]) * (1.0 - h2 ** 2)
    dz1 = (dz2 @ q["W2"]) * (1.0 - h1 ** 2)
    g = np.zeros(TRAIN)
    g[W1:B1] = (dz1.T @ xn).ravel()
    g[B1:W2] = dz1.sum(0)
    g[W2:B2] = (dz2.T @ h1).ravel()
    g[B2:W3] = dz2.sum(0)
    g[W3:B3] = (dy.T @ h2).ravel()
    g[B3:TRAIN] = dy.sum(0)
    return L, g


def random_params(seed, scale_in=None, scale_out=None):
    """Xavier-uniform weights, small random biases, given or unit scales."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS)
    for off, fo, fi in ((W1, NH, NIN), (W2, NH, NH), (W3, NOUT, NH)):
        a = np.sqrt(6.0 / (fi + fo))
        p[off:off + fo * fi] = rng.uniform(-a, a, fo * fi)
    p[B1:W2] = rng.uniform(-0.1, 0.1, NH)
    p[B2:W3] = rng.uniform(-0.1, 0.1, NH)
    p[B3:TRAIN] = rng.uniform(-0.1, 0.1, NOUT)
    p[XM:XS] = rng.uniform(-0.5, 0.5, NIN)
    p[XS:YM] = rng.uniform(0.5, 2.0, NIN) if scale_in is None else scale_in
    p[YM:YS] = rng.uniform(-0.5, 0.5, NOUT)
    p[YS:PARAMS] = rng.uniform(0.5, 2.0, NOUT) if scale_out is None else scale_out
    return p
