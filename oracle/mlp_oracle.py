"""CPU (float64, numpy) restatement of the planner MLP -- TEST INFRASTRUCTURE
ONLY: the reference result for the GPU forward/backward kernels in csrc/fcc_mlp.hip.

Network: x_n = (x - in_mean) / in_std;  h1 = relu(x_n W1 + b1);
h2 = relu(h1 W2 + b2);  y = h2 W3 + b3  (normalised outputs).
Loss: mean over samples and the 2 outputs of (y - t_n)^2 with
t_n = (t - out_mean) / out_std.
Flat parameter layout: W1[6][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2]
in_mean[6] in_std[6] out_mean[2] out_std[2].
"""
import numpy as np

NI, NH, NO = 6, 64, 2
SIZES = [("W1", (NI, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NO)), ("b3", (NO,)),
         ("in_mean", (NI,)), ("in_std", (NI,)), ("out_mean", (NO,)), ("out_std", (NO,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)
N_TRAIN = sum(int(np.prod(s)) for n, s in SIZES[:6])


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def forward(p, x):
    P = unpack(p)
    xn = (np.asarray(x, np.float64) - P["in_mean"]) / P["in_std"]
    z1 = xn @ P["W1"] + P["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ P["W2"] + P["b2"]
    h2 = np.maximum(z2, 0)
    y = h2 @ P["W3"] + P["b3"]
    return y, (xn, z1, h1, z2, h2)


def loss_and_grad(p, x, t):
    """Returns (loss, grad over the full flat vector; zero for normalisation)."""
    P = unpack(p)
    y, (xn, z1, h1, z2, h2) = forward(p, x)
    tn = (np.asarray(t, np.float64) - P["out_mean"]) / P["out_std"]
    n = y.shape[0]
    d = y - tn
    loss = float(np.sum(d * d) / (n * NO))
    dy = 2.0 * d / (n * NO)
    g = {}
    g["W3"] = h2.T @ dy
    g["b3"] = dy.sum(0)
    dz2 = (dy @ P["W3"].T) * (z2 > 0)
    g["W2"] = h1.T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ P["W2"].T) * (z1 > 0)
    g["W1"] = xn.T @ dz1
    g["b1"] = dz1.sum(0)
    flat = np.zeros(N_PARAMS)
    o = 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        if name in g:
            flat[o:o + k] = g[name].reshape(-1)
        o += k
    return loss, flat


def loss_only(p, x, t):
    P = unpack(p)
    y, _ = forward(p, x)
    tn = (np.asarray(t, np.float64) - P["out_mean"]) / P["out_std"]
    return float(np.mean((y - tn) ** 2))


def random_params(rng, x=None, t=None):
    p = np.zeros(N_PARAMS)
    P = unpack(p)
    o = 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shape[0] + shape[1]))
            p[o:o + k] = rng.uniform(-lim, lim, k)
        elif name.startswith("b"):
            p[o:o + k] = rng.uniform(-0.1, 0.1, k)
        elif name == "in_mean":
            p[o:o + k] = x.mean(0) if x is not None else 0.0
        elif name == "in_std":
            p[o:o + k] = x.std(0) + 0.5 if x is not None else 1.0
        elif name == "out_mean":
            p[o:o + k] = t.mean(0) if t is not None else 0.0
        elif name == "out_std":
            p[o:o + k] = t.std(0) + 0.5 if t is not None else 1.0
        o += k
    return p
