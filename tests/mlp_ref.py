"""FP64 CPU restatement of the surrogate network (the checker of the HIP
training and inference kernels; test infrastructure only): forward pass and
backpropagation of the mean squared error, in the parameter layout of
xtsrcmaps_amd/csrc/xts_mlp.hpp.  tests/test_mlp_ref.py pins the gradient by
central finite differences."""
import numpy as np

H, IN = 64, 4
OFF = {"W1": 0, "b1": 256, "W2": 320, "b2": 4416, "W3": 4480, "b3": 8576, "w4": 8640,
       "b4": 8704}
NPARAM = 8705


def unpack(p):
    p = np.asarray(p, np.float64)
    return dict(W1=p[0:256].reshape(H, IN), b1=p[256:320], W2=p[320:4416].reshape(H, H),
                b2=p[4416:4480], W3=p[4480:8576].reshape(H, H), b3=p[8576:8640],
                w4=p[8640:8704], b4=p[8704])


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p, norm, x, raw=True):
    """Predictions for rows x [B, >= 4]: y = y' scale + offset (raw) or y'."""
    q = unpack(p)
    norm = np.asarray(norm, np.float64)
    xn = (np.asarray(x, np.float64)[:, :IN] - norm[:4]) * norm[4:8]
    z1 = xn @ q["W1"].T + q["b1"]
    h1 = z1 * _sig(z1)
    z2 = h1 @ q["W2"].T + q["b2"]
    h2 = z2 * _sig(z2)
    z3 = h2 @ q["W3"].T + q["b3"]
    h3 = z3 * _sig(z3)
    yp = h3 @ q["w4"] + q["b4"]
    if raw:
        return yp * norm[8] + norm[9]
    return yp, (xn, z1, h1, z2, h2, z3, h3)


def loss_and_grad(p, norm, x, y):
    """Mean over the batch of (y' - (y - offset) / scale)^2 and its gradient."""
    q = unpack(p)
    norm = np.asarray(norm, np.float64)
    yp, (xn, z1, h1, z2, h2, z3, h3) = forward(p, norm, x, raw=False)
    yt = (np.asarray(y, np.float64) - norm[9]) / norm[8]
    B = len(yt)
    e = yp - yt
    loss = np.mean(e * e)
    dy = 2.0 * e / B
    dsilu = lambda z: _sig(z) * (1.0 + z * (1.0 - _sig(z)))
    g = np.zeros(NPARAM)
    g[OFF["w4"]:OFF["w4"] + H] = h3.T @ dy
    g[OFF["b4"]] = dy.sum()
    d3 = np.outer(dy, q["w4"]) * dsilu(z3)
    g[OFF["W3"]:OFF["W3"] + H * H] = (d3.T @ h2).ravel()
    g[OFF["b3"]:OFF["b3"] + H] = d3.sum(0)
    d2 = (d3 @ q["W3"]) * dsilu(z2)
    g[OFF["W2"]:OFF["W2"] + H * H] = (d2.T @ h1).ravel()
    g[OFF["b2"]:OFF["b2"] + H] = d2.sum(0)
    d1 = (d2 @ q["W2"]) * dsilu(z1)
    g[OFF["W1"]:OFF["W1"] + H * IN] = (d1.T @ xn).ravel()
    g[OFF["b1"]:OFF["b1"] + H] = d1.sum(0)
    return loss, g


def init_params(seed=0):
    """Glorot-uniform weights, zero biases (the product's own initialiser is
    xtsrcmaps_amd.surrogate.init_params; tests use either)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM)
    for name, (fo, fi) in (("W1", (H, IN)), ("W2", (H, H)), ("W3", (H, H)), ("w4", (1, H))):
        a = np.sqrt(6.0 / (fo + fi))
        p[OFF[name]:OFF[name] + fo * fi] = rng.uniform(-a, a, fo * fi)
    return p
