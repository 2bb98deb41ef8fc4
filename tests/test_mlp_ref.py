"""The FP64 backpropagation of tests/mlp_ref.py (the checker of the GPU
training kernels) against central finite differences of its own forward
pass: a dropped term, a wrong sign or a transposed weight in the backward
pass fails by orders of magnitude."""
import numpy as np

import mlp_ref as mr


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = mr.init_params(7) + rng.normal(0, 0.05, mr.NPARAM)
    norm = np.array([0.1, 0.3, 0.1, 0.0, 3.0, 2.0, 4.0, 50.0, 0.02, 0.01])
    x = rng.normal(size=(37, 4)) * [0.2, 0.3, 0.2, 0.01] + [0.1, 0.3, 0.1, 0.0]
    y = rng.uniform(0, 0.05, 37)
    loss, g = mr.loss_and_grad(p, norm, x, y)
    idx = np.concatenate([rng.choice(256, 12, replace=False), 256 + np.arange(0, 64, 9),
                          320 + rng.choice(4096, 20, replace=False), 4416 + np.arange(0, 64, 11),
                          4480 + rng.choice(4096, 20, replace=False), 8576 + np.arange(3),
                          8640 + rng.choice(64, 8, replace=False), [8704]])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mr.loss_and_grad(pp, norm, x, y)[0] - mr.loss_and_grad(pm, norm, x, y)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1e-3, abs(g).max()), (k, fd, g[k])


def test_forward_normalisation():
    p = mr.init_params(1)
    norm = np.array([1, 2, 3, 4, 0.5, 0.5, 0.5, 0.5, 3.0, 7.0])
    x = np.ones((5, 4))
    y = mr.forward(p, norm, x)
    yp, _ = mr.forward(p, norm, x, raw=False)
    np.testing.assert_allclose(y, yp * 3.0 + 7.0, rtol=1e-15)
