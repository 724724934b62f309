"""The FP64 MLP reference: its gradient against central finite differences."""
import numpy as np

import oracle


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    P = rng.normal(0, 0.3, 4545)
    x = rng.normal(0, 1, (40, 4))
    t = rng.normal(0, 1, 40)
    L, g = oracle.mlp_loss_grad(P, x, t)
    idx = rng.choice(4545, 60, replace=False)
    idx = np.concatenate([idx, [0, 255, 256, 319, 320, 4415, 4416, 4479, 4480, 4544]])
    h = 1e-6
    for k in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += h
        Pm[k] -= h
        fd = (oracle.mlp_loss_grad(Pp, x, t)[0] - oracle.mlp_loss_grad(Pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])
    assert np.isclose(L, np.mean((oracle.mlp_forward(P, x) - t) ** 2), rtol=1e-14)


def test_mlp_forward_closed_form():
    rng = np.random.default_rng(4)
    P = rng.normal(0, 0.3, 4545)
    x = rng.normal(0, 1, (5, 4))
    W1, b1 = P[:256].reshape(64, 4), P[256:320]
    W2, b2 = P[320:4416].reshape(64, 64), P[4416:4480]
    W3, b3 = P[4480:4544], P[4544]
    ref = np.tanh(np.tanh(x @ W1.T + b1) @ W2.T + b2) @ W3 + b3
    assert np.allclose(oracle.mlp_forward(P, x), ref, rtol=1e-13, atol=1e-13)
