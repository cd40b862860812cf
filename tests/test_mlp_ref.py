"""The numpy network reference against central finite differences, so that it
can serve as the checker of the device backward kernel."""
import numpy as np

from oracle import mlp


def test_reference_gradient_matches_finite_differences():
    n_in, H = 5, 7
    rng = np.random.default_rng(0)
    nparam = (n_in * H + H) + 2 * (H * H + H) + (H + 1)
    blob = np.concatenate([rng.normal(size=n_in), rng.uniform(0.5, 2, n_in),
                           rng.normal(0, 0.5, nparam)])
    X = rng.normal(size=(11, n_in))
    t = rng.normal(size=11)
    loss, g = mlp.grad(blob, X, t, n_in, H)
    eps = 1e-6
    for k in rng.choice(nparam, 25, replace=False):
        bp = blob.copy()
        bm = blob.copy()
        bp[2 * n_in + k] += eps
        bm[2 * n_in + k] -= eps
        fd = (mlp.grad(bp, X, t, n_in, H)[0] - mlp.grad(bm, X, t, n_in, H)[0]) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-7 + 1e-6 * abs(fd)
