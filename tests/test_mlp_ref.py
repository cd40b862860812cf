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


def test_descriptor_reference_closed_forms():
    """Dimer at distance r along z in a large box: every neighbour sum has one
    term, so each feature has a closed form; a periodic image beyond the
    cutoff adds nothing."""
    r, rc = 4.0, 15.0
    xyz = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, r]])
    omega = np.array([0.0, 0.5])
    aw = np.array([[2.0, 3.0], [1.5, 2.5]])  # (grid point, atom)
    hirsh, rvdw, beta = np.array([0.7, 0.9]), np.array([3.0, 3.5]), 0.83
    images = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 40.0], [0.0, 0.0, -40.0]])
    X = mlp.descriptors(xyz, images, aw.reshape(-1), hirsh, rvdw, omega, beta, rc).reshape(2, 2, 16)
    fc = 0.5 * (np.cos(np.pi * r / rc) + 1.0)
    fsr = 1.0 - 1.0 / (1.0 + np.exp(-6.0 * (r / (beta * 6.5) - 1.0)))
    for f in range(2):
        for i, j in ((0, 1), (1, 0)):
            x = X[f, i]
            ai, aj = aw[f, i], aw[f, j]
            assert x[0] == np.log(ai) and x[1] == np.log1p(omega[f])
            assert x[2] == hirsh[i] and x[3] == rvdw[i] / 4
            mu = 2.0 + 1.5 * np.arange(8)
            assert np.allclose(x[4:12], 0.1 * aj * np.exp(-0.5 * (r - mu) ** 2) * fc)
            assert np.isclose(x[12], ai * aj * fc / r ** 3)
            # traceless (3 zz - I) / r^3 along z: Frobenius norm sqrt(4 + 1 + 1) / r^3
            assert np.isclose(x[13], ai * aj * fc * np.sqrt(6.0) / r ** 3)
            assert np.isclose(x[14], fc)
            assert np.isclose(x[15], ai * aj * fc * fsr / r ** 3)
