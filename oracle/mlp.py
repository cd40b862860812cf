"""CPU (numpy, float64) restatement of the learned form's network.

TEST INFRASTRUCTURE ONLY: the checker for the device forward/backward kernels
(mbd_amd/csrc/net.hip).  Blob layout: n_in means, n_in standard deviations,
then W1 b1 W2 b2 W3 b3 W4 b4 with each W column-major (fo x fi); tanh hidden
layers, identity output, loss mean((y - t)^2).
"""
import numpy as np


def unpack(blob, n_in, hidden):
    b = np.asarray(blob, dtype=np.float64)
    mean, std = b[:n_in], b[n_in:2 * n_in]
    off = 2 * n_in
    dims = [n_in, hidden, hidden, hidden, 1]
    Ws, bs = [], []
    for l in range(4):
        fi, fo = dims[l], dims[l + 1]
        Ws.append(b[off:off + fo * fi].reshape(fi, fo).T)
        off += fo * fi
        bs.append(b[off:off + fo])
        off += fo
    return mean, std, Ws, bs


def forward(blob, X, n_in, hidden):
    mean, std, Ws, bs = unpack(blob, n_in, hidden)
    h = (np.asarray(X, dtype=np.float64) - mean) / std
    acts = [h]
    for l in range(4):
        z = h @ Ws[l].T + bs[l]
        h = np.tanh(z) if l < 3 else z
        acts.append(h)
    return h[:, 0], acts


def grad(blob, X, t, n_in, hidden):
    """Loss and its gradient w.r.t. the parameters (blob order, no norm)."""
    mean, std, Ws, bs = unpack(blob, n_in, hidden)
    y, acts = forward(blob, X, n_in, hidden)
    B = len(y)
    d = y - np.asarray(t, dtype=np.float64)
    loss = np.mean(d * d)
    delta = (2.0 * d / B)[:, None]
    gW, gb = [None] * 4, [None] * 4
    for l in range(3, -1, -1):
        gW[l] = delta.T @ acts[l]
        gb[l] = delta.sum(axis=0)
        if l > 0:
            delta = (delta @ Ws[l]) * (1.0 - acts[l] ** 2)
    parts = []
    for l in range(4):
        parts.append(gW[l].T.reshape(-1))
        parts.append(gb[l])
    return loss, np.concatenate(parts)


def descriptors(xyz, images, alpha_w, hirsh, rvdw, omega, beta, rc=15.0):
    """Per-(grid point, atom) network inputs, brute force over all `images`
    (translation vectors, bohr) with no pruning or neighbour list; rows are
    ordered s = f * N + i as the device writes them.  Features: log alpha_i,
    log(1 + w), hirshfeld_i, R_i / 4, eight radial Gaussians
    sum_j alpha_j / 10 exp(-(r - mu_m)^2 / 2) fc (mu_m = 2 + 1.5 m), the
    dipole-field sums alpha_i sum_j alpha_j fc / r^3 and its traceless tensor
    norm, the coordination sum_j fc, and the short-range-damped
    alpha_i sum_j alpha_j fc (1 - 1 / (1 + exp(-6 (r / (beta (R_i + R_j)) - 1)))) / r^3,
    with fc = (cos(pi r / rc) + 1) / 2 for r < rc."""
    xyz = np.asarray(xyz, dtype=np.float64)
    N = len(xyz)
    nf = len(omega)
    aw = np.asarray(alpha_w, dtype=np.float64).reshape(nf, N)
    X = np.zeros((nf, N, 16))
    mus = 2.0 + 1.5 * np.arange(8)
    for i in range(N):
        d = (xyz[i][None, None, :] - xyz[None, :, :]) + np.asarray(images)[:, None, :]
        d = d.reshape(-1, 3)
        j = np.tile(np.arange(N), len(images))
        r = np.linalg.norm(d, axis=1)
        m = (r > 0) & (r < rc)
        d, j, r = d[m], j[m], r[m]
        fc = 0.5 * (np.cos(np.pi * r / rc) + 1.0)
        u = d / r[:, None]
        fsr = 1.0 - 1.0 / (1.0 + np.exp(-6.0 * (r / (beta * (rvdw[i] + rvdw[j])) - 1.0)))
        gau = np.exp(-0.5 * (r[:, None] - mus[None, :]) ** 2) * fc[:, None]
        for f in range(nf):
            a = aw[f, j]
            w3 = a * fc / r ** 3
            Q = np.einsum("p,pa,pb->ab", w3, u, u) * 3.0 - np.eye(3) * w3.sum()
            X[f, i, 0] = np.log(aw[f, i])
            X[f, i, 1] = np.log1p(omega[f])
            X[f, i, 2] = hirsh[i]
            X[f, i, 3] = rvdw[i] / 4.0
            X[f, i, 4:12] = 0.1 * (a[:, None] * gau).sum(0)
            X[f, i, 12] = aw[f, i] * w3.sum()
            X[f, i, 13] = aw[f, i] * np.linalg.norm(Q)
            X[f, i, 14] = fc.sum()
            X[f, i, 15] = aw[f, i] * (w3 * fsr).sum()
    return X.reshape(nf * N, 16)
