"""The learned form's network kernels and its effect on the iterative solve.

Forward and backward passes against the numpy reference (oracle/mlp.py,
itself pinned by finite differences); descriptor invariants; device training
reduces the loss; and the A/B comparison of PCG iterations with and without
the trained network's warm start (reported, and bounded for correctness)."""
import os

import numpy as np
import pytest

from mbd_amd import engine, generators, learned
from oracle import mlp
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _blob(rng, X):
    return learned.init_blob(X, seed=int(rng.integers(1 << 30)))


def test_forward_and_gradient_match_reference():
    n_in, H, _ = engine.net_layout()
    rng = np.random.default_rng(1)
    X = rng.normal(size=(300, n_in)).astype(np.float32)
    t = rng.normal(size=300).astype(np.float32)
    blob = _blob(rng, X)
    y = engine.net_forward(blob, X)
    y_ref, _ = mlp.forward(blob, X, n_in, H)
    assert np.allclose(y, y_ref, rtol=1e-4, atol=1e-4)
    g, loss = engine.net_grad(blob, X, t)
    loss_ref, g_ref = mlp.grad(blob, X, t, n_in, H)
    assert loss == pytest.approx(loss_ref, rel=1e-4)
    assert np.abs(g - g_ref).max() <= 1e-4 * np.abs(g_ref).max() + 1e-6


def test_descriptors_invariants():
    eng = engine.MbdEngine(0)
    # an isolated atom: no neighbour terms
    eng.set_system(["C"], np.zeros((1, 3)), [0.8], None, False)
    X, aw = eng.net_descriptors()
    assert X.shape == (21, 16)
    assert np.allclose(X[:, 4:], 0.0)
    assert np.allclose(X[:, 0], np.log(aw), rtol=1e-6)
    # permutation covariance of a molecule
    geo, st = generators.ring()
    xyz, _ = geo.to_bohr()
    eng.set_system(geo.names, xyz, geo.hirshfeld, None, False)
    X1, _ = eng.net_descriptors()
    p = np.random.default_rng(0).permutation(12)
    eng.set_system([geo.names[i] for i in p], xyz[p], geo.hirshfeld[p], None, False)
    X2, _ = eng.net_descriptors()
    idx = (np.arange(21)[:, None] * 12 + p[None, :]).reshape(-1)
    assert np.allclose(X2, X1[idx], rtol=1e-5, atol=1e-6)
    eng.close()


def _images(lat_bohr, periodic, st):
    if not periodic:
        return np.zeros((1, 3))
    a, _, _ = O.image_ranges(lat_bohr, True, st.mbd_scs_vacuum_axis, st.mbd_scs_dip_cutoff,
                             st.mbd_cfdm_dip_cutoff, st.mbd_supercell_cutoff)
    L = np.asarray(lat_bohr).reshape(3, 3)
    return np.array([i * L[0] + j * L[1] + k * L[2]
                     for i in range(-a[0], a[0] + 1) for j in range(-a[1], a[1] + 1)
                     for k in range(-a[2], a[2] + 1)])


@pytest.mark.parametrize("case", ["ring", "periodic64", "sheet_vacuum"])
def test_descriptors_match_reference(case):
    """Neighbour-list descriptor kernel against the brute-force numpy sums
    over the whole SCS image box (no pruning): the pruned image set and the
    per-atom lists must not lose a pair."""
    if case == "ring":
        geo, st = generators.ring()
    elif case == "periodic64":
        geo, st = generators.random_periodic(64, 9.5, seed=11)
    else:
        geo, st = generators.hex_sheet(4, 4, seed=3)
    xyz, lat = geo.to_bohr()
    eng = engine.MbdEngine(0)
    eng.set_geometry(geo, st)
    X, aw = eng.net_descriptors()
    eng.close()
    N = geo.n_atoms
    r0 = np.array([O.vdw_param(nm)[2] for nm in geo.names])
    rvdw = np.cbrt(np.asarray(geo.hirshfeld)) * r0
    omega, _ = O.casimir_grid()
    want = mlp.descriptors(xyz, _images(lat, geo.periodic, st), aw, geo.hirshfeld, rvdw, omega,
                           O.beta(st.flag_xc))
    scale = np.abs(want).max(axis=0) + 1e-30
    err = (np.abs(X - want) / scale).max(axis=0)
    assert err.max() <= 2e-6, err
    assert X.shape == (21 * N, 16)


def test_device_training_reduces_loss():
    eng = engine.MbdEngine(0)
    rng = np.random.default_rng(2)
    Xs, ys = [], []
    while len(Xs) < 6:
        geo = generators.chain_molecule(rng, 20)
        xy = learned.sample(eng, geo, generators.Settings())
        if xy is not None:
            Xs.append(xy[0])
            ys.append(xy[1])
    X = np.concatenate(Xs)
    y = np.concatenate(ys)
    blob0 = learned.init_blob(X)
    blob, hist = engine.net_train(blob0, X, y, epochs=20, batch=256, lr=2e-3)
    assert hist[-1] < 0.2 * hist[0]
    eng.close()


@pytest.mark.skipif(not os.path.exists(learned.DEFAULT_WEIGHTS), reason="no trained weights")
def test_trained_network_warm_start_ab():
    """Held-out periodic cell: the learned form with and without the network.
    Both must match the oracle.  With the atom-cluster preconditioner the
    iteration is started from the cluster solve M^-1 B; the network's
    prediction only corrects each atom's diagonal block to the predicted
    alpha_SCS.  Measured: the same iteration count with and without it
    (12 here, 14 on a 1024-atom cell, 15 at C3) -- the network must not cost
    iterations, and it does not save any."""
    blob = learned.load()
    geo, st = generators.random_periodic(384, 38.0 * (384 / 4096) ** (1 / 3), seed=999)
    eng = engine.MbdEngine(0)
    eng.set_geometry(geo, st)
    X, aw = eng.net_descriptors()
    _, y = learned.sample(eng, geo, st)  # a physical held-out system
    rmse = float(np.sqrt(np.mean((engine.net_forward(blob, X) - y) ** 2)))
    off = eng.run("learned")
    eng.net_set_weights(blob)
    on = eng.run("learned")
    eng.close()
    xyz, lat = geo.to_bohr()
    want = O.mbd_run(O.system_from(geo.names, xyz, geo.hirshfeld, lat, True, st))
    print(f"held-out rmse(log ratio)={rmse:.3e} pcg iters without={off.diag['refine_iters']} "
          f"with={on.diag['refine_iters']}")
    assert abs(on.e_mbd - want.e_mbd) <= 1e-6 and abs(off.e_mbd - want.e_mbd) <= 1e-6
    assert on.diag["fallback_used"] == 0
    assert rmse < 0.1
    assert on.diag["refine_iters"] <= off.diag["refine_iters"]
