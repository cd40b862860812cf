"""Two-stage eigen reduction (dense -> band of half-width 32 -> tridiagonal,
csrc/sbr.hip) through the C ABI, forced with MBD_EIG_METHOD=2:

* eigenvalues against numpy's eigvalsh for orders that exercise a first panel
  narrower than 32 (n % 32 != 0), one to three TSQR tree levels (panel heights
  above 256 and 2048 rows), n < 3 (no bulge-chasing sweep) and n < 32 (no
  panel at all);
* structured matrices whose panels are exactly zero or rank deficient
  (diagonal, block diagonal, identity plus rank one), where reflectors with
  tau = 0 and the sign choice of the Householder reconstruction matter;
* E_MBD of a 512-atom periodic cell equal to the one-stage reduction's within
  1e-10 Ha, and the eager and graph-replayed launch forms bit-identical;
* the look-ahead schedule (a panel pair's SYR2K on a second stream while the
  next panel is updated and factored) bit-identical to the in-order schedule:
  every element of the trailing matrix gets the same products in the same
  order, only on another stream.
"""
import numpy as np
import pytest

from mbd_amd import engine, generators

pytestmark = pytest.mark.gpu


def _check(A, monkeypatch, rel=1e-12):
    monkeypatch.setenv("MBD_EIG_METHOD", "2")
    w = engine.device_sym_eigvals(A)
    ref = np.linalg.eigvalsh(A)
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(w - ref).max() / scale
    print(f"n={A.shape[0]} max|dl|/|l|max={err:.2e}", flush=True)
    assert np.all(np.diff(w) >= 0)
    assert err <= rel * np.sqrt(A.shape[0])


@pytest.mark.parametrize("n", [1, 2, 3, 5, 31, 32, 33, 64, 65, 100, 257, 300, 517, 1000, 2113,
                               3001])
def test_two_stage_eigvals_random(n, monkeypatch):
    rng = np.random.default_rng(1000 + n)
    M = rng.standard_normal((n, n))
    A = (M + M.T) / 2 + np.diag(rng.uniform(1, 3, n))
    _check(A, monkeypatch)


def test_two_stage_diagonal(monkeypatch):
    rng = np.random.default_rng(5)
    _check(np.diag(rng.uniform(-2, 2, 300)), monkeypatch)


def test_two_stage_block_diagonal(monkeypatch):
    rng = np.random.default_rng(6)
    A = np.zeros((400, 400))
    for a, b in ((0, 150), (150, 151), (151, 400)):
        M = rng.standard_normal((b - a, b - a))
        A[a:b, a:b] = M + M.T
    _check(A, monkeypatch)


def test_two_stage_identity_plus_rank_one(monkeypatch):
    rng = np.random.default_rng(7)
    u = rng.standard_normal(350)
    _check(np.eye(350) + np.outer(u, u), monkeypatch)


@pytest.mark.parametrize("n", [130, 517, 1000, 2113])
def test_two_stage_lookahead_bit_identical(n, monkeypatch):
    rng = np.random.default_rng(4000 + n)
    M = rng.standard_normal((n, n))
    A = (M + M.T) / 2
    monkeypatch.setenv("MBD_EIG_METHOD", "2")
    monkeypatch.setenv("MBD_SBR_LOOKAHEAD", "0")
    w0 = engine.device_sym_eigvals(A)
    monkeypatch.setenv("MBD_SBR_LOOKAHEAD", "1")
    w1 = engine.device_sym_eigvals(A)
    print(f"n={n} max|w_la - w_inorder|={np.abs(w1 - w0).max():.2e}", flush=True)
    assert np.array_equal(w0, w1)


def test_two_stage_energy_matches_one_stage(monkeypatch):
    geo, st = generators.random_periodic(512, 19.0, seed=31, rc_scs=20.0, rc_sc=12.0,
                                         rc_cfdm=30.0)
    with engine.MbdEngine(0) as eng:
        eng.set_geometry(geo, st)
        monkeypatch.setenv("MBD_EIG_METHOD", "1")
        one = eng.run("fp64")
        monkeypatch.setenv("MBD_EIG_METHOD", "2")
        two = eng.run("fp64")
    print(f"E one-stage {one.e_mbd:.12f} two-stage {two.e_mbd:.12f}", flush=True)
    assert abs(one.e_mbd - two.e_mbd) <= 1e-10
    assert one.diag["n_negative"] == two.diag["n_negative"]
    assert two.diag["info_eig"] == 0


def test_two_stage_forms_agree(monkeypatch):
    """Eager launches and the captured graph (several replays, with other
    eager work in between) give bit-identical energies."""
    geo, st = generators.random_periodic(400, 17.5, seed=78, rc_scs=20.0, rc_sc=12.0,
                                         rc_cfdm=30.0)
    monkeypatch.setenv("MBD_EIG_METHOD", "2")
    with engine.MbdEngine(0) as g, engine.MbdEngine(0) as other:
        g.set_geometry(geo, st)
        other.set_geometry(*generators.ring())
        monkeypatch.setenv("MBD_TRD_GRAPH", "0")
        eager = g.run("fp64")
        monkeypatch.setenv("MBD_TRD_GRAPH", "1")  # the captured graph (not the default at this order)
        for _ in range(3):
            r = g.run("fp64")
            other.run("mixed")
            assert r.e_mbd == eager.e_mbd
            assert r.diag["n_negative"] == eager.diag["n_negative"]
        monkeypatch.setenv("MBD_SBR_LOOKAHEAD", "0")
        inorder = g.run("fp64")
        assert inorder.e_mbd == eager.e_mbd
