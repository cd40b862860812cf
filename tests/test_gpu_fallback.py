"""Fallbacks of the SCS solve when a relay matrix is not positive definite.

The reference factors every relay matrix with PDPOTRF('L') and ignores its
info (UTILS.f90:469-473): on a matrix that is not positive definite it solves
with a broken factor.  This engine reports the FP64 Cholesky info and solves
that grid point by LU with partial pivoting instead, so its alpha_SCS is the
exact solution of the reference's own linear system A X = B.

* device LU (csrc/lu.hip) against numpy.linalg.solve on general, symmetric
  indefinite and pivot-forcing matrices, orders across the 32-column panels;
* an array of atomic chains with a 1.3 A period: the self-image terms of the
  diagonal blocks (sigma_i instead of sqrt(2) sigma_i, UTILS.f90:580-600)
  make the relay matrix indefinite at the low grid points (a polarisation
  catastrophe along each chain).  Every form must report the FP64 Cholesky
  info, re-solve those points, and return alpha_SCS(i w_k) = tr(X_i)/3 of
  numpy.linalg.solve on the oracle's relay matrix, at every grid point.
"""
import numpy as np
import pytest

from mbd_amd import engine, io
from mbd_amd.engine import MbdEngine
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _stacked_identity(n):
    return np.tile(np.eye(3), (n // 3, 1))


@pytest.mark.parametrize("n", [3, 30, 33, 96, 255])
def test_lu_solve_matches_numpy(n):
    rng = np.random.default_rng(300 + n)
    mats = [rng.standard_normal((n, n)),                    # general
            (lambda M: M + M.T)(rng.standard_normal((n, n))),  # symmetric indefinite
            np.diag(rng.uniform(-1, 1, n)) + np.tril(rng.standard_normal((n, n)), -1) * 3.0]
    A = np.stack(mats)
    X, info = engine.device_lu_solve(A)
    B = _stacked_identity(n)
    for b in range(A.shape[0]):
        ref = np.linalg.solve(A[b], B)
        err = np.abs(X[b] - ref).max() / np.abs(ref).max()
        cond = np.linalg.cond(A[b])
        print(f"n={n} matrix {b} info={info[b]} cond={cond:.1e} rel err={err:.2e}", flush=True)
        assert info[b] == 0
        assert err <= 1e-14 * cond * n


def test_lu_reports_zero_pivot():
    n = 42
    rng = np.random.default_rng(9)
    A = rng.standard_normal((n, n))
    A[:, 17] = 0.0  # column 18 has no pivot: dgetrf's info = 18
    _, info = engine.device_lu_solve(A)
    assert info[0] == 18


def _chain_array(m=6, period=1.3, seed=3):
    rng = np.random.default_rng(seed)
    xyz = np.array([[rng.uniform(0, period), 3.0 * i + rng.normal(0, 0.1),
                     3.0 * j + rng.normal(0, 0.1)] for i in range(m) for j in range(m)])
    n = len(xyz)
    names = [str(x) for x in rng.choice(["C", "N", "O"], n)]
    h = rng.uniform(0.7, 0.95, n)
    lat = np.array([[period, 0, 0], [0, 3.0 * m, 0], [0, 0, 3.0 * m]])
    st = io.Settings(mbd_scs_dip_cutoff=20.0, mbd_supercell_cutoff=5.0, mbd_cfdm_dip_cutoff=20.0)
    return io.Geometry(names, xyz, h, lat), st


@pytest.fixture(scope="module")
def chains():
    geo, st = _chain_array()
    xyz, lat = geo.to_bohr()
    s = O.system_from(geo.names, xyz, geo.hirshfeld, lat, geo.periodic, st)
    N = len(geo.names)
    alpha = np.zeros((21, N))
    neg = []
    for k in range(21):
        A = O.relay_matrix(s, k)
        neg.append(int((np.linalg.eigvalsh(A) < 0).sum()))
        X = np.linalg.solve(A, _stacked_identity(3 * N))
        alpha[k] = [np.trace(X[3 * i:3 * i + 3]) / 3.0 for i in range(N)]
    bad = [k for k in range(21) if neg[k] > 0]
    print(f"chain array: N={N}, grid points with an indefinite relay matrix: {bad}", flush=True)
    assert len(bad) >= 4 and O.scs_all(s)[4] != 0  # the oracle's Cholesky fails too
    return geo, st, alpha, bad


@pytest.mark.parametrize("mode", ["fp64", "mixed", "learned"])
def test_indefinite_relay_matrix_solved_by_lu(chains, mode):
    geo, st, alpha_ref, bad = chains
    with MbdEngine(0) as eng:
        eng.set_geometry(geo, st)
        aeff, cpol, _, _, _, diag = eng.scs(mode)
    got = np.vstack([aeff[None], cpol.T])  # (21, N)
    err = np.abs(got - alpha_ref).max(axis=1) / np.abs(alpha_ref).max(axis=1)
    print(f"{mode}: info_scs={diag['info_scs']} fallback_used={diag['fallback_used']} "
          f"max rel err per point={np.array2string(err, precision=1)}", flush=True)
    assert diag["info_scs"] > 0
    assert diag["fallback_used"] >= len(bad)
    tol = 1e-9 if mode == "fp64" else 1e-5
    assert np.all(err[bad] <= 1e-9)  # the indefinite points are solved in FP64 in every form
    assert np.all(err <= tol)
