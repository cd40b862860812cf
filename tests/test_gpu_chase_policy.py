"""The bulge chase on one XCD (band accesses through that XCD's L2) and on
every XCD (device-coherent accesses) run the same arithmetic in the same
order, so the eigenvalues must agree bit for bit; a stale read anywhere in
the single-XCD form would show up as a difference.  Orders cover one to
several hundred sweeps per workgroup-claim round and a first panel narrower
than 32."""
import numpy as np
import pytest

from mbd_amd import engine

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [40, 517, 2113, 4099])
def test_chase_policies_bit_identical(n, monkeypatch):
    rng = np.random.default_rng(7000 + n)
    M = rng.standard_normal((n, n))
    A = (M + M.T) / 2
    monkeypatch.setenv("MBD_EIG_METHOD", "2")
    monkeypatch.setenv("MBD_SBR_CHASE_XCC", "0")
    w0 = engine.device_sym_eigvals(A)
    monkeypatch.setenv("MBD_SBR_CHASE_XCC", "1")
    w1 = engine.device_sym_eigvals(A)
    ref = np.linalg.eigvalsh(A)
    err = np.abs(w1 - ref).max() / np.abs(ref).max()
    print(f"n={n} max|w_xcc - w_dev|={np.abs(w1 - w0).max():.2e} rel err vs numpy={err:.2e}",
          flush=True)
    assert np.array_equal(w0, w1)
    assert err <= 1e-12 * np.sqrt(n)
