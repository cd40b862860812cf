"""Host-side mirror of the reference's MBD@rsSCS interface over libmbd_hip.so.

``MbdEngine.run`` is ``MBD_at_rsSCS(ene)`` (``UTILS.f90:153-349``): the SCS
solve at the 21 Casimir-Polder points, the C6 sums and the coupled-QHO energy,
all on the GPU.  ``scs`` and ``energy`` are the two halves
(``get_screened_alpha`` over the frequency loop, ``UTILS.f90:239-315``, and
``get_mbd_rSCS``, ``UTILS.f90:700-879``).  Unknown elements raise ``KeyError``
before any device call, where the reference warns and STOPs
(``UTILS.f90:1693-1699``).
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import MbdBatchDiag, MbdDiag, dptr, fptr, iptr

MODES = {"fp64": _lib.MODE_FP64, "mixed": _lib.MODE_MIXED, "learned": _lib.MODE_LEARNED}


def free_atom_params(symbol: str):
    """(C6, alpha0, R0) of an element as the reference's REAL(4) table holds them."""
    c6, a, r = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    if _lib.lib().mbd_free_atom_params(symbol.encode(), ctypes.byref(c6), ctypes.byref(a),
                                       ctypes.byref(r)) != 0:
        raise KeyError(f"VdW parameters not defined for atom: {symbol}")
    return c6.value, a.value, r.value


def beta(flag_xc: int) -> float:
    return _lib.lib().mbd_beta(int(flag_xc))


def casimir_grid():
    w = np.zeros(21)
    g = np.zeros(21)
    _lib.lib().mbd_casimir_grid(dptr(w), dptr(g))
    return w, g


@dataclasses.dataclass
class MbdResult:
    e_mbd: float
    alpha_eff: np.ndarray
    coupled_pol: np.ndarray   # (N, 20): alpha_SCS(i w_k), k = 1..20
    molpol_eig: np.ndarray    # (21, 3) ascending eigenvalues of the 3x3 tensor
    c6_eff: np.ndarray
    c6_mol: float
    diag: dict


def _diag_dict(d: MbdDiag) -> dict:
    return {"info_scs": d.info_scs, "fallback_used": d.fallback_used,
            "refine_iters": d.refine_iters, "n_negative": d.n_negative, "info_eig": d.info_eig,
            "supercell": list(d.supercell), "residual": d.residual, "t_scs_ms": d.t_scs_ms,
            "t_mbd_ms": d.t_mbd_ms, "t_total_ms": d.t_total_ms}


class MbdError(RuntimeError):
    pass


class MbdEngine:
    """One device handle (one HIP stream); not re-entrant."""

    def __init__(self, device: int = 0):
        self._L = _lib.lib()
        self._h = ctypes.c_void_p()
        rc = self._L.mbd_create(device, ctypes.byref(self._h))
        if rc != 0:
            raise MbdError(f"mbd_create failed ({rc})")
        self.n_atoms = 0

    def close(self):
        if self._h:
            self._L.mbd_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self._L.mbd_last_error(self._h).decode()
            raise MbdError(f"{what} failed ({rc}) {msg}")

    # -- system ---------------------------------------------------------------
    def set_system(self, names: Sequence[str], xyz_bohr, hirshfeld, lattice_bohr=None,
                   periodic: bool = False, vacuum=(False, False, False), flag_xc: int = 1,
                   rc_scs: float = 120.0, rc_cfdm: float = 100.0, rc_supercell: float = 25.0):
        a0, c6, r0 = [], [], []
        for nm in names:
            c, a, r = free_atom_params(nm)
            c6.append(c)
            a0.append(a)
            r0.append(r)
        self.set_system_params(np.asarray(xyz_bohr), np.asarray(hirshfeld), np.array(a0),
                               np.array(c6), np.array(r0), lattice_bohr, periodic, vacuum,
                               beta(flag_xc), rc_scs, rc_cfdm, rc_supercell)

    def set_system_params(self, xyz_bohr, hirshfeld, alpha0, c6_0, r0, lattice_bohr=None,
                          periodic=False, vacuum=(False, False, False), beta_=0.83,
                          rc_scs=120.0, rc_cfdm=100.0, rc_supercell=25.0):
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
        n = xyz.shape[0]
        h = np.ascontiguousarray(hirshfeld, dtype=np.float64)
        a0 = np.ascontiguousarray(alpha0, dtype=np.float64)
        c6 = np.ascontiguousarray(c6_0, dtype=np.float64)
        r0 = np.ascontiguousarray(r0, dtype=np.float64)
        lat = np.ascontiguousarray(np.zeros((3, 3)) if lattice_bohr is None else lattice_bohr,
                                   dtype=np.float64).reshape(3, 3)
        vac = np.array([int(bool(v)) for v in vacuum], dtype=np.int32)
        rc = self._L.mbd_set_system(self._h, n, dptr(xyz), dptr(h), dptr(a0), dptr(c6), dptr(r0),
                                    dptr(lat), int(bool(periodic)), iptr(vac), float(beta_),
                                    float(rc_scs), float(rc_cfdm), float(rc_supercell))
        self._check(rc, "mbd_set_system")
        self.n_atoms = n

    def set_geometry(self, geometry, settings):
        """From parsed geometry.in / settings.in (mbd_amd.io)."""
        xyz, lat = geometry.to_bohr()
        self.set_system(geometry.names, xyz, geometry.hirshfeld, lat, geometry.periodic,
                        settings.mbd_scs_vacuum_axis, settings.flag_xc,
                        settings.mbd_scs_dip_cutoff, settings.mbd_cfdm_dip_cutoff,
                        settings.mbd_supercell_cutoff)

    # -- compute --------------------------------------------------------------
    def run(self, mode: str = "fp64") -> MbdResult:
        N = self.n_atoms
        e = ctypes.c_double()
        c6m = ctypes.c_double()
        aeff, c6e = np.zeros(N), np.zeros(N)
        cpol = np.zeros(20 * N)
        mp = np.zeros(63)
        d = MbdDiag()
        rc = self._L.mbd_run(self._h, MODES[mode], ctypes.byref(e), dptr(aeff), dptr(cpol),
                             dptr(mp), dptr(c6e), ctypes.byref(c6m), ctypes.byref(d))
        self._check(rc, "mbd_run")
        return MbdResult(e.value, aeff, cpol.reshape(N, 20), mp.reshape(21, 3), c6e, c6m.value,
                         _diag_dict(d))

    def run_energy_only(self, mode: str = "fp64") -> float:
        """mbd_run without copying the per-atom arrays back (benchmark path)."""
        e = ctypes.c_double()
        d = MbdDiag()
        rc = self._L.mbd_run(self._h, MODES[mode], ctypes.byref(e), None, None, None, None,
                             None, ctypes.byref(d))
        self._check(rc, "mbd_run")
        return e.value, _diag_dict(d)

    def scs(self, mode: str = "fp64"):
        N = self.n_atoms
        c6m = ctypes.c_double()
        aeff, c6e = np.zeros(N), np.zeros(N)
        cpol = np.zeros(20 * N)
        mp = np.zeros(63)
        d = MbdDiag()
        rc = self._L.mbd_scs(self._h, MODES[mode], dptr(aeff), dptr(cpol), dptr(mp), dptr(c6e),
                             ctypes.byref(c6m), ctypes.byref(d))
        self._check(rc, "mbd_scs")
        return aeff, cpol.reshape(N, 20), mp.reshape(21, 3), c6e, c6m.value, _diag_dict(d)

    def energy(self, alpha_eff, c6_eff):
        N = self.n_atoms
        a = np.ascontiguousarray(alpha_eff, dtype=np.float64)
        c = np.ascontiguousarray(c6_eff, dtype=np.float64)
        e = ctypes.c_double()
        om, rf = np.zeros(N), np.zeros(N)
        d = MbdDiag()
        rc = self._L.mbd_energy(self._h, dptr(a), dptr(c), ctypes.byref(e), dptr(om), dptr(rf),
                                ctypes.byref(d))
        self._check(rc, "mbd_energy")
        return e.value, om, rf, _diag_dict(d)

    # -- batched molecules -------------------------------------------------------
    def batch_set(self, molecules, settings=None):
        """Many non-periodic molecules (mbd_amd.io.Geometry, Angstrom) for the
        batched kernel; settings gives xc and the SCS / CFDM cutoffs."""
        from .io import BOHR, Settings
        st = settings or Settings()
        offs = [0]
        xyz, hir, a0, c6, r0 = [], [], [], [], []
        for g in molecules:
            if g.periodic:
                raise ValueError("the batched form takes non-periodic molecules")
            for nm in g.names:
                c, a, r = free_atom_params(nm)
                c6.append(c)
                a0.append(a)
                r0.append(r)
            xyz.append(np.asarray(g.coords_ang, dtype=np.float64) / BOHR)
            hir.append(np.asarray(g.hirshfeld, dtype=np.float64))
            offs.append(offs[-1] + g.n_atoms)
        self.batch_set_params(np.array(offs, dtype=np.int32), np.concatenate(xyz),
                              np.concatenate(hir), np.array(a0), np.array(c6), np.array(r0),
                              beta(st.flag_xc), st.mbd_scs_dip_cutoff, st.mbd_cfdm_dip_cutoff)

    def batch_set_params(self, offsets, xyz_bohr, hirshfeld, alpha0, c6_0, r0, beta_,
                         rc_scs=120.0, rc_cfdm=100.0):
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        arrs = [np.ascontiguousarray(x, dtype=np.float64)
                for x in (xyz_bohr, hirshfeld, alpha0, c6_0, r0)]
        rc = self._L.mbd_batch_set(self._h, len(off) - 1, iptr(off), *[dptr(x) for x in arrs],
                                   float(beta_), float(rc_scs), float(rc_cfdm))
        self._check(rc, "mbd_batch_set")
        self._batch_off = off

    def batch_tune(self, mode: str = "mixed", reps: int = 3):
        """Times every eligible launch shape; returns (best index, times in ms)."""
        t = np.zeros(batch_num_configs())
        best = ctypes.c_int()
        self._check(self._L.mbd_batch_tune(self._h, MODES[mode], reps, dptr(t), ctypes.byref(best)),
                    "mbd_batch_tune")
        return best.value, t

    def batch_set_config(self, index: int):
        """Force launch-shape candidate `index` (-1: default)."""
        self._check(self._L.mbd_batch_set_config(self._h, int(index)), "mbd_batch_set_config")

    def batch_run(self, mode: str = "mixed", full: bool = True) -> dict:
        off = self._batch_off
        nmol, na = len(off) - 1, int(off[-1])
        e = np.zeros(nmol)
        out = {"energy": e}
        d = MbdBatchDiag()
        if full:
            ae, c6, c6m, aa = np.zeros(na), np.zeros(na), np.zeros(nmol), np.zeros(21 * na)
            st = np.zeros(4 * nmol, dtype=np.int32)
            rc = self._L.mbd_batch_run(self._h, MODES[mode], dptr(e), dptr(ae), dptr(c6),
                                       dptr(c6m), dptr(aa), iptr(st), ctypes.byref(d))
            out.update(alpha_eff=ae, c6_eff=c6, c6_mol=c6m, alpha_all=aa.reshape(21, na),
                       stat=st.reshape(nmol, 4))
        else:
            rc = self._L.mbd_batch_run(self._h, MODES[mode], dptr(e), None, None, None, None, None,
                                       ctypes.byref(d))
        self._check(rc, "mbd_batch_run")
        out["diag"] = {f: getattr(d, f) for f, _ in MbdBatchDiag._fields_ if f != "phase_ms"}
        out["diag"]["phase_ms"] = dict(zip(("relay_build", "factor", "solve_refine", "qho_build",
                                            "tridiag", "bisect_energy"), list(d.phase_ms)))
        out["offsets"] = off
        return out

    # -- learned form -----------------------------------------------------------
    def net_set_weights(self, blob):
        """Load network weights (float32 blob, see net_layout); None unloads."""
        if blob is None:
            self._check(self._L.mbd_net_set_weights(self._h, None, 0), "mbd_net_set_weights")
            return
        b = np.ascontiguousarray(blob, dtype=np.float32)
        self._check(self._L.mbd_net_set_weights(self._h, fptr(b), b.size), "mbd_net_set_weights")

    def net_descriptors(self):
        """(X, alpha_w): descriptors (21 N, n_in) and bare alpha_i(i w_k) (21 N),
        sample s = k N + i."""
        n_in, _, _ = net_layout()
        S = 21 * self.n_atoms
        X = np.zeros(n_in * S, dtype=np.float32)
        aw = np.zeros(S)
        self._check(self._L.mbd_net_descriptors(self._h, fptr(X), dptr(aw)), "mbd_net_descriptors")
        return X.reshape(n_in, S).T.copy(), aw

    # -- instrumentation -------------------------------------------------------
    def profile_enable(self, on: bool = True):
        self._check(self._L.mbd_profile_enable(self._h, int(on)), "mbd_profile_enable")

    def profile_read(self) -> dict:
        """{kernel class: {launches, ms, bytes, flops}} since profile_enable."""
        out = {}
        cls = 0
        buf = np.zeros(4)
        name = ctypes.c_char_p()
        while self._L.mbd_profile_read(self._h, cls, dptr(buf), ctypes.byref(name)) == 0:
            out[name.value.decode()] = {"launches": int(buf[0]), "ms": float(buf[1]),
                                        "bytes": float(buf[2]), "flops": float(buf[3])}
            cls += 1
        return out

    # -- test hooks -----------------------------------------------------------
    def relay_matrix(self, k: int, mode: str = "fp64") -> np.ndarray:
        n = 3 * self.n_atoms
        A = np.zeros(n * n)
        self._check(self._L.mbd_debug_relay_matrix(self._h, MODES[mode], k, dptr(A)),
                    "mbd_debug_relay_matrix")
        return A.reshape(n, n).T.copy()

    def cfdm_matrix(self, alpha_eff, omega, reff, ncell: int) -> np.ndarray:
        nm = 3 * self.n_atoms * ncell
        C = np.zeros(nm * nm)
        a = np.ascontiguousarray(alpha_eff, dtype=np.float64)
        o = np.ascontiguousarray(omega, dtype=np.float64)
        r = np.ascontiguousarray(reff, dtype=np.float64)
        self._check(self._L.mbd_debug_cfdm_matrix(self._h, dptr(a), dptr(o), dptr(r), dptr(C)),
                    "mbd_debug_cfdm_matrix")
        return C.reshape(nm, nm).T.copy()


def device_cholesky(A_batch: np.ndarray, device: int = 0):
    """Lower Cholesky factors of a (batch, n, n) stack on the device."""
    A = np.asarray(A_batch, dtype=np.float64)
    if A.ndim == 2:
        A = A[None]
    b, n, _ = A.shape
    src = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))  # column-major per matrix
    L = np.zeros_like(src)
    info = np.zeros(b, dtype=np.int32)
    rc = _lib.lib().mbd_debug_cholesky(device, n, b, dptr(src), dptr(L), iptr(info))
    if rc != 0:
        raise MbdError(f"mbd_debug_cholesky failed ({rc})")
    return np.tril(np.transpose(L, (0, 2, 1))), info


def device_lu_solve(A_batch: np.ndarray, device: int = 0):
    """X = A^-1 [I; I; ...] (n x 3) for a (batch, n, n) stack with the device
    LU (partial pivoting); returns (X (batch, n, 3), info (batch,))."""
    A = np.asarray(A_batch, dtype=np.float64)
    if A.ndim == 2:
        A = A[None]
    b, n, _ = A.shape
    src = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))  # column-major per matrix
    X = np.zeros((b, 3, n))
    info = np.zeros(b, dtype=np.int32)
    rc = _lib.lib().mbd_debug_lu_solve(device, n, b, dptr(src), dptr(X), iptr(info))
    if rc != 0:
        raise MbdError(f"mbd_debug_lu_solve failed ({rc})")
    return np.transpose(X, (0, 2, 1)), info


def device_sym_eigvals(A: np.ndarray, device: int = 0) -> np.ndarray:
    n = A.shape[0]
    src = np.asfortranarray(np.asarray(A, dtype=np.float64))
    w = np.zeros(n)
    rc = _lib.lib().mbd_debug_sym_eigvals(device, n, src.ctypes.data_as(_lib._D), dptr(w))
    if rc != 0:
        raise MbdError(f"mbd_debug_sym_eigvals failed ({rc})")
    return w


def device_gemm_nt(A, B, C, alpha=1.0, beta_=0.0, fp32=False, device: int = 0):
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    Cf = np.asfortranarray(C, dtype=np.float64).copy(order="F")
    m, k = A.shape
    n = B.shape[0]
    rc = _lib.lib().mbd_debug_gemm_nt(device, int(fp32), m, n, k, alpha,
                                      A.ctypes.data_as(_lib._D), B.ctypes.data_as(_lib._D),
                                      beta_, Cf.ctypes.data_as(_lib._D))
    if rc != 0:
        raise MbdError(f"mbd_debug_gemm_nt failed ({rc})")
    return Cf


def device_copy(n_bytes: int, device: int = 0) -> float:
    """Plain device-to-device copy of n_bytes with the library's streaming
    kernel; returns its duration in ms."""
    ms = ctypes.c_double()
    rc = _lib.lib().mbd_debug_copy(device, int(n_bytes), ctypes.byref(ms))
    if rc != 0:
        raise MbdError(f"mbd_debug_copy failed ({rc})")
    return ms.value


def batch_num_configs() -> int:
    n = 0
    while _lib.lib().mbd_batch_config(n, None, None, None) == 0:
        n += 1
    return n


def batch_config(index: int):
    nb, nt, pc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    if _lib.lib().mbd_batch_config(index, ctypes.byref(nb), ctypes.byref(nt), ctypes.byref(pc)):
        raise IndexError(index)
    return {"nb": nb.value, "nt": nt.value, "per_cu": pc.value}


def net_layout():
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_long()
    _lib.lib().mbd_net_layout(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value


def _feat_major(X):
    """(B, n_in) samples -> the ABI's sample-fastest layout X[s + d B]."""
    return np.ascontiguousarray(np.asarray(X, dtype=np.float32).T)


def net_forward(blob, X, device: int = 0):
    b = np.ascontiguousarray(blob, dtype=np.float32)
    Xf = _feat_major(X)
    B = Xf.shape[1]
    y = np.zeros(B, dtype=np.float32)
    rc = _lib.lib().mbd_net_forward(device, fptr(b), B, fptr(Xf), fptr(y))
    if rc != 0:
        raise MbdError(f"mbd_net_forward failed ({rc})")
    return y


def net_grad(blob, X, t, device: int = 0):
    b = np.ascontiguousarray(blob, dtype=np.float32)
    Xf = _feat_major(X)
    B = Xf.shape[1]
    tt = np.ascontiguousarray(t, dtype=np.float32)
    _, _, nb = net_layout()
    n_in = Xf.shape[0]
    g = np.zeros(nb - 2 * n_in, dtype=np.float32)
    loss = ctypes.c_float()
    rc = _lib.lib().mbd_net_grad(device, fptr(b), B, fptr(Xf), fptr(tt), fptr(g), ctypes.byref(loss))
    if rc != 0:
        raise MbdError(f"mbd_net_grad failed ({rc})")
    return g, loss.value


def net_train(blob, X, t, epochs=30, batch=1024, lr=1e-3, seed=1, device: int = 0):
    """Train on the device; returns (trained blob, per-epoch losses)."""
    b = np.array(blob, dtype=np.float32, copy=True)
    Xf = _feat_major(X)
    S = Xf.shape[1]
    tt = np.ascontiguousarray(t, dtype=np.float32)
    hist = np.zeros(epochs, dtype=np.float32)
    rc = _lib.lib().mbd_net_train(device, S, fptr(Xf), fptr(tt), epochs, batch, lr, seed,
                                  fptr(b), fptr(hist))
    if rc != 0:
        raise MbdError(f"mbd_net_train failed ({rc})")
    return b, hist
