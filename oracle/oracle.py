"""ctypes view of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's CPU-baseline leg, always as the checker.  The product path
(mbd_amd) never imports this module.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
import subprocess
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    if force and os.path.exists(_LIB_PATH):
        os.remove(_LIB_PATH)
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_vdw_param.argtypes = [ctypes.c_char_p, _D, _D, _D]
        L.or_vdw_param.restype = ctypes.c_int
        L.or_beta.argtypes = [ctypes.c_int]
        L.or_beta.restype = ctypes.c_double
        L.or_casimir_grid.argtypes = [_D, _D]
        L.or_image_ranges.argtypes = [_D, ctypes.c_int, _I, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, _I, _I, _I]
        L.or_alpha_sigma.argtypes = [ctypes.c_double] * 4 + [_D, _D]
        L.or_sr_tensor.argtypes = [_D, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, _D]
        L.or_lr_tensor.argtypes = [_D, ctypes.c_double, ctypes.c_double, ctypes.c_double, _D]
        L.or_relay_matrix.argtypes = [ctypes.c_int, _D, _D, _I, ctypes.c_double, ctypes.c_double,
                                      _D, _D, _D, _D]
        L.or_cfdm_matrix.argtypes = [ctypes.c_int, _D, _D, _I, _I, ctypes.c_double,
                                     ctypes.c_double, _D, _D, _D, _D]
        L.or_cholesky.argtypes = [ctypes.c_int, _D, ctypes.c_int]
        L.or_cholesky.restype = ctypes.c_int
        L.or_cholesky_solve.argtypes = [ctypes.c_int, _D, ctypes.c_int, ctypes.c_int, _D,
                                        ctypes.c_int]
        L.or_sym_eigvals.argtypes = [ctypes.c_int, _D, ctypes.c_int, _D]
        L.or_sym_eigvals.restype = ctypes.c_int
        L.or_tridiag_eigvals.argtypes = [ctypes.c_int, _D, _D, _D]
        L.or_sym3_eigvals.argtypes = [_D, _D]
        L.or_mbd_run.argtypes = ([ctypes.c_int, _D, _D, _D, _D, _D, _D, ctypes.c_int, _I,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double]
                                 + [_D] * 8 + [_I, _I, _I])
        L.or_mbd_run.restype = ctypes.c_int
        L.or_scs_all.argtypes = [ctypes.c_int, _D, _D, _D, _D, _D, _D, ctypes.c_int, _I,
                                 ctypes.c_double, ctypes.c_double, _D, _D, _D, _D, _D, _I]
        L.or_time_scs_frequency.argtypes = [ctypes.c_int, _D, _D, _D, _D, _D, _D, ctypes.c_int,
                                            _I, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.or_time_scs_frequency.restype = ctypes.c_double
        L.or_time_eigvals.argtypes = [ctypes.c_int, ctypes.c_ulonglong]
        L.or_time_eigvals.restype = ctypes.c_double
        L.or_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_D)


def _pi(a: np.ndarray):
    return a.ctypes.data_as(_I)


def vdw_param(symbol: str):
    c6, a, r = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    rc = lib().or_vdw_param(symbol.encode(), ctypes.byref(c6), ctypes.byref(a), ctypes.byref(r))
    if rc != 0:
        raise KeyError(symbol)
    return c6.value, a.value, r.value


def beta(flag_xc: int) -> float:
    return lib().or_beta(flag_xc)


def casimir_grid():
    w = np.zeros(21)
    g = np.zeros(21)
    lib().or_casimir_grid(_p(w), _p(g))
    return w, g


def image_ranges(lat_bohr, periodic, vacuum, rc_scs, rc_cfdm, rc_sc):
    lat = np.ascontiguousarray(np.asarray(lat_bohr, dtype=np.float64).reshape(3, 3))
    vac = np.array([int(v) for v in vacuum], dtype=np.int32)
    a, b, c = (np.zeros(3, dtype=np.int32) for _ in range(3))
    lib().or_image_ranges(_p(lat), int(periodic), _pi(vac), rc_scs, rc_cfdm, rc_sc,
                          _pi(a), _pi(b), _pi(c))
    return a, b, c


def alpha_sigma(h, c6, a0, w):
    a, s = ctypes.c_double(), ctypes.c_double()
    lib().or_alpha_sigma(h, c6, a0, w, ctypes.byref(a), ctypes.byref(s))
    return a.value, s.value


def sr_tensor(d, r, sgm, rvdw, beta_):
    d = np.ascontiguousarray(d, dtype=np.float64)
    T = np.zeros(9)
    lib().or_sr_tensor(_p(d), r, sgm, rvdw, beta_, _p(T))
    return T.reshape(3, 3).T.copy()


def lr_tensor(d, r, rvdw, beta_):
    d = np.ascontiguousarray(d, dtype=np.float64)
    T = np.zeros(9)
    lib().or_lr_tensor(_p(d), r, rvdw, beta_, _p(T))
    return T.reshape(3, 3).T.copy()


def cholesky(A: np.ndarray):
    """Lower Cholesky of a symmetric matrix; returns (L, info)."""
    n = A.shape[0]
    F = np.asfortranarray(A.copy())
    info = lib().or_cholesky(n, F.ctypes.data_as(_D), n)
    return np.tril(F), info


def sym_eigvals(A: np.ndarray) -> np.ndarray:
    n = A.shape[0]
    F = np.asfortranarray(A.copy())
    w = np.zeros(n)
    lib().or_sym_eigvals(n, F.ctypes.data_as(_D), n, _p(w))
    return w


def tridiag_eigvals(d, e) -> np.ndarray:
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e if len(e) else np.zeros(1), dtype=np.float64)
    w = np.zeros(len(d))
    lib().or_tridiag_eigvals(len(d), _p(d), _p(e), _p(w))
    return w


def sym3_eigvals(P) -> np.ndarray:
    P = np.asfortranarray(np.asarray(P, dtype=np.float64))
    w = np.zeros(3)
    lib().or_sym3_eigvals(P.ctypes.data_as(_D), _p(w))
    return w


@dataclasses.dataclass
class OracleSystem:
    """Inputs of one evaluation, in bohr, with free-atom data resolved."""
    xyz: np.ndarray          # (N, 3) bohr
    hirsh: np.ndarray
    alpha0: np.ndarray
    c6_0: np.ndarray
    r0: np.ndarray
    lattice: np.ndarray      # (3, 3) bohr, row k = vector k
    periodic: bool
    vacuum: tuple
    beta: float
    rc_scs: float            # Angstrom
    rc_cfdm: float
    rc_sc: float

    @property
    def n_atoms(self):
        return self.xyz.shape[0]


def system_from(names, xyz_bohr, hirsh, lattice_bohr, periodic, settings) -> OracleSystem:
    a0, c6, r0 = [], [], []
    for nm in names:
        c, a, r = vdw_param(nm)
        c6.append(c); a0.append(a); r0.append(r)
    return OracleSystem(np.ascontiguousarray(xyz_bohr, dtype=np.float64),
                        np.ascontiguousarray(hirsh, dtype=np.float64),
                        np.array(a0), np.array(c6), np.array(r0),
                        np.ascontiguousarray(lattice_bohr, dtype=np.float64).reshape(3, 3),
                        bool(periodic), tuple(bool(v) for v in settings.mbd_scs_vacuum_axis),
                        beta(settings.flag_xc), settings.mbd_scs_dip_cutoff,
                        settings.mbd_cfdm_dip_cutoff, settings.mbd_supercell_cutoff)


@dataclasses.dataclass
class OracleResult:
    e_mbd: float
    alpha_eff: np.ndarray
    coupled_pol: np.ndarray  # (N, 20): alpha_SCS(i w_k), k = 1..20
    molpol_eig: np.ndarray   # (21, 3)
    c6_mol: float
    c6_eff: np.ndarray
    omega: np.ndarray
    reff: np.ndarray
    supercell: np.ndarray
    nimag: int
    info: int


def mbd_run(s: OracleSystem) -> OracleResult:
    N = s.n_atoms
    vac = np.array([int(v) for v in s.vacuum], dtype=np.int32)
    e = ctypes.c_double()
    c6m = ctypes.c_double()
    aeff, c6e, om, ref = (np.zeros(N) for _ in range(4))
    cpol = np.zeros(20 * N)
    mp = np.zeros(63)
    sc = np.zeros(3, dtype=np.int32)
    nimag, info = ctypes.c_int(), ctypes.c_int()
    lat = np.ascontiguousarray(s.lattice, dtype=np.float64)
    lib().or_mbd_run(N, _p(s.xyz), _p(s.hirsh), _p(s.alpha0), _p(s.c6_0), _p(s.r0), _p(lat),
                     int(s.periodic), _pi(vac), s.beta, s.rc_scs, s.rc_cfdm, s.rc_sc,
                     ctypes.byref(e), _p(aeff), _p(cpol), _p(mp), ctypes.byref(c6m), _p(c6e),
                     _p(om), _p(ref), _pi(sc), ctypes.byref(nimag), ctypes.byref(info))
    return OracleResult(e.value, aeff, cpol.reshape(N, 20), mp.reshape(21, 3), c6m.value,
                        c6e, om, ref, sc, nimag.value, info.value)


def scs_all(s: OracleSystem):
    """SCS half only: (alpha_eff, coupled_pol (N, 20), c6_eff, c6_mol, info)."""
    N = s.n_atoms
    vac = np.array([int(v) for v in s.vacuum], dtype=np.int32)
    aeff, c6e = np.zeros(N), np.zeros(N)
    cpol = np.zeros(20 * N)
    c6m, info = ctypes.c_double(), ctypes.c_int()
    lat = np.ascontiguousarray(s.lattice, dtype=np.float64)
    lib().or_scs_all(N, _p(s.xyz), _p(s.hirsh), _p(s.alpha0), _p(s.c6_0), _p(s.r0), _p(lat),
                     int(s.periodic), _pi(vac), s.beta, s.rc_scs, _p(aeff), _p(cpol), None,
                     ctypes.byref(c6m), _p(c6e), ctypes.byref(info))
    return aeff, cpol.reshape(N, 20), c6e, c6m.value, info.value


def relay_matrix(s: OracleSystem, freq_index: int) -> np.ndarray:
    """Full symmetric A(i w_k) for one grid point, as built by the oracle."""
    N = s.n_atoms
    w, _ = casimir_grid()
    aw = np.zeros(N); sg = np.zeros(N)
    for i in range(N):
        aw[i], sg[i] = alpha_sigma(s.hirsh[i], s.c6_0[i], s.alpha0[i], w[freq_index])
    rv = np.power(s.hirsh, 0.33333333333333333333) * s.r0
    scs, _, _ = image_ranges(s.lattice, s.periodic, s.vacuum, s.rc_scs, s.rc_cfdm, s.rc_sc)
    A = np.zeros(9 * N * N)
    lat = np.ascontiguousarray(s.lattice, dtype=np.float64)
    lib().or_relay_matrix(N, _p(s.xyz), _p(lat), _pi(scs), s.rc_scs, s.beta, _p(aw), _p(sg),
                          _p(rv), _p(A))
    return A.reshape(3 * N, 3 * N).T.copy()


def cfdm_matrix(s: OracleSystem, alpha_eff, omega, reff) -> np.ndarray:
    N = s.n_atoms
    _, sc, cf = image_ranges(s.lattice, s.periodic, s.vacuum, s.rc_scs, s.rc_cfdm, s.rc_sc)
    nm = 3 * N * int(np.prod(sc))
    C = np.zeros(nm * nm)
    lat = np.ascontiguousarray(s.lattice, dtype=np.float64)
    a = np.ascontiguousarray(alpha_eff, dtype=np.float64)
    o = np.ascontiguousarray(omega, dtype=np.float64)
    r = np.ascontiguousarray(reff, dtype=np.float64)
    lib().or_cfdm_matrix(N, _p(s.xyz), _p(lat), _pi(sc), _pi(cf), s.rc_cfdm, s.beta, _p(a),
                         _p(o), _p(r), _p(C))
    return C.reshape(nm, nm).T.copy()


def time_scs_frequency(s: OracleSystem, freq_index: int = 10) -> float:
    vac = np.array([int(v) for v in s.vacuum], dtype=np.int32)
    lat = np.ascontiguousarray(s.lattice, dtype=np.float64)
    return lib().or_time_scs_frequency(s.n_atoms, _p(s.xyz), _p(s.hirsh), _p(s.alpha0),
                                       _p(s.c6_0), _p(s.r0), _p(lat), int(s.periodic), _pi(vac),
                                       s.beta, s.rc_scs, freq_index)


def time_eigvals(n: int, seed: int = 7) -> float:
    return lib().or_time_eigvals(n, seed)


def num_threads() -> int:
    return lib().or_num_threads()
