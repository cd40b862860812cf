"""The no-allocation batched entry point (mbd_batch_run_noalloc) against the
allocating one (mbd_batch_run) on the same molecules, through the C ABI with
device arrays: energies, alpha_SCS at every grid point, C6 and molecular C6
bit for bit, and the per-molecule statistics equal.  Two calls in a row, the
second with fewer molecules: the library's scratch keeps the first call's
contents, so everything that must start at zero has to be cleared by the
call itself."""
import ctypes

import numpy as np
import pytest

from mbd_amd import generators, io
from mbd_amd.engine import MODES, MbdEngine, beta, free_atom_params

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers for the test through the HIP runtime the library itself
    links (one runtime per process)."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0
        self.bufs.append(p)
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2) == 0
        return out

    def sync(self):
        assert self.rt.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _mols(count, seed):
    rng = np.random.default_rng(seed)
    return [generators.chain_molecule(rng, int(k)) for k in rng.integers(6, 31, count)]


def _device_inputs(hip, mols):
    offs, xyz, hir, a0, c6, r0 = [0], [], [], [], [], []
    for g in mols:
        for nm in g.names:
            c, a, r = free_atom_params(nm)
            c6.append(c)
            a0.append(a)
            r0.append(r)
        xyz.append(np.asarray(g.coords_ang, dtype=np.float64) / io.BOHR)
        hir.append(np.asarray(g.hirshfeld, dtype=np.float64))
        offs.append(offs[-1] + g.n_atoms)
    arrs = dict(off=hip.put(np.array(offs, dtype=np.int32)),
                xyz=hip.put(np.concatenate(xyz).astype(np.float64)),
                hir=hip.put(np.concatenate(hir).astype(np.float64)),
                a0=hip.put(np.array(a0)), c6=hip.put(np.array(c6)), r0=hip.put(np.array(r0)))
    return arrs, offs[-1], max(g.n_atoms for g in mols)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_noalloc_matches_allocating_batch(mode):
    st = io.Settings()
    hip = _Hip()
    with MbdEngine(0) as eng:
        for call, (count, seed) in enumerate(((40, 11), (9, 12))):
            mols = _mols(count, seed)
            eng.batch_set(mols, st)
            want = eng.batch_run(mode)
            inp, na, amax = _device_inputs(hip, mols)
            assert amax <= 32
            host = dict(e=np.full(count, -7.0), aa=np.full(21 * na, -7.0), c6=np.full(na, -7.0),
                        c6m=np.full(count, -7.0), st=np.full(4 * count, -7, dtype=np.int32))
            out = {k: hip.put(v) for k, v in host.items()}
            hip.sync()
            rc = eng._L.mbd_batch_run_noalloc(
                eng._h, None, MODES[mode], count, na, amax, inp["off"], inp["xyz"], inp["hir"],
                inp["a0"], inp["c6"], inp["r0"], float(beta(st.flag_xc)),
                float(st.mbd_scs_dip_cutoff), float(st.mbd_cfdm_dip_cutoff), out["e"], out["aa"],
                out["c6"], out["c6m"], out["st"])
            assert rc == 0, eng._L.mbd_last_error(eng._h)
            got = {k: hip.get(out[k], host[k]) for k in host}
            print(f"call {call} ({mode}, {count} molecules): max |dE| = "
                  f"{np.abs(got['e'] - want['energy']).max():.2e}", flush=True)
            assert np.array_equal(got["e"], want["energy"])
            assert np.array_equal(got["aa"].reshape(21, na), want["alpha_all"])
            assert np.array_equal(got["c6"], want["c6_eff"])
            assert np.array_equal(got["c6m"], want["c6_mol"])
            assert np.array_equal(got["st"].reshape(count, 4), want["stat"])
    hip.free()


def test_noalloc_rejects_large_molecules():
    rng = np.random.default_rng(3)
    mols = [generators.chain_molecule(rng, 40)]
    hip = _Hip()
    with MbdEngine(0) as eng:
        inp, na, amax = _device_inputs(hip, mols)
        buf = hip.put(np.zeros(21 * na))
        sti = hip.put(np.zeros(4, dtype=np.int32))
        rc = eng._L.mbd_batch_run_noalloc(
            eng._h, None, MODES["mixed"], 1, na, amax, inp["off"], inp["xyz"], inp["hir"],
            inp["a0"], inp["c6"], inp["r0"], 0.83, 120.0, 100.0, buf, buf, buf, buf, sti)
        assert rc == 1005  # MBD_ERR_TOOBIG
    hip.free()
