"""Batched form: many non-periodic molecules in one persistent launch.

Each molecule's energy, alpha_eff, C6 and molecular C6 are checked against
the oracle (FP64 form within 1e-8 Ha, mixed form within 1e-6 Ha and C6 within
1e-5 relative), the 512-molecule batch against its own FP64 form, graph
replays against each other, every launch shape against the default one (FP64)
and against the FP64 form (mixed), and the autotuner's choice in the mixed
form."""
import numpy as np
import pytest

from mbd_amd import engine, generators, io
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _oracle(g, st):
    xyz = np.asarray(g.coords_ang) / io.BOHR
    return O.mbd_run(O.system_from(g.names, xyz, g.hirshfeld, np.zeros((3, 3)), False, st))


def _small_batch():
    rng = np.random.default_rng(7)
    mols = [io.Geometry(["C"], np.zeros((1, 3)), np.array([0.8])),
            io.Geometry(["N", "N"], np.array([[0.0, 0, 0], [0, 0, 1.1]]), np.array([0.7, 0.9]))]
    ring, _ = generators.ring()
    mols.append(io.Geometry(ring.names, ring.coords_ang, ring.hirshfeld))
    for n in (6, 9, 14, 23, 31, 40, 57):
        mols.append(generators.chain_molecule(rng, n))
    return mols


@pytest.mark.parametrize("mode,tol_e,tol_c6", [("fp64", 1e-8, 1e-9), ("mixed", 1e-6, 1e-5)])
def test_batch_matches_oracle(mode, tol_e, tol_c6):
    mols = _small_batch()
    st = io.Settings()
    print(f"batch {mode}: {len(mols)} molecules, sizes {[m.n_atoms for m in mols]}", flush=True)
    with engine.MbdEngine(0) as eng:
        eng.batch_set(mols, st)
        out = eng.batch_run(mode)
    off = out["offsets"]
    for k, g in enumerate(mols):
        want = _oracle(g, st)
        sl = slice(off[k], off[k + 1])
        assert abs(out["energy"][k] - want.e_mbd) <= tol_e, (k, out["energy"][k], want.e_mbd)
        assert np.allclose(out["alpha_eff"][sl], want.alpha_eff, rtol=tol_c6, atol=0)
        assert np.allclose(out["c6_eff"][sl], want.c6_eff, rtol=tol_c6, atol=0)
        assert out["c6_mol"][k] == pytest.approx(want.c6_mol, rel=tol_c6)
        assert out["stat"][k, 0] == 0
    assert abs(out["energy"][0]) <= 1e-14  # one atom: E = (3/2) w - (3/2) w
    d = out["diag"]
    print(f"diag {d}")
    assert d["n_failed"] == 0 and d["max_residual"] <= (1e-12 if mode == "mixed" else 1e-10)


def test_batch_512_mixed_vs_fp64_and_spot_oracle():
    mols, st = generators.chain_molecules()
    print("batch: 512 chain molecules, mixed and fp64", flush=True)
    with engine.MbdEngine(0) as eng:
        eng.batch_set(mols, st)
        m = eng.batch_run("mixed")
        f = eng.batch_run("fp64")
    print(f"mixed {m['diag']}\nfp64 {f['diag']}")
    assert m["diag"]["n_failed"] == 0 and f["diag"]["n_failed"] == 0
    de = np.abs(m["energy"] - f["energy"])
    assert de.max() <= 1e-6
    assert np.max(np.abs(m["c6_eff"] / f["c6_eff"] - 1)) <= 1e-5
    assert np.max(np.abs(m["alpha_eff"] / f["alpha_eff"] - 1)) <= 1e-5
    sizes = np.array([g.n_atoms for g in mols])
    for k in (int(np.argmax(sizes)), int(np.argmin(sizes)), int(np.argsort(sizes)[256])):
        want = _oracle(mols[k], st)
        assert abs(f["energy"][k] - want.e_mbd) <= 1e-8, k
        assert abs(m["energy"][k] - want.e_mbd) <= 1e-6, k


def test_graph_replays_are_identical():
    mols = _small_batch()
    with engine.MbdEngine(0) as eng, engine.MbdEngine(0) as other:
        eng.batch_set(mols, io.Settings())
        geo, st = generators.ring()
        other.set_geometry(geo, st)
        runs = []
        for _ in range(3):
            runs.append(eng.batch_run("mixed"))
            other.run("fp64")  # eager work between replays
    for r in runs[1:]:
        assert np.array_equal(r["energy"], runs[0]["energy"])
        assert np.array_equal(r["alpha_all"], runs[0]["alpha_all"])


def test_every_launch_shape_matches_default():
    """Each candidate launch shape, forced in turn, against the default one
    (FP64 form: only the panel width changes the rounding)."""
    mols, st = generators.chain_molecules(n_mol=64, seed=11)
    with engine.MbdEngine(0) as eng:
        eng.batch_set(mols, st)
        ref = eng.batch_run("fp64")
        worst = {}
        for c in range(engine.batch_num_configs()):
            cfg = engine.batch_config(c)
            print(f"shape {c} {cfg}", flush=True)
            eng.batch_set_config(c)
            got = eng.batch_run("fp64")
            assert got["diag"]["config"] == c
            rel = np.abs(got["energy"] - ref["energy"]) / np.maximum(1.0, np.abs(ref["energy"]))
            worst[c] = float(rel.max())
            print(f"shape {c} max rel dE {worst[c]:.3e} at molecule {int(rel.argmax())} "
                  f"(E={ref['energy'][rel.argmax()]:.6e}) t={got['diag']['t_ms']:.2f} ms", flush=True)
    assert max(worst.values()) <= 1e-10, worst


def test_tuner_mixed_over_every_shape():
    """Every launch shape in the mixed form, forced in turn, against the FP64
    form within the mixed tolerance (1 uHa), then the autotuner in the mixed
    form: it times every shape, keeps the fastest, and the run after it uses
    that shape."""
    mols, st = generators.chain_molecules(n_mol=64, seed=11)
    with engine.MbdEngine(0) as eng:
        eng.batch_set(mols, st)
        ref = eng.batch_run("fp64")
        for c in range(engine.batch_num_configs()):
            print(f"mixed shape {c} {engine.batch_config(c)}", flush=True)
            eng.batch_set_config(c)
            got = eng.batch_run("mixed")
            assert got["diag"]["config"] == c
            de = float(np.abs(got["energy"] - ref["energy"]).max())
            print(f"mixed shape {c} max |dE| {de:.3e} t={got['diag']['t_ms']:.2f} ms", flush=True)
            assert de <= 1e-6
            assert got["diag"]["fallbacks"] == 0
        best, t = eng.batch_tune("mixed", reps=2)
        print(f"tuned: {best} times {t}", flush=True)
        assert np.all(t > 0)
        assert best == int(np.argmin(t))
        got = eng.batch_run("mixed")
        assert got["diag"]["config"] == best
        assert float(np.abs(got["energy"] - ref["energy"]).max()) <= 1e-6


def test_batch_rejects_bad_input():
    with engine.MbdEngine(0) as eng:
        big = io.Geometry(["C"] * 600, np.arange(1800, dtype=float).reshape(600, 3) * 0.5,
                          np.full(600, 0.8))
        with pytest.raises(engine.MbdError, match="1005"):
            eng.batch_set([big])
        with pytest.raises(engine.MbdError):
            eng.batch_set_params(np.array([0, 2, 2], dtype=np.int32), np.zeros((2, 3)),
                                 np.ones(2), np.ones(2), np.ones(2), np.ones(2), 0.83)
        e = np.zeros(1)
        assert eng._L.mbd_batch_run(eng._h, 1, engine.dptr(e), None, None, None, None, None,
                                    None) == 1002  # no batch set
        assert eng._L.mbd_batch_run(eng._h, 2, engine.dptr(e), None, None, None, None, None,
                                    None) == -2  # learned mode is not batched
