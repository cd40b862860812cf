"""Seeded synthetic systems for the benchmark configurations.

Each generator returns ``(Geometry, Settings)`` in the units of geometry.in
(Angstrom).  The species-count rule rounds each fraction of ``n`` and gives the
remainder to the largest fractional parts (ties in listed order), so 1024 atoms
at C:H:N:O = 40:40:10:10 are 410/410/102/102 and 4096 are 1638/1638/410/410.

* ring      12-atom hexagonal ring molecule (C at 1.39 A, H at 2.48 A).
* random_periodic  uniform random atoms in a periodic cube, rejection-sampled
                   to a minimum-image separation; Hirshfeld ratios U(0.60, 0.95).
* chain_molecules  random-walk chains with H caps (batched form).
* hex_sheet        2D periodic hexagonal sheet with vacuum axis z and jitter.

The periodic-sheet construction generalises the reference's
``examples/process.py`` (an i x j replication of a periodic cell).
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import numpy as np

from .io import Geometry, Settings


def species_counts(n: int, fractions: Sequence[Tuple[str, float]]):
    raw = [(s, n * f) for s, f in fractions]
    base = [int(math.floor(x)) for _, x in raw]
    rem = n - sum(base)
    order = sorted(range(len(raw)), key=lambda i: (-(raw[i][1] - base[i]), i))
    for i in order[:rem]:
        base[i] += 1
    return [(raw[i][0], base[i]) for i in range(len(raw))]


def ring() -> Tuple[Geometry, Settings]:
    names, xyz, h = [], [], []
    for k in range(6):
        a = math.radians(60.0 * k)
        names.append("C")
        xyz.append([1.39 * math.cos(a), 1.39 * math.sin(a), 0.0])
        h.append(0.806)
    for k in range(6):
        a = math.radians(60.0 * k)
        names.append("H")
        xyz.append([2.48 * math.cos(a), 2.48 * math.sin(a), 0.0])
        h.append(0.602)
    return Geometry(names, np.array(xyz), np.array(h), None), Settings()


def _place_random(rng, n, box, dmin, max_tries=10000):
    """Sequential random insertion in a periodic cube with a cell list."""
    ncell = max(1, int(box // dmin))
    cw = box / ncell
    grid = {}
    pts = np.zeros((n, 3))
    for i in range(n):
        for _ in range(max_tries):
            p = rng.uniform(0.0, box, 3)
            c = tuple((p // cw).astype(int) % ncell)
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        key = ((c[0] + dx) % ncell, (c[1] + dy) % ncell, (c[2] + dz) % ncell)
                        for j in grid.get(key, ()):
                            d = p - pts[j]
                            d -= box * np.round(d / box)
                            if d @ d < dmin * dmin:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                pts[i] = p
                grid.setdefault(c, []).append(i)
                break
        else:
            raise RuntimeError("random placement failed")
    return pts


def random_periodic(n_atoms: int = 1024, box: float = 24.0, seed: int = 2024,
                    dmin: float = 1.0, rc_scs: float = 40.0, rc_sc: float = 25.0,
                    rc_cfdm: float = 100.0,
                    fractions=(("C", 0.4), ("H", 0.4), ("N", 0.1), ("O", 0.1))):
    rng = np.random.default_rng(seed)
    counts = species_counts(n_atoms, fractions)
    names: List[str] = []
    for s, c in counts:
        names += [s] * c
    names = list(np.array(names)[rng.permutation(n_atoms)])
    xyz = _place_random(rng, n_atoms, box, dmin)
    h = rng.uniform(0.60, 0.95, n_atoms)
    lat = np.diag([box, box, box])
    st = Settings(flag_xc=1, mbd_supercell_cutoff=rc_sc, mbd_cfdm_dip_cutoff=rc_cfdm,
                  mbd_scs_dip_cutoff=rc_scs, mbd_scs_vacuum_axis=(False, False, False),
                  given=(True,) * 5 + (False, False))
    return Geometry([str(x) for x in names], xyz, h, lat), st


def chain_molecule(rng, n_heavy: int):
    """Random-walk backbone of C atoms with steps U(1.1, 1.5) A, isotropic
    directions, rejection below 1.0 A, and an H cap at each end."""
    pts = [np.zeros(3)]
    while len(pts) < n_heavy:
        for _ in range(1000):
            v = rng.standard_normal(3)
            v /= np.linalg.norm(v)
            p = pts[-1] + rng.uniform(1.1, 1.5) * v
            if all(np.linalg.norm(p - q) >= 1.0 for q in pts[:-1]):
                pts.append(p)
                break
        else:
            raise RuntimeError("chain growth failed")
    caps = []
    for end, prev in ((pts[0], pts[1] if len(pts) > 1 else None),
                      (pts[-1], pts[-2] if len(pts) > 1 else None)):
        for _ in range(1000):
            v = rng.standard_normal(3) if prev is None else (end - prev)
            v = v / np.linalg.norm(v)
            v = v + 0.3 * rng.standard_normal(3)
            v /= np.linalg.norm(v)
            p = end + 1.09 * v
            if all(np.linalg.norm(p - q) >= 1.0 for q in pts + caps if q is not end):
                caps.append(p)
                break
        else:
            raise RuntimeError("cap placement failed")
    names = ["C"] * len(pts) + ["H"] * len(caps)
    xyz = np.array(pts + caps)
    h = rng.uniform(0.6, 0.95, len(names))
    return Geometry(names, xyz, h, None)


def chain_molecules(n_mol: int = 512, lo: int = 8, hi: int = 256, seed: int = 4):
    """Batched form: n_i = round(exp(U(ln lo, ln hi))) atoms per molecule
    (chain of n_i - 2 C plus two H caps)."""
    rng = np.random.default_rng(seed)
    mols = []
    for _ in range(n_mol):
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        n = max(n, 3)
        mols.append(chain_molecule(rng, n - 2))
    return mols, Settings()


def hex_sheet(nx: int = 64, ny: int = 32, seed: int = 5, cc: float = 1.42, jitter: float = 0.05,
              vacuum: float = 20.0, rc_scs: float = 120.0, rc_sc: float = 25.0,
              rc_cfdm: float = 100.0, hlo: float = 0.78, hhi: float = 0.84):
    """2-atom hexagonal primitive cell (a = cc sqrt 3) replicated nx x ny, with
    vacuum axis z."""
    rng = np.random.default_rng(seed)
    a = cc * math.sqrt(3.0)
    a1 = np.array([a, 0.0, 0.0])
    a2 = np.array([0.5 * a, 0.5 * math.sqrt(3.0) * a, 0.0])
    basis = [np.zeros(3), (a1 + a2) / 3.0]
    xyz = []
    for i in range(nx):
        for j in range(ny):
            for b in basis:
                xyz.append(i * a1 + j * a2 + b)
    xyz = np.array(xyz) + rng.normal(0.0, jitter, (len(xyz), 3))
    n = len(xyz)
    lat = np.array([nx * a1, ny * a2, [0.0, 0.0, vacuum]])
    h = rng.uniform(hlo, hhi, n)
    st = Settings(flag_xc=1, mbd_supercell_cutoff=rc_sc, mbd_cfdm_dip_cutoff=rc_cfdm,
                  mbd_scs_dip_cutoff=rc_scs, mbd_scs_vacuum_axis=(False, False, True),
                  given=(True,) * 5 + (False, False))
    return Geometry(["C"] * n, xyz, h, lat), st


CONFIGS = {
    "c1_ring": lambda: ring(),
    "c2_random_1024": lambda: random_periodic(1024, 24.0, seed=2024),
    "c3_random_4096": lambda: random_periodic(4096, 38.0, seed=4096),
    "c5_hex_sheet_4096": lambda: hex_sheet(64, 32),
}


def write_config(name: str, outdir: str) -> List[str]:
    """Write a configuration as pmbd.x inputs under ``outdir``: geometry.in and
    settings.in, or for the batched form mol_XXX.in (one per molecule) and a
    shared settings.in.  Returns the written paths."""
    import os

    from .io import write_geometry, write_settings
    os.makedirs(outdir, exist_ok=True)
    paths = []
    if name == "c4_chains_512":
        mols, st = chain_molecules()
        for k, g in enumerate(mols):
            p = os.path.join(outdir, f"mol_{k:03d}.in")
            write_geometry(p, g.names, g.coords_ang, g.hirshfeld, g.lattice_ang)
            paths.append(p)
    else:
        g, st = CONFIGS[name]()
        p = os.path.join(outdir, "geometry.in")
        write_geometry(p, g.names, g.coords_ang, g.hirshfeld, g.lattice_ang)
        paths.append(p)
    p = os.path.join(outdir, "settings.in")
    write_settings(p, st)
    paths.append(p)
    return paths


if __name__ == "__main__":
    import sys
    if len(sys.argv) != 3 or sys.argv[1] not in list(CONFIGS) + ["c4_chains_512"]:
        print(f"usage: python -m mbd_amd.generators "
              f"{{{'|'.join(list(CONFIGS) + ['c4_chains_512'])}}} OUTDIR", file=sys.stderr)
        sys.exit(2)
    for path in write_config(sys.argv[1], sys.argv[2]):
        print(path)
