"""Input files of pmbd.x: geometry.in and settings.in.

Semantics follow the reference reader, ``MBD_rsSCS.f90:66-135``:

* geometry: first line ``n_atoms``; one skipped line; ``n_atoms`` records
  ``name x y z hirshfeld`` (list-directed, so blanks, tabs and commas all
  separate); then, if the first token of the next line contains
  ``lattice_vector``, three records ``lattice_vector ax ay az`` whose vectors are
  the lattice columns and the system is periodic (``MBD_rsSCS.f90:80-88``).
* settings: each line is matched by substring, in the reference's order (``xc``,
  ``mbd_supercell_cutoff``, ``mbd_cfdm_dip_cutoff``, ``mbd_scs_dip_cutoff``,
  ``mbd_scs_vacuum_axis``, ``nprow``, ``npcol``); the line is then re-read as
  ``keyword value(s)``.  ``nprow``/``npcol`` are accepted and ignored by the
  single-GPU engine.
* defaults: ``UTILS.f90:78-93`` (xc 1, supercell 25 A, cfdm 100 A, scs 120 A,
  no vacuum axis).

Coordinates and lattice are kept in Angstrom here; ``to_bohr`` performs the
reference's division by ``bohr = 0.52917721`` (``MBD_rsSCS.f90:211-212``).
"""
from __future__ import annotations

import dataclasses
import re
from typing import List, Optional, Sequence

import numpy as np

BOHR = 0.52917721  # UTILS.f90:85

_SEP = re.compile(r"[\s,]+")


def _tokens(line: str) -> List[str]:
    return [t for t in _SEP.split(line.strip()) if t]


def _fortran_real(tok: str) -> float:
    """A Fortran list-directed REAL token (``100.d0``, ``1.5E+01``, ``25``)."""
    t = tok.strip().rstrip(",").replace("d", "e").replace("D", "e")
    if t.endswith("e"):
        t = t[:-1]
    return float(t)


def _fortran_logical(tok: str) -> bool:
    """List-directed LOGICAL: optional '.', then T or F; the rest is ignored."""
    t = tok.strip().lstrip(".")
    if not t:
        raise ValueError("empty logical")
    c = t[0].upper()
    if c == "T":
        return True
    if c == "F":
        return False
    raise ValueError(f"bad logical {tok!r}")


@dataclasses.dataclass
class Geometry:
    names: List[str]
    coords_ang: np.ndarray  # (N, 3)
    hirshfeld: np.ndarray  # (N,)
    lattice_ang: Optional[np.ndarray] = None  # (3, 3), row k = lattice vector k

    @property
    def n_atoms(self) -> int:
        return len(self.names)

    @property
    def periodic(self) -> bool:
        return self.lattice_ang is not None

    def to_bohr(self):
        """(coords, lattice) divided by bohr as the reference does."""
        xyz = self.coords_ang / BOHR
        lat = (self.lattice_ang / BOHR) if self.lattice_ang is not None else np.zeros((3, 3))
        return xyz, lat


@dataclasses.dataclass
class Settings:
    flag_xc: int = 1
    mbd_supercell_cutoff: float = 25.0
    mbd_cfdm_dip_cutoff: float = 100.0
    mbd_scs_dip_cutoff: float = 120.0
    mbd_scs_vacuum_axis: Sequence[bool] = (False, False, False)
    nprow: int = 1
    npcol: int = 1
    given: Sequence[bool] = (False,) * 7  # input_settings(1:7)


def parse_geometry_text(text: str) -> Geometry:
    lines = text.splitlines()
    n = int(_tokens(lines[0])[0])
    names, xyz, hir = [], [], []
    pos = 2
    for _ in range(n):
        t = _tokens(lines[pos])
        pos += 1
        names.append(t[0][:2])
        xyz.append([_fortran_real(v) for v in t[1:4]])
        hir.append(_fortran_real(t[4]))
    lattice = None
    if pos < len(lines):
        t = _tokens(lines[pos])
        if t and "lattice_vector" in t[0]:
            lat = np.zeros((3, 3))
            for k in range(3):
                tk = _tokens(lines[pos + k])
                lat[k] = [_fortran_real(v) for v in tk[1:4]]
            lattice = lat
    return Geometry(names, np.array(xyz, dtype=np.float64), np.array(hir, dtype=np.float64), lattice)


def read_geometry(path: str) -> Geometry:
    with open(path) as f:
        return parse_geometry_text(f.read())


_KEYS = ["xc", "mbd_supercell_cutoff", "mbd_cfdm_dip_cutoff", "mbd_scs_dip_cutoff",
         "mbd_scs_vacuum_axis", "nprow", "npcol"]


def parse_settings_text(text: str) -> Settings:
    s = Settings()
    given = [False] * 7
    for line in text.splitlines():
        for idx, key in enumerate(_KEYS):
            if key in line:
                t = _tokens(line)
                try:
                    if idx == 0:
                        s.flag_xc = int(t[1])
                    elif idx == 1:
                        s.mbd_supercell_cutoff = _fortran_real(t[1])
                    elif idx == 2:
                        s.mbd_cfdm_dip_cutoff = _fortran_real(t[1])
                    elif idx == 3:
                        s.mbd_scs_dip_cutoff = _fortran_real(t[1])
                    elif idx == 4:
                        s.mbd_scs_vacuum_axis = tuple(_fortran_logical(v) for v in t[1:4])
                    elif idx == 5:
                        s.nprow = int(t[1])
                    elif idx == 6:
                        s.npcol = int(t[1])
                except (IndexError, ValueError):
                    # The reference's READ ... IOSTAT ends its settings loop on
                    # a malformed record (MBD_rsSCS.f90:102).
                    s.given = tuple(given)
                    return s
                given[idx] = True
                break
    s.given = tuple(given)
    return s


def read_settings(path: Optional[str]) -> Settings:
    if path is None:
        return Settings()
    with open(path) as f:
        return parse_settings_text(f.read())


def write_geometry(path: str, names, coords_ang, hirshfeld, lattice_ang=None) -> None:
    """Write a geometry.in with 17 significant digits (round-trips float64)."""
    with open(path, "w") as f:
        f.write(f"{len(names)}\n\n")
        for nm, (x, y, z), h in zip(names, coords_ang, hirshfeld):
            f.write(f"{nm:<2s} {x:.17g} {y:.17g} {z:.17g} {h:.17g}\n")
        if lattice_ang is not None:
            for v in lattice_ang:
                f.write(f"lattice_vector {v[0]:.17g} {v[1]:.17g} {v[2]:.17g}\n")


def write_settings(path: str, s: Settings) -> None:
    vac = " ".join(".true." if v else ".false." for v in s.mbd_scs_vacuum_axis)
    with open(path, "w") as f:
        f.write(f"xc                    {s.flag_xc}\n")
        f.write(f"mbd_cfdm_dip_cutoff   {s.mbd_cfdm_dip_cutoff!r}\n")
        f.write(f"mbd_supercell_cutoff  {s.mbd_supercell_cutoff!r}\n")
        f.write(f"mbd_scs_dip_cutoff    {s.mbd_scs_dip_cutoff!r}\n")
        f.write(f"mbd_scs_vacuum_axis   {vac}\n")
