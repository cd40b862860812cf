"""Input-file semantics (MBD_rsSCS.f90:66-135) and the seeded generators of
the benchmark configurations: writer -> reader round trips are exact, the
species-count rule, minimum separations, parameter ranges and lattices."""
import math
import os

import numpy as np
import pytest

from mbd_amd import generators, io
from tests.conftest import GOLDEN


def test_reads_reference_inputs():
    g = io.read_geometry(os.path.join(GOLDEN, "C6H6.in"))
    assert g.n_atoms == 12 and g.periodic  # isolated molecule in a 2000 A box
    assert np.array_equal(g.lattice_ang, 2000.0 * np.eye(3))
    assert sorted(set(g.names)) == ["C", "H"]
    g = io.read_geometry(os.path.join(GOLDEN, "graphene_111.in"))
    assert g.periodic and g.lattice_ang.shape == (3, 3)
    s = io.read_settings(os.path.join(GOLDEN, "settings.in"))
    assert s.flag_xc in (1, 2, 3)


def test_settings_substring_semantics():
    s = io.parse_settings_text("xc 2\nmbd_scs_dip_cutoff 80.d0\nmbd_scs_vacuum_axis .false. .F. .TRUE.\n")
    assert s.flag_xc == 2 and s.mbd_scs_dip_cutoff == 80.0
    assert s.mbd_scs_vacuum_axis == (False, False, True)
    assert s.mbd_supercell_cutoff == 25.0 and s.mbd_cfdm_dip_cutoff == 100.0
    # a malformed record ends the loop: later lines are not read
    s = io.parse_settings_text("xc\nmbd_scs_dip_cutoff 80\n")
    assert s.mbd_scs_dip_cutoff == 120.0
    # list-directed separators
    g = io.parse_geometry_text("1\n\nC, 0.5,\t1.0 2.0 , 0.8\n")
    assert np.array_equal(g.coords_ang[0], [0.5, 1.0, 2.0]) and g.hirshfeld[0] == 0.8


@pytest.mark.parametrize("name", ["c1_ring", "c2_random_1024", "c5_hex_sheet_4096"])
def test_write_read_round_trip_is_exact(name, tmp_path):
    g, st = generators.CONFIGS[name]()
    generators.write_config(name, str(tmp_path))
    g2 = io.read_geometry(str(tmp_path / "geometry.in"))
    st2 = io.read_settings(str(tmp_path / "settings.in"))
    assert g2.names == g.names
    assert np.array_equal(g2.coords_ang, g.coords_ang)
    assert np.array_equal(g2.hirshfeld, g.hirshfeld)
    if g.periodic:
        assert np.array_equal(g2.lattice_ang, g.lattice_ang)
    for f in ("flag_xc", "mbd_supercell_cutoff", "mbd_cfdm_dip_cutoff", "mbd_scs_dip_cutoff",
              "mbd_scs_vacuum_axis"):
        assert getattr(st2, f) == getattr(st, f)


def test_species_counts_rule():
    fr = [("C", 0.4), ("H", 0.4), ("N", 0.1), ("O", 0.1)]
    assert generators.species_counts(1024, fr) == [("C", 410), ("H", 410), ("N", 102), ("O", 102)]
    assert generators.species_counts(4096, fr) == [("C", 1638), ("H", 1638), ("N", 410), ("O", 410)]


def _min_image_dmin(xyz, box):
    d = xyz[:, None, :] - xyz[None, :, :]
    d -= box * np.round(d / box)
    r = np.sqrt((d ** 2).sum(-1))
    r[np.diag_indices(len(xyz))] = np.inf
    return r.min()


def test_random_periodic_c2():
    g, st = generators.CONFIGS["c2_random_1024"]()
    assert g.n_atoms == 1024 and np.allclose(g.lattice_ang, 24.0 * np.eye(3))
    assert _min_image_dmin(g.coords_ang, 24.0) >= 1.0
    assert g.hirshfeld.min() >= 0.60 and g.hirshfeld.max() <= 0.95
    assert (g.coords_ang >= 0).all() and (g.coords_ang < 24.0).all()
    assert (st.mbd_scs_dip_cutoff, st.mbd_supercell_cutoff, st.mbd_cfdm_dip_cutoff) == (40.0, 25.0, 100.0)
    g2, _ = generators.CONFIGS["c2_random_1024"]()
    assert np.array_equal(g.coords_ang, g2.coords_ang)  # seeded


def test_chain_molecules_c4():
    mols, st = generators.chain_molecules()
    assert len(mols) == 512
    n = np.array([m.n_atoms for m in mols])
    assert n.min() >= 8 and n.max() <= 256
    # log-uniform: the median of n is near sqrt(8 * 256) = 45
    assert 30 <= np.median(n) <= 65
    for m in mols[:64]:
        x = m.coords_ang
        r = np.sqrt(((x[:, None] - x[None]) ** 2).sum(-1)) + 10 * np.eye(len(x))
        assert r.min() >= 1.0 - 1e-12
        c = x[[i for i, s in enumerate(m.names) if s == "C"]]
        steps = np.linalg.norm(np.diff(c, axis=0), axis=1)
        assert steps.min() >= 1.1 - 1e-12 and steps.max() <= 1.5 + 1e-12
        assert m.names.count("H") == 2 and not m.periodic
        assert m.hirshfeld.min() >= 0.6 and m.hirshfeld.max() <= 0.95


def test_hex_sheet_c5():
    g, st = generators.CONFIGS["c5_hex_sheet_4096"]()
    a = 1.42 * math.sqrt(3.0)
    assert g.n_atoms == 4096
    assert np.allclose(np.linalg.norm(g.lattice_ang, axis=1), [64 * a, 32 * a, 20.0])
    assert st.mbd_scs_vacuum_axis == (False, False, True) and st.mbd_scs_dip_cutoff == 120.0
    assert 0.78 <= g.hirshfeld.min() and g.hirshfeld.max() <= 0.84
    assert abs(np.std(g.coords_ang[:, 2]) - 0.05) < 0.005  # jitter sigma on a flat sheet
