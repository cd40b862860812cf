"""Extract the reference's shipped example inputs and printed results into
fixtures (data only): copies ``examples/*.in`` and ``settings.in`` and parses the
three golden outputs (``C6H6.out``, ``graphene_111.out``, ``PBBA.out``) into
``golden.json``: E_MBD (Ha), system C6, per-atom (C6, alpha), the alpha(i w)
table and the supercell dimensions.

Run in the build container only (it reads the reference tree):
    python tests/golden/make_golden.py /path/to/reference
"""
import json
import os
import re
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


RESULT_LINE = re.compile(r"(\| MBD@rsSCS energy|C6 coefficient|\| ATOM |super cell of dimension|"
                         r"\| containing|^\s*\|\s+\d+\.\d{6}\s+\S+E)")


def parse_out(path):
    res = {"atoms": {}, "table": [], "lines": []}
    for line in open(path):
        if RESULT_LINE.search(line):
            res["lines"].append(line.rstrip("\n"))
        m = re.search(r"MBD@rsSCS energy\s*:\s*(\S+)\s*Ha\s+(\S+)\s*eV", line)
        if m:
            res["e_mbd_ha"] = float(m.group(1))
            res["e_mbd_ev"] = float(m.group(2))
        m = re.search(r"C6 coefficient\s*:\s*(\S+)", line)
        if m:
            res["c6_mol"] = float(m.group(1))
        m = re.search(r"\| ATOM\s+(\d+)\s+(\S+)\s+(\S+)\s+(\S+)", line)
        if m:
            res["atoms"][int(m.group(1))] = [m.group(2), float(m.group(3)), float(m.group(4))]
        m = re.search(r"^\s*\|\s+(\d+\.\d{6})\s+(\S+E\S+)\s+(\S+E\S+)\s+(\S+E\S+)\s+(\S+E\S+)", line)
        if m:
            res["table"].append([float(m.group(i)) for i in range(1, 6)])
        m = re.search(r"super cell of dimension\s+(\d+) X\s+(\d+) X\s+(\d+)", line)
        if m:
            res["supercell"] = [int(m.group(i)) for i in range(1, 4)]
    return res


def main(ref):
    ex = os.path.join(ref, "examples")
    for f in ["C6H6.in", "graphene_111.in", "PBBA.in", "H2O.in", "TiOCHF.in"]:
        shutil.copy(os.path.join(ex, f), os.path.join(HERE, f))
    shutil.copy(os.path.join(ref, "settings.in"), os.path.join(HERE, "settings.in"))
    golden = {}
    for name, settings in [("C6H6", None), ("graphene_111", "settings.in"), ("PBBA", None)]:
        r = parse_out(os.path.join(ex, name + ".out"))
        r["geometry"] = name + ".in"
        r["settings"] = settings
        # the settings the C6H6 run echoed (C6H6.out:20-22): supercell 25,
        # cfdm 100, scs 100 Angstrom; it is effectively isolated either way
        golden[name] = r
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(golden, f, indent=1, sort_keys=True)
    # the free-atom table as the reference's decimal literals
    # (SELECT CASE in get_vdw_param): {symbol: [alpha, C6, R0]}
    table, cur = {}, None
    for line in open(os.path.join(ref, "UTILS.f90")):
        if re.match(r"\s*case default", line):
            cur = None
            continue
        m = re.match(r"\s*case\('(\w+)'\)", line)
        if m:
            cur = m.group(1)
            table[cur] = [None, None, None]
            continue
        m = re.match(r"\s*(alpha|C6|R0)=([0-9.]+)\s*$", line)
        if m and cur is not None:
            table[cur][["alpha", "C6", "R0"].index(m.group(1))] = float(m.group(2))
    with open(os.path.join(HERE, "vdw_table.json"), "w") as f:
        json.dump(table, f, indent=0)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
