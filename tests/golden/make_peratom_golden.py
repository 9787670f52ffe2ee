"""Writes tests/golden/peratom.json: per-atom energies and virials of the reference's own FP64 engine for the C/H/O fixtures.

Run in the build container only (it runs oracle/_ref/ref_peratom, built from the reference's sources by
`make -f oracle/ref.mk`):  python3 tests/golden/make_peratom_golden.py

For each C/H/O case of tests/golden/fixtures.json (same inputs, stand-in force field, golden charges) it keeps, per owned atom in
case order, the reference's per-atom energy (its ev_tally attribution: bond, triple-bond and C2 energies half to each end,
torsion half to each atom of the central bond, H-bond a third to each atom of the triple, valence to the centre, lone pair and
over/under to the atom, non-bonded half to each atom of the pair) and per-atom virial (xx yy zz xy xz yz), images folded onto
their owners, and the bond table the LAMMPS analysis commands read (the reference's BondTable: per owned atom its bonds as
[partner atom index in case order, corrected BO] in the reference's order, total bond order and lone pairs). It checks that the
per-atom energies add up to the case's 13 reference terms.
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
TOOL = os.path.join(REPO, "oracle", "_ref", "ref_peratom")
FF = os.path.join(HERE, "ffield.cho_standin")
sys.path.insert(0, HERE)
from make_golden import case_text  # noqa: E402


def main():
    if not os.path.exists(TOOL):
        sys.exit("build the reference engine first: make -f oracle/ref.mk")
    fx = json.load(open(os.path.join(HERE, "fixtures.json")))
    out = {"ffield": "ffield.cho_standin", "cases": {}}
    for case in fx["cases"]:
        if not case["id"].startswith("cho_"):
            continue
        path = "/tmp/_peratom_case.txt"
        with open(path, "w") as fh:
            fh.write(case_text(case["origin"], case["vectors"], case["periodic"], case["positions"], case["types"], case["charges"]))
        res = subprocess.run([TOOL, FF, ",".join(case["elements"]), path], capture_output=True, text=True, check=True)
        n = len(case["types"])
        e = [None] * n
        v = [None] * n
        tbo = [None] * n
        nlp = [None] * n
        bonds = [[] for _ in range(n)]
        for line in res.stdout.splitlines():
            p = line.split()
            if p[0] == "eatom":
                e[int(p[1]) - 1] = float(p[2])
            elif p[0] == "vatom":
                v[int(p[1]) - 1] = [float(t) for t in p[2:8]]
            elif p[0] == "bondatom":
                tbo[int(p[1]) - 1] = float(p[2])
                nlp[int(p[1]) - 1] = float(p[3])
            elif p[0] == "bond":
                bonds[int(p[1]) - 1].append([int(p[2]) - 1, float(p[3])])
        assert all(a is not None for a in e + v + tbo + nlp), case["id"]
        total = sum(case["ref_energy"].values())
        assert abs(sum(e) - total) <= 1e-9 * max(1.0, abs(total)), (case["id"], sum(e), total)
        out["cases"][case["id"]] = {"eatom": e, "vatom": v, "bonds": bonds, "total_bo": tbo, "nlp": nlp}
    with open(os.path.join(HERE, "peratom.json"), "w") as fh:
        json.dump(out, fh, separators=(",", ":"))
    print("wrote", len(out["cases"]), "cases")


if __name__ == "__main__":
    main()
