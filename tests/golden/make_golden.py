"""Writes tests/golden/fixtures.json: golden vectors for the oracle and the HIP engine.

Run in the build container only (it reads /root/reference and runs the reference's own FP64 engine built by
`make -f oracle/ref.mk`):  python3 tests/golden/make_golden.py

For each of the reference's 58 fixtures (tests/fixtures/cases + tests/fixtures/reference in the reference repository) it keeps
the inputs (LAMMPS positions, types, cell, periodicity, converged charges) and the LAMMPS pair tallies vdw.oo/og/self, which
depend only on geometry and the 10 A cutoff. For the 47 C/H/O fixtures it also runs the reference engine
(reaxmetal_energy_tool, backend cpu64) with the stand-in C/H/O force field tests/golden/ffield.cho_standin and the golden
charges, and keeps the 13 energy terms and the gradient folded onto the owned atoms.
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("REAX_REFERENCE", "/root/reference")
TOOL = os.path.join(REPO, "oracle", "_ref", "reaxmetal_energy_tool")
FF = os.path.join(HERE, "ffield.cho_standin")
TERMS = ("e_bond", "e_ov", "e_un", "e_lp", "e_ang", "e_pen", "e_coa", "e_hb", "e_tor", "e_con", "e_vdW", "e_ele", "e_pol")


def cell_of(case):
    cell = case["cell"]
    if "vectors" in cell:
        return list(cell["origin"]), [list(v) for v in cell["vectors"]]
    lo, hi = cell["lo"], cell["hi"]
    return list(lo), [[hi[0] - lo[0], 0, 0], [0, hi[1] - lo[1], 0], [0, 0, hi[2] - lo[2]]]


def case_text(origin, vectors, periodic, pos, types, charges):
    s = "origin %r %r %r\n" % tuple(origin)
    for k, v in zip("abc", vectors):
        s += "%s %r %r %r\n" % (k, *v)
    s += "periodic %d %d %d\n" % tuple(int(p) for p in periodic)
    s += "atoms %d\n" % len(pos)
    for i, (t, x) in enumerate(zip(types, pos)):
        s += "%d %d %r %r %r\n" % (i + 1, t, *x)
    s += "charges " + " ".join(repr(q) for q in charges) + "\n"
    return s


def main():
    if not os.path.exists(TOOL):
        sys.exit("build the reference engine first: make -f oracle/ref.mk")
    out = {"ffield": "ffield.cho_standin", "terms": TERMS, "cases": []}
    names = sorted(f[:-5] for f in os.listdir(os.path.join(REF, "tests/fixtures/cases")) if f.endswith(".json"))
    for name in names:
        case = json.load(open(os.path.join(REF, "tests/fixtures/cases", name + ".json")))
        gold = json.load(open(os.path.join(REF, "tests/fixtures/reference", name + ".json")))
        origin, vectors = cell_of(case)
        tal = gold.get("tallies", {})
        rec = {
            "id": name, "elements": case["elements"], "origin": origin, "vectors": vectors, "periodic": case["periodic"],
            "positions": gold["positions"], "types": gold["types"], "charges": gold["charges"],
            "lammps_tallies": {k: tal.get(k, 0) for k in ("vdw.oo", "vdw.og", "vdw.self")},
            # the reference's conditioning class: well-conditioned under its extended definition (no near-linear angle X1, no
            # near-threshold product X2, no pre-registered W-flag); only these carry the tight FP64 bar
            "well_conditioned": bool(gold.get("conditioning", {}).get("well_conditioned_extended", False)),
            "conditioning_flags": list(gold.get("conditioning", {}).get("extended_flags", [])) +
                                  list(gold.get("conditioning", {}).get("fails_pre_registered", [])),
        }
        if name.startswith("cho_"):
            path = "/tmp/_golden_case.txt"
            with open(path, "w") as fh:
                fh.write(case_text(origin, vectors, case["periodic"], gold["positions"], gold["types"], gold["charges"]))
            res = subprocess.run([TOOL, "--ffield", FF, "--elements", ",".join(case["elements"]), "--grad", path],
                                 capture_output=True, text=True, check=True)
            e, grad = {}, []
            for line in res.stdout.splitlines():
                p = line.split()
                if p[0] in TERMS:
                    e[p[0]] = float(p[1])
                elif p[0] == "grad":
                    grad.append([float(v) for v in p[2:5]])
            rec["ref_energy"] = e
            rec["ref_grad"] = grad
        out["cases"].append(rec)
    with open(os.path.join(HERE, "fixtures.json"), "w") as fh:
        json.dump(out, fh, separators=(",", ":"))
    print("wrote", len(out["cases"]), "cases")


if __name__ == "__main__":
    main()
