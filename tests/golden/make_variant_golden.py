"""Writes tests/golden/variants.json: golden vectors of the reference's own reader and FP64 engine (oracle/_ref, built by
`make -f oracle/ref.mk`) for what the C/H/O fixtures of tests/golden/fixtures.json leave open.

* dumps: for the stand-in field, the carbon field (tests/golden/ffield.C2013) and every variant of
  tests/golden/ffield_variants.py, the SHA-256 of the reference reader's canonical table dump (plain and portable), or its
  rejection;
* census: the reference engine's per-atom decision census (reaxmetal_energy_tool --census) on the 41 C/H/O fixtures with the
  stand-in field;
* variants: the 13 energy terms and the gradient of the reference engine with each accepted variant on a subset of the C/H/O
  fixtures (inner wall, shielding + inner wall, low-gradient dispersion, triple-bond stabilisation for every pair, no C2);
* carbon: small periodic random-carbon cells made by the engine's own generator (reaxamd.systems.random_carbon, the method
  of the reference's examples/graphitization/make_box.py), charges zero (checkqeq no), with the carbon field: energy terms,
  gradient and census.

Run in the build container only (it needs /root/reference and oracle/_ref):  python tests/golden/make_variant_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(REPO, "metal-reax_amd"))
import ffield_variants as fv   # noqa: E402
from make_golden import TERMS, case_text   # noqa: E402

DUMP = os.path.join(REPO, "oracle", "_ref", "reaxmetal_ffield_dump")
TOOL = os.path.join(REPO, "oracle", "_ref", "reaxmetal_energy_tool")
CARBON = os.path.join(HERE, "ffield.C2013")
CENSUS = ("bonds", "angle_sets", "torsions", "hbonds", "nonbonded", "sbo_region", "lp_trunc")
# C/H/O fixtures the variants are evaluated on: condensed water, carbon-oxygen molecules (triple-bond stabilisation), conjugated
# and strained carbon (C2 correction, torsions), and a periodic carbon crystal
VARIANT_CASES = ("cho_water_box_8", "cho_co2_pert", "cho_methanol_pert", "cho_benzene_pert", "cho_butadiene_pert",
                 "cho_cyclopropane_pert", "cho_water_dimer_pert", "cho_diamond_triclinic", "cho_graphene")
# random-carbon cells: atoms, density (g/cm^3), seed
CARBON_CELLS = ((64, 1.0, 1), (128, 1.0, 2), (96, 2.0, 3), (216, 1.0, 4))


def tmpfile(name, text):
    path = os.path.join("/tmp", "_variant_" + name)
    with open(path, "w") as fh:
        fh.write(text)
    return path


def dump_record(text, lgvdw, tag):
    path = tmpfile(tag + ".ff", text)
    rec = {"lgvdw": lgvdw}
    for port in (False, True):
        r = subprocess.run([DUMP] + (["--lgvdw"] if lgvdw else []) + (["--portable"] if port else []) + [path], capture_output=True, text=True)
        if r.returncode == 2:
            rec["rejected"] = True
            return rec
        r.check_returncode()
        rec["sha256_portable" if port else "sha256"] = hashlib.sha256(r.stdout.encode()).hexdigest()
    return rec


def run_tool(ff_path, lgvdw, elements, origin, vectors, periodic, pos, types, charges):
    path = tmpfile("case.txt", case_text(origin, vectors, periodic, pos, types, charges))
    r = subprocess.run([TOOL, "--ffield", ff_path, "--elements", ",".join(elements), "--grad", "--census"] + (["--lgvdw"] if lgvdw else []) +
                       [path], capture_output=True, text=True, check=True)
    e, grad, cen = {}, [], {}
    for line in r.stdout.splitlines():
        p = line.split()
        if p[0] in TERMS:
            e[p[0]] = float(p[1])
        elif p[0] == "grad":
            grad.append([float(v) for v in p[2:5]])
        elif p[0] == "census":
            cen[p[1]] = [int(v) for v in p[2:]]
    return e, grad, cen


def main():
    if not os.path.exists(TOOL):
        sys.exit("build the reference engine first: make -f oracle/ref.mk")
    fixtures = json.load(open(os.path.join(HERE, "fixtures.json")))
    out = {"dumps": {}, "census": {}, "variants": {}, "carbon": []}
    out["dumps"]["standin"] = dump_record(fv.standin_text(), False, "standin")
    out["dumps"]["C2013"] = dump_record(open(CARBON).read(), False, "C2013")
    for name, lg in list(fv.ACCEPTED.items()) + list(fv.REJECTED.items()):
        out["dumps"][name] = dump_record(fv.variant(name), lg, name)
    standin = os.path.join(HERE, "ffield.cho_standin")
    for c in fixtures["cases"]:
        if "ref_energy" not in c:
            continue
        _, _, cen = run_tool(standin, False, c["elements"], c["origin"], c["vectors"], c["periodic"], c["positions"], c["types"], c["charges"])
        out["census"][c["id"]] = [cen[k] for k in CENSUS]
    for name, lg in fv.ACCEPTED.items():
        path = tmpfile(name + ".ff", fv.variant(name))
        out["variants"][name] = {}
        for c in fixtures["cases"]:
            if c["id"] not in VARIANT_CASES:
                continue
            e, g, _ = run_tool(path, lg, c["elements"], c["origin"], c["vectors"], c["periodic"], c["positions"], c["types"], c["charges"])
            out["variants"][name][c["id"]] = {"energy": e, "grad": g}
    from reaxamd import systems
    for n, dens, seed in CARBON_CELLS:
        w = systems.random_carbon(n, density=dens, seed=seed)
        L = float(w.hi[0])
        vec = [[L, 0.0, 0.0], [0.0, L, 0.0], [0.0, 0.0, L]]
        e, g, cen = run_tool(CARBON, False, ["C"], [0.0, 0.0, 0.0], vec, [True, True, True], w.x.tolist(), [1] * n, [0.0] * n)
        out["carbon"].append({"id": "carbon_%d_%g_%d" % (n, dens, seed), "n": n, "density": dens, "seed": seed, "L": L,
                              "positions": w.x.tolist(), "energy": e, "grad": g, "census": [cen[k] for k in CENSUS]})
    with open(os.path.join(HERE, "variants.json"), "w") as fh:
        json.dump(out, fh, separators=(",", ":"))
    print("wrote variants.json: %d dumps, %d census cases, %d variants, %d carbon cells" % (
        len(out["dumps"]), len(out["census"]), len(out["variants"]), len(out["carbon"])))


if __name__ == "__main__":
    main()
