"""Writes tests/golden/pvector.json: for each of the reference's 58 LAMMPS golden results, the 13 energy terms
(energy_data_fields) and the 14 pvector slots LAMMPS reported for the same evaluation (energy.slots), both taken as data from
the reference's fixtures (tests/fixtures/reference/*.json). They pin rxa_pvector's term-to-slot map (the reference's
src/core/energy_terms.cpp:35-72), including eng_coul's split (Q-31: ep = e_ele, eqeq = e_pol)."""
import glob
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
TERMS = ("e_bond", "e_ov", "e_un", "e_lp", "e_ang", "e_pen", "e_coa", "e_hb", "e_tor", "e_con", "e_vdW", "e_ele", "e_pol")
SLOTS = ("eb", "ea", "elp", "emol", "ev", "epen", "ecoa", "ehb", "et", "eco", "ew", "ep", "efi", "eqeq")

out = []
for path in sorted(glob.glob(os.path.join(REF, "tests", "fixtures", "reference", "*.json"))):
    d = json.load(open(path))
    if not d.get("valid", True) or "energy_data_fields" not in d:
        continue
    out.append({"id": d["case_id"], "terms": [d["energy_data_fields"][k] for k in TERMS],
                "slots": [d["energy"]["slots"][k] for k in SLOTS]})
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pvector.json")
json.dump({"terms": TERMS, "slots": SLOTS, "cases": out}, open(dst, "w"), indent=None)
print("wrote %d cases to %s" % (len(out), dst))
