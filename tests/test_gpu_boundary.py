"""The drop-in boundary beyond the per-step force call:

* rxa_bond_table (what fix reaxff/bonds and fix reaxff/species read; the reference's BondTable, include/reaxmetal/bonded.hpp:32-37,
  published at plugin/adapter/pair_reaxff_metal.cpp:330-368) against the reference engine's own table on the 41 C/H/O fixtures
  (tests/golden/peratom.json): per owned atom the same partners, corrected bond orders, total bond order and lone pairs; FP64 and
  mixed (FP64 bond-order chain) within 1e-9, FP32 within 1e-4 (bond orders are of order 1);
* rxa_set_atoms_ghosted with LAMMPS-style owned + ghost arrays gives bit for bit what rxa_set_atoms gives with the owned atoms,
  and refuses more than one rank;
* a compiled C99 caller of include/reaxamd.h (tests/abi_caller.c, built by __graft_entry__.build) gives bit for bit what the
  Python binding gives.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import case_inputs, lammps_ghosts, new_engine

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
from make_golden import case_text   # noqa: E402


def peratom_golden():
    with open(os.path.join(REPO, "tests", "golden", "peratom.json")) as fh:
        return json.load(fh)["cases"]


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("mixed", 1e-9), ("fp32", 1e-4)])
def test_bond_table_matches_reference(golden, ffield_text, prec, tol):
    ref = peratom_golden()
    nb = 0
    for case in golden["cases"]:
        if case["id"] not in ref:
            continue
        x, types, tags, lo, hi, tilt, per, q, _ = case_inputs(case)
        e = new_engine(ffield_text, case["elements"], prec)
        e.set_box(lo, hi, tilt, per)
        e.set_atoms(types, tags, x)
        e.set_charges(q)
        e.compute(rebuild=True)
        bt = e.bond_table()
        e.close()
        r = ref[case["id"]]
        for i, bonds in enumerate(r["bonds"]):
            a, b = bt["start"][i], bt["start"][i + 1]
            got = sorted(zip(bt["nbr"][a:b].tolist(), bt["bo"][a:b].tolist()))
            want = sorted((p, v) for p, v in bonds)
            assert [p for p, _ in got] == [p for p, _ in want], (prec, case["id"], i)
            assert np.allclose([v for _, v in got], [v for _, v in want], rtol=tol, atol=tol), (prec, case["id"], i)
            nb += b - a
        assert np.allclose(bt["total_bo"], r["total_bo"], rtol=tol, atol=tol), (prec, case["id"])
        assert np.allclose(bt["nlp"], r["nlp"], rtol=tol, atol=tol), (prec, case["id"])
    assert nb > 5000


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_ghosted_input_equals_owned_input(golden, ffield_text, prec):
    from reaxamd import ReaxError
    periodic = [c for c in golden["cases"] if "ref_energy" in c and all(c["periodic"])]
    assert len(periodic) >= 5
    for case in periodic[:8]:
        x, types, tags, lo, hi, tilt, per, q, _ = case_inputs(case)
        T, G, X = lammps_ghosts(x, types, tags, lo, hi, tilt, per, shell=10.0)
        out = []
        for ghosted in (False, True):
            e = new_engine(ffield_text, case["elements"], prec)
            e.set_box(lo, hi, tilt, per)
            if ghosted:
                e.set_atoms_ghosted(len(types), T, G, X)
            else:
                e.set_atoms(types, tags, x)
            e.set_charges(q)
            out.append(e.compute(rebuild=True))
            if ghosted:
                with pytest.raises(ReaxError):
                    e.set_atoms_ghosted(len(types), T, G, X, nranks=4)
            e.close()
        assert out[0]["energy"] == out[1]["energy"], case["id"]
        assert np.array_equal(out[0]["grad"], out[1]["grad"]) and np.array_equal(out[0]["virial"], out[1]["virial"]), case["id"]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_compiled_c_caller_equals_binding(golden, ffield_text, tmp_path, prec):
    from reaxamd import TERMS, PVECTOR
    exe = os.path.join(REPO, "tests", "abi_caller")
    assert os.path.exists(exe), "tests/abi_caller is built by __graft_entry__.build()"
    ff = os.path.join(REPO, "tests", "golden", "ffield.cho_standin")
    for cid in ("cho_water_box_8", "cho_diamond_triclinic", "cho_graphene"):
        case = next(c for c in golden["cases"] if c["id"] == cid)
        path = tmp_path / (cid + ".txt")
        path.write_text(case_text(case["origin"], case["vectors"], case["periodic"], case["positions"], case["types"], case["charges"]))
        print("abi_caller", cid, prec, flush=True)
        res = subprocess.run([exe, ff, ",".join(case["elements"]), str(path), prec], capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stderr
        e_c, pv_c, vir_c, g_c = {}, {}, {}, []
        for line in res.stdout.splitlines():
            p = line.split()
            if p[0] == "energy":
                e_c[TERMS[int(p[1])]] = float(p[2])
            elif p[0] == "pvector":
                pv_c[PVECTOR[int(p[1])]] = float(p[2])
            elif p[0] == "virial":
                vir_c[int(p[1])] = float(p[2])
            elif p[0] == "grad":
                g_c.append([float(v) for v in p[2:5]])
        x, types, tags, lo, hi, tilt, per, q, _ = case_inputs(case)
        e = new_engine(ffield_text, case["elements"], prec)
        e.set_box(lo, hi, tilt, per)
        e.set_atoms(types, tags, x)
        e.set_charges(q)
        r = e.compute(rebuild=True)
        e.close()
        assert e_c == r["energy"], cid
        assert np.array_equal(np.array(g_c), r["grad"]), cid
        assert np.array_equal(np.array([vir_c[k] for k in range(6)]), r["virial"]), cid
        assert abs(sum(pv_c.values()) - sum(r["energy"].values())) <= 1e-9 * max(1.0, abs(sum(r["energy"].values())))
