"""HIP engine vs the reference's own FP64 engine on what the C/H/O fixtures leave open (golden vectors in
tests/golden/variants.json, written by tests/golden/make_variant_golden.py):

* force-field variants of the stand-in (tests/golden/ffield_variants.py): the van der Waals inner wall (method 2), shielding
  plus inner wall (method 3), low-gradient dispersion (lgvdw yes), triple-bond stabilisation of every pair, no C2 correction,
  on nine C/H/O fixtures;
* periodic random-carbon cells with the carbon field of the reference's SI PDF (tests/golden/ffield.C2013), charges zero
  (checkqeq no): energy terms, forces and the decision census;
* the 30,000-atom random-carbon box (30,000 C at 1.0 g/cm^3, 84.265 A, dmin 1.2 A, seed 1; the carbon workload): one force
  call in each precision against the FP64 oracle (the reference engine's restatement, pinned on the cells above), and its
  census.

Bars as tests/test_gpu_parity.py: FP64 each term within 1e-9 relative to max(1, |E|) and gradient components within 1e-7
kcal/mol/A (fixtures the reference flags as ill-conditioned: 1e-7 and 0.1, see there); mixed |dE|/atom <= 1e-4 kcal/mol,
force max <= 4e-3 and RMS <= 9e-4 kcal/mol/A; FP32 1e-3, 5e-2, 5e-3.

Census: FP64 and mixed (whose gates and SBO branch are FP64) take the oracle's decision at every atom. FP32 takes its gates in
single precision: its SBO branch may differ only at atoms whose FP64 SBO lies within SBO_WINDOW of a branch threshold (0, 1, 2),
and every other category must match.
"""
import json
import os
import sys

import numpy as np
import pytest

from helpers import case_inputs, new_engine

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import ffield_variants as fv   # noqa: E402

TERMS = ("e_bond", "e_ov", "e_un", "e_lp", "e_ang", "e_pen", "e_coa", "e_hb", "e_tor", "e_con", "e_vdW", "e_ele", "e_pol")
BARS = {"mixed": (1e-4, 4e-3, 9e-4), "fp32": (1e-3, 5e-2, 5e-3)}
SBO_WINDOW = 1e-5   # FP32 rounding of SBO, a sum of terms of order 1 from a single-precision bond-order chain


def sbo_flips_outside_window(base, got, sbo, window):
    """atoms whose SBO branch differs although their FP64 SBO is farther than `window` from every threshold"""
    sbo = np.asarray(sbo)
    near = np.min(np.abs(sbo[:, None] - np.array([0.0, 1.0, 2.0])[None, :]), axis=1) <= window
    return int(((np.asarray(base) != np.asarray(got)) & ~near).sum()), int(((np.asarray(base) != np.asarray(got)) & near).sum())


@pytest.fixture(scope="module")
def variants():
    with open(os.path.join(HERE, "golden", "variants.json")) as fh:
        return json.load(fh)


def carbon_text():
    with open(os.path.join(HERE, "golden", "ffield.C2013")) as fh:
        return fh.read()


def check(r, e_ref, g_ref, n, prec, tight, label):
    g_ref = np.asarray(g_ref)
    if prec == "fp64":
        e_tol, g_tol = (1e-9, 1e-7) if tight else (1e-7, 1e-1)
        for k in TERMS:
            assert abs(r["energy"][k] - e_ref[k]) <= e_tol * max(1.0, abs(e_ref[k])), (label, k, r["energy"][k], e_ref[k])
        assert np.abs(r["grad"] - g_ref).max() <= g_tol, (label, np.abs(r["grad"] - g_ref).max())
        return
    e_tol, f_max, f_rms = BARS[prec]
    assert abs(sum(r["energy"].values()) - sum(e_ref.values())) / n <= e_tol, label
    if tight:
        df = r["grad"] - g_ref
        assert np.abs(df).max() <= f_max and np.sqrt((df ** 2).mean()) <= f_rms, (label, np.abs(df).max())


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed", "fp32"])
def test_force_field_variants(golden, variants, prec):
    by_id = {c["id"]: c for c in golden["cases"]}
    for name, lg in fv.ACCEPTED.items():
        text = fv.variant(name)
        for cid, ref in variants["variants"][name].items():
            case = by_id[cid]
            x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(case)
            e = new_engine(text, case["elements"], prec, extra=["lgvdw", "yes"] if lg else [])
            e.set_box(lo, hi, tilt, per)
            e.set_atoms(types, tags, x)
            e.set_charges(q)
            r = e.compute(rebuild=True)
            e.close()
            check(r, ref["energy"], ref["grad"], len(types), prec, case["well_conditioned"], (name, cid, prec))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed", "fp32"])
def test_random_carbon_cells(variants, prec):
    from oracle.oracle import Oracle
    from reaxamd import Engine
    ff = carbon_text()
    O = Oracle(ff)
    for c in variants["carbon"]:
        n, L = c["n"], c["L"]
        e = new_engine(ff, ["C"], prec, extra=["checkqeq", "no"])
        e.set_box([0.0, 0.0, 0.0], [L, L, L])
        e.set_atoms(np.ones(n, dtype=np.int32), np.arange(1, n + 1), np.array(c["positions"]))
        r = e.compute(rebuild=True, census=True)
        cen = e.census()
        e.close()
        check(r, c["energy"], c["grad"], n, prec, True, (c["id"], prec))
        for k, cat in enumerate(Engine.CENSUS):
            if cat == "sbo_region" and prec == "fp32":
                L3 = [0.0, 0.0, 0.0, L, 0.0, 0.0, 0.0, L, 0.0, 0.0, 0.0, L]
                ref = O.compute(np.array(c["positions"]), np.zeros(n, dtype=np.int32), np.arange(1, n + 1), L3, [1, 1, 1], q=np.zeros(n))
                assert ref["census"]["sbo_region"].tolist() == c["census"][k]   # the oracle's branch is the reference's
                out, near = sbo_flips_outside_window(c["census"][k], cen[k], ref["sbo"], SBO_WINDOW)
                assert out == 0, (c["id"], prec, cat, out, near)
            else:
                assert cen[k].tolist() == c["census"][k], (c["id"], prec, cat)


@pytest.mark.gpu
def test_random_carbon_30k(ffield_text):
    """One force call of the 30,000-atom carbon workload in every precision against the FP64 oracle, and its census."""
    from oracle.oracle import CENSUS, Oracle
    from reaxamd import systems
    w = systems.random_carbon(30000, density=1.0, dmin=1.2, seed=1)
    ff = carbon_text()
    O = Oracle(ff)
    q = np.zeros(w.n)
    ref = O.compute(w.x, np.zeros(w.n, dtype=np.int32), w.tags, w.box12(), w.periodic, q=q)
    base = np.array([ref["census"][k] for k in CENSUS])
    print("30k carbon: oracle bonds/atom %.2f, angle sets %d, torsions %d" % (base[0].mean(), base[1].sum(), base[2].sum()), flush=True)
    for prec in ("fp64", "mixed", "fp32"):
        e = new_engine(ff, ["C"], prec, extra=["checkqeq", "no"])
        e.set_box(w.lo, w.hi)
        e.set_atoms(w.types, w.tags, w.x)
        r = e.compute(rebuild=True, census=True)
        cen = e.census()
        e.close()
        df = r["grad"] - ref["grad"]
        de = abs(sum(r["energy"].values()) - sum(ref["energy"].values())) / w.n
        flips = {k: int((base[i] != cen[i]).sum()) for i, k in enumerate(CENSUS)}
        sbo_out, sbo_near = sbo_flips_outside_window(base[CENSUS.index("sbo_region")], cen[CENSUS.index("sbo_region")], ref["sbo"],
                                                     SBO_WINDOW)
        print("30k carbon %s: |dE|/atom %.2e  F max %.2e  RMS %.2e  flipped atoms %s (SBO branch: %d outside the %.0e window, %d "
              "inside)" % (prec, de, np.abs(df).max(), np.sqrt((df ** 2).mean()), flips, sbo_out, SBO_WINDOW, sbo_near), flush=True)
        if prec == "fp64":
            for k in TERMS:
                assert abs(r["energy"][k] - ref["energy"][k]) <= 1e-9 * max(1.0, abs(ref["energy"][k])), k
            assert np.abs(df).max() <= 1e-6
            assert not any(flips.values())
        else:
            e_tol, f_max, f_rms = BARS[prec]
            assert de <= e_tol and np.abs(df).max() <= f_max and np.sqrt((df ** 2).mean()) <= f_rms
            if prec == "mixed":   # FP64 gates and SBO branch
                assert not any(flips.values()), flips
            else:   # gates in FP32: only rounding events near a threshold, under the reference's sanity rate (1e-3 of decisions)
                assert sbo_out == 0, (sbo_out, sbo_near)
                # decisions: the counts of the counted categories; one per atom for the two per-atom branches (SBO branch,
                # lone-pair truncation), whose values are branch indices rather than counts. SBO flips inside the rounding
                # window are excluded (asserted above to be the only ones).
                rate_flips = dict(flips, sbo_region=sbo_out)
                for i, k in enumerate(CENSUS):
                    dec = w.n if k in ("sbo_region", "lp_trunc") else int(base[i].sum())
                    assert rate_flips[k] <= 1e-3 * max(1, dec), (k, rate_flips[k], dec)
