"""Per-atom energies and virials (rxa_compute with RXA_COMPUTE_PERATOM, then rxa_peratom) on the reference's C/H/O fixtures.

Golden vectors: the reference's own FP64 engine with its per-atom tallies on (tests/golden/make_peratom_golden.py, through
oracle/ref_peratom.cpp): its ev_tally attribution, images folded onto owners.

Tolerances (written here):
  fp64  : per-atom energy within 1e-9 of max(1, |e|) on the well-conditioned fixtures and 1e-7 on the flagged ones (the bars of
          the energy terms in test_gpu_parity.py: same formulas in FP64, differences are summation order and libm last bits).
  mixed : per-atom energy within 5e-4 kcal/mol (single-precision terms of up to ~200 kcal/mol per atom, FP64 sums; measured 8.5e-5).
  fp32  : per-atom energy within 5e-4 kcal/mol as well (FP32 bond orders too; measured 1.1e-4).
  every precision: the per-atom energies add up to the 13 terms (1e-9 relative) and the per-atom virials to the virial of the
  same call (1e-9 relative in FP64; 1e-6 with single-precision terms, where the shares and the call's virial group the same
  FP32 derivatives differently); a per-atom call leaves energies and gradient bitwise unchanged.
Per-atom virial (the reference's split: valence thirds, torsion quarters, H-bond thirds, the bond-order derivative pass per
bond, non-bonded half pairs), each component against the golden vatom:
  fp64  : within 1e-8 of max(1, the case's largest |component|) on the well-conditioned fixtures and 1e-6 on the flagged ones
          (FP64 sums of the same contributions in another order);
  mixed, fp32 : within 2e-4 of max(1, the case's largest |component|) (single-precision term derivatives; measured 9.8e-6
                and 2.9e-5).
  The fixtures the reference flags as ill-conditioned (near-collinear angles, products at a threshold) are held to the energy
  bars only in single-precision terms, as in the parity tests: their angle forces are large cancelling terms that FP32 resolves
  only to a few digits (the water dimer's total virial moves by 0.2 of 76 kcal/mol in mixed).
"""
import json
import os

import numpy as np
import pytest

from helpers import load_case, new_engine

HERE = os.path.dirname(os.path.abspath(__file__))


def _golden():
    with open(os.path.join(HERE, "golden", "peratom.json")) as fh:
        return json.load(fh)


def test_peratom_golden_adds_up_to_the_reference_terms(golden):
    ref = _golden()["cases"]
    n = 0
    for case in golden["cases"]:
        if case["id"] not in ref:
            continue
        e = ref[case["id"]]["eatom"]
        assert len(e) == len(case["types"])
        total = sum(case["ref_energy"].values())
        assert abs(sum(e) - total) <= 1e-9 * max(1.0, abs(total)), case["id"]
        n += 1
    assert n >= 40


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed", "fp32"])
def test_peratom_matches_reference_tallies(golden, ffield_text, precision):
    ref = _golden()["cases"]
    worst = worst_v = 0.0
    for case in golden["cases"]:
        if case["id"] not in ref:
            continue
        print("case", case["id"], precision, flush=True)
        e = new_engine(ffield_text, case["elements"], precision)
        load_case(e, case)
        plain = e.compute(rebuild=True)
        r = e.compute(rebuild=True, peratom=True)
        assert r["energy"] == plain["energy"] and np.array_equal(r["grad"], plain["grad"])
        etot = sum(r["energy"].values())
        assert abs(r["eatom"].sum() - etot) <= 1e-9 * max(1.0, abs(etot)), case["id"]
        d = np.abs(r["eatom"] - np.array(ref[case["id"]]["eatom"]))
        if precision == "fp64":
            tol = (1e-9 if case["well_conditioned"] else 1e-7) * np.maximum(1.0, np.abs(np.array(ref[case["id"]]["eatom"])))
        else:
            tol = 5e-4
        worst = max(worst, d.max())
        assert (d <= tol).all(), (case["id"], int(np.argmax(d)), d.max())
        if precision != "fp64" and not case["well_conditioned"]:
            e.close()
            continue   # flagged fixtures: the energy bar only in single-precision terms, as the parity tests
        vs = r["vatom"].sum(axis=0)
        stol = 1e-9 if precision == "fp64" else 1e-6
        assert np.abs(vs - r["virial"]).max() <= stol * max(1.0, np.abs(r["virial"]).max()), case["id"]
        vref = np.array(ref[case["id"]]["vatom"])
        scale = max(1.0, np.abs(vref).max())
        dv = np.abs(r["vatom"] - vref).max() / scale
        if precision == "fp64":
            vtol = 1e-8 if case["well_conditioned"] else 1e-6
        else:
            vtol = 2e-4
        worst_v = max(worst_v, dv)
        assert dv <= vtol, (case["id"], dv, np.unravel_index(np.argmax(np.abs(r["vatom"] - vref)), vref.shape))
        e.close()
    print("worst per-atom energy difference", precision, worst, "virial (relative to the case's largest)", worst_v)


@pytest.mark.gpu
def test_peratom_needs_a_tallying_call(golden, ffield_text):
    from reaxamd import ReaxError, _p, lib
    case = next(c for c in golden["cases"] if c["id"] == "cho_water_box_8")
    e = new_engine(ffield_text, case["elements"], "fp64")
    load_case(e, case)
    e.compute(rebuild=True, peratom=True)
    e.compute()   # a call without the flag leaves no per-atom tallies
    ea = np.zeros(e.n)
    with pytest.raises(ReaxError):
        e._chk(lib().rxa_peratom(e.h, _p(ea), None))
    e.close()
