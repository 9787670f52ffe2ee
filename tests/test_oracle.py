"""Pins the FP64 oracle (oracle/reaxff_oracle.cpp) independently of the HIP engine it checks.

* against the reference's own FP64 engine (golden vectors written by tests/golden/make_golden.py from oracle/_ref): 13 energy terms
  and gradients on 41 C/H/O fixtures;
* against LAMMPS itself where the force field does not matter: the pair tallies vdw.oo/og/self of all 58 reference goldens (image
  expansion, far list, owner-computes counting);
* against mathematics: central finite differences of the energy, translation invariance, supercell invariance, and the charge
  equilibration against a dense explicit-image EEM solve (numpy) - a closed form independent of the CG and of the half-matrix
  bookkeeping.
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import VOIGT, case_inputs, strain_virial
from oracle.oracle import Oracle, TERMS

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def oracle_eval(O, case, q=True):
    x, types, tags, lo, hi, tilt, per, qq, box12 = case_inputs(case)
    return O.compute(x, types - 1, tags, box12, per, q=qq if q else None)


def test_oracle_matches_reference_engine(golden, ffield_text):
    O = Oracle(ffield_text)
    cases = [c for c in golden["cases"] if "ref_energy" in c]
    assert len(cases) == 41
    for c in cases:
        r = oracle_eval(O, c)
        for k in TERMS:
            ref = c["ref_energy"][k]
            assert abs(r["energy"][k] - ref) <= 1e-13 * max(1.0, abs(ref)), (c["id"], k)
        assert np.abs(r["grad"] - np.array(c["ref_grad"])).max() <= 1e-11, c["id"]


def test_pair_tallies_match_lammps(golden, ffield_text):
    """vdw.oo / vdw.og / vdw.self of LAMMPS (instrumented stock pair_style reaxff) on all 58 fixtures; they depend only on the
    geometry and the 10 A cut-off, so the stand-in force field reproduces them exactly."""
    O = Oracle(ffield_text)
    for c in golden["cases"]:
        x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(c)
        r = O.compute(x, np.zeros(len(types), dtype=np.int32), tags, box12, per, q=np.zeros(len(types)))
        got = (r["counts"]["pairs_oo"], r["counts"]["pairs_og"], r["counts"]["pairs_self"])
        t = c["lammps_tallies"]
        assert got == (t["vdw.oo"], t["vdw.og"], t["vdw.self"]), c["id"]


def test_forces_are_the_energy_gradient(golden, ffield_text):
    """Central differences (h = 1e-5 A) of the total energy at fixed charges on well-conditioned fixtures."""
    O = Oracle(ffield_text)
    ids = ["cho_water", "cho_methane_pert", "cho_benzene_pert", "cho_butadiene_pert", "cho_methanol_pert", "cho_water_dimer_pert",
           "cho_cyclopropane_pert", "cho_co2_pert", "cho_water_box_8"]
    h = 1e-5
    for c in golden["cases"]:
        if c["id"] not in ids:
            continue
        x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(c)
        g = O.compute(x, types - 1, tags, box12, per, q=q)["grad"]
        rng = np.random.default_rng(3)
        atoms = rng.choice(len(types), size=min(4, len(types)), replace=False)
        for a in atoms:
            for d in range(3):
                xp, xm = x.copy(), x.copy()
                xp[a, d] += h
                xm[a, d] -= h
                ep = sum(O.compute(xp, types - 1, tags, box12, per, q=q)["energy"].values())
                em = sum(O.compute(xm, types - 1, tags, box12, per, q=q)["energy"].values())
                fd = (ep - em) / (2 * h)
                assert abs(fd - g[a, d]) <= 1e-4 * max(1.0, abs(g[a, d])), (c["id"], a, d, fd, g[a, d])


def test_translation_invariance(golden, ffield_text):
    O = Oracle(ffield_text)
    for cid in ("cho_water_box_8", "cho_diamond_triclinic", "cho_methanol"):
        c = next(k for k in golden["cases"] if k["id"] == cid)
        x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(c)
        a = O.compute(x, types - 1, tags, box12, per, q=q)
        b = O.compute(x + np.array([0.3137, -1.271, 2.05]), types - 1, tags, box12, per, q=q)
        ea, eb = sum(a["energy"].values()), sum(b["energy"].values())
        assert abs(ea - eb) <= 2e-11 * max(1.0, abs(ea)), cid
        assert np.abs(a["grad"] - b["grad"]).max() <= 1e-8, cid


def test_supercell_invariance(golden, ffield_text):
    """An exact 2x2x2 replica of the perturbed diamond cell (8 atoms, 3.567 A: every bond, angle and torsion crosses the boundary and
    meets self images) has the same energy per atom and the same force on every image of an atom (the reference measured <= 1.5e-10)."""
    O = Oracle(ffield_text)
    c = next(k for k in golden["cases"] if k["id"] == "cho_diamond_1x1x1")
    x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(c)
    V = np.array(box12[3:]).reshape(3, 3)
    shifts = [np.array([a, b, d]) @ V for a in (0, 1) for b in (0, 1) for d in (0, 1)]
    X = np.concatenate([x + s for s in shifts])
    T = np.tile(types - 1, 8)
    Q = np.tile(q, 8)
    box2 = list(box12[:3]) + list((2 * V).reshape(-1))
    r1 = O.compute(x, types - 1, tags, box12, per, q=q)
    r8 = O.compute(X, T, np.arange(1, len(T) + 1), box2, per, q=Q)
    e1 = sum(r1["energy"].values()) / len(types)
    e8 = sum(r8["energy"].values()) / len(T)
    assert abs(e1 - e8) < 1e-9 * abs(e1)
    assert np.abs(r8["grad"] - np.tile(r1["grad"], (8, 1))).max() < 1e-8


def _dense_eem(x, types, box12, per, chi, eta, gam, swa, swb):
    """Explicit-image EEM matrix (diag eta + every image within swb), dense; closed form of the stock fix's matrix."""
    n = len(types)
    o = np.array(box12[:3])
    V = np.array(box12[3:]).reshape(3, 3)
    d1 = swb - swa
    d7 = d1 ** 7
    a2, a3, b2, b3 = swa ** 2, swa ** 3, swb ** 2, swb ** 3
    T = [(-35 * a3 * b2 * b2 + 21 * a2 * b3 * b2 - 7 * swa * b3 * b3 + b3 * b3 * swb) / d7, 140 * a3 * b3 / d7,
         -210 * (a3 * b2 + a2 * b3) / d7, 140 * (a3 * swb + 3 * a2 * b2 + swa * b3) / d7,
         -35 * (a3 + 9 * a2 * swb + 9 * swa * b2 + b3) / d7, 84 * (a2 + 3 * swa * swb + b2) / d7, -70 * (swa + swb) / d7, 20 / d7]
    h = np.abs(np.linalg.det(V)) / np.linalg.norm(np.cross(V[[1, 2, 0]], V[[2, 0, 1]]), axis=1)
    sm = [int(np.ceil(swb / h[d])) + 1 if per[d] else 0 for d in range(3)]
    A = np.diag(eta[types].astype(float))
    for s0 in range(-sm[0], sm[0] + 1):
        for s1 in range(-sm[1], sm[1] + 1):
            for s2 in range(-sm[2], sm[2] + 1):
                sh = np.array([s0, s1, s2]) @ V
                dv = x[None, :, :] + sh - x[:, None, :]
                r = np.sqrt((dv ** 2).sum(-1))
                m = (r <= swb) & (r > 1e-12)
                rr = np.where(m, r, 1.0)
                tap = np.polyval(T[::-1], rr)
                shld = (gam[types][:, None] * gam[types][None, :]) ** -1.5
                A += np.where(m, 14.4 * tap / np.cbrt(rr ** 3 + shld), 0.0)
    return A


def test_qeq_matches_dense_eem(golden, ffield_text):
    O = Oracle(ffield_text)
    chi, eta, gam = O.qeq_params()
    for cid in ("cho_water_box_8", "cho_water_cell6.2_hbond_self_image", "cho_methanol", "cho_methane_cell6.0", "cho_polyethylene_period2.553"):
        c = next(k for k in golden["cases"] if k["id"] == cid)
        x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(c)
        t0 = types - 1
        A = _dense_eem(x, t0, box12, per, chi, eta, gam, 0.0, 10.0)
        s = np.linalg.solve(A, -chi[t0])
        t = np.linalg.solve(A, -np.ones(len(t0)))
        qd = s - s.sum() / t.sum() * t
        ro = O.qeq(x, t0, tags, box12, per, tol=1e-12, imax=1000)
        assert np.abs(ro["q"] - qd).max() < 1e-9, cid
        v = np.random.default_rng(1).normal(size=len(t0))
        assert np.abs(O.qeq_matvec(x, t0, tags, box12, per, v) - A @ v).max() < 1e-11 * np.abs(A @ v).max(), cid


@pytest.mark.skipif(not os.path.exists(os.path.join(REPO, "oracle", "_ref", "reaxmetal_energy_tool")), reason="reference engine not built here")
def test_oracle_matches_live_reference_build(golden, ffield_text, tmp_path):
    """When the reference engine is built (oracle/ref.mk, build container only), re-run one fixture live."""
    c = next(k for k in golden["cases"] if k["id"] == "cho_diamond_triclinic")
    x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(c)
    path = tmp_path / "case.txt"
    lines = ["origin %r %r %r" % tuple(c["origin"])] + ["%s %r %r %r" % (k, *v) for k, v in zip("abc", c["vectors"])]
    lines += ["periodic %d %d %d" % tuple(int(p) for p in per), "atoms %d" % len(types)]
    lines += ["%d %d %r %r %r" % (i + 1, int(t), *(float(v) for v in xx)) for i, (t, xx) in enumerate(zip(types, x))]
    lines += ["charges " + " ".join(repr(float(v)) for v in q)]
    path.write_text("\n".join(lines) + "\n")
    ff = os.path.join(REPO, "tests", "golden", "ffield.cho_standin")
    out = subprocess.run([os.path.join(REPO, "oracle", "_ref", "reaxmetal_energy_tool"), "--ffield", ff, "--elements", "C,H,O", str(path)],
                         capture_output=True, text=True, check=True).stdout
    ref = {l.split()[0]: float(l.split()[1]) for l in out.splitlines() if l.split()[0] in TERMS}
    r = Oracle(ffield_text).compute(x, types - 1, tags, box12, per, q=q)
    for k in TERMS:
        assert abs(r["energy"][k] - ref[k]) <= 1e-13 * max(1.0, abs(ref[k]))


def test_strain_virial_of_an_open_molecule(golden, ffield_text):
    """The strain derivative used to pin the device virial, checked where the virial has a second closed form: for a molecule
    with no periodic images W = -sum_i x_i (x) dE/dx_i (origin-independent because the forces sum to zero)."""
    O = Oracle(ffield_text)
    for cid in ("cho_methanol_pert", "cho_benzene_pert"):
        c = next(k for k in golden["cases"] if k["id"] == cid)
        x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(c)
        assert not any(per)
        g = O.compute(x, types - 1, tags, box12, per, q=q)["grad"]
        direct = np.array([-(x[:, a] * g[:, b]).sum() for a, b in VOIGT])
        fd = strain_virial(O, x, types - 1, tags, box12, per, q)
        assert np.abs(fd - direct).max() <= 1e-5 * np.abs(direct).max() + 1e-5, (cid, fd, direct)


def test_oracle_bond_table_matches_reference(golden, ffield_text):
    """The oracle's bond table (what fix reaxff/bonds and fix reaxff/species read) against the reference engine's BondTable on the
    41 C/H/O fixtures (tests/golden/peratom.json, written by the reference's own engine): per owned atom the same partners in the
    same order, corrected bond orders, total bond order and lone pairs within 1e-12."""
    import json
    from helpers import case_inputs
    with open(os.path.join(REPO, "tests", "golden", "peratom.json")) as fh:
        ref = json.load(fh)["cases"]
    O = Oracle(ffield_text)
    nb = 0
    for case in golden["cases"]:
        if case["id"] not in ref:
            continue
        x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(case)
        bt = O.bond_table(x, types - 1, tags, box12, per)
        r = ref[case["id"]]
        for i, bonds in enumerate(r["bonds"]):
            a, b = bt["start"][i], bt["start"][i + 1]
            assert bt["owner"][a:b].tolist() == [p for p, _ in bonds], (case["id"], i)
            assert np.allclose(bt["bo"][a:b], [v for _, v in bonds], rtol=1e-12, atol=1e-13), (case["id"], i)
            nb += b - a
        assert np.allclose(bt["total_bo"], r["total_bo"], rtol=1e-12, atol=1e-13), case["id"]
        assert np.allclose(bt["nlp"], r["nlp"], rtol=1e-12, atol=1e-13), case["id"]
    assert nb > 5000
