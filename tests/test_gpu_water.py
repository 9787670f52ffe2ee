"""The 648-atom water workload (216 H2O at 1.0 g/cm^3, seeded orientations) through the whole step: charges by the stock
fix's solve, then all 13 energy terms and forces, device vs oracle.

Tolerances: charges 1e-10 e in FP64 and mixed precision (both solves to a relative residual of 1e-12; mixed mode refines its
FP32-valued solve in FP64); fp64 1e-9 relative per energy term, 1e-6 kcal/mol/A force components; mixed |dE|/atom <= 1e-4,
F max <= 4e-3, RMS <= 9e-4 kcal/mol/A against the FP64 oracle.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from reaxamd import systems

TERMS = ("e_bond", "e_ov", "e_un", "e_lp", "e_ang", "e_pen", "e_coa", "e_hb", "e_tor", "e_con", "e_vdW", "e_ele", "e_pol")


def run_engine(sysm, ffield_text, precision):
    from reaxamd import Engine
    e = Engine(0)
    e.pair_style(["NULL", "precision", precision])
    e.pair_coeff(ffield_text, ["C", "H", "O"])
    e.set_box(sysm.lo, sysm.hi)
    e.set_atoms(sysm.types, sysm.tags, sysm.x)
    e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
    q, it, rel = e.qeq_charges(imax=1000, tol=1e-12)
    r = e.compute()
    st = e.stats()
    e.close()
    return q, it, r, st


@pytest.mark.gpu
def test_water_648_fp64_and_mixed(ffield_text):
    w = systems.water_box(6, seed=1)
    O = Oracle(ffield_text)
    ro = O.qeq(w.x, w.types - 1, w.tags, w.box12(), w.periodic, tol=1e-12, imax=1000)
    ref = O.compute(w.x, w.types - 1, w.tags, w.box12(), w.periodic, q=ro["q"])
    q, it, r, st = run_engine(w, ffield_text, "fp64")
    print("iters", it, (ro["iters_s"], ro["iters_t"]), "stats", st, flush=True)
    assert np.abs(q - ro["q"]).max() <= 1e-10
    for k in TERMS:
        a, b = r["energy"][k], ref["energy"][k]
        assert abs(a - b) <= 1e-9 * max(1.0, abs(b)), (k, a, b)
    assert np.abs(r["grad"] - ref["grad"]).max() < 1e-6
    qm, itm, rm, _ = run_engine(w, ffield_text, "mixed")
    print("charges: fp64 %.2e e, mixed %.2e e" % (np.abs(q - ro["q"]).max(), np.abs(qm - ro["q"]).max()))
    assert np.abs(qm - ro["q"]).max() <= 1e-10
    etot, eref = sum(rm["energy"].values()), sum(ref["energy"].values())
    df = rm["grad"] - ref["grad"]
    print("mixed: dE/atom %.3e  Fmax %.3e  Frms %.3e" % (abs(etot - eref) / w.n, np.abs(df).max(), np.sqrt((df ** 2).mean())))
    assert abs(etot - eref) / w.n <= 1e-4
    assert np.abs(df).max() <= 4e-3
    assert np.sqrt((df ** 2).mean()) <= 9e-4
    # net force vanishes (translation invariance), to the solve's charge noise
    assert np.abs(r["grad"].sum(0)).max() < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_water_648_nvt_run_per_step_against_oracle(ffield_text, precision):
    """The workload itself: 100 NVT steps at 300 K (dt 0.25 fs, Nose-Hoover chain, charges every step by the stock fix's
    warm-started solve at tol 1e-6) from seeded Maxwell-Boltzmann velocities. Every 10 steps, at the run's own positions and
    charges: the device's 13 energy terms and gradient against the oracle's (the bars of the single-step test), the run's
    potential energy against the oracle's total, and the charges of a tight solve at those positions (a second context, both
    sides to a relative residual of 1e-12) within 1e-10 e of the oracle's."""
    from reaxamd import Engine
    w = systems.water_box(6, seed=1)
    O = Oracle(ffield_text)
    masses = np.array([0.0] + [systems.MASS[el] for el in systems.ELEMENTS])
    e = Engine(0)
    e.pair_style(["NULL", "precision", precision])
    e.pair_coeff(ffield_text, ["C", "H", "O"])
    e.set_box(w.lo, w.hi)
    e.set_atoms(w.types, w.tags, w.x)
    e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
    e.md_setup(systems.maxwell_boltzmann(w.masses(), 300.0, seed=7), masses, dt=0.25, t_target=300.0, t_damp=25.0, skin=2.0,
               qeq_every=1, qeq_imax=200, qeq_tol=1e-6)
    worst = {"dE": 0.0, "F": 0.0, "q": 0.0}
    for block in range(10):
        th = e.md_run(10)
        x, _, q = e.state()
        r = e.compute()
        ref = O.compute(x, w.types - 1, w.tags, w.box12(), w.periodic, q=q)
        scale = sum(abs(v) for v in ref["energy"].values())   # the per-term bars, summed
        assert abs(th["pe"] - sum(ref["energy"].values())) <= (1e-9 * scale if precision == "fp64" else 1e-4 * w.n), block
        df = r["grad"] - ref["grad"]
        if precision == "fp64":
            for k in TERMS:
                assert abs(r["energy"][k] - ref["energy"][k]) <= 1e-9 * max(1.0, abs(ref["energy"][k])), (block, k)
            assert np.abs(df).max() < 1e-6, block
        else:
            assert abs(sum(r["energy"].values()) - sum(ref["energy"].values())) / w.n <= 1e-4, block
            assert np.abs(df).max() <= 4e-3 and np.sqrt((df ** 2).mean()) <= 9e-4, block
        tight = Engine(0)
        tight.pair_style(["NULL", "precision", precision])
        tight.pair_coeff(ffield_text, ["C", "H", "O"])
        tight.set_box(w.lo, w.hi)
        tight.set_atoms(w.types, w.tags, x)
        tight.qeq_setup(0.0, 10.0, tight.extract("chi"), tight.extract("eta"), tight.extract("gamma"))
        qt, _, _ = tight.qeq_charges(imax=1000, tol=1e-12)
        tight.close()
        ro = O.qeq(x, w.types - 1, w.tags, w.box12(), w.periodic, tol=1e-12, imax=1000)
        assert np.abs(qt - ro["q"]).max() <= 1e-10, block
        worst["dE"] = max(worst["dE"], abs(th["pe"] - sum(ref["energy"].values())) / w.n)
        worst["F"] = max(worst["F"], float(np.abs(df).max()))
        worst["q"] = max(worst["q"], float(np.abs(qt - ro["q"]).max()))
    e.close()
    print(precision, "100 NVT steps, worst over the 10 samples:", worst, "temperature %.1f K" % th["temp"], flush=True)
