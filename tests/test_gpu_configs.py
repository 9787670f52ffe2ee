"""The workloads themselves against the FP64 oracle, beyond the small fixtures:

* the 24,000-atom water box (8000 H2O at 1.0 g/cm^3): one step in each precision against the oracle at the same positions,
  with each side's charges from its own solve to a relative residual of 1e-12 (FP32 with the fix's 'resident' refinement);
* packed independent systems (the batched workload's generator: water boxes and CH4:O2 = 1:2 gas, 100-3000 atoms): each
  system's charges, energy terms and gradient from one packed context against the oracle run on that system alone;
* served requests (the serving workload's generator, fewer steps): each request's returned energies and forces against the
  oracle at the returned positions and charges, and the returned charges against the oracle's solve there.

Bars (the target's, per atom and per component, kcal/mol and kcal/mol/A): FP64 each term within 1e-9 relative to
max(1, |E|), gradient within 1e-6; mixed and FP32 |dE|/atom <= 1e-4, force max <= 4e-3, RMS <= 9e-4; charges from tight solves
within 1e-10 e in every precision.
"""
import numpy as np
import pytest

from oracle.oracle import TERMS, Oracle
from reaxamd import Engine, serving, systems

BARS = (1e-4, 4e-3, 9e-4)


def engine(ff, prec):
    e = Engine(0)
    e.pair_style(["NULL", "precision", prec])
    e.pair_coeff(ff, list(systems.ELEMENTS))
    return e


def compare(r_energy, r_grad, ref, n, prec, label):
    df = r_grad - ref["grad"]
    de = abs(sum(r_energy.values()) - sum(ref["energy"].values())) / n
    if prec == "fp64":
        for k in TERMS:
            assert abs(r_energy[k] - ref["energy"][k]) <= 1e-9 * max(1.0, abs(ref["energy"][k])), (label, k)
        assert np.abs(df).max() <= 1e-6, (label, np.abs(df).max())
    else:
        assert de <= BARS[0], (label, de)
        assert np.abs(df).max() <= BARS[1], (label, np.abs(df).max())
        assert np.sqrt((df ** 2).mean()) <= BARS[2], (label, np.sqrt((df ** 2).mean()))
    return de, np.abs(df).max(), np.sqrt((df ** 2).mean())


@pytest.mark.gpu
def test_water_24k_one_step_every_precision(ffield_text):
    w = systems.water_box(20, seed=1)
    O = Oracle(ffield_text)
    ro = O.qeq(w.x, w.types - 1, w.tags, w.box12(), w.periodic, tol=1e-12, imax=1000)
    ref = O.compute(w.x, w.types - 1, w.tags, w.box12(), w.periodic, q=ro["q"])
    print("24k oracle: %d/%d iterations" % (ro["iters_s"], ro["iters_t"]), flush=True)
    for prec in ("fp64", "mixed", "fp32"):
        e = engine(ffield_text, prec)
        e.set_box(w.lo, w.hi)
        e.set_atoms(w.types, w.tags, w.x)
        e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
        if prec == "fp32":
            e.qeq_refine(1)   # the fix's 'resident' keyword
        q, it, rel = e.qeq_charges(imax=1000, tol=1e-12)
        r = e.compute()
        e.close()
        dq = np.abs(q - ro["q"]).max()
        de, fm, fr = compare(r["energy"], r["grad"], ref, w.n, prec, prec)
        print("24k %s: charges %.1e e, |dE|/atom %.2e, F max %.2e, RMS %.2e, iterations %s" % (prec, dq, de, fm, fr, it), flush=True)
        assert dq <= 1e-10, (prec, dq)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_packed_systems_against_oracle(ffield_text, prec):
    pool = serving.batch_systems(64, nmin=100, nmax=3000, seed=1)
    pick = [s for s in pool if s.n <= 1500][:8]
    assert len({s.name.split("_")[0] for s in pick}) == 2   # water and gas
    e = engine(ffield_text, prec)
    e.set_systems(pick)
    e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
    q, it, rel = e.qeq_charges(imax=1000, tol=1e-12)
    r = e.compute()
    res = e.system_results()
    e.close()
    O = Oracle(ffield_text)
    a = 0
    for k, s in enumerate(pick):
        sl = slice(a, a + s.n)
        a += s.n
        ro = O.qeq(s.x, s.types - 1, s.tags, s.box12(), s.periodic, tol=1e-12, imax=1000)
        ref = O.compute(s.x, s.types - 1, s.tags, s.box12(), s.periodic, q=ro["q"])
        dq = np.abs(q[sl] - ro["q"]).max()
        de, fm, fr = compare(dict(zip(TERMS, res["energy"][k])), r["grad"][sl], ref, s.n, prec, (prec, s.name))
        print("packed %s %s (%d atoms): charges %.1e e, |dE|/atom %.2e, F max %.2e" % (prec, s.name, s.n, dq, de, fm), flush=True)
        assert dq <= 1e-10, (prec, s.name, dq)


@pytest.mark.gpu
def test_served_results_against_oracle(ffield_text):
    reqs = serving.serving_requests(6, nmin=200, nmax=5000, smin=10, smax=40, seed=5)
    srv = serving.Server(ffield_text, slots=3, precision="mixed", packed=True)
    done = srv.run(reqs)
    srv.close()
    O = Oracle(ffield_text)
    for rq in done:
        s, res = rq.system, rq.result
        x = np.asarray(res["positions"])
        ref = O.compute(x, s.types - 1, s.tags, s.box12(), s.periodic, q=res["charges"])
        de, fm, fr = compare(res["energy"], res["grad"], ref, s.n, "mixed", (rq.rid, s.name))
        # the charges a tol-1e-6 solve returned, against the oracle's tight solve at the returned positions
        ro = O.qeq(x, s.types - 1, s.tags, s.box12(), s.periodic, tol=1e-12, imax=1000)
        dq = np.abs(res["charges"] - ro["q"]).max()
        print("served %d %s (%d atoms, %d steps): |dE|/atom %.2e, F max %.2e, charges vs tight solve %.1e e" % (
            rq.rid, s.name, s.n, rq.steps, de, fm, dq), flush=True)
        assert dq <= 1e-4, (rq.rid, dq)
