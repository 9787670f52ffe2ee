"""Numerical robustness of the reduced-precision modes on configurations that MD actually reaches.

water648_md_step153.npz holds positions and charges of the 648-atom water box (216 H2O, water_box(6, seed=1)) after 153 steps
of NVT MD in mixed precision (Maxwell-Boltzmann velocities at 300 K with seed 2, dt 0.25 fs, Nose-Hoover t_damp 25 fs, QEq
every step at tol 1e-6). Two oxygens there (atoms 132 and 135) are joined by a weak O-O bond (BO 0.0014) that is the central
bond of torsions, and both are under-coordinated in Delta_boc, so exp(-p_tor3 (Dj + Dk)) in the torsion's f11 switching
function (p_tor3 = 10) is large. The FP64 terms are ordinary there; the quotient-rule derivative of f11 written with its
uncancelled a^2 terms overflows FP32 and turned the gradient of 17 atoms into NaN. Each mode is checked against the FP64 oracle at the same charges, to the parity bars of
test_gpu_parity.py (mixed: |dE|/atom <= 1e-4, F max <= 4e-3, RMS <= 9e-4; fp32: 1e-3, 5e-2, 5e-3 kcal/mol/A).

The MD test runs the same start in fp64 and mixed for 400 steps, past that configuration: every coordinate stays finite and
the mixed potential energy tracks the FP64 one (the trajectories separate by chaos only slowly over 100 fs).
"""
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from reaxamd import Engine, systems

HERE = os.path.dirname(os.path.abspath(__file__))


def _engine(ff, w, prec):
    e = Engine(0)
    e.pair_style(["NULL", "precision", prec])
    e.pair_coeff(ff, list(systems.ELEMENTS))
    e.set_box(w.lo, w.hi)
    e.set_atoms(w.types, w.tags, w.x)
    return e


@pytest.mark.gpu
@pytest.mark.parametrize("prec,e_tol,f_max,f_rms", [("fp64", 1e-9, 1e-6, 1e-7), ("mixed", 1e-4, 4e-3, 9e-4), ("fp32", 1e-3, 5e-2, 5e-3)])
def test_undercoordinated_torsion_state(ffield_text, prec, e_tol, f_max, f_rms):
    w = systems.water_box(6, seed=1)
    d = np.load(os.path.join(HERE, "golden", "water648_md_step153.npz"))
    x, q = d["x"], d["q"]
    ref = Oracle(ffield_text).compute(x, w.types - 1, w.tags, w.box12(), w.periodic, q=q)
    e = _engine(ffield_text, w, prec)
    e.set_positions(x)
    e.set_charges(q)
    r = e.compute(rebuild=True)
    e.close()
    g = r["grad"]
    assert np.isfinite(g).all(), np.argwhere(~np.isfinite(g).all(1)).ravel()
    df = g - ref["grad"]
    de = abs(sum(r["energy"].values()) - sum(ref["energy"].values())) / w.n
    print("%s: dE/atom %.3e  Fmax %.3e  Frms %.3e" % (prec, de, np.abs(df).max(), np.sqrt((df ** 2).mean())), flush=True)
    assert de <= e_tol
    assert np.abs(df).max() <= f_max
    assert np.sqrt((df ** 2).mean()) <= f_rms


@pytest.mark.gpu
def test_mixed_md_stays_finite_and_tracks_fp64(ffield_text):
    w = systems.water_box(6, seed=1)
    masses = np.array([0.0] + [systems.MASS[el] for el in systems.ELEMENTS])
    v = systems.maxwell_boltzmann(w.masses(), 300.0, seed=2)
    pe = {}
    for prec in ("fp64", "mixed"):
        e = _engine(ffield_text, w, prec)
        e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
        e.md_setup(v, masses, dt=0.25, t_target=300.0, t_damp=25.0, skin=2.0, qeq_every=1, qeq_imax=200, qeq_tol=1e-6)
        th = [e.md_run(100) for _ in range(4)]
        x, vel, qq = e.state()
        e.close()
        assert np.isfinite(x).all() and np.isfinite(vel).all() and np.isfinite(qq).all(), prec
        pe[prec] = np.array([t["pe"] for t in th]) / w.n
        print(prec, "pe/atom per 100 steps", np.round(pe[prec], 5), "T", round(th[-1]["temp"], 1), flush=True)
    assert np.abs(pe["mixed"] - pe["fp64"]).max() < 0.05
