"""The non-bonded pass over the charge-matrix rows against the same pass over the skin list, in every precision.

Within a step whose charge matrix belongs to the current positions, the rows list, in skin-list order, exactly the pairs within
the matrix cutoff (10 A = nonb_cut here), with a 16-bit key that reproduces each pair's displacement (canonical end, shift
vector). Lane l of the wave takes the same pairs in both passes, so energies, gradient and virial must agree bit for bit, on a
box smaller than twice the cutoff (several images of one atom in a row) and on a gas in cell mode, and so must a short MD
trajectory. RXA_NB_ROWS=0 forces the skin-list pass.
"""
import os

import numpy as np
import pytest

from reaxamd import Engine, systems


def _engine(ff, w, prec, rows):
    os.environ["RXA_NB_ROWS"] = "1" if rows else "0"
    try:
        e = Engine(0)
    finally:
        os.environ.pop("RXA_NB_ROWS", None)
    e.pair_style(["NULL", "precision", prec])
    e.pair_coeff(ff, list(systems.ELEMENTS))
    e.set_box(w.lo, w.hi)
    e.set_atoms(w.types, w.tags, w.x)
    e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
    return e


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed", "fp32"])
@pytest.mark.parametrize("which", ["water648", "gas1053"])
def test_rows_pass_equals_skin_pass(ffield_text, prec, which):
    w = systems.water_box(6, seed=1) if which == "water648" else systems.ch4_o2_gas(1053, density=0.3, seed=5)
    out = []
    for rows in (True, False):
        e = _engine(ffield_text, w, prec, rows)
        q, it, rel = e.qeq_charges(imax=200, tol=1e-8)
        r = e.compute()
        e.close()
        out.append((q, r))
    (qa, ra), (qb, rb) = out
    assert np.array_equal(qa, qb)
    for k in ra["energy"]:
        assert ra["energy"][k] == rb["energy"][k], (k, ra["energy"][k], rb["energy"][k])
    assert np.array_equal(ra["grad"], rb["grad"])
    assert np.array_equal(ra["virial"], rb["virial"])


@pytest.mark.gpu
def test_rows_pass_md_trajectory(ffield_text):
    w = systems.water_box(6, seed=1)
    masses = np.array([0.0] + [systems.MASS[el] for el in systems.ELEMENTS])
    v = systems.maxwell_boltzmann(w.masses(), 300.0, seed=2)
    res = []
    for rows in (True, False):
        e = _engine(ffield_text, w, "mixed", rows)
        e.md_setup(v, masses, dt=0.25, t_target=300.0, t_damp=25.0, skin=2.0, qeq_every=1, qeq_imax=200, qeq_tol=1e-6)
        th = e.md_run(60)
        res.append((e.state(), th))
        e.close()
    (xa, va, qa), tha = res[0]
    (xb, vb, qb), thb = res[1]
    assert np.array_equal(xa, xb) and np.array_equal(va, vb) and np.array_equal(qa, qb)
    assert tha["pe"] == thb["pe"]
