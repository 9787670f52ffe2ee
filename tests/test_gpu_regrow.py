"""Bond-slot capacity regrow (the reference forces tiny list capacities to exercise its growth path, tests/
test_metal_emulation.cpp:439). A context whose bond list starts at 1 or 2 slots per atom (RXA_BOND_CAP) overflows on the first
force call, drains, grows the slot arrays and redoes the bond-order chain; its results must equal, bit for bit, those of a context
that never overflowed (the default 12 slots): on carbon cells (up to 4-6 bonds per atom) and on fixtures of every kind, and over
an MD run of the 648-atom water box with charges in the loop, where the growth happens inside md_run."""
import os

import numpy as np
import pytest

from helpers import case_inputs, new_engine
from reaxamd import systems


class bond_cap:
    def __init__(self, cap):
        self.cap = cap

    def __enter__(self):
        os.environ["RXA_BOND_CAP"] = str(self.cap)

    def __exit__(self, *a):
        os.environ.pop("RXA_BOND_CAP", None)


def one_call(ffield_text, case, prec):
    x, types, tags, lo, hi, tilt, per, q, _ = case_inputs(case)
    e = new_engine(ffield_text, case["elements"], prec)
    e.set_box(lo, hi, tilt, per)
    e.set_atoms(types, tags, x)
    e.set_charges(q)
    r = e.compute(rebuild=True)
    st = e.stats()
    e.close()
    return r, st


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_regrow_from_tiny_capacity_is_bitwise_neutral(golden, ffield_text, prec):
    grown = 0
    for case in golden["cases"]:
        if "ref_energy" not in case:
            continue
        base, st0 = one_call(ffield_text, case, prec)
        with bond_cap(1):
            r, st = one_call(ffield_text, case, prec)
        grown += st["maxb"] > 1
        for k in base["energy"]:
            assert r["energy"][k] == base["energy"][k], (case["id"], k)
        assert np.array_equal(r["grad"], base["grad"]), case["id"]
        assert np.array_equal(r["virial"], base["virial"]), case["id"]
    assert grown > 10   # most fixtures have an atom with two or more bonds


@pytest.mark.gpu
def test_regrow_inside_md_run(ffield_text):
    w = systems.water_box(6, seed=1)
    v = systems.maxwell_boltzmann(w.masses(), 300.0, seed=2)
    masses = np.array([0.0] + [systems.MASS[el] for el in systems.ELEMENTS])
    out = []
    for cap in (None, 1):
        def run():
            e = new_engine(ffield_text, list(systems.ELEMENTS), "mixed")
            e.set_box(w.lo, w.hi)
            e.set_atoms(w.types, w.tags, w.x)
            e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
            e.md_setup(v, masses, dt=0.25, t_target=300.0, t_damp=25.0, skin=2.0, qeq_every=1, qeq_imax=200, qeq_tol=1e-6)
            th = e.md_run(30)
            x, vv, q = e.state()
            cap_now = e.stats()["maxb"]
            e.close()
            return th, x, vv, q, cap_now
        if cap is None:
            out.append(run())
        else:
            with bond_cap(cap):
                out.append(run())
    (th0, x0, v0, q0, c0), (th1, x1, v1, q1, c1) = out
    assert c0 == 12 and c1 >= 2
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1) and np.array_equal(q0, q1)
    assert th0["pe"] == th1["pe"]
