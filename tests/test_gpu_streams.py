"""The bonded chain on its own stream, overlapping the charge solve and the non-bonded pass, against everything on one stream.

The bonded kernels (bond list, BO correction, per-atom, angle, torsion, H-bond terms, BO-derivative pass) read no charges, and
the charge solve and the non-bonded pass read nothing the bonded chain writes; the gather joins the two. Run concurrently or in
sequence, every kernel computes the same numbers, so an MD trajectory (with neighbour rebuilds and the charge solve every step)
and a single force call must agree bit for bit. RXA_OVERLAP=0 puts the bonded chain on the main stream.
"""
import os

import numpy as np
import pytest

from reaxamd import Engine, systems


def _engine(ff, w, overlap, prec="mixed"):
    os.environ["RXA_OVERLAP"] = "1" if overlap else "0"
    try:
        e = Engine(0)
    finally:
        os.environ.pop("RXA_OVERLAP", None)
    e.pair_style(["NULL", "precision", prec])
    e.pair_coeff(ff, list(systems.ELEMENTS))
    e.set_box(w.lo, w.hi)
    e.set_atoms(w.types, w.tags, w.x)
    e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
    return e


@pytest.mark.gpu
def test_overlapped_md_equals_serial(ffield_text):
    w = systems.water_box(8, seed=3)   # 1536 atoms
    masses = np.array([0.0] + [systems.MASS[el] for el in systems.ELEMENTS])
    v = systems.maxwell_boltzmann(w.masses(), 300.0, seed=4)
    out = []
    for ovl in (True, False):
        e = _engine(ffield_text, w, ovl)
        e.md_setup(v, masses, dt=0.25, t_target=300.0, t_damp=25.0, skin=0.5, qeq_every=1, qeq_imax=200, qeq_tol=1e-6)
        th = e.md_run(80)
        x, vel, q = e.state()
        r = e.compute()
        out.append((x, vel, q, th, r))
        e.close()
    a, b = out
    assert a[3]["rebuilds"] >= 1
    for k in range(3):
        assert np.array_equal(a[k], b[k])
    assert a[3]["pe"] == b[3]["pe"] and a[3]["temp"] == b[3]["temp"]
    assert np.array_equal(a[4]["grad"], b[4]["grad"]) and np.array_equal(a[4]["virial"], b[4]["virial"])


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed", "fp32"])
def test_full_size_runs_repeat_bitwise(ffield_text, prec):
    """The 24,000-atom water system of the throughput measurement: two fresh contexts (and, in mixed precision, a third with
    everything on one stream) run the same 80 steps, neighbour rebuilds and the charge solve included, to the same bits."""
    w = systems.water_box(20, seed=1)   # 24,000 atoms
    masses = np.array([0.0] + [systems.MASS[el] for el in systems.ELEMENTS])
    v = systems.maxwell_boltzmann(w.masses(), 300.0, seed=2)
    runs = [True, True] + ([False] if prec == "mixed" else [])
    out = []
    for ovl in runs:
        print("24k", prec, "overlap" if ovl else "one stream", flush=True)
        e = _engine(ffield_text, w, ovl, prec)
        e.md_setup(v, masses, dt=0.25, t_target=300.0, t_damp=25.0, skin=0.5, qeq_every=1, qeq_imax=200, qeq_tol=1e-6)
        th = e.md_run(80)
        out.append((*e.state(), th))
        e.close()
    assert out[0][3]["rebuilds"] >= 1
    for o in out[1:]:
        for k in range(3):
            assert np.array_equal(out[0][k], o[k])
        assert out[0][3]["pe"] == o[3]["pe"] and out[0][3]["temp"] == o[3]["temp"]
