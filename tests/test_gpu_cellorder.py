"""Cell-list neighbour search with atoms kept in cell order on the device (5184-atom water, 37.3 A cube: three list cut-offs per
edge, so the cell path runs and the engine re-orders its atoms at every rebuild), checked at the boundary, which keeps the
caller's order:

* one FP64 step against the oracle (per-term 1e-9 relative, gradient 1e-6 kcal/mol/A);
* the same atoms handed over in a shuffled order give the same energies (1e-10 relative) and, per atom, the same gradient and
  bond table;
* an MD run with frequent rebuilds (skin 0.1 A) and a charge solve every step tracks the same run without the re-ordering
  (RXA_ATOM_SORT=0): positions within 1e-7 A after 40 steps in FP64 (only summation orders differ);
* charges and the charge history follow the atoms when the caller re-orders them (same iteration count as without).
"""
import os

import numpy as np
import pytest

from oracle.oracle import Oracle, TERMS


def water(nside=12):
    from reaxamd import systems
    return systems.water_box(nside, seed=3)


def engine(ff, w, prec="fp64", order=None):
    import reaxamd
    from reaxamd import systems
    e = reaxamd.Engine(0)
    e.pair_style(["NULL", "precision", prec])
    e.pair_coeff(ff, list(systems.ELEMENTS))
    e.set_box(w.lo, w.hi)
    o = np.arange(w.n) if order is None else order
    e.set_atoms(w.types[o], w.tags[o], w.x[o])
    e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
    return e


@pytest.mark.gpu
def test_cell_order_step_matches_oracle(ffield_text):
    w = water()
    print("cell order: %d atoms, box %.2f A" % (w.n, w.hi[0]), flush=True)
    O = Oracle(ffield_text)
    q = np.tile([-0.8, 0.4, 0.4], w.n // 3)
    ref = O.compute(w.x, w.types - 1, w.tags, w.box12(), w.periodic, q=q)
    e = engine(ffield_text, w)
    e.set_charges(q)
    r = e.compute()
    for k in TERMS:
        assert abs(r["energy"][k] - ref["energy"][k]) <= 1e-9 * max(1.0, abs(ref["energy"][k])), k
    assert np.abs(r["grad"] - ref["grad"]).max() < 1e-6
    assert e.stats()["rebuilds"] >= 1
    e.close()


@pytest.mark.gpu
def test_caller_order_does_not_matter(ffield_text):
    w = water()
    rng = np.random.default_rng(7)
    order = rng.permutation(w.n)
    q = np.tile([-0.8, 0.4, 0.4], w.n // 3)
    a = engine(ffield_text, w)
    a.set_charges(q)
    ra = a.compute()
    ta = a.bond_table()
    b = engine(ffield_text, w, order=order)
    b.set_charges(q[order])
    rb = b.compute()
    tb = b.bond_table()
    for k in TERMS:
        assert abs(ra["energy"][k] - rb["energy"][k]) <= 1e-10 * max(1.0, abs(ra["energy"][k])), k
    assert np.abs(ra["grad"][order] - rb["grad"]).max() < 1e-9
    # bond tables: caller index u of b is atom order[u] of a
    for u in range(0, w.n, 97):
        i = order[u]
        na = sorted(ta["nbr"][ta["start"][i]:ta["start"][i + 1]].tolist())
        nb = sorted(order[tb["nbr"][tb["start"][u]:tb["start"][u + 1]]].tolist())
        assert na == nb, u
    a.close()
    b.close()


def _md(ffield_text, w, sort, steps):
    from reaxamd import systems
    old = os.environ.get("RXA_ATOM_SORT")
    os.environ["RXA_ATOM_SORT"] = "1" if sort else "0"
    try:
        e = engine(ffield_text, w)
    finally:
        if old is None:
            del os.environ["RXA_ATOM_SORT"]
        else:
            os.environ["RXA_ATOM_SORT"] = old
    v = systems.maxwell_boltzmann(w.masses(), 300.0, seed=4)
    masses = np.array([0.0] + [systems.MASS[el] for el in systems.ELEMENTS])
    e.md_setup(v, masses, dt=0.25, t_target=300.0, t_damp=25.0, skin=0.1, qeq_every=1, qeq_imax=200, qeq_tol=1e-10)
    th = e.md_run(steps)
    x, vv, q = e.state()
    e.close()
    return th, x, vv, q


@pytest.mark.gpu
def test_md_with_reordering_tracks_fixed_order(ffield_text):
    w = water()
    ta, xa, va, qa = _md(ffield_text, w, True, 40)
    tb, xb, vb, qb = _md(ffield_text, w, False, 40)
    print("rebuilds", ta["rebuilds"], tb["rebuilds"], "max |dx|", np.abs(xa - xb).max(), flush=True)
    assert ta["rebuilds"] >= 3
    assert np.abs(xa - xb).max() < 1e-7
    assert np.abs(va - vb).max() < 1e-7
    assert np.abs(qa - qb).max() < 1e-8


@pytest.mark.gpu
def test_charge_history_follows_caller_reorder(ffield_text):
    w = water()
    e = engine(ffield_text, w)
    for _ in range(3):
        q0, it0, _ = e.qeq_charges(imax=200, tol=1e-8)
    rng = np.random.default_rng(11)
    order = rng.permutation(w.n)
    e.set_atoms(w.types[order], w.tags[order], w.x[order])   # the same atoms, new caller order
    q1, it1, _ = e.qeq_charges(imax=200, tol=1e-8)
    f = engine(ffield_text, w)
    for _ in range(3):
        f.qeq_charges(imax=200, tol=1e-8)
    q2, it2, _ = f.qeq_charges(imax=200, tol=1e-8)
    assert np.abs(q1 - q2[order]).max() < 1e-7
    assert abs(int(it1[0]) - int(it2[0])) <= 1 and abs(int(it1[1]) - int(it2[1])) <= 1
    e.close()
    f.close()
