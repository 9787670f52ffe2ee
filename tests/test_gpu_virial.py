"""The device virial (non-bonded pairs + every bonded term) against the strain derivative of the oracle's energy.

For a homogeneous deformation x_a -> x_a + eps x_b of every position and cell vector, W_ab = -dE/d eps at fixed charges (the
virial LAMMPS' virial_fdotr_compute would return for the same forces on an explicit image set). Central differences of the FP64
oracle (eps = 1e-5) pin all six components independently of the engine's own force assembly (midpoint-referenced torsion halves,
half pair virials of the bond-order derivative forces). Tolerance: 1e-5 of the largest component + 1e-3 kcal/mol in FP64
(finite-difference noise of the energy), 2e-3 relative in mixed precision.
"""
import numpy as np
import pytest

from helpers import case_inputs, load_case, new_engine, strain_virial

from oracle.oracle import Oracle

IDS = ("cho_water_box_8", "cho_diamond_triclinic", "cho_water_cell6.2_hbond_self_image", "cho_methane_cell6.0",
       "cho_polyethylene_period2.553")


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_virial_is_the_strain_derivative(golden, ffield_text, prec):
    O = Oracle(ffield_text)
    cases = {c["id"]: c for c in golden["cases"]}
    for cid in IDS:
        c = cases[cid]
        x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(c)
        wo = strain_virial(O, x, types - 1, tags, box12, per, q)
        e = new_engine(ffield_text, c["elements"], prec)
        load_case(e, c)
        wd = e.compute()["virial"]
        e.close()
        scale = np.abs(wo).max()
        print(cid, prec, "device", np.round(wd, 4), "oracle", np.round(wo, 4), flush=True)
        tol = 1e-5 * scale + 1e-3 if prec == "fp64" else 2e-3 * scale + 1e-2
        assert np.abs(wd - wo).max() <= tol, (cid, wd, wo)


@pytest.mark.gpu
def test_virial_cell_mode_gas(ffield_text):
    """The same check through the cell-list path (atoms re-ordered on the device): a 1053-atom CH4/O2 gas in a 37.4 A cube.
    A condensed water box of that size cannot serve here: the H-bond term has a hard 7.5 A cut-off, and with thousands of
    donor-acceptor pairs some cross it under a 1e-5 strain, so the finite difference no longer measures a derivative."""
    from reaxamd import systems
    import reaxamd
    w = systems.ch4_o2_gas(1050, density=0.3, seed=5)
    assert w.hi[0] > 36.5   # three cells of the 12 A list cut-off per edge
    q = np.where(w.types == 2, 0.05, np.where(w.types == 1, -0.2, 0.0))
    O = Oracle(ffield_text)
    wo = strain_virial(O, w.x, w.types - 1, w.tags, w.box12(), w.periodic, q)
    e = reaxamd.Engine(0)
    e.pair_style(["NULL", "precision", "fp64"])
    e.pair_coeff(ffield_text, list(systems.ELEMENTS))
    e.set_box(w.lo, w.hi)
    e.set_atoms(w.types, w.tags, w.x)
    e.set_charges(q)
    wd = e.compute()["virial"]
    assert e.stats()["rebuilds"] >= 1
    e.close()
    print("gas %d" % w.n, np.round(wd, 4), np.round(wo, 4), flush=True)
    assert np.abs(wd - wo).max() <= 1e-5 * np.abs(wo).max() + 1e-3
