"""Systems above 65,535 atoms: the charge matrix and the non-bonded rows switch to 32-bit columns.

A 73,167-atom water box (24,389 H2O, above the 16-bit column range): mixed and FP32 precision (FP32 matrix values, the
non-bonded pass over the charge-matrix rows, 32-bit columns) against FP64 (FP64 matrix values, the skin-list non-bonded pass) on the same
positions, both with their charges solved to tol 1e-8. The oracle takes minutes at this size, so the FP64 mode (which matches
the oracle to 1e-9 on the golden fixtures) is the reference here. Tolerances (written here): the reduced-precision bars of the
parity test (mixed: |dE| per atom <= 1e-4 kcal/mol, force max <= 4e-3 and RMS <= 9e-4 kcal/mol/A; FP32: 1e-3, 5e-2, 5e-3) and
charges within 2e-6 e (measured, mixed: 2.3e-7 e, 3.7e-6 kcal/mol per atom, 3.2e-4 and 5.5e-5 kcal/mol/A; FP32: 2.3e-7, 6.0e-6,
6.7e-4, 7.6e-5; as at 24,000 atoms with 16-bit columns). Two runs of the same precision agree bit for bit.
"""
import numpy as np
import pytest

from reaxamd import Engine, systems


def _run(ff, w, prec):
    e = Engine(0)
    e.pair_style(["NULL", "precision", prec])
    e.pair_coeff(ff, list(systems.ELEMENTS))
    e.set_box(w.lo, w.hi)
    e.set_atoms(w.types, w.tags, w.x)
    e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
    q, iters, rel = e.qeq_charges(imax=400, tol=1e-8)
    r = e.compute(rebuild=True)
    e.close()
    return q, r


@pytest.mark.gpu
@pytest.mark.parametrize("prec,e_tol,f_max,f_rms", [("mixed", 1e-4, 4e-3, 9e-4), ("fp32", 1e-3, 5e-2, 5e-3)])
def test_wide_columns_match_fp64(ffield_text, prec, e_tol, f_max, f_rms):
    w = systems.water_box(29, seed=5)
    n = len(w.types)
    assert n > 65535
    print("wide", n, "atoms", prec, flush=True)
    q64, r64 = _run(ffield_text, w, "fp64")
    qm, rm = _run(ffield_text, w, prec)
    qm2, rm2 = _run(ffield_text, w, prec)
    assert np.array_equal(qm, qm2) and np.array_equal(rm["grad"], rm2["grad"]) and rm["energy"] == rm2["energy"]
    dq = np.abs(qm - q64).max()
    de = abs(sum(rm["energy"].values()) - sum(r64["energy"].values())) / n
    df = np.abs(rm["grad"] - r64["grad"])
    print("wide %s dq %.3e dE/atom %.3e Fmax %.3e Frms %.3e" % (prec, dq, de, df.max(), np.sqrt((df ** 2).mean())), flush=True)
    assert dq <= 2e-6
    assert de <= e_tol
    assert df.max() <= f_max and np.sqrt((df ** 2).mean()) <= f_rms
