"""Charge equilibration on the device vs the oracle's restatement of stock fix qeq/reaxff.

Tolerances: matvec 1e-12 relative (FP64 both sides, summation order differs); charges 1e-10 e in every precision with both
solves run to a relative residual of 1e-12 (the reference's bar for its resident solve: charges equal to the stock fix's to
1e-10 e, docs/USER_MANUAL.md:99). FP64 follows the stock loop (iteration counts within 1 of the oracle's); mixed mode, and FP32
mode with the fix's 'resident' keyword, multiply with FP32 matrix values inside the CG but form the initial and the verification
residuals with the matrix to about 48 bits and restart from the true residual (FP64 iterative refinement). Without the
refinement the mixed solve converges to the FP32-rounded matrix's charges, about 1e-7 e away, which the same bar rejects.
"""
import numpy as np
import pytest

from helpers import case_inputs, load_case, new_engine

from oracle.oracle import Oracle


def charged_cases(golden):
    return [c for c in golden["cases"] if c["id"].startswith("cho_") and np.abs(np.array(c["charges"])).max() > 1e-6]


def setup_qeq(e, swb=10.0):
    chi, eta, gam = e.extract("chi"), e.extract("eta"), e.extract("gamma")
    e.qeq_setup(0.0, swb, chi, eta, gam)
    return chi, eta, gam


@pytest.mark.gpu
def test_qeq_matvec_matches_oracle(golden, ffield_text):
    O = Oracle(ffield_text)
    rng = np.random.default_rng(5)
    for case in charged_cases(golden):
        e = new_engine(ffield_text, case["elements"], "fp64")
        load_case(e, case)
        setup_qeq(e)
        x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(case)
        v = rng.normal(size=len(types))
        y = e.qeq_matvec(v)
        yo = O.qeq_matvec(x, types - 1, tags, box12, per, v)
        assert np.abs(y - yo).max() <= 1e-12 * max(1.0, np.abs(yo).max()), case["id"]
        e.close()


# precision and refinement mode (-1: the precision's default, refined in mixed mode; 1: the fix's 'resident' keyword)
MODES = (("fp64", -1), ("mixed", -1), ("fp32", 1))


@pytest.mark.gpu
@pytest.mark.parametrize("prec,refine", MODES)
def test_qeq_charges_match_oracle(golden, ffield_text, prec, refine):
    O = Oracle(ffield_text)
    worst = 0.0
    for case in charged_cases(golden):
        e = new_engine(ffield_text, case["elements"], prec)
        e.qeq_refine(refine)
        load_case(e, case)
        setup_qeq(e)
        x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(case)
        qd, it, rel = e.qeq_charges(imax=1000, tol=1e-12)
        ro = O.qeq(x, types - 1, tags, box12, per, swa=0.0, swb=10.0, tol=1e-12, imax=1000)
        err = np.abs(qd - ro["q"]).max()
        worst = max(worst, err)
        print(prec, case["id"], it, (ro["iters_s"], ro["iters_t"]), "rel", rel, "dq %.2e" % err, flush=True)
        assert ro["rel_s"] <= 1e-12 and ro["rel_t"] <= 1e-12, case["id"]
        assert rel[0] <= 1e-12 and rel[1] <= 1e-12, (case["id"], rel)
        assert err <= 1e-10, (case["id"], err)
        if prec == "fp64":
            assert abs(int(it[0]) - ro["iters_s"]) <= 1 and abs(int(it[1]) - ro["iters_t"]) <= 1, case["id"]
        assert abs(qd.sum()) < 1e-9
        e.close()
    print(prec, "largest charge difference %.2e e" % worst)


@pytest.mark.gpu
def test_refinement_is_what_meets_the_bar(golden, ffield_text):
    """The same mixed-precision solve with the refinement switched off converges to the charges of the FP32-rounded matrix:
    the 1e-10 e bar above then fails, so that test measures the refinement and not a loose comparison."""
    case = next(c for c in golden["cases"] if c["id"] == "cho_water_box_8")
    O = Oracle(ffield_text)
    x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(case)
    ro = O.qeq(x, types - 1, tags, box12, per, swa=0.0, swb=10.0, tol=1e-12, imax=1000)
    errs = {}
    for mode in (0, -1):
        e = new_engine(ffield_text, case["elements"], "mixed")
        e.qeq_refine(mode)
        load_case(e, case)
        setup_qeq(e)
        qd, it, rel = e.qeq_charges(imax=1000, tol=1e-12)
        errs[mode] = np.abs(qd - ro["q"]).max()
        e.close()
    print("mixed, refinement off %.2e e, on %.2e e" % (errs[0], errs[-1]))
    assert errs[0] > 1e-9 and errs[-1] <= 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_large_maxiter_keeps_the_matrix(golden, ffield_text, prec):
    """An iteration cap beyond the engine's initial scalar capacity (2046) grows only the CG scalars and flags: the matrix,
    the diagonal and the per-atom vectors stay as built. qeq_matvec then qeq_solve (the adapter's order), and MD with the charge
    solve, give the same finite results with maxiter 3000 as with 200 (both converge in far fewer iterations)."""
    case = next(c for c in golden["cases"] if c["id"] == "cho_water_box_8")
    x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(case)
    n = len(types)
    b = np.concatenate([-e_chi(ffield_text, case, types), -np.ones(n)])
    out = {}
    for imax in (200, 3000):
        e = new_engine(ffield_text, case["elements"], prec)
        load_case(e, case)
        setup_qeq(e)
        y = e.qeq_matvec(np.ones(n))
        xs, it, rel, st = e.qeq_solve(b, np.zeros(2 * n), imax=imax, tol=1e-10)
        out[imax] = (y, xs, it, st)
        e.close()
    for k in range(2):
        assert np.all(np.isfinite(out[3000][k])) and np.array_equal(out[200][k], out[3000][k]), k
    assert np.array_equal(out[200][2], out[3000][2]) and list(out[3000][3]) == [1, 1]
    from reaxamd import systems
    masses = np.array([0.0] + [systems.MASS[el] for el in case["elements"]])
    v = np.zeros((n, 3))
    md = {}
    for imax in (200, 3000):
        e = new_engine(ffield_text, case["elements"], prec)
        load_case(e, case)
        setup_qeq(e)
        e.md_setup(v, masses, dt=0.25, t_target=300.0, t_damp=25.0, skin=1.0, qeq_every=1, qeq_imax=imax, qeq_tol=1e-8)
        e.md_run(5)
        md[imax] = e.state()
        e.close()
    for k in range(3):
        assert np.all(np.isfinite(md[3000][k])) and np.array_equal(md[200][k], md[3000][k]), k


def e_chi(ffield_text, case, types):
    """chi of each atom as the fix holds it (LAMMPS type order of the case)"""
    e = new_engine(ffield_text, case["elements"], "fp64")
    chi = e.extract("chi")
    e.close()
    return chi[types]


@pytest.mark.gpu
def test_qeq_history_extrapolation(golden, ffield_text):
    """Five solves at one geometry: the device's history extrapolation (stock init_matvec: s = 4(s0+s2) - (6 s1 + s3),
    t = t2 + 3 (t0 - t1), zero history at start) gives the same guesses, iteration counts and charges as the oracle fed with the
    same extrapolated guesses; with four history entries at one geometry the extrapolation reproduces the solution up to the
    solves' own tolerance, so the fifth solve needs far fewer iterations than the cold first one."""
    case = next(c for c in golden["cases"] if c["id"] == "cho_water_box_8")
    e = new_engine(ffield_text, case["elements"], "fp64")
    load_case(e, case)
    setup_qeq(e)
    x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(case)
    O = Oracle(ffield_text)
    sh, th = [], []
    first = None
    for step in range(5):
        qd, it, rel = e.qeq_charges(imax=200, tol=1e-8)
        z = np.zeros(len(types))
        hs = sh + [z] * (4 - len(sh))
        ht = th + [z] * (4 - len(th))
        s0 = 4 * (hs[0] + hs[2]) - (6 * hs[1] + hs[3]) if sh else None
        t0 = ht[2] + 3 * (ht[0] - ht[1]) if th else None
        ro = O.qeq(x, types - 1, tags, box12, per, tol=1e-8, imax=200, s0=s0, t0=t0)
        sh, th = [ro["s"]] + sh[:3], [ro["t"]] + th[:3]
        print(step, it, (ro["iters_s"], ro["iters_t"]), flush=True)
        assert abs(int(it[0]) - ro["iters_s"]) <= 1 and abs(int(it[1]) - ro["iters_t"]) <= 1
        assert np.abs(qd - ro["q"]).max() < 1e-7
        if first is None:
            first = it.copy()
    assert it[0] <= first[0] // 2 and it[1] <= first[1] // 2
    e.close()


@pytest.mark.gpu
def test_refinement_bound_keeps_the_true_residual(ffield_text):
    """Warm-started mixed solves at the stock tolerance, where the refinement's bound (max |x - x0| times the M^-1-norm of the
    FP32 remainders' row sums) stands in for the check's two-float product: the charges still meet the stock stopping test
    measured with the oracle's FP64 matrix, sqrt(r.M^-1 r)/|b| <= tol, for both systems. The starts lie 1e-4 to 1e-2 (relative)
    from the tight solution, from well inside the bound to where the check has to run."""
    from reaxamd import systems
    w = systems.water_box(6, seed=3)
    O = Oracle(ffield_text)
    e = new_engine(ffield_text, ["C", "H", "O"], "mixed")
    e.set_box(w.lo, w.hi)
    e.set_atoms(w.types, w.tags, w.x)
    chi, eta, gam = setup_qeq(e)
    n = w.n
    chi_a, eta_a = chi[w.types], eta[w.types]
    b = np.concatenate([-chi_a, -np.ones(n)])
    xt, it, rel, st = e.qeq_solve(b, np.zeros(2 * n), imax=1000, tol=1e-12)
    rng = np.random.default_rng(11)
    params = (chi[1:], eta[1:], gam[1:])
    for scale in (1e-4, 1e-3, 1e-2):
        x0 = xt * (1.0 + scale * rng.normal(size=2 * n))
        x, it, rel, st = e.qeq_solve(b, x0, imax=200, tol=1e-6)
        for s in range(2):
            xs, bs = x[s * n:(s + 1) * n], b[s * n:(s + 1) * n]
            r = bs - O.qeq_matvec(w.x, w.types - 1, w.tags, w.box12(), w.periodic, xs, params=params)
            true_rel = np.sqrt((r * r / eta_a).sum()) / np.linalg.norm(bs)
            print("start %.0e, system %d: iterations %d, reported %.2e, true %.2e" % (scale, s, it[s], rel[s], true_rel), flush=True)
            assert true_rel <= 1e-6 and rel[s] <= 1e-6, (scale, s, true_rel, rel[s])
    e.close()
