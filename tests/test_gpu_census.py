"""Decision and list parity with the FP64 oracle: the reference's decision census (include/reaxmetal/census.hpp:12-20). Per owned
atom, the count of each kind of discrete decision: the bond-order cut-off (bonds), the three-body gates (angle_sets), the four-body
gates (torsions), the H-bond gate (hbonds), the non-bonded cut-off (nonbonded), the SBO branch (sbo_region) and the lone-pair
truncation (lp_trunc). The kernels that take the decisions record them (rxa_compute with RXA_COMPUTE_CENSUS). An atom whose count
differs from the oracle's is a flip. Mirrors the reference's tests/python/test_decision_census.py:1-87 and
tests/python/test_threshold_scan.py:36-86:

* the 41 C/H/O fixtures as given: no flip in any precision;
* seeded random displacements of them (sigma 0.05 and 0.15 A, two samples each): no flip in FP64; in mixed and FP32 the flip
  rates are reported, each category under the reference's sanity limit (1e-3 of its decisions: more would be a logic
  difference, not a rounding event);
* threshold scans along distance families (a C-C pair, the last bond of C-C-C and C-C-C-C chains, a water dimer's O...H): the
  oracle's threshold located by bisection, the device's decisions at +-10^-k A from it (k = 3..9): FP64 takes the oracle's
  decision everywhere outside 1e-7 A; mixed and FP32 may differ from it farther out only where the energy does not jump there;
* LAMMPS' own pair tallies (vdw.oo + vdw.og + vdw.self, instrumented stock pair_style reaxff) of all 58 reference fixtures
  equal the device's counted non-bonded pairs;
* the 24,000-atom water box: no flip in FP64.
"""
import math

import numpy as np
import pytest

from helpers import case_inputs, new_engine
from oracle.oracle import Oracle
from reaxamd import Engine, systems

CATS = Engine.CENSUS


def cho_cases(golden):
    return [c for c in golden["cases"] if "ref_energy" in c]


def oracle_census(O, x, types0, tags, box12, per, q):
    r = O.compute(x, types0, tags, box12, per, q=q)
    return np.array([r["census"][c] for c in CATS]), sum(r["energy"].values())


class Device:
    """One context per (system, precision), re-used for displaced positions of the same atoms."""

    def __init__(self, ffield_text, elements, prec, lo, hi, tilt, per, types, tags, x, q):
        self.e = new_engine(ffield_text, elements, prec)
        self.e.set_box(lo, hi, tilt, per)
        self.e.set_atoms(types, tags, x)
        self.q = np.asarray(q, dtype=np.float64)

    def census(self, x=None):
        if x is not None:
            self.e.set_positions(x)
        self.e.set_charges(self.q)
        r = self.e.compute(rebuild=True, census=True)
        return self.e.census(), sum(r["energy"].values())

    def close(self):
        self.e.close()


def flips(base, other):
    """per category: decisions (the oracle's total count), atoms whose count differs"""
    return {c: (int(base[k].sum()), int((base[k] != other[k]).sum())) for k, c in enumerate(CATS)}


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed", "fp32"])
def test_census_matches_oracle_on_fixtures(golden, ffield_text, prec):
    O = Oracle(ffield_text)
    total = {c: 0 for c in CATS}
    for case in cho_cases(golden):
        x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(case)
        base, _ = oracle_census(O, x, types - 1, tags, box12, per, q)
        d = Device(ffield_text, case["elements"], prec, lo, hi, tilt, per, types, tags, x, q)
        got, _ = d.census()
        d.close()
        for k, c in enumerate(CATS):
            total[c] += int(base[k].sum())
            bad = np.nonzero(base[k] != got[k])[0]
            assert bad.size == 0, (prec, case["id"], c, bad[:8].tolist(), base[k][bad[:8]].tolist(), got[k][bad[:8]].tolist())
    print(prec, "decisions per category on the fixtures:", total)
    # every category is exercised (lp_trunc sums (int)(Delta_e / 2), which is negative at over-coordinated atoms)
    assert all(total[c] != 0 for c in CATS)


@pytest.mark.gpu
def test_census_flip_rates_under_displacements(golden, ffield_text):
    O = Oracle(ffield_text)
    rng = np.random.default_rng(20261009)
    stats = {}   # (sigma, prec, category) -> [decisions, flipped atoms]
    for case in cho_cases(golden):
        x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(case)
        samples = [(sg, x + rng.normal(0.0, sg, size=x.shape)) for sg in (0.05, 0.15) for _ in range(2)]
        bases = [oracle_census(O, xs, types - 1, tags, box12, per, q)[0] for _, xs in samples]
        for prec in ("fp64", "mixed", "fp32"):
            d = Device(ffield_text, case["elements"], prec, lo, hi, tilt, per, types, tags, samples[0][1], q)
            for (sigma, xs), base in zip(samples, bases):
                got, _ = d.census(xs)
                for c, (dec, fl) in flips(base, got).items():
                    st = stats.setdefault((sigma, prec, c), [0, 0])
                    st[0] += dec
                    st[1] += fl
                    if prec == "fp64":
                        assert fl == 0, (case["id"], sigma, c)
            d.close()
    for sigma in (0.05, 0.15):
        for prec in ("fp64", "mixed", "fp32"):
            row = []
            for c in CATS:
                dec, fl = stats[(sigma, prec, c)]
                row.append("%s %d/%d" % (c, fl, dec))
                assert dec == 0 or fl / dec <= 1e-3, (sigma, prec, c, fl, dec)
            print("sigma %.2f %-5s flipped atoms / decisions: %s" % (sigma, prec, ", ".join(row)))


def _families():
    th = math.radians(109.5)
    ct, st = math.cos(math.pi - th), math.sin(math.pi - th)

    def pair(d):
        return [1, 1], [[0, 0, 0], [d, 0, 0]], [0.1, -0.1]

    def chain3(d):
        return [1, 1, 1], [[0, 0, 0], [1.5, 0, 0], [1.5 + d * ct, d * st, 0]], [0.0] * 3

    def chain4(d):
        return [1, 1, 1, 1], [[0, 0, 0], [1.5, 0, 0], [1.5 + 1.5 * ct, 1.5 * st, 0], [1.5 + 1.5 * ct + d, 1.5 * st, 0.4]], [0.0] * 4

    def dimer(d):   # a water molecule (O at the origin, O-H along x) and a second one whose O sits d beyond the first one's H
        return ([3, 2, 2, 3, 2, 2], [[0, 0, 0], [0.96, 0, 0], [-0.24, 0.93, 0], [0.96 + d, 0, 0], [1.92 + d, 0, 0], [0.72 + d, 0.93, 0]],
                [-0.8, 0.4, 0.4, -0.8, 0.4, 0.4])

    return [("C-C pair", pair, 0.8, 10.6), ("C-C-C, last bond", chain3, 1.0, 4.5), ("C-C-C-C, last bond", chain4, 1.0, 4.5),
            ("water dimer O...H", dimer, 1.2, 8.0)]


@pytest.mark.gpu
def test_threshold_scan(ffield_text):
    O = Oracle(ffield_text)
    lo, hi, tilt, per = np.array([-30.0] * 3), np.array([30.0] * 3), np.zeros(3), (False, False, False)
    box12 = [-30.0, -30.0, -30.0, 60.0, 0, 0, 0, 60.0, 0, 0, 0, 60.0]
    rows, fails = [], []
    for name, fam, d0, d1 in _families():
        t0, x0, q0 = fam(d0)
        types = np.array(t0, dtype=np.int32)
        tags = np.arange(1, len(types) + 1, dtype=np.int64)
        q = np.array(q0)
        orc = lambda d: oracle_census(O, np.array(fam(d)[1], dtype=np.float64), types - 1, tags, box12, per, q)
        devs = {p: Device(ffield_text, ["C", "H", "O"], p, lo, hi, tilt, per, types, tags, np.array(x0, dtype=np.float64), q)
                for p in ("fp64", "mixed", "fp32")}
        nstep = int(round((d1 - d0) / 0.05))
        grid = [d0 + k * (d1 - d0) / nstep for k in range(nstep + 1)]
        prev = None
        for d in grid:
            cen = orc(d)[0]
            if prev is not None and not np.array_equal(cen, prev[1]):
                a, b = prev[0], d
                for _ in range(60):
                    mid = 0.5 * (a + b)
                    if np.array_equal(orc(mid)[0], prev[1]):
                        a = mid
                    else:
                        b = mid
                ds = 0.5 * (a + b)
                changed = [k for k in range(len(CATS)) if not np.array_equal(cen[k], prev[1][k])]
                jump = orc(ds + 1e-9)[1] - orc(ds - 1e-9)[1]
                window = {p: 0.0 for p in devs}
                for k in range(3, 10):
                    for sgn in (-1, 1):
                        dd = ds + sgn * 10.0 ** -k
                        base = orc(dd)[0]
                        xs = np.array(fam(dd)[1], dtype=np.float64)
                        for p, dv in devs.items():
                            got = dv.census(xs)[0]
                            if any(not np.array_equal(got[c], base[c]) for c in changed):
                                window[p] = max(window[p], 10.0 ** -k)
                rows.append((name, ds, "+".join(CATS[c] for c in changed), jump, window))
                if window["fp64"] > 1e-7:
                    fails.append("%s d*=%.9f: FP64 takes another decision %.0e A from the threshold" % (name, ds, window["fp64"]))
                for p in ("mixed", "fp32"):
                    if window[p] >= 1e-4 and abs(jump) > 1e-6:
                        fails.append("%s d*=%.9f: %s flips %.0e A from the threshold (energy jump %.1e)" % (name, ds, p, window[p], jump))
            prev = (d, cen)
        for dv in devs.values():
            dv.close()
    for name, ds, ch, jump, window in rows:
        print("%-20s d*=%13.9f %-24s jump %10.3e  window fp64 %s mixed %s fp32 %s" % (
            name, ds, ch, jump, *["%.0e" % window[p] if window[p] else "none" for p in ("fp64", "mixed", "fp32")]))
    assert len(rows) >= 8
    assert not fails, fails


@pytest.mark.gpu
def test_pair_counts_match_lammps_tallies(golden, ffield_text):
    """vdw.oo + vdw.og + vdw.self depend only on the geometry and the 10 A cut-off: every atom typed as carbon."""
    for case in golden["cases"]:
        x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(case)
        ones = np.ones(len(types), dtype=np.int32)
        d = Device(ffield_text, ["C"], "fp64", lo, hi, tilt, per, ones, tags, x, np.zeros(len(types)))
        got, _ = d.census()
        d.close()
        t = case["lammps_tallies"]
        assert int(got[CATS.index("nonbonded")].sum()) == t["vdw.oo"] + t["vdw.og"] + t["vdw.self"], case["id"]


@pytest.mark.gpu
def test_census_24k_water(ffield_text):
    """The 24,000-atom water box of the throughput measurement (charges do not enter any decision)."""
    w = systems.water_box(20, seed=1)
    O = Oracle(ffield_text)
    q = np.zeros(w.n)
    base, _ = oracle_census(O, w.x, w.types - 1, w.tags, w.box12(), w.periodic, q)
    for prec in ("fp64", "mixed", "fp32"):
        d = Device(ffield_text, list(systems.ELEMENTS), prec, w.lo, w.hi, np.zeros(3), w.periodic, w.types, w.tags, w.x, q)
        got, _ = d.census()
        d.close()
        fl = flips(base, got)
        print("24k", prec, "flipped atoms / decisions:", ", ".join("%s %d/%d" % (c, f, n) for c, (n, f) in fl.items()), flush=True)
        if prec == "fp64":
            assert all(f == 0 for _, f in fl.values())
        else:
            assert all(n == 0 or f / n <= 1e-3 for n, f in fl.values())
