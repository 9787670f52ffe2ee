"""Several independent systems packed in one context (rxa_set_systems) against one context per system.

FP64, charges solved to 1e-10: per-system energies within 1e-9 relative per term, gradients within 1e-7 kcal/mol/A, charges
within 1e-8 e and the same stock iteration counts (+-1), for water and CH4/O2 gas of uneven sizes (one box smaller than twice
the list cut-off, so several images of an atom are listed). The packed search and solve sum in other orders than the single
path, hence the tolerances. Then the packed path's own invariance: a request run alone in a pack of one gives, bit for bit,
what it gives packed with others (its neighbour rebuilds, charge solve, thermostat and sums are its own), and a short packed
MD run stays with single-context runs of the same systems.
"""
import numpy as np
import pytest

from reaxamd import Engine, serving, systems

TERMS = ("e_bond", "e_ov", "e_un", "e_lp", "e_ang", "e_pen", "e_coa", "e_hb", "e_tor", "e_con", "e_vdW", "e_ele", "e_pol")


def _mix():
    return [systems.water_box(3, seed=5), systems.ch4_o2_gas(300, density=0.3, seed=6), systems.water_box(5, seed=7),
            systems.ch4_o2_gas(150, density=0.3, seed=8)]


def _engine(ff, prec):
    e = Engine(0)
    e.pair_style(["NULL", "precision", prec])
    e.pair_coeff(ff, list(systems.ELEMENTS))
    return e


@pytest.mark.gpu
def test_packed_step_matches_single_contexts(ffield_text):
    ws = _mix()
    e = _engine(ffield_text, "fp64")
    e.set_systems(ws)
    e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
    q, it, rel = e.qeq_charges(imax=500, tol=1e-10)
    r = e.compute()
    per = e.system_results()
    e.close()
    a = 0
    for k, w in enumerate(ws):
        f = _engine(ffield_text, "fp64")
        f.set_box(w.lo, w.hi)
        f.set_atoms(w.types, w.tags, w.x)
        f.qeq_setup(0.0, 10.0, f.extract("chi"), f.extract("eta"), f.extract("gamma"))
        q1, it1, _ = f.qeq_charges(imax=500, tol=1e-10)
        r1 = f.compute()
        f.close()
        b = a + w.n
        print(k, w.n, "iters", it1, (per["iters_s"][k], per["iters_t"][k]), flush=True)
        assert abs(int(per["iters_s"][k]) - int(it1[0])) <= 1 and abs(int(per["iters_t"][k]) - int(it1[1])) <= 1
        assert np.abs(q[a:b] - q1).max() < 1e-8
        for t, name in enumerate(TERMS):
            x, y = per["energy"][k][t], r1["energy"][name]
            assert abs(x - y) <= 1e-9 * max(1.0, abs(y)), (k, name, x, y)
        assert np.abs(r["grad"][a:b] - r1["grad"]).max() < 1e-7
        assert np.abs(per["virial"][k] - r1["virial"]).max() <= 1e-8 * max(1.0, np.abs(r1["virial"]).max())
        a = b


def _requests(ws, steps=25):
    return [serving.Request(rid=k, system=w, steps=steps, t_target=2500.0 if k % 2 else 300.0, dt=0.1, seed=50 + k)
            for k, w in enumerate(ws)]


@pytest.mark.gpu
def test_packed_alone_equals_in_the_pack(ffield_text):
    ws = _mix()
    e = _engine(ffield_text, "mixed")
    mixed = serving.run_packed(e, ffield_text, _requests(ws))
    for k in (0, 1, 3):
        alone = serving.run_packed(e, ffield_text, [_requests(ws)[k]])[0]
        m = mixed[k]
        for name in TERMS:
            assert alone["energy"][name] == m["energy"][name], (k, name)
        for key in ("grad", "charges", "positions"):
            assert np.array_equal(alone[key], m[key]), (k, key)
        assert alone["thermo"]["temp"] == m["thermo"]["temp"]
    e.close()


@pytest.mark.gpu
def test_packed_md_tracks_single_contexts(ffield_text):
    ws = _mix()
    reqs = _requests(ws, steps=20)
    e = _engine(ffield_text, "fp64")
    pk = serving.run_packed(e, ffield_text, reqs)
    e.close()
    for k, req in enumerate(reqs):
        f = serving.new_engine(ffield_text, "fp64")
        one = serving.run_one(f, ffield_text, req, "fp64")
        f.close()
        dx = np.abs(pk[k]["positions"] - one["positions"]).max()
        de = abs(sum(pk[k]["energy"].values()) - sum(one["energy"].values())) / req.system.n
        print(k, req.system.n, "dx %.2e  dE/atom %.2e" % (dx, de), flush=True)
        assert dx < 1e-6 and de < 1e-6
