"""Launch configurations and the autotuner (rxa_launch_config, rxa_autotune).

Every configuration the autotuner chooses between computes the same sums in the same order (lanes per atom: the atom sums run
in slot order through the group; the angle kernels' larger group adds exact zeros ahead of the same butterfly; the second stream
only changes when kernels run; the non-bonded pass over the charge-matrix rows takes the skin-list kernel's lane assignment),
so the results must agree bit for bit: one force call on the water box in FP64 and mixed, and 12 MD steps, under every setting
against the default. The autotuner itself must return valid choices and a time for each of its 20 trials (every value timed twice).
"""
import numpy as np
import pytest

from reaxamd import Engine, systems

CONFIGS = [dict(lanes=1), dict(lanes=2), dict(lanes=8), dict(group_multiplier=2), dict(overlap=0), dict(rows=0),
           dict(lanes=8, group_multiplier=2, overlap=0, rows=0)]


def water_engine(ffield_text, prec, nside=6):
    w = systems.water_box(nside, seed=4)
    e = Engine(0)
    e.pair_style(["NULL", "precision", prec])
    e.pair_coeff(ffield_text, list(systems.ELEMENTS))
    e.set_box(w.lo, w.hi)
    e.set_atoms(w.types, w.tags, w.x)
    e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
    return e, w


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_launch_configs_give_identical_results(ffield_text, prec):
    e, w = water_engine(ffield_text, prec)
    e.qeq_charges(imax=200, tol=1e-10)
    base = e.compute(rebuild=True)
    for cfg in CONFIGS:
        e.launch_config(lanes=4, group_multiplier=1, overlap=1, rows=1)
        got = e.launch_config(**cfg)
        r = e.compute(rebuild=True)
        print(prec, got, flush=True)
        assert r["energy"] == base["energy"], (prec, cfg)
        assert np.array_equal(r["grad"], base["grad"]), (prec, cfg)
        assert np.array_equal(r["virial"], base["virial"]), (prec, cfg)
    e.close()


@pytest.mark.gpu
def test_md_identical_under_every_launch_config(ffield_text):
    masses = np.array([0.0] + [systems.MASS[el] for el in systems.ELEMENTS])
    out = []
    for cfg in (dict(), CONFIGS[-1], dict(lanes=2, overlap=0)):
        e, w = water_engine(ffield_text, "mixed")
        e.launch_config(**cfg)
        v = systems.maxwell_boltzmann(w.masses(), 300.0, seed=5)
        e.md_setup(v, masses, dt=0.25, t_target=300.0, t_damp=25.0, skin=2.0, qeq_every=1, qeq_imax=200, qeq_tol=1e-6)
        th = e.md_run(12)
        x, vel, q = e.state()
        out.append((th["pe"], x, vel, q))
        e.close()
    for k in (1, 2):
        assert out[k][0] == out[0][0]
        for a, b in zip(out[k][1:], out[0][1:]):
            assert np.array_equal(a, b)


@pytest.mark.gpu
def test_autotune_returns_valid_choices(ffield_text):
    masses = np.array([0.0] + [systems.MASS[el] for el in systems.ELEMENTS])
    e, w = water_engine(ffield_text, "mixed")
    v = systems.maxwell_boltzmann(w.masses(), 300.0, seed=6)
    e.md_setup(v, masses, dt=0.25, t_target=300.0, t_damp=25.0, skin=2.0, qeq_every=1, qeq_imax=200, qeq_tol=1e-6)
    tuned = e.autotune(steps=3)
    print(tuned, flush=True)
    assert tuned["lanes_per_atom"] in (1, 2, 4, 8) and tuned["angle_group_multiplier"] in (1, 2)
    assert len(tuned["trial_ms_per_step"]) == 20 and all(t > 0 for t in tuned["trial_ms_per_step"])
    cfg = e.launch_config()
    assert cfg["lanes_per_atom"] == tuned["lanes_per_atom"] and cfg["overlap"] == tuned["overlap"]
    th = e.md_run(5)   # the run continues under the chosen configuration
    assert np.isfinite(th["pe"])
    e.close()
