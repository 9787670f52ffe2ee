"""Shared helpers of the test suite: fixture -> engine / oracle inputs."""
import numpy as np


def case_inputs(case):
    """positions, LAMMPS types (1-based, order of case['elements']), tags, LAMMPS box (lo, hi, tilt), periodic, charges."""
    x = np.array(case["positions"], dtype=np.float64)
    types = np.array(case["types"], dtype=np.int32)
    tags = np.arange(1, len(types) + 1, dtype=np.int64)
    o = np.array(case["origin"], dtype=np.float64)
    v = np.array(case["vectors"], dtype=np.float64)
    lo = o
    hi = o + np.array([v[0][0], v[1][1], v[2][2]])
    tilt = np.array([v[1][0], v[2][0], v[2][1]])
    box12 = list(o) + list(v.reshape(-1))
    return x, types, tags, lo, hi, tilt, case["periodic"], np.array(case["charges"]), box12


def new_engine(ffield_text, elements, precision="fp64", extra=()):
    from reaxamd import Engine
    e = Engine(0)
    e.pair_style(["NULL", "precision", precision] + list(extra))
    e.pair_coeff(ffield_text, elements)
    return e


def load_case(engine, case):
    x, types, tags, lo, hi, tilt, per, q, _ = case_inputs(case)
    engine.set_box(lo, hi, tilt, per)
    engine.set_atoms(types, tags, x)
    engine.set_charges(q)
