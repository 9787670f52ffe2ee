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


VOIGT = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))   # xx yy zz xy xz yz


def strained(x, box12, a, b, eps):
    """positions and LAMMPS-form cell (origin + vectors as rows) after x_a -> x_a + eps x_b (a <= b keeps the LAMMPS form)"""
    F = np.eye(3)
    F[a, b] += eps
    o = np.array(box12[:3]) @ F.T
    V = np.array(box12[3:]).reshape(3, 3) @ F.T
    return x @ F.T, list(o) + list(V.reshape(-1))


def strain_virial(O, x, types, tags, box12, per, q, eps=1e-5):
    """W_ab = -dE/d eps_ab of the oracle's energy at fixed charges (central differences), Voigt order xx yy zz xy xz yz"""
    w = np.zeros(6)
    for k, (a, b) in enumerate(VOIGT):
        e = []
        for s in (eps, -eps):
            xs, bs = strained(x, box12, a, b, s)
            e.append(sum(O.compute(xs, types, tags, bs, per, q=q)["energy"].values()))
        w[k] = -(e[0] - e[1]) / (2 * eps)
    return w
