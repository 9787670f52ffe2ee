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


def lammps_ghosts(x, types, tags, lo, hi, tilt, per, shell):
    """Owned atoms followed by LAMMPS-style ghosts: every periodic image (shifts in {-1, 0, 1}^3 along periodic directions) whose
    fractional coordinates lie within `shell` (A, measured along each face normal) outside the cell, in a shuffled order.
    Returns (types, tags, x) of length nlocal + nghost."""
    x = np.asarray(x, dtype=np.float64)
    V = np.array([[hi[0] - lo[0], 0, 0], [tilt[0], hi[1] - lo[1], 0], [tilt[1], tilt[2], hi[2] - lo[2]]], dtype=np.float64)
    f = (x - np.asarray(lo)) @ np.linalg.inv(V)
    hgt = abs(np.linalg.det(V)) / np.linalg.norm(np.cross(V[[1, 2, 0]], V[[2, 0, 1]]), axis=1)
    w = shell / hgt
    gx, gt, gg = [], [], []
    rng = np.random.default_rng(11)
    for s in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij")).reshape(3, -1).T:
        if not s.any() or any(s[d] != 0 and not per[d] for d in range(3)):
            continue
        fs = f + s
        keep = np.all((fs >= -w) & (fs < 1 + w), axis=1)
        gx.append(x[keep] + s @ V)
        gt.append(np.asarray(types)[keep])
        gg.append(np.asarray(tags)[keep])
    gx, gt, gg = np.concatenate(gx), np.concatenate(gt), np.concatenate(gg)
    order = rng.permutation(len(gg))
    return (np.concatenate([types, gt[order]]).astype(np.int32), np.concatenate([tags, gg[order]]).astype(np.int64),
            np.concatenate([x, gx[order]]))
