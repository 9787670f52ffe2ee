"""Precision study of the three modes (FP64, mixed, FP32) on the water workloads.

For each mode: the error of one step against the FP64 oracle (energy per atom, largest and RMS force error, largest charge
error; charges from each side's own charge solve at tol 1e-10), then an NVE run after a short NVT equilibration, reporting the
drift of the total energy (linear fit, kcal/mol per atom per ps) and its RMS fluctuation.

    python tests/precision_study.py [--nside 6|20] [--nvt-steps N] [--nve-steps N] [--chunk N] [--modes fp64,mixed,fp32]

A measurement script beside the tests: the oracle (oracle/, test infrastructure) is its checker only.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "metal-reax_amd"))
from reaxamd import Engine, systems   # noqa: E402


def engine(ff, w, prec):
    e = Engine(0)
    e.pair_style(["NULL", "precision", prec])
    e.pair_coeff(ff, list(systems.ELEMENTS))
    e.set_box(w.lo, w.hi)
    e.set_atoms(w.types, w.tags, w.x)
    e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
    return e


def step_error(ff, w, precs):
    sys.path.insert(0, REPO)
    from oracle.oracle import Oracle
    O = Oracle(ff)
    t0 = time.perf_counter()
    ro = O.qeq(w.x, w.types - 1, w.tags, w.box12(), w.periodic, tol=1e-10, imax=500)
    ref = O.compute(w.x, w.types - 1, w.tags, w.box12(), w.periodic, q=ro["q"])
    print("oracle step: %.1f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    eref = sum(ref["energy"].values())
    out = {}
    for p in precs:
        e = engine(ff, w, p)
        q, it, rel = e.qeq_charges(imax=500, tol=1e-10)
        r = e.compute()
        e.close()
        df = r["grad"] - ref["grad"]
        out[p] = {"energy_err_per_atom": abs(sum(r["energy"].values()) - eref) / w.n,
                  "force_err_max": float(np.abs(df).max()), "force_err_rms": float(np.sqrt((df ** 2).mean())),
                  "charge_err_max": float(np.abs(q - ro["q"]).max())}
    return out


def nve_drift(ff, w, prec, nvt_steps, nve_steps, chunk, dt=0.25):
    e = engine(ff, w, prec)
    v = systems.maxwell_boltzmann(w.masses(), 300.0, seed=2)
    masses = np.array([0.0] + [systems.MASS[el] for el in systems.ELEMENTS])
    e.md_setup(v, masses, dt=dt, t_target=300.0, t_damp=25.0, skin=2.0, qeq_every=1, qeq_imax=200, qeq_tol=1e-6)
    e.md_run(nvt_steps)
    x, vel, q = e.state()
    e.close()
    e = engine(ff, w, prec)
    e.set_positions(x)
    e.md_setup(vel, masses, dt=dt, t_target=300.0, t_damp=0.0, skin=2.0, qeq_every=1, qeq_imax=200, qeq_tol=1e-6)
    t, etot = [], []
    t0 = time.perf_counter()
    for k in range(nve_steps // chunk):
        th = e.md_run(chunk)
        t.append((k + 1) * chunk * dt / 1000.0)
        etot.append((th["pe"] + th["ke"]) / w.n)
        if (k + 1) % max(1, 10000 // chunk) == 0:
            print("  %s NVE %d steps, E/atom %.6f" % (prec, (k + 1) * chunk, etot[-1]), file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    e.close()
    t, etot = np.array(t), np.array(etot)
    slope = float(np.polyfit(t, etot, 1)[0]) if len(t) > 1 else 0.0
    return {"nve_steps": nve_steps, "ps": float(t[-1]), "drift_kcal_mol_atom_ps": slope,
            "rms_fluct_kcal_mol_atom": float(etot.std()), "ms_per_step": 1e3 * wall / nve_steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nside", type=int, default=6)
    ap.add_argument("--nve-steps", type=int, default=2000)
    ap.add_argument("--nvt-steps", type=int, default=200)
    ap.add_argument("--chunk", type=int, default=50)
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--modes", default="fp64,mixed,fp32")
    args = ap.parse_args()
    ff = systems.load_ffield(os.path.join(REPO, "tests", "golden", "ffield.cho_standin"))
    w = systems.water_box(args.nside, seed=1)
    precs = tuple(args.modes.split(","))
    res = {"atoms": w.n}
    if not args.no_oracle:
        res["step_error_vs_fp64_oracle"] = step_error(ff, w, precs)
    res["nve"] = {}
    for p in precs:
        res["nve"][p] = nve_drift(ff, w, p, args.nvt_steps, args.nve_steps, args.chunk)
        print(p, json.dumps(res["nve"][p]), file=sys.stderr, flush=True)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
