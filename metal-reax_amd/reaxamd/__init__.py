"""Python binding of the MI355X ReaxFF engine (libreaxamd.so, C ABI in include/reaxamd.h).

The engine is native (host C++ + HIP kernels for gfx950); this module only marshals arrays through ctypes. It never falls back to
a CPU path: if the shared library or the GPU is missing, every call raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "libreaxamd.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "reaxamd.h")

TERMS = ("e_bond", "e_ov", "e_un", "e_lp", "e_ang", "e_pen", "e_coa", "e_hb", "e_tor", "e_con", "e_vdW", "e_ele", "e_pol")
PVECTOR = ("eb", "ea", "elp", "emol", "ev", "epen", "ecoa", "ehb", "et", "eco", "ew", "ep", "efi", "eqeq")
PRECISIONS = ("fp64", "mixed", "fp32")

_LIB = None


class ReaxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("rxa error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Loads libreaxamd.so (built by `make -C metal-reax_amd`); raises if it is missing."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libreaxamd.so is not built (run: make -C metal-reax_amd)")
        L = ctypes.CDLL(LIB_PATH)
        P, D, I, I64, LONG, S = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int64, ctypes.c_long, ctypes.c_char_p
        PP = ctypes.POINTER(ctypes.c_char_p)
        sig = {
            "rxa_abi_version": (I, []),
            "rxa_create": (I, [ctypes.POINTER(P), I]),
            "rxa_destroy": (None, [P]),
            "rxa_last_error": (S, [P]),
            "rxa_pair_style": (I, [P, I, PP, S]),
            "rxa_pair_coeff": (I, [P, S, I, PP, P]),
            "rxa_extract": (I, [P, S, P]),
            "rxa_set_systems": (I, [P, I, P, P, P, P, P, P]),
            "rxa_system_results": (I, [P, P, P, P]),
            "rxa_md_system_targets": (I, [P, P, P]),
            "rxa_cutoffs": (I, [P, P]),
            "rxa_table_dump": (LONG, [P, I, S, LONG]),
            "rxa_set_box": (I, [P, P, P, P, P]),
            "rxa_set_atoms": (I, [P, I64, P, P, P]),
            "rxa_set_atoms_ghosted": (I, [P, I64, I64, P, P, P, I]),
            "rxa_check_ghosts": (I, [P, P, P, P, I64, I64, P, P, P, I, S, LONG]),
            "rxa_set_positions": (I, [P, P]),
            "rxa_set_charges": (I, [P, P]),
            "rxa_get_charges": (I, [P, P]),
            "rxa_compute": (I, [P, I, P, P, P]),
            "rxa_peratom": (I, [P, P, P]),
            "rxa_profile_every": (I, [P, I]),
            "rxa_pvector": (None, [P, P]),
            "rxa_census": (I, [P, P]),
            "rxa_bond_table": (LONG, [P, P, P, P, P, P]),
            "rxa_qeq_setup": (I, [P, D, D, P, P, P]),
            "rxa_qeq_matvec": (I, [P, P, P]),
            "rxa_qeq_matvec_noalloc": (I, [P, P, P]),
            "rxa_qeq_solve": (I, [P, P, P, I, D, P, P, P, P]),
            "rxa_qeq_charges": (I, [P, I, D, I, P, P, P]),
            "rxa_qeq_reset_history": (I, [P]),
            "rxa_qeq_refine": (I, [P, I]),
            "rxa_md_setup": (I, [P, P, P, D, D, D, D, I, I, D]),
            "rxa_md_run": (I, [P, I, P]),
            "rxa_get_state": (I, [P, P, P, P]),
            "rxa_md_state": (I, [P, P, P, P]),
            "rxa_autotune": (I, [P, I, P, P, I]),
            "rxa_launch_config": (I, [P, P, P]),
            "rxa_md_resume": (I, [P, P, P, D, D, D, D, I, I, D, P, P, P, P, P]),
            "rxa_stats": (I, [P, P]),
            "rxa_qeq_stats": (I, [P, P]),
            "rxa_set_positions_device": (I, [P, P]),
            "rxa_stream": (I, [P, ctypes.POINTER(P)]),
            "rxa_stream_copy": (I, [P, I64, I]),
            "rxa_profile": (I, [P, I]),
            "rxa_profile_read": (I, [P, P, P]),
            "rxa_ffield_table_dump": (I, [S, I, I, S, LONG, ctypes.POINTER(LONG), S, LONG]),
            "rxa_check_pair_style": (I, [I, PP, S, S, LONG]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _strv(items):
    arr = (ctypes.c_char_p * len(items))(*[s.encode() for s in items])
    return arr


def ffield_table_dump(text, lgvdw=False, portable=False):
    """Canonical table dump of a force field (host-only; no device)."""
    L = lib()
    n = ctypes.c_long(0)
    err = ctypes.create_string_buffer(4096)
    rc = L.rxa_ffield_table_dump(text.encode(), int(lgvdw), int(portable), None, 0, ctypes.byref(n), err, 4096)
    if rc:
        raise ReaxError(rc, err.value.decode())
    buf = ctypes.create_string_buffer(n.value + 1)
    L.rxa_ffield_table_dump(text.encode(), int(lgvdw), int(portable), buf, n.value + 1, ctypes.byref(n), err, 4096)
    return buf.value.decode()


def check_pair_style(args, control_text=""):
    """Parses `pair_style reaxff/metal` arguments (host-only); returns the effective settings as a dict."""
    L = lib()
    buf = ctypes.create_string_buffer(16384)
    rc = L.rxa_check_pair_style(len(args), _strv(list(args)), control_text.encode(), buf, 16384)
    if rc:
        raise ReaxError(rc, buf.value.decode())
    out = {"notices": []}
    for line in buf.value.decode().splitlines():
        k, _, v = line.partition("=")
        if k == "notice":
            out["notices"].append(v)
        else:
            out[k] = v
    return out


def check_ghosts(lo, hi, tilt, periodic, nlocal, types, tags, x, nranks=1):
    """The ghost checks of rxa_set_atoms_ghosted without a context: (status, message)."""
    lo, hi, tilt = (np.ascontiguousarray(v, dtype=np.float64) for v in (lo, hi, tilt))
    per = np.ascontiguousarray([int(bool(v)) for v in periodic], dtype=np.int32)
    t = np.ascontiguousarray(types, dtype=np.int32)
    g = np.ascontiguousarray(tags, dtype=np.int64)
    xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    err = ctypes.create_string_buffer(512)
    rc = lib().rxa_check_ghosts(_p(lo), _p(hi), _p(tilt), _p(per), int(nlocal), t.size - int(nlocal), _p(t), _p(g), _p(xx), int(nranks),
                                err, 512)
    return rc, err.value.decode()


def box_from_lammps(lo, hi, tilt=(0.0, 0.0, 0.0)):
    """origin + lattice vectors (12 floats) of a LAMMPS box, the layout the oracle takes."""
    a = [hi[0] - lo[0], 0.0, 0.0]
    b = [tilt[0], hi[1] - lo[1], 0.0]
    c = [tilt[1], tilt[2], hi[2] - lo[2]]
    return list(lo) + a + b + c


class Engine:
    """One engine context: force field + atoms + one HIP stream on one GPU."""

    def __init__(self, device=0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.rxa_create(ctypes.byref(h), int(device))
        if rc:
            raise ReaxError(rc, "cannot create a context on device %d (no GPU?)" % device)
        self.h = h
        self.n = 0
        self.ntypes = 0

    def close(self):
        if getattr(self, "h", None):
            lib().rxa_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc:
            raise ReaxError(rc, lib().rxa_last_error(self.h).decode())

    # ---- pair_style / pair_coeff
    def pair_style(self, args, control_text=None):
        args = list(args)
        self._chk(lib().rxa_pair_style(self.h, len(args), _strv(args), None if control_text is None else control_text.encode()))

    def pair_coeff(self, ffield_text, elements):
        self.ntypes = len(elements)
        m = np.zeros(len(elements), dtype=np.int32)
        self._chk(lib().rxa_pair_coeff(self.h, ffield_text.encode(), len(elements), _strv(list(elements)), _p(m)))
        return m

    def extract(self, name):
        out = np.zeros(self.ntypes + 1)
        self._chk(lib().rxa_extract(self.h, name.encode(), _p(out)))
        return out

    def cutoffs(self):
        out = np.zeros(4)
        self._chk(lib().rxa_cutoffs(self.h, _p(out)))
        return dict(zip(("nonb_cut", "bond_cut", "hbond_cut", "required_shell"), out.tolist()))

    def table_dump(self, portable=False):
        L = lib()
        n = L.rxa_table_dump(self.h, int(portable), None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        L.rxa_table_dump(self.h, int(portable), buf, n + 1)
        return buf.value.decode()

    # ---- atoms
    def set_box(self, lo, hi, tilt=(0.0, 0.0, 0.0), periodic=(True, True, True)):
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        hi = np.ascontiguousarray(hi, dtype=np.float64)
        ti = np.ascontiguousarray(tilt, dtype=np.float64)
        pe = np.ascontiguousarray([int(bool(p)) for p in periodic], dtype=np.int32)
        self._chk(lib().rxa_set_box(self.h, _p(lo), _p(hi), _p(ti), _p(pe)))

    def set_atoms(self, types, tags, x):
        t = np.ascontiguousarray(types, dtype=np.int32)
        g = np.ascontiguousarray(tags, dtype=np.int64)
        xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        self._chk(lib().rxa_set_atoms(self.h, t.size, _p(t), _p(g), _p(xx)))
        self.n = t.size
        self.nsys = 1

    def set_atoms_ghosted(self, nlocal, types, tags, x, nranks=1):
        """LAMMPS' owned + ghost arrays (length nlocal + nghost); the ghosts are checked and the owned atoms loaded."""
        t = np.ascontiguousarray(types, dtype=np.int32)
        g = np.ascontiguousarray(tags, dtype=np.int64)
        xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        self._chk(lib().rxa_set_atoms_ghosted(self.h, int(nlocal), t.size - int(nlocal), _p(t), _p(g), _p(xx), int(nranks)))
        self.n = int(nlocal)
        self.nsys = 1

    def set_systems(self, systems):
        """Packs independent systems (objects with types, tags, x, lo, hi and optionally periodic) into this context."""
        counts = np.array([len(w.types) for w in systems], dtype=np.int64)
        boxes = np.zeros((len(systems), 9))
        per = np.ones((len(systems), 3), dtype=np.int32)
        for k, w in enumerate(systems):
            boxes[k, 0:3] = w.lo
            boxes[k, 3:6] = w.hi
            if getattr(w, "periodic", None) is not None:
                per[k] = [int(bool(v)) for v in w.periodic]
        t = np.ascontiguousarray(np.concatenate([w.types for w in systems]), dtype=np.int32)
        g = np.ascontiguousarray(np.concatenate([w.tags for w in systems]), dtype=np.int64)
        xx = np.ascontiguousarray(np.concatenate([np.asarray(w.x).reshape(-1, 3) for w in systems]), dtype=np.float64).reshape(-1)
        self._chk(lib().rxa_set_systems(self.h, len(systems), _p(counts), _p(boxes), _p(per), _p(t), _p(g), _p(xx)))
        self.n = t.size
        self.nsys = len(systems)
        self.counts = counts

    def system_results(self):
        S = self.nsys
        e = np.zeros((S, len(TERMS)))
        v = np.zeros((S, 6))
        th = np.zeros((S, 4))
        self._chk(lib().rxa_system_results(self.h, _p(e), _p(v), _p(th)))
        return {"energy": e, "virial": v, "ke": th[:, 0], "temp": th[:, 1], "iters_s": th[:, 2].astype(int), "iters_t": th[:, 3].astype(int)}

    def md_system_targets(self, t_target=None, t_damp=None):
        tt = None if t_target is None else np.ascontiguousarray(t_target, dtype=np.float64)
        td = None if t_damp is None else np.ascontiguousarray(t_damp, dtype=np.float64)
        self._chk(lib().rxa_md_system_targets(self.h, _p(tt), _p(td)))

    def set_positions(self, x):
        xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        self._chk(lib().rxa_set_positions(self.h, _p(xx)))

    def set_charges(self, q):
        qq = np.ascontiguousarray(q, dtype=np.float64)
        self._chk(lib().rxa_set_charges(self.h, _p(qq)))

    def charges(self):
        q = np.zeros(self.n)
        self._chk(lib().rxa_get_charges(self.h, _p(q)))
        return q

    # ---- evaluation
    def compute(self, rebuild=False, want_grad=True, peratom=False, census=False):
        """Energies, gradient and virial; with peratom=True also "eatom" (n) and "vatom" (n, 6) of the same call (rxa_peratom);
        census=True records the decision census of the call (read it with census())."""
        e = np.zeros(13)
        g = np.zeros(3 * self.n) if want_grad else None
        v = np.zeros(6)
        flags = (1 if rebuild else 0) | (2 if peratom else 0) | (4 if census else 0)
        self._chk(lib().rxa_compute(self.h, flags, _p(e), _p(g), _p(v)))
        r = {"energy": dict(zip(TERMS, e.tolist())), "grad": None if g is None else g.reshape(self.n, 3), "virial": v}
        if peratom:
            ea = np.zeros(self.n)
            va = np.zeros(6 * self.n)
            self._chk(lib().rxa_peratom(self.h, _p(ea), _p(va)))
            r["eatom"] = ea
            r["vatom"] = va.reshape(self.n, 6)
        return r

    CENSUS = ("bonds", "angle_sets", "torsions", "hbonds", "nonbonded", "sbo_region", "lp_trunc")

    def census(self):
        """Per-atom decision census (7, n) in the order of CENSUS: that of the last compute(census=True), else of a force call
        made now at the current positions and charges."""
        out = np.zeros(7 * self.n, dtype=np.int32)
        self._chk(lib().rxa_census(self.h, _p(out)))
        return out.reshape(7, self.n)

    def bond_table(self):
        L = lib()
        nb = L.rxa_bond_table(self.h, None, None, None, None, None)
        if nb < 0:
            self._chk(int(nb))
        start = np.zeros(self.n + 1, dtype=np.int32)
        nbr = np.zeros(max(nb, 1), dtype=np.int32)
        bo = np.zeros(max(nb, 1))
        tbo = np.zeros(self.n)
        nlp = np.zeros(self.n)
        L.rxa_bond_table(self.h, _p(start), _p(nbr), _p(bo), _p(tbo), _p(nlp))
        return {"start": start, "nbr": nbr[:nb], "bo": bo[:nb], "total_bo": tbo, "nlp": nlp}

    # ---- charge equilibration
    def qeq_setup(self, swa, swb, chi, eta, gamma):
        c, e, g = (np.ascontiguousarray(v, dtype=np.float64) for v in (chi, eta, gamma))
        self._chk(lib().rxa_qeq_setup(self.h, float(swa), float(swb), _p(c), _p(e), _p(g)))

    def qeq_matvec(self, x):
        xx = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        self._chk(lib().rxa_qeq_matvec(self.h, _p(xx), _p(y)))
        return y

    def qeq_solve(self, b, x0, imax=200, tol=1e-6):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        x = np.zeros(2 * self.n)
        it = np.zeros(2, dtype=np.int32)
        rel = np.zeros(2)
        st = np.zeros(2, dtype=np.int32)
        self._chk(lib().rxa_qeq_solve(self.h, _p(b), _p(x0), int(imax), float(tol), _p(x), _p(it), _p(rel), _p(st)))
        return x, it, rel, st

    def qeq_charges(self, imax=200, tol=1e-6, strict=False):
        q = np.zeros(self.n)
        it = np.zeros(2, dtype=np.int32)
        rel = np.zeros(2)
        self._chk(lib().rxa_qeq_charges(self.h, int(imax), float(tol), int(strict), _p(q), _p(it), _p(rel)))
        return q, it, rel

    def qeq_reset_history(self):
        self._chk(lib().rxa_qeq_reset_history(self.h))

    def qeq_refine(self, mode):
        """FP64 iterative refinement of the charge solve: -1 the precision's default (on in mixed mode), 0 off, 1 on (the fix's
        'resident' keyword)."""
        self._chk(lib().rxa_qeq_refine(self.h, int(mode)))

    # ---- MD
    def md_setup(self, v, mass_by_type, dt=0.25, t_target=300.0, t_damp=25.0, skin=2.0, qeq_every=1, qeq_imax=200, qeq_tol=1e-6):
        vv = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        m = np.ascontiguousarray(mass_by_type, dtype=np.float64)
        self._chk(lib().rxa_md_setup(self.h, _p(vv), _p(m), float(dt), float(t_target), float(t_damp), float(skin), int(qeq_every),
                                      int(qeq_imax), float(qeq_tol)))

    def md_run(self, steps):
        th = np.zeros(8)
        self._chk(lib().rxa_md_run(self.h, int(steps), _p(th)))
        return dict(zip(("pe", "ke", "temp", "rebuilds", "qeq_iters_mean", "iters_s", "iters_t", "max_disp"), th.tolist()))

    def autotune(self, steps=20):
        """Times launch configurations on the running MD and keeps the fastest (rxa_autotune); each trial advances the run by
        2 + 2 x steps steps; each value is timed twice (in order, then in reverse) and scored by the sum. Returns the choices
        and the trials' ms per step."""
        ch = np.zeros(5, dtype=np.int64)
        ms = np.zeros(64)
        self._chk(lib().rxa_autotune(self.h, int(steps), _p(ch), _p(ms), 64))
        return {"lanes_per_atom": int(ch[0]), "angle_group_multiplier": int(ch[1]), "overlap": bool(ch[2]),
                "nonbonded_rows": bool(ch[3]), "trial_ms_per_step": ms[:int(ch[4])].tolist()}

    def launch_config(self, lanes=None, group_multiplier=None, overlap=None, rows=None):
        """Reads, and with arguments sets, the launch configuration (rxa_launch_config)."""
        cur = np.zeros(4, dtype=np.int32)
        self._chk(lib().rxa_launch_config(self.h, None, _p(cur)))
        new = cur.copy()
        for k, v in enumerate((lanes, group_multiplier, overlap, rows)):
            if v is not None:
                new[k] = int(v)
        if not np.array_equal(new, cur):
            self._chk(lib().rxa_launch_config(self.h, _p(new), None))
        return {"lanes_per_atom": int(new[0]), "angle_group_multiplier": int(new[1]), "overlap": bool(new[2]),
                "nonbonded_rows": bool(new[3])}

    def md_state(self):
        """Charge history (s and t, each (4, n), newest first) and thermostat chains ((systems, 32)): with state(), what a run
        continues from (rxa_md_state)."""
        ns = getattr(self, "nsys", 1)
        sh, th, nh = np.zeros(4 * self.n), np.zeros(4 * self.n), np.zeros(32 * ns)
        self._chk(lib().rxa_md_state(self.h, _p(sh), _p(th), _p(nh)))
        return {"s_hist": sh.reshape(4, -1), "t_hist": th.reshape(4, -1), "nh": nh.reshape(ns, 32)}

    def md_resume(self, v, mass_by_type, keep, q, mdstate, dt=0.25, t_target=300.0, t_damp=25.0, skin=2.0, qeq_every=1, qeq_imax=200,
                  qeq_tol=1e-6):
        """rxa_md_setup where the systems with keep[k] continue from charges q, and the history and chains of mdstate (the
        arrays of md_state(), laid out for this context's systems); the others start fresh (rxa_md_resume)."""
        ns = getattr(self, "nsys", 1)
        vv = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        m = np.ascontiguousarray(mass_by_type, dtype=np.float64)
        kp = np.ascontiguousarray(keep, dtype=np.int32).reshape(-1)
        qq = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        sh = np.ascontiguousarray(mdstate["s_hist"], dtype=np.float64).reshape(-1)
        th = np.ascontiguousarray(mdstate["t_hist"], dtype=np.float64).reshape(-1)
        nh = np.ascontiguousarray(mdstate["nh"], dtype=np.float64).reshape(-1)
        if len(kp) != ns or len(qq) != self.n or len(sh) != 4 * self.n or len(th) != 4 * self.n or len(nh) != 32 * ns:
            raise ValueError("md_resume: arrays do not match the context's systems")
        self._chk(lib().rxa_md_resume(self.h, _p(vv), _p(m), float(dt), float(t_target), float(t_damp), float(skin), int(qeq_every),
                                       int(qeq_imax), float(qeq_tol), _p(kp), _p(qq), _p(sh), _p(th), _p(nh)))

    def state(self):
        x = np.zeros(3 * self.n)
        v = np.zeros(3 * self.n)
        q = np.zeros(self.n)
        self._chk(lib().rxa_get_state(self.h, _p(x), _p(v), _p(q)))
        return x.reshape(-1, 3), v.reshape(-1, 3), q

    def stats(self):
        out = np.zeros(12, dtype=np.int64)
        self._chk(lib().rxa_stats(self.h, _p(out)))
        keys = ("n", "maxnb", "maxb", "maxhc", "skin_entries", "bond_slots", "hbond_entries", "h_entries", "rebuilds", "precision",
                "bond_candidates", "hb_candidates")
        return dict(zip(keys, out.tolist()))

    def qeq_stats(self):
        out = np.zeros(4, dtype=np.int64)
        self._chk(lib().rxa_qeq_stats(self.h, _p(out)))
        return dict(zip(("solves", "full_checks", "restarts", "batches"), out.tolist()))

    PHASES = ("neighbor", "bond_lists", "bo_correct", "bonded_terms", "nonbonded", "dbond_gather", "qeq_matrix", "qeq_spmv", "qeq_cg",
              "integrate")

    def profile(self, enable=True, phases=None, every=1):
        """Device timing of the engine's phases (HIP events on its stream): off, every phase, or only the named phases; with
        every > 1 only every every-th launch sequence of a phase is timed (profile_read scales the timed mean to all of them)."""
        mask = 0
        if enable:
            mask = 1 if phases is None else sum(1 << self.PHASES.index(p) for p in phases)
            if mask == 1 and phases is not None:   # only phase 0: bit 0 alone would read as "every phase"
                mask = 1 | (1 << 31)
        self._chk(lib().rxa_profile_every(self.h, int(every)))
        self._chk(lib().rxa_profile(self.h, int(np.int32(np.uint32(mask)))))

    def profile_read(self):
        ms = np.zeros(10)
        cnt = np.zeros(10, dtype=np.int64)
        self._chk(lib().rxa_profile_read(self.h, _p(ms), _p(cnt)))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(self.PHASES)}

    def stream_copy(self, nbytes, reps=10):
        """`reps` launches of a plain copy kernel over nbytes (reads nbytes, writes nbytes): memory-counter calibration."""
        self._chk(lib().rxa_stream_copy(self.h, int(nbytes), int(reps)))

    def stream_handle(self):
        h = ctypes.c_void_p()
        self._chk(lib().rxa_stream(self.h, ctypes.byref(h)))
        return h.value
