"""ctypes binding of the FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py as the checker
of the HIP engine. The product (metal-reax_amd/) never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

TERMS = ("e_bond", "e_ov", "e_un", "e_lp", "e_ang", "e_pen", "e_coa", "e_hb", "e_tor", "e_con", "e_vdW", "e_ele", "e_pol")
CENSUS = ("bonds", "angle_sets", "torsions", "hbonds", "nonbonded", "sbo_region", "lp_trunc")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run: make -C oracle)")
        L = ctypes.CDLL(path)
        P, D, I, L64 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_long
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_create.restype = P
        L.orc_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, I]
        L.orc_destroy.argtypes = [P]
        L.orc_table_dump.restype = L64
        L.orc_table_dump.argtypes = [P, I, ctypes.c_char_p, L64]
        L.orc_num_types.argtypes = [P]
        L.orc_match.argtypes = [P, ctypes.c_char_p]
        L.orc_qeq_params.argtypes = [P, P, P, P]
        L.orc_cutoffs.argtypes = [P, P]
        L.orc_compute.argtypes = [P, L64, P, P, P, P, P, P, D, I, P, P, P, P, P, P]
        L.orc_qeq.argtypes = [P, L64, P, P, P, P, P, P, P, P, D, D, D, I, P, P, P, P, P, P]
        L.orc_qeq_matvec.argtypes = [P, L64, P, P, P, P, P, P, P, P, D, D, P, P]
        L.orc_set_threads.argtypes = [I]
        L.orc_max_threads.restype = I
        L.orc_bond_table.restype = L64
        L.orc_bond_table.argtypes = [P, L64, P, P, P, P, P, I, P, P, P, L64, P, P]
        _LIB = L
    return _LIB


def omp_threads():
    """the threads the OpenMP build would use for all host cores (OMP_NUM_THREADS, else the cores)"""
    return int(lib().orc_max_threads())


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Oracle:
    """One force field (+ control file) loaded into the oracle."""

    def __init__(self, ffield_text, control_text="", lgvdw=False):
        L = lib()
        self.h = L.orc_create(ffield_text.encode(), control_text.encode(), int(lgvdw))
        if not self.h:
            raise ValueError(L.orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_threads(self, threads):
        """OpenMP threads of the oracle's parallel loops, for every handle of this process (1: the serial oracle)."""
        lib().orc_set_threads(int(threads))

    def table_dump(self, portable=False):
        L = lib()
        n = L.orc_table_dump(self.h, int(portable), None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        L.orc_table_dump(self.h, int(portable), buf, n + 1)
        return buf.value.decode()

    def match(self, element):
        return lib().orc_match(self.h, element.encode())

    def num_types(self):
        return lib().orc_num_types(self.h)

    def qeq_params(self):
        n = self.num_types()
        chi, eta, gam = (np.zeros(n) for _ in range(3))
        lib().orc_qeq_params(self.h, _p(chi), _p(eta), _p(gam))
        return chi, eta, gam

    def cutoffs(self):
        out = np.zeros(7)
        lib().orc_cutoffs(self.h, _p(out))
        return dict(zip(("nonb_cut", "bond_cut", "hbond_cut", "thb_cut", "thb_cutsq", "bo_cut", "nonb_low"), out))

    @staticmethod
    def _sys(x, types, tags, box, periodic):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        t = np.ascontiguousarray(types, dtype=np.int32)
        g = np.ascontiguousarray(tags, dtype=np.int64)
        b = np.ascontiguousarray(box, dtype=np.float64).reshape(12)
        p = np.ascontiguousarray([int(bool(v)) for v in periodic], dtype=np.int32)
        return x, t, g, b, p

    def compute(self, x, types, tags, box, periodic, q=None, shell=0.0, enobonds=True):
        """FP64 energies (dict), gradient dE/dx (n,3) folded onto owners, counts, census, total bond order, the SBO value per atom."""
        x, t, g, b, p = self._sys(x, types, tags, box, periodic)
        n = t.size
        qq = None if q is None else np.ascontiguousarray(q, dtype=np.float64)
        e = np.zeros(13)
        grad = np.zeros(3 * n)
        counts = np.zeros(8, dtype=np.int64)
        census = np.zeros(7 * n, dtype=np.int32)
        tbo = np.zeros(n)
        sbo = np.zeros(n)
        rc = lib().orc_compute(self.h, n, _p(x), _p(t), _p(g), _p(b), _p(p), _p(qq), float(shell), int(enobonds), _p(e), _p(grad),
                               _p(counts), _p(census), _p(tbo), _p(sbo))
        if rc:
            raise RuntimeError(lib().orc_last_error().decode())
        return {
            "energy": dict(zip(TERMS, e.tolist())),
            "grad": grad.reshape(n, 3),
            "counts": dict(zip(("pairs_oo", "pairs_og", "pairs_self", "bonds", "nall", "angle_sets", "torsions", "hbonds"), counts.tolist())),
            "census": {k: census[i * n:(i + 1) * n].copy() for i, k in enumerate(CENSUS)},
            "total_bo": tbo,
            "sbo": sbo,
        }

    def bond_table(self, x, types, tags, box, periodic, enobonds=True):
        """The owned atoms' bond table (the reference's BondTable, partners as owned atom indices): start, owner, bo, total_bo, nlp."""
        x, t, g, b, p = self._sys(x, types, tags, box, periodic)
        n = t.size
        start = np.zeros(n + 1, dtype=np.int32)
        tbo, nlp = np.zeros(n), np.zeros(n)
        cap = 64 * n
        owner, bo = np.zeros(cap, dtype=np.int32), np.zeros(cap)
        nb = lib().orc_bond_table(self.h, n, _p(x), _p(t), _p(g), _p(b), _p(p), int(enobonds), _p(start), _p(owner), _p(bo), cap, _p(tbo), _p(nlp))
        if nb < 0 or nb > cap:
            raise RuntimeError(lib().orc_last_error().decode())
        return {"start": start, "owner": owner[:nb], "bo": bo[:nb], "total_bo": tbo, "nlp": nlp}

    def qeq(self, x, types, tags, box, periodic, swa=0.0, swb=10.0, tol=1e-6, imax=200, s0=None, t0=None, params=None):
        x, t, g, b, p = self._sys(x, types, tags, box, periodic)
        n = t.size
        chi, eta, gam = params if params is not None else self.qeq_params()
        chi, eta, gam = (np.ascontiguousarray(v, dtype=np.float64) for v in (chi, eta, gam))
        s0 = None if s0 is None else np.ascontiguousarray(s0, dtype=np.float64)
        t0 = None if t0 is None else np.ascontiguousarray(t0, dtype=np.float64)
        q, s, tt, info = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(4)
        rc = lib().orc_qeq(self.h, n, _p(x), _p(t), _p(g), _p(b), _p(p), _p(chi), _p(eta), _p(gam), float(swa), float(swb), float(tol), int(imax),
                           _p(s0), _p(t0), _p(q), _p(s), _p(tt), _p(info))
        if rc:
            raise RuntimeError(lib().orc_last_error().decode())
        return {"q": q, "s": s, "t": tt, "iters_s": int(info[0]), "iters_t": int(info[1]), "rel_s": info[2], "rel_t": info[3]}

    def qeq_matvec(self, x, types, tags, box, periodic, v, swa=0.0, swb=10.0, params=None):
        x, t, g, b, p = self._sys(x, types, tags, box, periodic)
        n = t.size
        chi, eta, gam = params if params is not None else self.qeq_params()
        chi, eta, gam = (np.ascontiguousarray(w, dtype=np.float64) for w in (chi, eta, gam))
        v = np.ascontiguousarray(v, dtype=np.float64)
        y = np.zeros(n)
        rc = lib().orc_qeq_matvec(self.h, n, _p(x), _p(t), _p(g), _p(b), _p(p), _p(chi), _p(eta), _p(gam), float(swa), float(swb), _p(v), _p(y))
        if rc:
            raise RuntimeError(lib().orc_last_error().decode())
        return y
