"""rxa_qeq_matvec_noalloc, the charge-matrix product that allocates nothing (device vectors in the caller's order, staging in
the library's file-scope scratch), against rxa_qeq_matvec on the same context: bit for bit, on two calls in a row, the second
with a smaller system; and its refusals (a matrix that is not current, null vectors)."""
import ctypes

import numpy as np
import pytest

import reaxamd
from helpers import new_engine
from reaxamd import systems


class DeviceBuffer:
    """n doubles of device memory from the HIP runtime the engine itself runs on (the caller's allocation, not the library's)."""

    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(8 * n)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(8 * self.n), 1) == 0   # host to device

    def get(self):
        a = np.empty(self.n)
        assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(8 * self.n), 2) == 0   # device to host
        return a

    def free(self):
        self.hip.hipFree(self.p)


def hip_runtime():
    reaxamd.lib()   # loads libreaxamd.so and, with it, its HIP runtime: the same library object is returned here
    hip = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def setup(e, w):
    e.set_box(w.lo, w.hi)
    e.set_atoms(w.types, w.tags, w.x)
    e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_matvec_equals_matvec(ffield_text, prec):
    L = reaxamd.lib()
    hip = hip_runtime()
    L.rxa_qeq_matvec_noalloc.restype = ctypes.c_int
    L.rxa_qeq_matvec_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    e = new_engine(ffield_text, list(systems.ELEMENTS), prec)
    rng = np.random.default_rng(7)
    for nside in (6, 4):   # 648 atoms, then 192 on the same context and the same scratch
        w = systems.water_box(nside, seed=nside)
        setup(e, w)
        v = rng.normal(size=w.n)
        y_ref = e.qeq_matvec(v)   # builds the matrix of these positions
        xd, yd = DeviceBuffer(hip, w.n), DeviceBuffer(hip, w.n)
        xd.put(v)
        yd.put(np.full(w.n, np.nan))
        print("noalloc matvec: %s, %d atoms" % (prec, w.n), flush=True)
        rc = L.rxa_qeq_matvec_noalloc(e.h, xd.p, yd.p)
        assert rc == 0, L.rxa_last_error(e.h)
        assert hip.hipDeviceSynchronize() == 0
        y = yd.get()
        assert np.array_equal(y, y_ref), np.abs(y - y_ref).max()
        # a stale matrix is refused, nothing is launched
        e.set_positions(w.x + 0.01)
        assert L.rxa_qeq_matvec_noalloc(e.h, xd.p, yd.p) == -4   # RXA_ESTATE
        assert L.rxa_qeq_matvec_noalloc(e.h, None, yd.p) != 0
        xd.free()
        yd.free()
    e.close()
