"""ctypes wrapper of the FP64 CPU oracle (oracle/sph_oracle.cpp).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product library never calls it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

_INT_FIELDS = {"n": 0, "nAlive": 1, "nLeaves": 2, "nInternal": 3, "jTotal": 4,
               "itersRun": 5, "unconverged": 6, "nNodes": 7}
_DTYPES = {"keys": np.uint64, "order": np.int32, "csTree": np.uint64, "leafLevel": np.int32,
           "layout": np.uint32, "particleLeaf": np.int32, "internalStart": np.uint64,
           "internalLevel": np.int32, "jOffset": np.int32, "jCount": np.int32,
           "jList": np.int32, "activePerIter": np.int32}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.orc_kernel.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.orc_kernel_consts.argtypes = [ctypes.c_int, _dp]
        L.orc_hilbert_encode.argtypes = [ctypes.c_int, _u32p, _u32p, _u32p, _u64p]
        L.orc_hilbert_decode.argtypes = [ctypes.c_int, _u64p, _u32p]
        L.orc_keys.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, _u64p]
        L.orc_open_box.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_node_geometry.argtypes = [ctypes.c_int, _u64p, _i32p, _dp, ctypes.c_int, _dp, _dp]
        L.orc_new.restype = ctypes.c_void_p
        L.orc_new.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_pos.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_set_vel.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_set_acc.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_set_thermo_sorted.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_set_h_sorted.argtypes = [ctypes.c_void_p, _dp]
        L.orc_set_gradh_sorted.argtypes = [ctypes.c_void_p, _dp]
        L.orc_build.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_get_int.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_get_int.restype = ctypes.c_int
        L.orc_get_box.argtypes = [ctypes.c_void_p, _dp]
        L.orc_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
        L.orc_get.restype = ctypes.c_int
        L.orc_get_size.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.orc_get_size.restype = ctypes.c_int
        L.orc_hmax_all.argtypes = [ctypes.c_void_p]
        L.orc_build_lists.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_density_solve.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_density_solve.restype = ctypes.c_int
        L.orc_density_solve_leaves.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_int, _i32p, ctypes.c_int]
        L.orc_density_solve_leaves.restype = ctypes.c_int
        L.orc_closure.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_force.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_force_leaves.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, ctypes.c_int]
        L.orc_brute_density.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp,
                                        ctypes.c_double, _dp, _dp]
        _LIB = L
    return _LIB


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def kernel(q, quintic=False):
    q = _c(q)
    w = np.empty_like(q)
    dw = np.empty_like(q)
    lib().orc_kernel(int(quintic), q.size, q, w, dw)
    return w, dw


def kernel_consts(quintic=False):
    out = np.zeros(4)
    lib().orc_kernel_consts(int(quintic), out)
    return {"rad": out[0], "rad2": out[1], "cnormk": out[2], "hfact": out[3]}


def hilbert_encode(ix, iy, iz):
    ix, iy, iz = _c(ix, np.uint32), _c(iy, np.uint32), _c(iz, np.uint32)
    k = np.empty(ix.size, np.uint64)
    lib().orc_hilbert_encode(ix.size, ix, iy, iz, k)
    return k


def hilbert_decode(keys):
    keys = _c(keys, np.uint64)
    out = np.empty(3 * keys.size, np.uint32)
    lib().orc_hilbert_decode(keys.size, keys, out)
    return out.reshape(-1, 3)


def keys(x, y, z, h, box6, periodic=False):
    x, y, z, h = _c(x), _c(y), _c(z), _c(h)
    k = np.empty(x.size, np.uint64)
    lib().orc_keys(x.size, x, y, z, h, _c(box6), int(periodic), k)
    return k


def open_box(x, y, z):
    out = np.zeros(6)
    lib().orc_open_box(len(x), _c(x), _c(y), _c(z), out)
    return out


def node_geometry(start, level, box6, periodic=False):
    start, level = _c(start, np.uint64), _c(level, np.int32)
    c = np.empty(3 * start.size)
    s = np.empty(3 * start.size)
    lib().orc_node_geometry(start.size, start, level, _c(box6), int(periodic), c, s)
    return c.reshape(-1, 3), s.reshape(-1, 3)


def brute_density(x, y, z, h, pmass, quintic=False):
    n = len(x)
    rho = np.empty(n)
    gradh = np.empty(n)
    lib().orc_brute_density(int(quintic), n, _c(x), _c(y), _c(z), _c(h), float(pmass), rho, gradh)
    return rho, gradh


def num_threads():
    return lib().orc_num_threads()


def set_threads(t):
    lib().orc_set_threads(int(t))


class Oracle:
    """One particle set through the FP64 CPU restatement, stage by stage."""

    def __init__(self, quintic=False, bucket=64):
        self._p = lib().orc_new(int(quintic), int(bucket))
        self.quintic = quintic

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_free(self._p)
            self._p = None

    def set_particles(self, x, y, z, h, v=None, a=None):
        n = len(x)
        lib().orc_set_pos(self._p, n, _c(x), _c(y), _c(z), _c(h))
        if v is not None:
            lib().orc_set_vel(self._p, _c(v[0]), _c(v[1]), _c(v[2]))
        if a is not None:
            lib().orc_set_acc(self._p, _c(a[0]), _c(a[1]), _c(a[2]))

    def set_thermo(self, pro2, cs, alpha, u):
        """THERMO slot in the caller's order, gathered into Hilbert order."""
        lib().orc_set_thermo_sorted(self._p, _c(pro2), _c(cs), _c(alpha), _c(u))

    def set_h_sorted(self, h):
        lib().orc_set_h_sorted(self._p, _c(h))

    def set_gradh_sorted(self, gradh):
        lib().orc_set_gradh_sorted(self._p, _c(gradh))

    def build(self, periodic_box=None):
        box = np.zeros(6) if periodic_box is None else _c(periodic_box)
        lib().orc_build(self._p, int(periodic_box is not None), box)

    def geti(self, name):
        return lib().orc_get_int(self._p, _INT_FIELDS[name])

    def box(self):
        out = np.zeros(6)
        lib().orc_get_box(self._p, out)
        return out

    def get(self, name):
        nm = name.encode()
        size = lib().orc_get_size(self._p, nm)
        out = np.empty(max(size, 0), dtype=_DTYPES.get(name, np.float64))
        got = lib().orc_get(self._p, nm, out.ctypes.data_as(ctypes.c_void_p))
        if got < 0:
            raise KeyError(name)
        return out[:got]

    def hmax_all(self):
        lib().orc_hmax_all(self._p)

    def build_lists(self, symmetric=False):
        lib().orc_build_lists(self._p, int(symmetric))

    def density_solve(self, pmass, tolh=1e-4, hfact=-1.0, max_iter=100, mode=0, with_gradients=False):
        return lib().orc_density_solve(self._p, float(pmass), float(tolh), float(hfact),
                                       int(max_iter), int(mode), int(with_gradients))

    def density_solve_leaves(self, pmass, leaves, tolh=1e-4, hfact=-1.0, max_iter=100):
        """FLAT-mode solve of the particles of the given (ascending) leaves only."""
        lv = np.ascontiguousarray(leaves, dtype=np.int32)
        return lib().orc_density_solve_leaves(self._p, float(pmass), float(tolh), float(hfact), int(max_iter),
                                              lv, lv.size)

    def closure(self, pmass):
        lib().orc_closure(self._p, float(pmass))

    def force_leaves(self, pmass, leaves, hfact=-1.0, beta=2.0, alphau=1.0, disc=0, pdv=1, shock=1):
        """Force pass for the particles of the given leaves only (h and gradh of all set)."""
        lv = np.ascontiguousarray(leaves, dtype=np.int32)
        lib().orc_force_leaves(self._p, float(pmass), float(hfact), float(beta), float(alphau), int(disc), int(pdv),
                               int(shock), lv, lv.size)

    def force(self, pmass, hfact=-1.0, beta=2.0, alphau=1.0, disc=0, pdv=1, shock=1):
        lib().orc_force(self._p, float(pmass), float(hfact), float(beta), float(alphau),
                        int(disc), int(pdv), int(shock))
