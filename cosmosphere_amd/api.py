"""Python mirror of the library's C ABI (include/cosmosphere.h), bound with ctypes.

The methods keep the reference's entry-point names, argument meaning and slot numbering
(src/dens_c_api.cu, src/force_c_api.cu, src/arrays_c_api.cu, src/pin_c_api.cu), so a test
written against this reads like a Phantom call sequence:

    lib = Library("cubic")
    lib.arrays_init(n)
    lib.upload(POS, np.stack([x, y, z]))
    lib.upload(HSML, h)
    lib.densityiterate(n, pmass, tolh=1e-4, hfact=1.2)
    h = lib.download(HSML, n)[0]

There is no fallback: if the gfx950 library is not built the constructor raises.
"""
import ctypes
import os

import numpy as np

POS, HSML, VEL, ACCEL, DENS_OUT, GRAD_OUT, THERMO, FORCE_OUT = range(1, 9)
NCOMP = {POS: 3, HSML: 1, VEL: 3, ACCEL: 3, DENS_OUT: 2, GRAD_OUT: 3, THERMO: 4, FORCE_OUT: 6}

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIBNAMES = {"cubic": "libcosmoSPHere.so", "quintic": "libcosmoSPHere_quintic.so"}

# every symbol include/cosmosphere.h declares
ABI_SYMBOLS = ["densityiterate_gpu_c", "cosmo_kernel_radius", "force_gpu_c", "cosmo_arrays_init",
               "cosmo_slot_ncomp", "cosmo_upload", "cosmo_upload_sorted", "cosmo_download",
               "cosmo_pin_host", "cosmo_unpin_host"]
EXT_SYMBOLS = ["cosmo_ext_solve_dens_h", "cosmo_ext_set_option", "cosmo_ext_get_stat",
               "cosmo_ext_get_array", "cosmo_ext_stash_inputs", "cosmo_ext_restore_inputs",
               "cosmo_ext_sync", "cosmo_batch_init", "cosmo_batch_solve", "cosmo_ext_autotune",
               "cosmo_batch_autotune", "cosmo_download_noalloc"]

_ARRAY_DTYPES = {"keys": np.uint64, "order": np.int32, "leafStart": np.uint64, "leafLevel": np.int32,
                 "leafSet": np.int32, "layout": np.uint32, "particleLeaf": np.int32,
                 "jOffset": np.uint32, "jCount": np.uint32, "jList": np.int32, "intStart": np.uint64,
                 "intLevel": np.int32, "nodeChild": np.int32, "leafNode": np.int32,
                 "aliveFirst": np.int32, "alive": np.int32}

_loaded = {}


def library_path(kernel="cubic"):
    return os.path.join(LIBDIR, LIBNAMES[kernel])


def _bind(L):
    c_dp = ctypes.POINTER(ctypes.c_double)
    L.densityiterate_gpu_c.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, c_dp, ctypes.c_double,
                                       ctypes.c_double]
    L.densityiterate_gpu_c.restype = None
    L.cosmo_kernel_radius.restype = ctypes.c_double
    L.force_gpu_c.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int]
    L.cosmo_arrays_init.argtypes = [ctypes.c_int]
    L.cosmo_slot_ncomp.argtypes = [ctypes.c_int]
    L.cosmo_slot_ncomp.restype = ctypes.c_int
    for f in (L.cosmo_upload, L.cosmo_upload_sorted, L.cosmo_download):
        f.argtypes = [ctypes.c_int, c_dp, ctypes.c_int]
    L.cosmo_download_noalloc.argtypes = [ctypes.c_int, c_dp, ctypes.c_int]
    L.cosmo_download_noalloc.restype = ctypes.c_int
    L.cosmo_pin_host.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    L.cosmo_unpin_host.argtypes = [ctypes.c_void_p]
    L.cosmo_ext_solve_dens_h.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         c_dp, ctypes.c_double, ctypes.c_double]
    L.cosmo_ext_solve_dens_h.restype = ctypes.c_int
    L.cosmo_ext_set_option.argtypes = [ctypes.c_char_p, ctypes.c_double]
    L.cosmo_ext_get_stat.argtypes = [ctypes.c_char_p]
    L.cosmo_ext_get_stat.restype = ctypes.c_double
    L.cosmo_ext_get_array.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_longlong]
    L.cosmo_ext_get_array.restype = ctypes.c_longlong
    L.cosmo_batch_init.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.cosmo_batch_solve.argtypes = [c_dp, ctypes.c_int, c_dp, ctypes.c_double, ctypes.c_double]
    L.cosmo_batch_solve.restype = ctypes.c_int
    L.cosmo_ext_autotune.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, c_dp, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double]
    L.cosmo_ext_autotune.restype = ctypes.c_int
    L.cosmo_batch_autotune.argtypes = [c_dp, ctypes.c_int, c_dp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    L.cosmo_batch_autotune.restype = ctypes.c_int
    return L


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Library:
    """One loaded kernel variant of the gfx950 library (process-wide device state)."""

    def __init__(self, kernel="cubic"):
        path = library_path(kernel)
        if not os.path.exists(path):
            raise RuntimeError("the gfx950 library %s is not built; run __graft_entry__.build()" % path)
        if kernel not in _loaded:
            _loaded[kernel] = _bind(ctypes.CDLL(path, mode=os.RTLD_LOCAL))
        self.L = _loaded[kernel]
        self.kernel = kernel
        self.path = path

    # ---- the reference entry points ----
    def kernel_radius(self):
        return self.L.cosmo_kernel_radius()

    def slot_ncomp(self, slot):
        return self.L.cosmo_slot_ncomp(int(slot))

    def arrays_init(self, n):
        self.L.cosmo_arrays_init(int(n))

    def _block(self, slot, data, n):
        a = np.ascontiguousarray(data, dtype=np.float64).reshape(-1)
        if a.size != NCOMP[slot] * n:
            raise ValueError("slot %d needs %d x %d doubles, got %d" % (slot, NCOMP[slot], n, a.size))
        return a

    def upload(self, slot, data, n=None):
        n = int(n if n is not None else np.asarray(data).shape[-1])
        a = self._block(slot, data, n)
        self.L.cosmo_upload(int(slot), _dptr(a), n)

    def upload_sorted(self, slot, data, n=None):
        n = int(n if n is not None else np.asarray(data).shape[-1])
        a = self._block(slot, data, n)
        self.L.cosmo_upload_sorted(int(slot), _dptr(a), n)

    def download(self, slot, n):
        out = np.empty(NCOMP[slot] * int(n))
        self.L.cosmo_download(int(slot), _dptr(out), int(n))
        return out.reshape(NCOMP[slot], int(n))

    def download_noalloc(self, slot, n):
        """cosmo_download_noalloc: as download, through the library's static device buffer."""
        out = np.empty(NCOMP[slot] * int(n))
        rc = self.L.cosmo_download_noalloc(int(slot), _dptr(out), int(n))
        if rc != 0:
            raise ValueError("cosmo_download_noalloc refused n=%d" % n)
        return out.reshape(NCOMP[slot], int(n))

    def densityiterate(self, n, pmass, periodic_box=None, tolh=1e-4, hfact=-1.0):
        box = np.ascontiguousarray(periodic_box if periodic_box is not None else np.zeros(6), dtype=np.float64)
        self.L.densityiterate_gpu_c(int(n), float(pmass), int(periodic_box is not None), _dptr(box), float(tolh),
                                    float(hfact))

    def force(self, n, pmass, beta=2.0, alphau=1.0, disc_viscosity=0, pdv_heating=1, shock_heating=1):
        self.L.force_gpu_c(int(n), float(pmass), float(beta), float(alphau), int(disc_viscosity), int(pdv_heating),
                           int(shock_heating))

    def pin_host(self, arr):
        self.L.cosmo_pin_host(arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes)

    def unpin_host(self, arr):
        self.L.cosmo_unpin_host(arr.ctypes.data_as(ctypes.c_void_p))

    # ---- extensions ----
    def solve_dens_h(self, n, pmass, mode=0, with_gradients=False, periodic_box=None, tolh=1e-4, hfact=-1.0):
        """The reference's solveDensH (include/density.hpp:85-91); returns iterations run."""
        box = np.ascontiguousarray(periodic_box if periodic_box is not None else np.zeros(6), dtype=np.float64)
        return self.L.cosmo_ext_solve_dens_h(int(n), float(pmass), int(mode), int(with_gradients),
                                             int(periodic_box is not None), _dptr(box), float(tolh), float(hfact))

    def set_option(self, name, value):
        self.L.cosmo_ext_set_option(name.encode(), float(value))

    def stat(self, name):
        return self.L.cosmo_ext_get_stat(name.encode())

    def array(self, name, count_hint=None):
        """A device array (Hilbert order where per particle) copied to host."""
        dt = np.dtype(_ARRAY_DTYPES.get(name, np.float64))
        count = self.L.cosmo_ext_get_array(name.encode(), None, 0)
        if count < 0:
            raise KeyError(name)
        out = np.empty(int(count), dtype=dt)
        if count:
            self.L.cosmo_ext_get_array(name.encode(), out.ctypes.data_as(ctypes.c_void_p), out.nbytes)
        return out

    PASS_FIELDS = ("leaves", "items", "iterating", "candidates", "pairs", "closure_candidates", "kernel_s",
                   "per_particle", "list_s", "compact_s", "list_entries", "max_list_log2",
                   "heavy_items", "split_items")

    def passes(self):
        """One row per NR pass of the last solve (columns PASS_FIELDS)."""
        return self.array("passes").reshape(-1, len(self.PASS_FIELDS))

    def stash_inputs(self):
        self.L.cosmo_ext_stash_inputs()

    def restore_inputs(self):
        self.L.cosmo_ext_restore_inputs()

    def sync(self):
        self.L.cosmo_ext_sync()

    # ---- ensemble ----
    def batch_init(self, counts):
        c = np.ascontiguousarray(counts, dtype=np.int32)
        self.L.cosmo_batch_init(int(c.size), c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return int(c.sum())

    def batch_solve(self, pmass, tolh=1e-4, hfact=-1.0, periodic_boxes=None):
        pm = np.ascontiguousarray(pmass, dtype=np.float64)
        bx = np.ascontiguousarray(periodic_boxes if periodic_boxes is not None else np.zeros(6), dtype=np.float64)
        return self.L.cosmo_batch_solve(_dptr(pm), int(periodic_boxes is not None), _dptr(bx), float(tolh),
                                        float(hfact))

    # ---- autotuner ----
    TUNE_OPTIONS = ("loop", "bucket", "mapping", "sparse", "tail_leaves", "loop_blocks_scale", "graph_passes", "lpt",
                    "cull", "units", "tile_depth")

    def autotune(self, n, pmass, tolh=1e-4, hfact=-1.0, periodic_box=None, budget_s=60.0):
        """Pick the solve's speed options for the uploaded set; returns the options chosen."""
        box = np.ascontiguousarray(periodic_box if periodic_box is not None else np.zeros(6), dtype=np.float64)
        self.L.cosmo_ext_autotune(int(n), float(pmass), int(periodic_box is not None), _dptr(box), float(tolh),
                                  float(hfact), float(budget_s))
        return self.tuned()

    def batch_autotune(self, pmass, tolh=1e-4, hfact=-1.0, periodic_boxes=None, budget_s=60.0):
        pm = np.ascontiguousarray(pmass, dtype=np.float64)
        bx = np.ascontiguousarray(periodic_boxes if periodic_boxes is not None else np.zeros(6), dtype=np.float64)
        self.L.cosmo_batch_autotune(_dptr(pm), int(periodic_boxes is not None), _dptr(bx), float(tolh), float(hfact),
                                    float(budget_s))
        return self.tuned()

    def tuned(self):
        out = {k: int(self.stat(k)) for k in self.TUNE_OPTIONS}
        for k in ("tune_trials", "tune_best_s", "tune_default_s", "tune_seconds", "tune_cached"):
            out[k] = self.stat(k)
        return out

    # ---- convenience: the Phantom call sequence of the reference driver (src/main.cu:118-141) ----
    def solve(self, x, y, z, h, pmass, tolh=1e-4, hfact=-1.0, v=None, a=None, gradients=True, periodic_box=None,
              mode=0):
        n = len(x)
        self.arrays_init(n)
        self.upload(POS, np.stack([x, y, z]))
        self.upload(HSML, h)
        self.upload(VEL, np.zeros((3, n)) if v is None else v)
        self.upload(ACCEL, np.zeros((3, n)) if a is None else a)
        if gradients and mode == 0:
            self.densityiterate(n, pmass, periodic_box, tolh, hfact)
        else:
            self.solve_dens_h(n, pmass, mode, gradients, periodic_box, tolh, hfact)
        out = {"h": self.download(HSML, n)[0]}
        d = self.download(DENS_OUT, n)
        out["rho"], out["gradh"] = d[0], d[1]
        if gradients:
            g = self.download(GRAD_OUT, n)
            out["divv"], out["xi"], out["ddivvdt"] = g[0], g[1], g[2]
        return out
