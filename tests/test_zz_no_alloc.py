"""cosmo_download_noalloc (no counterpart in the reference): the download that allocates nothing,
scattering through the library's static device buffer, against cosmo_download on every slot,
on two solves in a row, the second with fewer particles; sizes beyond the buffer are refused
before any device work."""
import ctypes

import numpy as np
import pytest

from cosmosphere_amd import Library, workloads
from cosmosphere_amd.api import ACCEL, DENS_OUT, FORCE_OUT, GRAD_OUT, HSML, POS, THERMO, VEL

pytestmark = pytest.mark.gpu
MAX = 4194304


def test_noalloc_download_equals_download():
    lib = Library("cubic")
    lib.set_option("precision", 0)
    lib.set_option("tree_only", 0)
    rng = np.random.default_rng(41)
    for npartx in (24, 14):          # the second call has fewer particles
        c = workloads.config2(npartx=npartx)
        n = c["n"]
        v = rng.normal(size=(3, n))
        lib.solve(c["x"], c["y"], c["z"], c["h"], c["pmass"], hfact=1.2, v=v)
        lib.upload_sorted(THERMO, np.stack([np.full(n, 0.5), np.ones(n), np.ones(n), np.ones(n)]))
        lib.force(n, c["pmass"])
        for slot in (POS, HSML, VEL, ACCEL, DENS_OUT, GRAD_OUT, THERMO, FORCE_OUT):
            ref = lib.download(slot, n)
            got = lib.download_noalloc(slot, n)
            np.testing.assert_array_equal(got, ref, err_msg="n=%d slot=%d" % (n, slot))
        np.testing.assert_array_equal(lib.download_noalloc(POS, n), np.stack([c["x"], c["y"], c["z"]]))


def test_noalloc_refuses_beyond_its_buffer():
    lib = Library("cubic")
    out = np.empty(1)
    # refused before the size, slot or ordering checks, so nothing else is touched
    assert lib.L.cosmo_download_noalloc(2, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), MAX + 1) == -1
