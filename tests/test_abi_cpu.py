"""The C-ABI library without a GPU: it loads, exports every symbol include/cosmosphere.h
declares (both kernel variants, .so and .a), answers the device-free queries, and aborts
with the reference's FATAL messages on misuse before touching a device
(src/dens_c_api.cu:60-67, src/force_c_api.cu:40-60, src/arrays_c_api.cu:54-82)."""
import os
import re
import subprocess
import sys

import pytest

from cosmosphere_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cosmosphere.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(?:void|int|double|long long)\s+(cosmo_\w+|densityiterate_gpu_c|force_gpu_c)\s*\(",
                                 text)))


def test_header_lists_the_reference_entry_points():
    fns = header_functions()
    for f in api.ABI_SYMBOLS:
        assert f in fns
    assert set(fns) == set(api.ABI_SYMBOLS) | set(api.EXT_SYMBOLS)


@pytest.mark.parametrize("kernel", ["cubic", "quintic"])
def test_shared_library_exports(kernel):
    out = subprocess.run(["nm", "-D", "--defined-only", api.library_path(kernel)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    for f in header_functions():
        assert f in exported, f


@pytest.mark.parametrize("kernel", ["cubic", "quintic"])
def test_static_archive_exports(kernel):
    lib = api.library_path(kernel)[:-3] + ".a"
    out = subprocess.run(["nm", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if " T " in line)
    for f in api.ABI_SYMBOLS:
        assert f in exported, f


def test_device_free_queries():
    cubic, quintic = api.Library("cubic"), api.Library("quintic")
    assert cubic.kernel_radius() == 2.0 and quintic.kernel_radius() == 3.0
    assert [cubic.slot_ncomp(s) for s in range(0, 10)] == [0, 3, 1, 3, 3, 2, 3, 4, 6, 0]


def _run(snippet):
    code = "import sys; sys.path.insert(0, %r)\nfrom cosmosphere_amd import api\nimport numpy as np\n" % ROOT + snippet
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("snippet,message", [
    ("api.Library('cubic').densityiterate(5, 0.2)",
     "FATAL: densityiterate_gpu_c called with n=5 but the device arrays are sized for 0"),
    ("api.Library('cubic').force(7, 0.1)",
     "FATAL: force_gpu_c called with n=7 but the device arrays are sized for 0"),
    ("api.Library('quintic').upload(1, np.zeros((3, 4)))",
     "FATAL: cosmo_upload(slot=1) with n=4 but the device arrays are sized for 0"),
    ("api.Library('cubic').download(5, 4)",
     "FATAL: cosmo_download(slot=5) with n=4 but the device arrays are sized for 0"),
    ("api.Library('cubic').solve_dens_h(3, 0.3)",
     "FATAL: cosmo_ext_solve_dens_h called with n=3 but the device arrays are sized for 0"),
])
def test_fatal_aborts_before_device_work(snippet, message):
    res = _run(snippet)
    assert res.returncode == -6, (res.returncode, res.stderr[-500:])
    assert message in res.stderr


def test_pin_is_best_effort():
    import numpy as np
    lib = api.Library("cubic")
    a = np.zeros(1000)
    lib.pin_host(a)
    lib.unpin_host(a)
