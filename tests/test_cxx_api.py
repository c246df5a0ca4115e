"""The C++ interface (include/cosmosphere.hpp): the reference's KernelMode, DensTimings,
solveDensH with the defaults of include/density.hpp:85-91, ForceTimings, gpuState and
computeForces (include/force.hpp:18-70).  A C++ caller (tests/native/cxx_api.cpp) is compiled
with the host compiler against the header and linked against libcosmoSPHere.a with the HIP
runtime, as Phantom links the archive; on the GPU it checks that solveDensH matches
cosmo_ext_solve_dens_h and computeForces matches force_gpu_c bit for bit."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cosmosphere_amd", "lib", "cxx_api_check")
ARCHIVE = os.path.join(ROOT, "cosmosphere_amd", "lib", "libcosmoSPHere.a")
SIGNATURES = [
    "solveDensH(int, double, KernelMode, bool, double const*, double, double)",
    "computeForces(GpuState&, double, double, double, bool, bool, bool, ForceTimings&)",
    "gpuState()",
]


@pytest.mark.parametrize("archive", ["libcosmoSPHere.a", "libcosmoSPHere_quintic.a"])
def test_archive_exports_the_reference_cxx_symbols(archive):
    out = subprocess.run(["nm", "-C", "--defined-only", os.path.join(ROOT, "cosmosphere_amd", "lib", archive)],
                         capture_output=True, text=True, check=True).stdout
    defined = set(line.split(" T ", 1)[1] for line in out.splitlines() if " T " in line)
    for sig in SIGNATURES:
        assert sig in defined, sig


def test_cxx_caller_links_with_the_host_compiler():
    assert os.path.exists(EXE), "built by __graft_entry__.build()"
    out = subprocess.run(["nm", "-C", "-u", EXE], capture_output=True, text=True, check=True).stdout
    assert "solveDensH" not in out and "computeForces" not in out   # resolved from the archive


@pytest.mark.gpu
@pytest.mark.parametrize("m", [12, 20])
def test_cxx_api_matches_c_entry_points(m):
    res = subprocess.run([EXE, str(m)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr[-1000:]
    lines = res.stdout.splitlines()
    assert "same=1" in lines[0] and ("n=%d" % m ** 3) in lines[0]
    assert "same=1" in lines[1]
