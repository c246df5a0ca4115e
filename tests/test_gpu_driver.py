"""The standalone driver (reference src/main.cu) end to end on the GPU: the lattice input and a
Fortran-record dump of the same particles give the same h file, which equals the library's own
WARP_PER_LEAF solve (the driver's second call) through the C ABI and the FP64 oracle's."""
import os
import subprocess

import numpy as np
import pytest

from cosmosphere_amd import Library, workloads
from gpu_common import rel_err
from oracle import oracle as O
from test_driver_cpu import DRIVER, write_dump

pytestmark = pytest.mark.gpu


def run(args, tmp_path):
    res = subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
    assert res.returncode == 0, res.stderr[-2000:]
    return res.stdout


def test_driver_lattice_and_dump(tmp_path):
    npartx = 12
    out = run(["lattice", str(npartx), "h_lattice.txt"], tmp_path)
    assert "--- Call 1 (flat-particle) ---" in out and "--- Call 2 (warp-per-leaf) ---" in out
    h_lat = np.loadtxt(tmp_path / "h_lattice.txt")
    x, y, z, dx = workloads.make_lattice(npartx)
    n = x.size
    assert h_lat.size == n
    h0 = np.full(n, 1.2 * dx)
    write_dump(str(tmp_path / "lattice.dat"), x, y, z, h0, 1.0 / n)
    run(["lattice.dat", "h_dump.txt"], tmp_path)
    h_dump = np.loadtxt(tmp_path / "h_dump.txt")
    np.testing.assert_array_equal(h_dump, h_lat)
    # the library through ctypes, same call (WARP_PER_LEAF, no gradients)
    lib = Library("cubic")
    lib.set_option("precision", 0)
    got = lib.solve(x, y, z, h0, 1.0 / n, tolh=1e-4, hfact=-1.0, gradients=False, mode=1)
    assert rel_err(h_lat, got["h"]) < 2e-15            # printed with 16 significant digits
    # the oracle's WARP_PER_LEAF solve
    o = O.Oracle(quintic=False)
    o.set_particles(x, y, z, h0)
    o.build()
    assert o.density_solve(1.0 / n, 1e-4, 1.2, 100, 1, False) == 0
    h_ref = np.empty(n)
    h_ref[o.get("order")] = o.get("h")
    assert rel_err(h_lat, h_ref) < 1e-10


def test_driver_jiggle_and_mixed(tmp_path):
    out = run(["lattice", "10", "no_output", "jiggle", "mixed"], tmp_path)
    assert "Jiggled" in out
    assert not (tmp_path / "h_cpp.txt").exists()
    iters = [int(line.split("(")[1].split()[0]) for line in out.splitlines() if "Time density kernel" in line]
    assert len(iters) == 2 and min(iters) >= 2
