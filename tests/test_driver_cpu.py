"""The standalone driver's dump reader (reference include/io.hpp:1-140, src/main.cu:143-171) on the
CPU: a Fortran unformatted sequential file -- records ngas (int32), pmass (float64), xyzh(4, ngas)
(float64, particle-interleaved), each framed by its byte length before and after -- written here
with numpy is read back exactly, and a file whose markers or length are wrong is refused with
exit status 1.  (--check-read touches no device.)"""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cosmosphere_amd", "lib", "density_hip")


def write_dump(path, x, y, z, h, pmass, corrupt=None):
    n = len(x)
    xyzh = np.stack([x, y, z, h], axis=1).astype("<f8").tobytes()   # x1 y1 z1 h1 x2 ...
    recs = [struct.pack("<i", n), struct.pack("<d", pmass), xyzh]
    with open(path, "wb") as f:
        for k, payload in enumerate(recs):
            head = tail = len(payload)
            if corrupt == "head" and k == 2:
                head += 8
            if corrupt == "tail" and k == 1:
                tail += 1
            f.write(struct.pack("<i", head))
            f.write(payload if not (corrupt == "short" and k == 2) else payload[:-16])
            if not (corrupt == "short" and k == 2):
                f.write(struct.pack("<i", tail))


def check_read(path):
    return subprocess.run([DRIVER, "--check-read", path], capture_output=True, text=True, timeout=60)


@pytest.fixture(scope="module")
def particles():
    rng = np.random.default_rng(5)
    n = 70001          # more than one read chunk
    return rng.random(n), rng.random(n), rng.random(n), 0.01 + 0.01 * rng.random(n), 1.0 / n


def test_dump_round_trip(tmp_path, particles):
    x, y, z, h, pmass = particles
    p = str(tmp_path / "dump.dat")
    write_dump(p, x, y, z, h, pmass)
    res = check_read(p)
    assert res.returncode == 0, res.stderr
    f = res.stdout.split()
    assert int(f[1]) == len(x)
    assert float(f[3]) == pmass
    sums = [float(v) for v in f[5:9]]
    first = [float(v) for v in f[10:14]]
    for got, want in zip(sums, (x.sum(), y.sum(), z.sum(), h.sum())):
        assert abs(got - want) <= 1e-12 * abs(want)     # summation order may differ
    assert first == [x[0], y[0], z[0], h[0]]            # %.17g round-trips a double


@pytest.mark.parametrize("corrupt,message", [("head", "xyzh record holds"), ("tail", "end marker"),
                                             ("short", "truncated")])
def test_dump_refused(tmp_path, particles, corrupt, message):
    x, y, z, h, pmass = particles
    p = str(tmp_path / ("bad_%s.dat" % corrupt))
    write_dump(p, x[:100], y[:100], z[:100], h[:100], pmass, corrupt=corrupt)
    res = check_read(p)
    assert res.returncode == 1
    assert "Error reading file" in res.stderr and message in res.stderr


def test_missing_file():
    res = check_read("/nonexistent/dump.dat")
    assert res.returncode == 1 and "cannot open" in res.stderr
