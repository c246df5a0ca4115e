"""The branch-free FP32 kernel form of the mixed pair loop (kernVD, common.hpp) against the
reference's piecewise form (kernWdW, include/kernel.hpp:35-78), both evaluated on the host from
the library's own header, for the cubic and quintic builds.

Pins: W = kWw V and -q dW/dq - 3W = kGp (q D) - kGw V to a few FP32 roundings of W(0) over
[0, R] (the inner piece has no cancellation); exact at q = 0 (the self pair every particle's sum
holds) and at the piece boundaries; no systematic bias (the mean error over the support is at
the level of single roundings)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "kernel_forms.cpp")


@pytest.fixture(scope="module", params=["cubic", "quintic"])
def forms(request, tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("forms") / ("kf_" + request.param))
    flags = ["-DCOSMO_KERNEL_QUINTIC"] if request.param == "quintic" else []
    subprocess.run([hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17", *flags, SRC, "-o", exe], check=True,
                   capture_output=True)
    radius = 3.0 if request.param == "quintic" else 2.0

    def run(q):
        out = subprocess.run([exe], input="\n".join("%.17g" % v for v in q), capture_output=True, text=True,
                             check=True).stdout
        return np.array([line.split() for line in out.strip().split("\n")], dtype=float).T

    return run, radius


def test_forms_agree_to_fp32(forms):
    run, radius = forms
    q = np.linspace(0.0, radius, 200001).astype(np.float32).astype(np.float64)
    w, g, wm, gm = run(q)
    # a few FP32 roundings of W(0): the inner piece is evaluated without cancellation
    assert np.max(np.abs(wm - w)) <= 3e-7 * w[0]
    assert np.max(np.abs(gm - g)) <= 3e-7 * np.abs(g).max()


def test_exact_at_self_pair_and_piece_boundaries(forms):
    run, radius = forms
    pts = np.arange(0.0, radius + 0.5, 1.0)
    w, g, wm, gm = run(pts)
    np.testing.assert_array_equal(wm, w)
    np.testing.assert_array_equal(gm, g)


def test_no_systematic_bias(forms):
    run, radius = forms
    # pair separations of a uniform neighbourhood: q^2 dq weighting over the support
    rng = np.random.default_rng(3)
    q = (radius * rng.random(200000) ** (1.0 / 3.0)).astype(np.float32).astype(np.float64)
    w, g, wm, gm = run(q)
    assert abs(np.sum(wm - w)) <= 1e-7 * np.sum(w)
    assert abs(np.sum(gm - g)) <= 1e-7 * np.sum(np.abs(g))
