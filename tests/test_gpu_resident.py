"""Resident Hilbert order across steps and the staged transfers.

The reference names the resident order and leaves it unbuilt (include/gpu_state.hpp:17-23).
Here an input slot uploaded after a single-set solve of the same size lands in that solve's
order, and the next solve re-sorts the nearly sorted keys incrementally (csrc/resort.hip).
The result must be the full sort's bit for bit: keys, order, tree, lists, h and rho -- each
step is solved twice, once from the resident order and once from the caller's order with
the resident path off, and the two are compared.  The transfer tests push blocks larger
than the page-locked ring through both directions (src/arrays_c_api.cu:120-193 semantics:
caller order in, caller order out)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from cosmosphere_amd import Library, workloads
from cosmosphere_amd.api import ACCEL, DENS_OUT, FORCE_OUT, HSML, POS, THERMO, VEL
from gpu_common import DEFAULT_ITEMS, dataset, periodic_lattice

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COMPARED = ("keys", "order", "leafStart", "leafLevel", "layout", "particleLeaf", "h", "rho", "gradh", "jOffset", "jList")


def _solve(lib, c, x, y, z, h, v, resident, tree_only):
    n = c["n"]
    lib.set_option("resident", int(resident))
    lib.upload(POS, np.stack([x, y, z]))
    lib.upload(HSML, h)
    lib.upload(VEL, v)
    lib.upload(ACCEL, np.zeros((3, n)))
    lib.solve_dens_h(n, c["pmass"], with_gradients=not tree_only, tolh=1e-4, hfact=c["hfact"])
    out = {k: lib.array(k) for k in (COMPARED[:6] if tree_only else COMPARED)}
    out["movers"] = lib.stat("resort_movers")
    if not tree_only:
        out["h_caller"] = lib.download(HSML, n)[0]
        out["rho_caller"] = lib.download(DENS_OUT, n)[0]
    return out


@pytest.mark.parametrize("name,tree_only", [("plummer_quintic", False), ("lattice", False), ("dead", False),
                                            ("duplicates", True)])
def test_incremental_resort_equals_full_sort(name, tree_only):
    c = dataset(name)
    n = c["n"]
    lib = Library(c["kernel"])
    for k, val in DEFAULT_ITEMS.items():
        lib.set_option(k, val)
    lib.set_option("precision", 0)
    lib.set_option("max_iter", 100)
    lib.set_option("tree_only", int(tree_only))
    lib.arrays_init(n)
    rng = np.random.default_rng(17)
    v = rng.normal(size=(3, n))
    h = c["h"].copy()
    try:
        # drifts of increasing size: from a few particles leaving their leaf to most of them
        for step, dt in enumerate((0.0, 2e-4, 2e-3, 2e-2)):
            x, y, z = c["x"] + v[0] * dt, c["y"] + v[1] * dt, c["z"] + v[2] * dt
            res = _solve(lib, c, x, y, z, h, v, True, tree_only)
            full = _solve(lib, c, x, y, z, h, v, False, tree_only)
            assert full["movers"] == -1
            if step > 0:
                assert 0 <= res["movers"] <= n, res["movers"]
            for k in full:
                if k != "movers":
                    np.testing.assert_array_equal(res[k], full[k], err_msg="%s step %d %s" % (name, step, k))
            if not tree_only:
                h = full["h_caller"]
    finally:
        lib.set_option("tree_only", 0)
        lib.set_option("resident", 1)


def test_incremental_resort_periodic_wrap():
    """A periodic box: drifting particles leave through one face and come back through the
    opposite one, so their keys jump across the whole curve.  Each step is solved from the
    resident order and again from the caller's order with the resident path off, as above."""
    c, box = periodic_lattice()
    n = c["n"]
    lib = Library(c["kernel"])
    for k, val in DEFAULT_ITEMS.items():
        lib.set_option(k, val)
    lib.set_option("precision", 0)
    lib.set_option("max_iter", 100)
    lib.set_option("tree_only", 0)
    lib.arrays_init(n)
    rng = np.random.default_rng(23)
    v = rng.normal(size=(3, n))
    lo, span = box[0::2], box[1::2] - box[0::2]
    h = c["h"].copy()

    def solve(x, y, z, resident):
        lib.set_option("resident", int(resident))
        lib.upload(POS, np.stack([x, y, z]))
        lib.upload(HSML, h)
        lib.upload(VEL, v)
        lib.upload(ACCEL, np.zeros((3, n)))
        lib.solve_dens_h(n, c["pmass"], with_gradients=True, periodic_box=box, tolh=1e-4, hfact=1.2)
        out = {k: lib.array(k) for k in COMPARED}
        out["movers"] = lib.stat("resort_movers")
        out["h_caller"] = lib.download(HSML, n)[0]
        return out

    try:
        for step, dt in enumerate((0.0, 3e-3, 3e-2)):
            x, y, z = ((np.stack([c["x"], c["y"], c["z"]]) + v * dt - lo[:, None]) % span[:, None]) + lo[:, None]
            res = solve(x, y, z, True)
            full = solve(x, y, z, False)
            assert full["movers"] == -1
            if step > 0:
                assert 0 < res["movers"] <= n, res["movers"]
            for k in full:
                if k != "movers":
                    np.testing.assert_array_equal(res[k], full[k], err_msg="periodic step %d %s" % (step, k))
            h = full["h_caller"]
    finally:
        lib.set_option("resident", 1)


def test_resident_order_survives_caller_permutation_and_deaths():
    """The caller reorders its particles and kills / revives some between steps: the resident
    upload is then just another permutation, and the re-sort still equals the full sort."""
    c = dataset("plummer_cubic")
    n = c["n"]
    lib = Library("cubic")
    for k, val in DEFAULT_ITEMS.items():
        lib.set_option(k, val)
    lib.set_option("precision", 0)
    lib.set_option("tree_only", 0)
    lib.arrays_init(n)
    rng = np.random.default_rng(23)
    v = rng.normal(size=(3, n))
    x, y, z, h = c["x"].copy(), c["y"].copy(), c["z"].copy(), c["h"].copy()
    _solve(lib, c, x, y, z, h, v, False, False)
    perm = rng.permutation(n)
    x, y, z, h, v = x[perm] + 1e-4 * v[0][perm], y[perm], z[perm], h[perm].copy(), v[:, perm]
    h[rng.choice(n, 30, replace=False)] *= -1.0      # deaths
    res = _solve(lib, c, x, y, z, h, v, True, False)
    assert res["movers"] >= 30
    full = _solve(lib, c, x, y, z, h, v, False, False)
    for k in full:
        if k != "movers":
            np.testing.assert_array_equal(res[k], full[k], err_msg=k)
    lib.set_option("resident", 1)


def test_large_transfers_round_trip():
    """Blocks of 48-144 MB (several passes round the page-locked ring) in both directions, raw,
    gathered and scattered; a tree-only solve provides the order."""
    n = 3_000_000
    rng = np.random.default_rng(3)
    pos = rng.random((3, n))
    lib = Library("cubic")
    lib.set_option("tree_only", 1)
    lib.set_option("resident", 1)
    try:
        lib.arrays_init(n)
        up0 = lib.stat("upload_s")
        lib.upload(POS, pos)
        lib.upload(HSML, np.full(n, 0.01))
        lib.upload(VEL, np.zeros((3, n)))
        lib.upload(ACCEL, np.zeros((3, n)))
        lib.solve_dens_h(n, 1.0 / n, with_gradients=False, tolh=1e-4, hfact=1.2)
        order = lib.array("order")
        assert np.array_equal(np.sort(order), np.arange(n))
        np.testing.assert_array_equal(lib.download(POS, n), pos)
        th = rng.normal(size=(4, n))
        lib.upload_sorted(THERMO, th)
        np.testing.assert_array_equal(lib.download(THERMO, n), th)
        # a raw upload of an output slot lands in caller order, the download scatters it
        f = rng.normal(size=(6, n))
        lib.upload(FORCE_OUT, f)
        got = lib.download(FORCE_OUT, n)
        np.testing.assert_array_equal(got[:, order], f)
        # a resident upload of positions comes back in the caller's order
        pos2 = rng.random((3, n))
        lib.upload(POS, pos2)
        np.testing.assert_array_equal(lib.download(POS, n), pos2)
        assert lib.stat("upload_s") > up0 and lib.stat("download_s") > 0
    finally:
        lib.set_option("tree_only", 0)


def test_stats_lines_report_measured_transfers():
    code = ("import sys; sys.path.insert(0, %r)\nimport numpy as np\nfrom cosmosphere_amd import Library, workloads\n"
            "c = workloads.config2(npartx=24)\nlib = Library('cubic')\n"
            "out = lib.solve(c['x'], c['y'], c['z'], c['h'], c['pmass'], hfact=1.2)\n"
            "lib.upload_sorted(7, np.ones((4, c['n'])))\nlib.force(c['n'], c['pmass'])\n" % ROOT)
    env = dict(os.environ, COSMO_DENS_STATS="1", COSMO_MEM="1")
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=180, env=env)
    assert res.returncode == 0, res.stderr[-400:]
    lines = res.stderr.splitlines()
    stats = [l for l in lines if l.startswith("COSMO_STATS")][0]
    force = [l for l in lines if l.startswith("COSMO_FORCE")][0]
    assert any(l.startswith("COSMO_MEM n=") and "host_peak=" in l for l in lines)

    def field(line, key):
        return float(line.split(key + "=")[1].split()[0])

    assert field(stats, "upload") > 0.0          # the four input slots before the solve
    assert field(force, "upload") > 0.0          # THERMO
    assert field(force, "download") > 0.0        # h, rho, gradients after the solve
