"""GPU parity of the force pass (src/force.cu:19-455) through the C ABI, Phantom's call
sequence: densityiterate -> cosmo_upload_sorted(THERMO) -> force_gpu_c -> cosmo_download.

Bit-exact: the symmetric neighbour-leaf lists.  FP64: accelerations, du/dt and div v within
1e-9 of the largest component (summation order and fused multiply-add differ), vsigmax
within 1e-12 relative.  Mixed: within 2e-6 of the largest component, with each tile's accepted
pairs shared over the lanes (the default) and with lanes that keep their own slots.  The oracle's own invariants (momentum / energy conservation) are
re-checked on the GPU output.
"""
import numpy as np
import pytest

from cosmosphere_amd import Library, workloads
from cosmosphere_amd.api import FORCE_OUT, HSML, THERMO
from gpu_common import periodic_lattice
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def setup(n=20000, seed=12, quintic=True, uniform_u=False):
    c = workloads.config3(n=n, seed=seed)
    if not quintic:
        c.update(kernel="cubic", quintic=False, hfact=1.2, h=np.full(n, 1.2 * n ** (-1.0 / 3.0)))
    idx = np.arange(n)
    v = np.stack([0.1 * workloads.gaussian(seed, 30 + k, idx) for k in range(3)])
    v[0] -= c["y"] - 0.5
    v[1] += c["x"] - 0.5
    u = np.ones(n) if uniform_u else 1.0 + 0.5 * workloads.uniform(seed, 40, idx)
    return c, v, u


def thermo(c, h, u, gamma=5.0 / 3.0):
    rhoh = c["pmass"] * (c["hfact"] / h) ** 3
    n = c["n"]
    return np.stack([(gamma - 1) * u / rhoh, np.sqrt(gamma * (gamma - 1) * u), np.ones(n), u])


@pytest.fixture
def _default_walk():
    yield
    for kernel in ("cubic", "quintic"):
        lib = Library(kernel)
        lib.set_option("list_heavy_steps", 0)
        lib.set_option("list_expand_leaves", -1)
        lib.set_option("sweep_split", 512)
        lib.set_option("force_rounds", 1)


@pytest.mark.parametrize("variant", [dict(), dict(disc=1), dict(pdv=0), dict(shock=0, quintic=False),
                                     dict(block=1), dict(sweep=4)])
def test_force_parity(variant, _default_walk):
    quintic = variant.get("quintic", True)
    c, v, u = setup(quintic=quintic)
    n = c["n"]
    disc, pdv, shock = variant.get("disc", 0), variant.get("pdv", 1), variant.get("shock", 1)
    lib = Library(c["kernel"])
    lib.set_option("precision", 0)
    lib.set_option("max_iter", 100)
    lib.set_option("tree_only", 0)
    if variant.get("block"):   # every list (gather and symmetric) through the block walk
        lib.set_option("list_heavy_steps", 1)
        lib.set_option("list_expand_leaves", 0)
    # leaves with more than `sweep` list entries: one wavefront per particle in the sweeps
    lib.set_option("sweep_split", variant.get("sweep", 512))
    out = lib.solve(c["x"], c["y"], c["z"], c["h"], c["pmass"], hfact=c["hfact"], v=v)
    th = thermo(c, out["h"], u)
    lib.upload_sorted(THERMO, th)
    lib.force(n, c["pmass"], beta=2.0, alphau=1.0, disc_viscosity=disc, pdv_heating=pdv, shock_heating=shock)
    f = lib.download(FORCE_OUT, n)
    if variant.get("sweep"):
        assert lib.stat("sweep_items") > n // 2   # most particles went through per-particle items

    o = O.Oracle(quintic=c["quintic"])
    o.set_particles(c["x"], c["y"], c["z"], c["h"], v=v, a=np.zeros((3, n)))
    o.build()
    assert o.density_solve(c["pmass"], 1e-4, c["hfact"], 100, 0, True) == 0
    o.set_thermo(*th)
    o.force(c["pmass"], c["hfact"], beta=2.0, alphau=1.0, disc=disc, pdv=pdv, shock=shock)
    order = o.get("order")

    def host(name):
        a = np.empty(n)
        a[order] = o.get(name)
        return a

    # symmetric lists, bit-exact (built from the GPU's converged h on both sides)
    o.set_h_sorted(lib.array("h"))
    o.hmax_all()
    o.build_lists(symmetric=True)
    np.testing.assert_array_equal(lib.array("jCount"), o.get("jCount"))
    np.testing.assert_array_equal(lib.array("jList"), o.get("jList"))

    fmax = np.abs(np.stack([host("fx"), host("fy"), host("fz")])).max()
    for k, name in enumerate(("fx", "fy", "fz")):
        assert np.max(np.abs(f[k] - host(name))) < 1e-9 * fmax, name
    du = host("du")
    assert np.max(np.abs(f[3] - du)) < 1e-9 * np.abs(du).max()
    vs = host("vsig")
    assert np.max(np.abs(f[4] - vs) / np.maximum(vs, 1e-300)) < 1e-12
    dv = host("divvF")
    assert np.max(np.abs(f[5] - dv)) < 1e-9 * np.abs(dv).max()
    # invariants on the GPU output
    m = c["pmass"]
    assert abs(f[0].sum()) < 1e-11 * np.abs(f[0]).sum()
    if pdv and shock and not disc:   # disc viscosity heats with its own form: not conservative
        power = v[0] * f[0] + v[1] * f[1] + v[2] * f[2] + f[3]
        assert abs((m * power).sum()) < 1e-10 * np.abs(m * power).sum()


def test_force_repeat_reuses_lists():
    """A second force call on the same tree (the leapfrog corrector) gives the same output."""
    c, v, u = setup(n=8000, seed=3)
    n = c["n"]
    lib = Library(c["kernel"])
    lib.set_option("precision", 0)
    out = lib.solve(c["x"], c["y"], c["z"], c["h"], c["pmass"], hfact=c["hfact"], v=v)
    lib.upload_sorted(THERMO, thermo(c, out["h"], u))
    lib.force(n, c["pmass"])
    f1 = lib.download(FORCE_OUT, n)
    lib.force(n, c["pmass"])
    f2 = lib.download(FORCE_OUT, n)
    np.testing.assert_array_equal(f1, f2)
    np.testing.assert_array_equal(lib.download(HSML, n)[0], out["h"])


def test_force_after_hsml_upload_rebuilds_lists():
    """New smoothing lengths uploaded between the solve and the force pass: the force pass builds
    its symmetric lists again from them (bit-exact against the oracle's at the new h) instead of
    reusing the lists of the solve's gradient sweep.  With count_list on, the force pass also
    counts its list candidates (every particle against every particle of its leaf's list leaves),
    which the oracle's lists fix exactly."""
    c, v, u = setup(n=8000, seed=4)
    n = c["n"]
    lib = Library(c["kernel"])
    lib.set_option("precision", 0)
    out = lib.solve(c["x"], c["y"], c["z"], c["h"], c["pmass"], hfact=c["hfact"], v=v)
    h2 = out["h"] * (1.0 + 0.2 * workloads.uniform(4, 50, np.arange(n)))
    lib.upload(HSML, h2)
    lib.upload_sorted(THERMO, thermo(c, h2, u))
    lib.set_option("count_list", 1)
    try:
        lib.force(n, c["pmass"])
    finally:
        lib.set_option("count_list", 0)
    o = O.Oracle(quintic=c["quintic"])
    o.set_particles(c["x"], c["y"], c["z"], c["h"], v=v, a=np.zeros((3, n)))
    o.build()
    order = o.get("order")
    o.set_h_sorted(h2[order])
    o.hmax_all()
    o.build_lists(symmetric=True)
    np.testing.assert_array_equal(lib.array("jList"), o.get("jList"))
    jo = o.get("jOffset").astype(np.int64)
    jl = o.get("jList").astype(np.int64)
    cnt = np.diff(o.get("layout").astype(np.int64))
    cs = np.concatenate([[0], np.cumsum(cnt[jl])])
    ref = int((cnt[:len(jo) - 1] * (cs[jo[1:]] - cs[jo[:-1]])).sum())
    print("force list candidates", lib.stat("force_list_candidates"), ref, "kernel", lib.stat("force_candidates"),
          flush=True)
    assert lib.stat("force_list_candidates") == ref
    assert 0 < lib.stat("force_candidates") <= ref


@pytest.mark.parametrize("variant", [dict(), dict(disc=1), dict(shock=0, quintic=False), dict(sweep=4),
                                     dict(rounds=0), dict(rounds=0, sweep=4)])
def test_force_mixed_vs_oracle(variant):
    """Mixed form of the force pass (FP32 pair algebra from FP64 differences, FP64 sums) against
    the FP64 oracle at the same h: accelerations, du/dt and div v within 2e-6 of their largest
    magnitude, vsigmax within 2e-6 relative; momentum and energy conserved to FP32 rounding.
    Both mappings: the tile's accepted pairs shared over the lanes (the default) and lanes that
    keep their own slots."""
    quintic = variant.get("quintic", True)
    c, v, u = setup(quintic=quintic)
    n = c["n"]
    disc, pdv, shock = variant.get("disc", 0), variant.get("pdv", 1), variant.get("shock", 1)
    lib = Library(c["kernel"])
    lib.set_option("precision", 1)
    lib.set_option("max_iter", 100)
    lib.set_option("tree_only", 0)
    lib.set_option("sweep_split", variant.get("sweep", 512))
    lib.set_option("force_rounds", variant.get("rounds", 1))
    try:
        out = lib.solve(c["x"], c["y"], c["z"], c["h"], c["pmass"], hfact=c["hfact"], v=v)
        th = thermo(c, out["h"], u)
        lib.upload_sorted(THERMO, th)
        lib.force(n, c["pmass"], beta=2.0, alphau=1.0, disc_viscosity=disc, pdv_heating=pdv, shock_heating=shock)
        f = lib.download(FORCE_OUT, n)
        if variant.get("sweep"):
            assert lib.stat("sweep_items") > n // 2
        h_sorted = lib.array("h")
        rho_sorted = lib.array("rho")
        gradh_sorted = lib.array("gradh")
    finally:
        lib.set_option("precision", 0)
        lib.set_option("sweep_split", 512)
        lib.set_option("force_rounds", 1)
    # the oracle at the GPU's own h and d(rho)/d(h), so only the force arithmetic differs
    o = O.Oracle(quintic=c["quintic"])
    o.set_particles(c["x"], c["y"], c["z"], c["h"], v=v, a=np.zeros((3, n)))
    o.build()
    assert o.density_solve(c["pmass"], 1e-4, c["hfact"], 100, 0, True) == 0
    o.set_h_sorted(h_sorted)
    o.closure(c["pmass"])
    assert np.max(np.abs(o.get("rho") - rho_sorted) / rho_sorted) < 1e-6
    o.set_thermo(*th)
    o.force(c["pmass"], c["hfact"], beta=2.0, alphau=1.0, disc=disc, pdv=pdv, shock=shock)
    order = o.get("order")

    def host(name):
        a = np.empty(n)
        a[order] = o.get(name)
        return a

    fmax = np.abs(np.stack([host("fx"), host("fy"), host("fz")])).max()
    errs = {}
    for k, name in enumerate(("fx", "fy", "fz")):
        errs[name] = np.max(np.abs(f[k] - host(name))) / fmax
    du = host("du")
    errs["du"] = np.max(np.abs(f[3] - du)) / np.abs(du).max()
    vs = host("vsig")
    errs["vsig"] = np.max(np.abs(f[4] - vs) / np.maximum(vs, 1e-300))
    dv = host("divvF")
    errs["divv"] = np.max(np.abs(f[5] - dv)) / np.abs(dv).max()
    print("mixed force errors", errs, flush=True)
    for k, e in errs.items():
        assert e < 2e-6, (k, e)
    m = c["pmass"]
    assert abs(f[0].sum()) < 1e-6 * np.abs(f[0]).sum()
    if pdv and shock and not disc:
        power = v[0] * f[0] + v[1] * f[1] + v[2] * f[2] + f[3]
        assert abs((m * power).sum()) < 1e-6 * np.abs(m * power).sum()


@pytest.mark.parametrize("precision,sweep", [(0, 512), (0, 4), (1, 512)])
def test_force_periodic(precision, sweep):
    """The force pass in a periodic box (nearest images in the pair terms and in the culling of
    list entries and of particles; with sweep=4 every leaf through per-particle items) against the
    FP64 oracle at the GPU's own h: FP64 within 1e-9 of the largest component, mixed within 2e-6."""
    c, box = periodic_lattice()
    n = c["n"]
    rng = np.random.default_rng(7)
    v = rng.normal(size=(3, n)) * 0.1
    u = 1.0 + 0.5 * rng.random(n)
    lib = Library("cubic")
    lib.set_option("precision", precision)
    lib.set_option("max_iter", 100)
    lib.set_option("tree_only", 0)
    lib.set_option("sweep_split", sweep)
    try:
        out = lib.solve(c["x"], c["y"], c["z"], c["h"], c["pmass"], hfact=1.2, v=v, periodic_box=box)
        th = thermo(c, out["h"], u)
        lib.upload_sorted(THERMO, th)
        lib.force(n, c["pmass"], beta=2.0, alphau=1.0, disc_viscosity=0, pdv_heating=1, shock_heating=1)
        f = lib.download(FORCE_OUT, n)
        if sweep < 512:
            assert lib.stat("sweep_items") > n // 2
        h_sorted = lib.array("h")
    finally:
        lib.set_option("precision", 0)
        lib.set_option("sweep_split", 512)
    o = O.Oracle(quintic=False)
    o.set_particles(c["x"], c["y"], c["z"], c["h"], v=v, a=np.zeros((3, n)))
    o.build(periodic_box=box)
    assert o.density_solve(c["pmass"], 1e-4, 1.2, 100, 0, True) == 0
    o.set_h_sorted(h_sorted)
    o.closure(c["pmass"])
    o.set_thermo(*th)
    o.force(c["pmass"], 1.2, beta=2.0, alphau=1.0, disc=0, pdv=1, shock=1)
    order = o.get("order")

    def host(name):
        a = np.empty(n)
        a[order] = o.get(name)
        return a

    tol = 1e-9 if precision == 0 else 2e-6
    fmax = np.abs(np.stack([host("fx"), host("fy"), host("fz")])).max()
    assert fmax > 0
    for k, name in enumerate(("fx", "fy", "fz")):
        assert np.max(np.abs(f[k] - host(name))) < tol * fmax, name
    du = host("du")
    assert np.max(np.abs(f[3] - du)) < tol * np.abs(du).max()
    dv = host("divvF")
    assert np.max(np.abs(f[5] - dv)) < tol * np.abs(dv).max()
    vs = host("vsig")
    assert np.max(np.abs(f[4] - vs) / np.maximum(vs, 1e-300)) < (1e-12 if precision == 0 else 2e-6)
