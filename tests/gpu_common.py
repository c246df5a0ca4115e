"""Shared helpers of the GPU parity tests: run one set through the library (C ABI) and the
FP64 oracle on identical inputs, and compare stage by stage."""
import numpy as np

from cosmosphere_amd import Library, workloads
from oracle import oracle as O


def periodic_lattice(npartx=12):
    """Jittered close-packed lattice filling a periodic box (cubic kernel, hfact 1.2)."""
    x, y, z, dx = workloads.make_lattice(npartx)
    dy = dx * np.sqrt(3.0) / 2.0
    dz = dx * np.sqrt(6.0) / 3.0
    ny = int(round((y.max() - (y.max() % dy)) / dy)) + 1
    nz = int(round(z.max() / dz)) + 1
    box = np.array([0.0, 1.0, 0.0, ny * dy, 0.0, nz * dz])
    rng = np.random.default_rng(5)
    x = (x + 0.1 * dx * rng.normal(size=x.size)) % 1.0
    y = (y + 0.1 * dx * rng.normal(size=x.size)) % box[3]
    z = (z + 0.1 * dx * rng.normal(size=x.size)) % box[5]
    n = x.size
    vol = box[1] * box[3] * box[5]
    c = workloads._pack(x, y, z, np.full(n, 1.2 * (vol / n) ** (1 / 3)), "cubic", "periodic lattice")
    c["pmass"] = 1.0 / n
    return c, box


def dataset(name):
    """Small versions of the benchmark inputs plus edge cases."""
    if name == "lattice":
        return workloads.config2(npartx=12)
    if name == "C1small":
        c = workloads.config1(npartx=20)
        return c
    if name == "plummer_cubic":
        c = workloads.config3(n=20000, seed=11)
        c.update(kernel="cubic", quintic=False, hfact=1.2)
        c["h"] = np.full(c["n"], 1.2 * c["n"] ** (-1.0 / 3.0))
        return c
    if name == "plummer_quintic":
        return workloads.config3(n=20000, seed=12)
    if name == "dead":
        c = workloads.config2(npartx=10)
        h = c["h"].copy()
        h[::17] = -h[::17]          # dead particles (h <= 0) are parked after the live ones
        c["h"] = h
        return c
    if name == "duplicates":
        # tree stages only: > 64 coincident particles have no Newton fixed point
        c = workloads.config2(npartx=10)
        x, y, z = c["x"].copy(), c["y"].copy(), c["z"].copy()
        x[:100] = 0.5; y[:100] = 0.5; z[:100] = 0.5   # > 64 keys in one level-21 cell
        c.update(x=x, y=y, z=z)
        return c
    if name == "tiny":
        c = workloads.config2(npartx=4)
        return c
    raise KeyError(name)


# tile_depth 64 (the library's default is 128): the device loop's 64-slot tiles sum each particle's
# pairs in the order of the host-driven loop, so the two are compared bit for bit
DEFAULT_ITEMS = {"mapping": 0, "sparse": 16, "closure": -1, "loop": 1, "graph": 1, "graph_passes": 4, "lpt": 1,
                 "loop_list_cap": 0, "cull": 1, "units": 1, "list_heavy_steps": 0, "list_expand_leaves": -1,
                 "split_len": -1, "sweep_split": 512, "tile_depth": 64, "force_rounds": 1,
                 "list_heavy_factor": 8, "grad_sym_lists": 1, "count_list": 0}
# every list through the block walk, every node expanded while the token buffer allows
FORCE_BLOCK_WALK = {"list_heavy_steps": 1, "list_expand_leaves": 0}


def gpu_run(c, precision=0, gradients=True, mode=0, max_iter=100, v=None, a=None, tree_only=False, **options):
    """Solve set c through the C ABI; `options` override the work-item and closure defaults."""
    lib = Library(c["kernel"])
    lib.set_option("precision", precision)
    lib.set_option("max_iter", max_iter)
    lib.set_option("tree_only", int(tree_only))
    for k, val in dict(DEFAULT_ITEMS, **options).items():
        lib.set_option(k, val)
    n = c["n"]
    lib.arrays_init(n)
    lib.upload(1, np.stack([c["x"], c["y"], c["z"]]))
    lib.upload(2, c["h"])
    lib.upload(3, np.zeros((3, n)) if v is None else v)
    lib.upload(4, np.zeros((3, n)) if a is None else a)
    iters = lib.solve_dens_h(n, c["pmass"], mode=mode, with_gradients=gradients, tolh=1e-4, hfact=c["hfact"])
    return lib, iters


def oracle_run(c, gradients=True, mode=0, max_iter=100, v=None, a=None):
    n = c["n"]
    o = O.Oracle(quintic=c["quintic"])
    o.set_particles(c["x"], c["y"], c["z"], c["h"], v=np.zeros((3, n)) if v is None else v,
                    a=np.zeros((3, n)) if a is None else a)
    o.build()
    unconv = o.density_solve(c["pmass"], 1e-4, c["hfact"], max_iter, mode, gradients)
    return o, unconv


def rel_err(a, b, floor=0.0):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(b), floor)
    den = np.where(den == 0, 1.0, den)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0
