"""GPU parity of the device-driven Newton-Raphson loop (loop.hip) against the host-driven one.

Both loops run the same passes with the same work items, lists and sums; the device-driven
loop only changes who decides (the control block instead of the host) and the order in which
wavefronts take items (a cost-ordered queue).  So h, rho and d(rho)/d(h) must be bit-identical,
the iteration count equal and the pair counters equal, whether the passes are replayed as a
HIP graph or launched directly, K passes per round trip, with or without the cost order, and
when the list buffer overflows mid-loop (the pass stops, the buffer grows, the graph is
captured again).  The host-driven loop is itself pinned to the FP64 oracle by
test_gpu_parity.py and test_gpu_items.py; one oracle check is repeated here.  Heavy items split
over several wavefronts (device loop only) add their partial sums in sub-item order, which
moves the last bits: those solves are compared with the oracle instead (test_split_items)."""
import numpy as np
import pytest

from cosmosphere_amd import Library
from gpu_common import DEFAULT_ITEMS, dataset, gpu_run, oracle_run, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _reset_options():
    yield
    for kernel in ("cubic", "quintic"):
        lib = Library(kernel)
        for k, v in DEFAULT_ITEMS.items():
            lib.set_option(k, v)


def solve(c, precision, **opts):
    lib, iters = gpu_run(c, precision=precision, gradients=False, **opts)
    out = {k: lib.array(k).copy() for k in ("h", "rho", "gradh")}
    out["iters"] = iters
    for k in ("pairs", "candidates", "closure_pairs", "graph_captures", "graph_replays"):
        out[k] = lib.stat(k)
    return out


def same(a, b):
    for k in ("h", "rho", "gradh"):
        assert np.array_equal(a[k], b[k]), k
    assert a["iters"] == b["iters"]
    assert a["pairs"] == b["pairs"]
    assert a["closure_pairs"] == b["closure_pairs"]


VARIANTS = [("graph_k4", {}), ("graph_k1", {"graph_passes": 1}), ("graph_k3", {"graph_passes": 3}),
            ("direct", {"graph": 0}), ("no_lpt", {"lpt": 0}), ("overflow", {"loop_list_cap": 64}),
            ("flat", {"mapping": 2}), ("whole", {"mapping": 1}), ("nocull", {"cull": 0}),
            ("nounits", {"units": 0}), ("block_walk", {"list_heavy_steps": 1, "list_expand_leaves": 0}),
            ("launch_rule", {"list_heavy_factor": 0}), ("early_handover", {"list_heavy_factor": 1})]


@pytest.mark.parametrize("label,opts", VARIANTS, ids=[v[0] for v in VARIANTS])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", ["plummer_cubic", "plummer_quintic", "dead", "C1small"])
def test_device_loop_bit_identical(name, precision, label, opts):
    c = dataset(name)
    print("case", name, precision, label, flush=True)
    base_opts = {k: v for k, v in opts.items() if k in ("mapping", "cull", "units")}
    if label == "block_walk":
        base_opts = {}   # the block walk's lists are the group walk's: the host loop keeps the default
    host = solve(c, precision, loop=0, **base_opts)
    # heavy items split over wavefronts add partial sums in another order (test_split_items)
    dev = solve(c, precision, loop=1, **dict({"split_len": 0}, **opts))
    same(dev, host)
    assert dev["graph_replays"] >= 1
    if label == "overflow":
        assert dev["graph_captures"] >= 1


@pytest.mark.parametrize("mapping", [0, 2])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", ["plummer_cubic", "plummer_quintic"])
def test_split_items(name, precision, mapping):
    """Heavy items split over wavefronts (here every heavy item with a list of more than 8
    entries): the sub-items' partial sums are added in sub-item order, so the solve is the
    oracle's to summation order: FP64 within 1e-10 and the same iterations and accepted pairs,
    mixed within 1e-6."""
    c = dataset(name)
    print("case split", name, precision, mapping, flush=True)
    lib, iters = gpu_run(c, precision=precision, gradients=False, mapping=mapping, split_len=8)
    assert lib.passes()[:, 13].sum() > 0   # items were split
    o, unconv = oracle_run(c, gradients=False)
    assert unconv == 0
    na = o.geti("nAlive")
    if precision == 0:
        assert iters == o.geti("itersRun")
        assert rel_err(lib.array("h")[:na], o.get("h")[:na]) < 1e-10
        assert rel_err(lib.array("rho")[:na], o.get("rho")[:na]) < 1e-10
        pairs = sum(o.get("pairsPerIter"))
        assert abs(lib.stat("pairs") - pairs) <= 1e-6 * pairs + 2
    else:
        assert rel_err(lib.array("h")[:na], o.get("h")[:na]) < 1e-6
        o.closure(c["pmass"])
        assert rel_err(lib.array("rho")[:na], o.get("rho")[:na]) < 1e-6
    # the host-driven loop does not split; the device loop without splits matches it bit for bit
    host, _ = gpu_run(c, precision=precision, gradients=False, mapping=mapping, loop=0)
    h_host = host.array("h").copy()
    dev, _ = gpu_run(c, precision=precision, gradients=False, mapping=mapping, split_len=0)
    assert np.array_equal(dev.array("h"), h_host)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", ["plummer_quintic", "C1small", "dead"])
def test_tile_depths(name, precision):
    """64- and 128-slot j-tiles (an option of the device-driven loop; 128 is the default) sum
    each particle's pairs in other lane orders: the same pairs (counters equal), h and rho
    against the oracle (FP64 1e-10, mixed 1e-6), the same iteration count."""
    c = dataset(name)
    print("case tile depth", name, precision, flush=True)
    shallow = solve(c, precision, loop=1, tile_depth=64)
    deep = solve(c, precision, loop=1, tile_depth=128)
    o, unconv = oracle_run(c, gradients=False, max_iter=c.get("max_iter", 100))
    assert unconv == 0
    na = o.geti("nAlive")
    if precision == 0:   # (the mixed form may stop a particle one iteration apart)
        assert deep["iters"] == shallow["iters"] == o.geti("itersRun")
        assert deep["pairs"] == shallow["pairs"] and deep["candidates"] == shallow["candidates"]
    tol = 1e-10 if precision == 0 else 1e-6
    for out in (shallow, deep):
        assert rel_err(out["h"][:na], o.get("h")[:na]) < tol


def test_device_loop_oracle_and_graph_reuse():
    c = dataset("plummer_quintic")
    lib, iters = gpu_run(c, precision=0, gradients=False)
    o, unconv = oracle_run(c, gradients=False)
    assert unconv == 0 and iters == o.geti("itersRun")
    # the pass records add up to the solve's counters
    p = lib.passes()
    assert p.shape[1] == len(lib.PASS_FIELDS)
    assert p[:, 3].sum() == lib.stat("candidates") and p[:, 4].sum() == lib.stat("pairs")
    assert p[0, 10] > 0 and (p[:, 8] >= 0).all()
    na = o.geti("nAlive")
    assert rel_err(lib.array("h")[:na], o.get("h")[:na]) < 1e-10
    assert rel_err(lib.array("rho")[:na], o.get("rho")[:na]) < 1e-10
    # repeated solves of the same inputs replay cached graphs: the particle arrays alternate
    # between two buffers from one solve to the next, so at most one more capture, then none
    h1 = lib.array("h").copy()
    lib2, iters2 = gpu_run(c, precision=0, gradients=False)
    caps = lib2.stat("graph_captures")
    for _ in range(3):
        lib3, iters3 = gpu_run(c, precision=0, gradients=False)
        assert iters3 == iters
        assert np.array_equal(lib3.array("h"), h1)
    assert lib3.stat("graph_captures") == caps


def test_device_loop_unconverged_cap():
    # the NR pass cap ends the device loop exactly where the host loop stops
    c = dataset("plummer_cubic")
    host = solve(c, 0, loop=0, max_iter=100)
    assert host["iters"] > 3
    import subprocess
    import sys
    import os
    code = ("import sys; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
            "from gpu_common import dataset, gpu_run\n"
            "c = dataset('plummer_cubic')\n"
            "gpu_run(c, precision=0, gradients=False, max_iter=3, loop=1)\n") % (
        os.path.dirname(os.path.dirname(os.path.abspath(__file__))), os.path.dirname(os.path.abspath(__file__)))
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert res.returncode != 0
    assert "FATAL: solveDensH: unconverged=" in res.stderr and "iters=3" in res.stderr


def test_fresh_process_first_solve_mixed():
    # regression: the first solve of a process, mixed precision, device-driven loop (the
    # density arguments once took a buffer's pointer before the loop first sized it)
    import os
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    code = ("import sys; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
            "from gpu_common import dataset, gpu_run\n"
            "c = dataset('plummer_quintic')\n"
            "lib, it = gpu_run(c, precision=1, gradients=False, loop=1)\n"
            "print('iters', it, lib.stat('graph_replays'))\n") % (os.path.dirname(here), here)
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-2000:]
    assert "iters" in res.stdout
