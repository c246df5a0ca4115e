"""GPU parity of the NR work-item mappings and of the folded closure.

Every mapping (whole-leaf items, per-particle items, the automatic switch at several
thresholds) and the j-leaf pruning that goes with it must leave the plain FP64 solve within
1e-10 of the oracle with the same iteration count and the same accepted pairs (pruning only
drops leaves without a pair inside the support), and the mixed solve within its 1e-6 bar.
With the closure on, rho is the oracle's closure at the returned h and h is unchanged.

Work units (mixed whole-leaf items): lanes over (particle, tile-slot subset) units instead of
one group of floor(64 / c) lanes per particle changes only the order in which each particle's
pair terms are summed, so the first pass (identical h and lists) has identical candidate and
accepted-pair counts with the units on and off, and both solves stay within the 1e-6 bar."""
import numpy as np
import pytest

from cosmosphere_amd import Library
from gpu_common import DEFAULT_ITEMS, dataset, gpu_run, oracle_run, rel_err

pytestmark = pytest.mark.gpu

MAPPINGS = [("whole", {"mapping": 1}), ("flat", {"mapping": 2}), ("auto4", {"sparse": 4}),
            ("auto64", {"sparse": 64}), ("nocull", {"cull": 0}), ("nocull_whole", {"cull": 0, "mapping": 1}),
            ("block_walk", {"list_heavy_steps": 1, "list_expand_leaves": 0})]


@pytest.fixture(autouse=True)
def _reset_options():
    yield
    for kernel in ("cubic", "quintic"):
        lib = Library(kernel)
        for k, v in DEFAULT_ITEMS.items():
            lib.set_option(k, v)


@pytest.mark.parametrize("label,opts", MAPPINGS, ids=[m[0] for m in MAPPINGS])
@pytest.mark.parametrize("name", ["plummer_cubic", "plummer_quintic", "dead"])
def test_mappings_plain_fp64(name, label, opts):
    c = dataset(name)
    print("case", name, label, flush=True)
    lib, iters = gpu_run(c, precision=0, gradients=False, **opts)
    o, unconv = oracle_run(c, gradients=False)
    assert unconv == 0 and iters == o.geti("itersRun")
    na = o.geti("nAlive")
    assert rel_err(lib.array("h")[:na], o.get("h")[:na]) < 1e-10
    assert rel_err(lib.array("rho")[:na], o.get("rho")[:na]) < 1e-10
    pairs = sum(o.get("pairsPerIter"))
    assert abs(lib.stat("pairs") - pairs) <= 1e-6 * pairs + 2


@pytest.mark.parametrize("label,opts", MAPPINGS[:2], ids=[m[0] for m in MAPPINGS[:2]])
def test_mappings_mixed(label, opts):
    c = dataset("plummer_quintic")
    print("case mixed", label, flush=True)
    lib, _ = gpu_run(c, precision=1, gradients=False, **opts)
    o, _ = oracle_run(c, gradients=False)
    na = o.geti("nAlive")
    assert rel_err(lib.array("h")[:na], o.get("h")[:na]) < 1e-6
    o.closure(c["pmass"])
    assert rel_err(lib.array("rho")[:na], o.get("rho")[:na]) < 1e-6


@pytest.mark.parametrize("mapping", [0, 1])
@pytest.mark.parametrize("name", ["plummer_cubic", "plummer_quintic", "dead", "C1small"])
def test_work_units_mixed(name, mapping):
    c = dataset(name)
    print("case units", name, mapping, flush=True)
    o, _ = oracle_run(c, gradients=False)
    na = o.geti("nAlive")
    o.closure(c["pmass"])
    first = {}
    for units in (1, 0):
        lib, _ = gpu_run(c, precision=1, gradients=False, mapping=mapping, units=units)
        assert rel_err(lib.array("h")[:na], o.get("h")[:na]) < 1e-6
        assert rel_err(lib.array("rho")[:na], o.get("rho")[:na]) < 1e-6
        p = lib.passes()
        first[units] = (p[0, 3], p[0, 4])
    assert first[0] == first[1]


@pytest.mark.parametrize("name", ["plummer_cubic", "plummer_quintic"])
def test_closure_plain_fp64(name):
    c = dataset(name)
    lib, iters = gpu_run(c, precision=0, gradients=False, closure=0)
    h0 = lib.array("h").copy()
    lib, iters1 = gpu_run(c, precision=0, gradients=False, closure=1)
    assert iters1 == iters
    assert rel_err(lib.array("h"), h0) < 1e-12   # the closure does not move h (items may sum in another order)
    assert lib.stat("closure_pairs") > 0
    o, _ = oracle_run(c, gradients=False)
    na = o.geti("nAlive")
    o.set_h_sorted(lib.array("h"))
    o.closure(c["pmass"])
    assert rel_err(lib.array("rho")[:na], o.get("rho")[:na]) < 1e-10
    g = o.get("gradh")[:na]
    assert rel_err(lib.array("gradh")[:na], g, floor=np.abs(g).max() * 1e-6) < 1e-9
