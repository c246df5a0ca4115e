"""Oracle pin: sort, live count, leaf tree, internal nodes (src/tree.cu:88-153).

  - order is numpy's stable argsort of the keys; nAlive counts keys != ~0;
  - the leaf tree equals an independent top-down construction in Python (split iff the
    node holds more than 64 live keys and is above level 21) -- the unique fixed point of
    Cornerstone's rebalance loop;
  - the leaves tile [0, 2^63) with aligned power-of-8 ranges, leaves hold <= 64 keys
    unless at level 21, every internal node holds > 64;
  - the library's level-free construction (treemodel: 65-key windows + closed-form leaf
    index) reproduces the same leaves and internal nodes.
"""
import numpy as np
import pytest

import treemodel
from cosmosphere_amd import workloads
from oracle import oracle as O

DEAD = np.uint64(0xFFFFFFFFFFFFFFFF)


def sets():
    out = {"lattice": workloads.config2(npartx=14), "plummer": workloads.config3(n=40000, seed=21),
           "jittered": workloads._pack(*workloads._jittered_lattice(3000, 9), "cubic", "jittered lattice")}
    x = np.zeros(400); y = np.zeros(400); z = np.zeros(400)
    x[:200] = 0.25; y[:200] = 0.25; z[:200] = 0.25
    x[200:] = np.linspace(0, 1, 200); y[200:] = np.linspace(0, 1, 200) ** 2; z[200:] = 0.3
    h = np.full(400, 0.05)
    h[::37] = -1.0
    out["duplicates_dead"] = workloads._pack(x, y, z, h, "cubic", "dup")
    return out


def topdown(keys, bucket=64):
    leaves, internal = [], []

    def rec(start, level, lo, hi):
        if hi - lo > bucket and level < 21:
            internal.append((level, start))
            span = 1 << (3 * (20 - level))
            a = lo
            for d in range(8):
                e = int(np.searchsorted(keys[lo:hi], np.uint64(start + (d + 1) * span), side="left")) + lo
                rec(start + d * span, level + 1, a, e)
                a = e
        else:
            leaves.append((start, level, hi - lo))

    rec(0, 0, 0, len(keys))
    return leaves, sorted(internal)


@pytest.mark.parametrize("name", list(sets().keys()))
def test_sort_and_tree(name):
    c = sets()[name]
    o = O.Oracle(quintic=False)
    o.set_particles(c["x"], c["y"], c["z"], c["h"])
    o.build()
    box = o.box()
    k_host = O.keys(c["x"], c["y"], c["z"], c["h"], box)
    order = np.argsort(k_host, kind="stable")
    np.testing.assert_array_equal(o.get("order"), order)
    keys = o.get("keys")
    np.testing.assert_array_equal(keys, k_host[order])
    na = int(np.count_nonzero(k_host != DEAD))
    assert o.geti("nAlive") == na

    leaves, internal = topdown(keys[:na])
    cs, lv, lay = o.get("csTree"), o.get("leafLevel"), o.get("layout")
    assert len(cs) == len(leaves) + 1 and int(cs[-1]) == 1 << 63
    np.testing.assert_array_equal(cs[:-1], np.array([s for s, _, _ in leaves], dtype=np.uint64))
    np.testing.assert_array_equal(lv, [l for _, l, _ in leaves])
    np.testing.assert_array_equal(np.diff(lay.astype(np.int64)), [cnt for _, _, cnt in leaves])
    # tiling and bucket invariants
    for (s, l, cnt), e in zip(leaves, list(cs[1:])):
        assert int(e) - s == 1 << (3 * (21 - l)) and s % (1 << (3 * (21 - l))) == 0
        assert cnt <= 64 or l == 21
    ist, il = o.get("internalStart"), o.get("internalLevel")
    assert [(int(a), int(b)) for b, a in zip(ist, il)] == internal
    assert len(leaves) == 7 * len(internal) + 1
    # the library's level-free construction
    ls, ll, im = treemodel.window_tree(keys[:na])
    assert [int(v) for v in ls] == [int(v) for v in cs[:-1]]
    assert list(ll) == list(lv)
    assert [(int(s), l) for s, l in im] == [(int(a), int(b)) for a, b in zip(ist, il)]
