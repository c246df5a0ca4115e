"""Oracle pin: neighbour-leaf lists (include/tree.cuh:228-389, src/force.cu:19-88).

  - order: a recursive Python restatement of the reference's LIFO walk -- accepted leaf
    children in octant order, then the accepted internal children in reverse octant order,
    each expanded recursively -- reproduces the oracle's CSR lists entry for entry;
  - superset: every particle j within R h_i of a particle i of the leaf lies in a listed
    leaf (brute force), so the density sums cannot miss a neighbour; the symmetric (force)
    list also covers every j within R h_j;
  - geometry: every particle lies inside its leaf's box.
"""
import numpy as np
import pytest

from cosmosphere_amd import workloads
from oracle import oracle as O


def solved(name):
    if name == "lattice":
        c = workloads.config2(npartx=12)
    else:
        c = workloads.config3(n=6000, seed=5)
    o = O.Oracle(quintic=c["quintic"])
    o.set_particles(c["x"], c["y"], c["z"], c["h"])
    o.build()
    assert o.density_solve(c["pmass"], 1e-4, c["hfact"], 100, 0, False) == 0
    return c, o


def tree_nodes(o):
    cs, lv = o.get("csTree"), o.get("leafLevel")
    ist, il = o.get("internalStart"), o.get("internalLevel")
    leaf_of = {(int(s), int(l)): i for i, (s, l) in enumerate(zip(cs[:-1], lv))}
    internal = {(int(s), int(l)) for s, l in zip(ist, il)}
    nodes = list(leaf_of.keys()) + list(internal)
    starts = np.array([s for s, _ in nodes], dtype=np.uint64)
    levels = np.array([l for _, l in nodes], dtype=np.int32)
    c, s = O.node_geometry(starts, levels, o.box())
    geo = {nd: (c[i], s[i]) for i, nd in enumerate(nodes)}
    return leaf_of, internal, geo


def dist2(ac, asz, bc, bsz):
    d = []
    for a in range(3):
        v = abs(float(bc[a]) - float(ac[a])) - float(asz[a]) - float(bsz[a])
        v += abs(v)
        v *= 0.5
        d.append(v)
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2]


def python_walk(o, leaf, leaf_of, internal, geo, hmax_leaf, rad, sym=False, hmax_node=None):
    inv = {v: k for k, v in leaf_of.items()}
    ic, isz = geo[inv[leaf]]
    hm = float(hmax_leaf[leaf])
    if not np.isfinite(hm) or hm <= 0:
        hm = 0.0
    radHi = rad * hm
    a_size = isz if sym else [float(isz[k]) + radHi for k in range(3)]

    def ok(nd):
        d2 = dist2(ic, a_size, *geo[nd])
        if sym:
            rs = max(radHi, rad * hmax_node[nd])
            return d2 <= rs * rs
        return d2 <= 0.0

    def visit(nd):
        s, l = nd
        span = 1 << (3 * (20 - l))
        kids = [(s + d * span, l + 1) for d in range(8)]
        out, inner = [], []
        for k in kids:
            if not ok(k):
                continue
            if k in internal:
                inner.append(k)
            else:
                out.append(leaf_of[k])
        for k in reversed(inner):
            out += visit(k)
        return out

    root = (0, 0)
    if not ok(root):
        return []
    if root not in internal:
        return [leaf_of[root]]
    return visit(root)


def hmax_nodes(o, leaf_of, internal):
    hl = o.get("hmaxLeaf")
    memo = {}

    def hm(nd):
        if nd in memo:
            return memo[nd]
        if nd in leaf_of:
            v = float(hl[leaf_of[nd]])
            v = v if (np.isfinite(v) and v > 0) else 0.0
        else:
            s, l = nd
            span = 1 << (3 * (20 - l))
            v = max(hm((s + d * span, l + 1)) for d in range(8))
        memo[nd] = v
        return v

    for nd in list(leaf_of) + list(internal):
        hm(nd)
    return memo


@pytest.mark.parametrize("name", ["lattice", "plummer"])
@pytest.mark.parametrize("sym", [False, True])
def test_list_order_matches_recursive_restatement(name, sym):
    c, o = solved(name)
    o.hmax_all()
    o.build_lists(symmetric=sym)
    leaf_of, internal, geo = tree_nodes(o)
    rad = O.kernel_consts(c["quintic"])["rad"]
    hn = hmax_nodes(o, leaf_of, internal) if sym else None
    jo, jl = o.get("jOffset"), o.get("jList")
    L = o.geti("nLeaves")
    for leaf in range(0, L, max(1, L // 60)):
        ref = python_walk(o, leaf, leaf_of, internal, geo, o.get("hmaxLeaf"), rad, sym, hn)
        assert list(jl[jo[leaf]:jo[leaf + 1]]) == ref


@pytest.mark.parametrize("name", ["lattice", "plummer"])
def test_lists_are_supersets_and_geometry(name):
    c, o = solved(name)
    o.hmax_all()
    rad = O.kernel_consts(c["quintic"])["rad"]
    x, y, z, h = o.get("x"), o.get("y"), o.get("z"), o.get("h")
    pl, lay = o.get("particleLeaf"), o.get("layout")
    leaf_of, internal, geo = tree_nodes(o)
    inv = {v: k for k, v in leaf_of.items()}
    for leaf in range(o.geti("nLeaves")):
        cc, ss = geo[inv[leaf]]
        for i in range(lay[leaf], lay[leaf + 1]):
            for a, v in enumerate((x[i], y[i], z[i])):
                assert abs(v - cc[a]) <= ss[a] * (1 + 1e-12)
    r2 = (x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2 + (z[:, None] - z[None, :]) ** 2
    for sym in (False, True):
        o.build_lists(symmetric=sym)
        jo, jl = o.get("jOffset"), o.get("jList")
        need = r2 < (rad * h[:, None]) ** 2
        if sym:
            need |= r2 < (rad * h[None, :]) ** 2
        for leaf in range(0, o.geti("nLeaves"), 3):
            listed = set(jl[jo[leaf]:jo[leaf + 1]].tolist())
            ii = np.arange(lay[leaf], lay[leaf + 1])
            if ii.size == 0:
                continue
            js = np.nonzero(need[ii].any(axis=0))[0]
            assert set(pl[js].tolist()) <= listed
