"""Oracle pin: periodic boundaries (include/tree.cuh:55-84 nearest image, :240-242 the periodic
box-box distance of the list walk; src/tree.cu:47-53 the given box).

  - superset: with a periodic box, every j within R h_i of a particle i of a leaf under the
    minimum image (brute force over all pairs) lies in a leaf of that leaf's gather list, and
    every j within R max(h_i, h_j) in its symmetric (force) list.  A box-box distance that
    dropped an image, or measured it from the wrong side, fails this where a clump straddles
    the box's faces and corner;
  - wrap invariance: shifting every particle by a fraction of the box and wrapping it back
    in changes keys, tree and lists but no minimum-image separation, so h, rho, the gradient
    outputs and the forces of each particle are unchanged to rounding (solved to a tight
    tolerance, so no convergence decision can flip).
"""
import numpy as np
import pytest

from oracle import oracle as O

BOX = np.array([0.0, 1.0, -0.2, 0.6, 0.1, 1.3])   # non-cubic, offset from the origin


def periodic_set(n=3000, seed=8):
    """Uniform background plus a clump straddling the box corner (wrapped)."""
    rng = np.random.default_rng(seed)
    lo, hi = BOX[0::2], BOX[1::2]
    L = hi - lo
    nb = n * 2 // 3
    p = lo + rng.random((nb, 3)) * L
    corner = np.array([BOX[0] + 0.01, BOX[3] - 0.01, BOX[5] - 0.01])
    c = corner + rng.normal(scale=0.05, size=(n - nb, 3))
    p = np.vstack([p, c])
    p = lo + np.mod(p - lo, L)
    h = np.full(n, 1.2 * (np.prod(L) / n) ** (1.0 / 3.0))
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), h, 1.0 / n


def min_image_r2(x, y, z):
    L = BOX[1::2] - BOX[0::2]
    r2 = 0.0
    for a, v in enumerate((x, y, z)):
        d = v[:, None] - v[None, :]
        d -= L[a] * np.rint(d / L[a])
        r2 = r2 + d * d
    return r2


@pytest.mark.parametrize("quintic", [False, True])
def test_periodic_lists_are_minimum_image_supersets(quintic):
    x, y, z, h, m = periodic_set()
    o = O.Oracle(quintic=quintic)
    o.set_particles(x, y, z, h)
    o.build(periodic_box=BOX)
    hfact = 1.0 if quintic else 1.2
    assert o.density_solve(m, 1e-4, hfact, 100, 0, False) == 0
    o.hmax_all()
    rad = O.kernel_consts(quintic)["rad"]
    xs, ys, zs, hs = o.get("x"), o.get("y"), o.get("z"), o.get("h")
    pl, lay = o.get("particleLeaf"), o.get("layout")
    r2 = min_image_r2(xs, ys, zs)
    L = o.geti("nLeaves")
    wrapped_pairs = 0
    for sym in (False, True):
        o.build_lists(symmetric=sym)
        jo, jl = o.get("jOffset"), o.get("jList")
        need = r2 < (rad * hs[:, None]) ** 2
        if sym:
            need |= r2 < (rad * hs[None, :]) ** 2
        for leaf in range(L):
            ii = np.arange(lay[leaf], lay[leaf + 1])
            if ii.size == 0:
                continue
            listed = set(jl[jo[leaf]:jo[leaf + 1]].tolist())
            js = np.nonzero(need[ii].any(axis=0))[0]
            assert set(pl[js].tolist()) <= listed, (sym, leaf)
    # the set does exercise the images: pairs that are neighbours only across a face
    direct = (xs[:, None] - xs[None, :]) ** 2 + (ys[:, None] - ys[None, :]) ** 2 + (zs[:, None] - zs[None, :]) ** 2
    wrapped_pairs = int(((r2 < (rad * hs[:, None]) ** 2) & ~(direct < (rad * hs[:, None]) ** 2)).sum())
    assert wrapped_pairs > 1000


def solve_in_caller_order(x, y, z, h, m, v, quintic, tolh=1e-11):
    n = x.size
    o = O.Oracle(quintic=quintic)
    o.set_particles(x, y, z, h, v=v, a=np.zeros((3, n)))
    o.build(periodic_box=BOX)
    hfact = 1.0 if quintic else 1.2
    assert o.density_solve(m, tolh, hfact, 100, 0, True) == 0
    order = o.get("order")
    hs = o.get("h")
    rhoh = m * (hfact / hs) ** 3
    pro2 = np.empty(n)
    pro2[:] = (2.0 / 3.0) / rhoh          # gamma 5/3, u = 1, in Hilbert order
    cs = np.full(n, np.sqrt(10.0 / 9.0))
    # THERMO is set in caller order: scatter the Hilbert-order values back first
    pro2_c = np.empty(n)
    pro2_c[order] = pro2
    o.set_thermo(pro2_c, cs, np.ones(n), np.ones(n))
    o.force(m, hfact, 2.0, 1.0, 0, 1, 1)
    out = {}
    for k in ("h", "rho", "gradh", "divv", "ddivvdt", "fx", "fy", "fz", "du"):
        v_ = np.empty(n)
        v_[order] = o.get(k)
        out[k] = v_
    return out


@pytest.mark.parametrize("quintic", [False, True])
def test_wrap_invariance(quintic):
    x, y, z, h, m = periodic_set(n=2500, seed=9)
    n = x.size
    rng = np.random.default_rng(10)
    v = rng.normal(scale=0.1, size=(3, n))
    base = solve_in_caller_order(x, y, z, h, m, v, quintic)
    lo, L = BOX[0::2], BOX[1::2] - BOX[0::2]
    shift = np.array([0.37, 0.61, 0.13]) * L
    xs, ys, zs = [lo[a] + np.mod(c + shift[a] - lo[a], L[a]) for a, c in enumerate((x, y, z))]
    moved = solve_in_caller_order(xs, ys, zs, h, m, v, quintic)
    for k in ("h", "rho"):
        assert np.max(np.abs(moved[k] - base[k]) / base[k]) < 1e-10, k
    for k in ("gradh", "divv", "ddivvdt", "fx", "fy", "fz", "du"):
        scale = np.abs(base[k]).max()
        assert np.max(np.abs(moved[k] - base[k])) < 1e-9 * scale, k
