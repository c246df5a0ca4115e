"""numpy model of the library's level-free leaf-tree construction (tree.hip).

Used by tests to pin the algorithm itself against the oracle's recursive canonical tree:
  internal nodes = for each sorted live key p, the levels l with
      commonDigits(k[p-1], k[p]) < l <= min(commonDigits(k[p], k[p+64]), 20)
  leaf index of a leaf (start s, level l) = 7 (A - a) + digit sum of s over levels 1..l,
      A = internal nodes with start < s, a = ancestors of the leaf with start < s.
"""
import numpy as np

MAXL = 21


def _bitlen(x):
    x = np.asarray(x, dtype=np.uint64)
    hb = np.zeros(x.shape, dtype=np.int64)
    cur = x.copy()
    for s in (32, 16, 8, 4, 2, 1):
        m = (cur >> np.uint64(s)) != 0
        hb[m] += s
        cur = np.where(m, cur >> np.uint64(s), cur)
    return np.where(x == 0, 0, hb + 1)


def common_digits(a, b):
    x = np.asarray(a, dtype=np.uint64) ^ np.asarray(b, dtype=np.uint64)
    clz = 64 - _bitlen(x)
    return np.where(x == 0, MAXL, (clz - 1) // 3)


def prefix(key, level):
    if level == 0:
        return 0
    low = (1 << (63 - 3 * level)) - 1
    return int(key) & ~low & ((1 << 63) - 1)


def window_tree(keys, bucket=64):
    """keys: sorted live keys of one set -> (leafStart list, leafLevel list, internal list)."""
    k = np.asarray(keys, dtype=np.uint64)
    n = k.size
    c = np.full(n, -1, dtype=np.int64)
    if n > bucket:
        c[: n - bucket] = common_digits(k[: n - bucket], k[bucket:])
    e = np.full(n, -1, dtype=np.int64)
    if n > 1:
        e[1:] = common_digits(k[:-1], k[1:])
    cmax = np.minimum(c, MAXL - 1)
    internal = []
    for p in np.nonzero(cmax > e)[0]:
        for lvl in range(e[p] + 1, cmax[p] + 1):
            internal.append((prefix(k[p], lvl), lvl))
    pre_starts = np.array([s for s, _ in internal], dtype=object)
    assert all(internal[i] <= internal[i + 1] for i in range(len(internal) - 1)), "pre-order not sorted"
    iset = set(internal)
    nleaf = 7 * len(internal) + 1
    leaf_start = [None] * nleaf
    leaf_level = [None] * nleaf
    if not internal:
        leaf_start[0], leaf_level[0] = 0, 0
    for (s0, lvl) in internal:
        cl = lvl + 1
        span = 1 << (3 * (MAXL - cl))
        for d in range(8):
            cs = s0 + d * span
            if cl <= MAXL - 1 and (cs, cl) in iset:
                continue
            A = int(np.searchsorted(np.array([s for s, _ in internal], dtype=np.uint64), np.uint64(cs), side="left"))
            a = sum(1 for t in range(cl) if cs & ((1 << (63 - 3 * t)) - 1))
            ds = sum((cs >> (63 - 3 * (t + 1))) & 7 for t in range(cl))
            idx = 7 * (A - a) + ds
            assert leaf_start[idx] is None, "leaf index collision"
            leaf_start[idx], leaf_level[idx] = cs, cl
    del pre_starts
    return leaf_start, leaf_level, sorted(internal, key=lambda t: (t[1], t[0]))
