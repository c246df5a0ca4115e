"""Synthetic particle sets for the five benchmark configurations.

Every generator draws from a counter-based RNG (splitmix64 over (seed, stream, index)),
so the CPU oracle and the GPU library are fed bit-identical inputs in any process.
Masses are m = 1/N, positions FP64.

  C1  close-packed lattice, npartx = 50 (173,850 particles), h0 = 1.2 N^-1/3, 1% of h
      scaled by (1 + U[-0.35, 0.35]), seed 1, cubic kernel
  C2  lattice npartx = 100 (1,403,000), Gaussian displacement sigma = 0.05 dx, seed 2
  C3  8,000,000-particle Plummer mixture: 64 spheres (centres U[0,1]^3, scale radii
      log-uniform 0.005-0.05, Dirichlet(1) masses) holding 80% of N, 20% uniform
      background, h0 = N^-1/3, seed 3, quintic kernel
  C4  256 independent sets, sizes log-uniform 1e4-5e5, even = jittered lattice, odd =
      Plummer mixture, seed 4 + b, cubic kernel
  C5  16,000,000-particle Plummer mixture, seed 5, with velocities N(0, 0.1^2) plus
      solid-body rotation about z, u = 1, gamma = 5/3, alpha = 1, beta = 2

The lattice restates the reference driver's makeLattice (src/main.cu:84-109).
"""
import math

import numpy as np

_U64 = np.uint64
_GOLD = _U64(0x9E3779B97F4A7C15)
_C1 = _U64(0xBF58476D1CE4E5B9)
_C2 = _U64(0x94D049BB133111EB)
_STEP = _U64(0xD1B54A32D192ED03)


def _mix(z):
    z = (z ^ (z >> _U64(30))) * _C1
    z = (z ^ (z >> _U64(27))) * _C2
    return z ^ (z >> _U64(31))


def _stream_key(seed, stream):
    with np.errstate(over="ignore"):
        return _mix(_U64((int(seed) << 32) + int(stream)) * _STEP + _GOLD)


def rand_u64(seed, stream, idx):
    """splitmix64 of (seed, stream, idx); idx is an integer array."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix(_stream_key(seed, stream) ^ (idx * _STEP + _GOLD))


def uniform(seed, stream, idx):
    """U[0, 1) with 53 random bits."""
    return (rand_u64(seed, stream, idx) >> _U64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform_open(seed, stream, idx):
    """U(0, 1), never 0 or 1."""
    return ((rand_u64(seed, stream, idx) >> _U64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def gaussian(seed, stream, idx):
    """Standard normal by Box-Muller from two uniform streams (2*stream, 2*stream+1)."""
    u1 = uniform_open(seed, 2 * stream, idx)
    u2 = uniform(seed, 2 * stream + 1, idx)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def make_lattice(npartx):
    """Close-packed lattice in the unit box, as src/main.cu:84-109 (same FP expressions)."""
    if npartx < 4:
        raise ValueError("npartx must be at least 4")
    dx = 1.0 / npartx
    dy = dx * math.sqrt(3.0) / 2.0
    dz = dx * math.sqrt(6.0) / 3.0
    ny = int(1.0 / dy)
    nz = int(1.0 / dz)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(npartx), indexing="ij")
    k, j, i = k.ravel(), j.ravel(), i.ravel()
    xo = ((j + k) % 2).astype(np.float64) * 0.5 * dx
    yo = (k % 2).astype(np.float64) * dy / 3.0
    x = i.astype(np.float64) * dx + xo
    y = j.astype(np.float64) * dy + yo
    z = k.astype(np.float64) * dz
    return x, y, z, dx


def _largest_remainder(total, frac):
    raw = total * frac
    base = np.floor(raw).astype(np.int64)
    rem = int(total - base.sum())
    if rem > 0:
        order = np.lexsort((np.arange(frac.size), -(raw - base)))
        base[order[:rem]] += 1
    return base


def plummer_mixture(n, seed, nsph=64, frac=0.8, amin=0.005, amax=0.05, rcut=10.0):
    """64 Plummer spheres holding 80% of the particles plus a 20% uniform background."""
    ks = np.arange(nsph)
    centres = np.stack([uniform(seed, 1, 3 * ks + a) for a in range(3)], axis=1)
    a = np.exp(math.log(amin) + (math.log(amax) - math.log(amin)) * uniform(seed, 2, ks))
    e = -np.log(uniform_open(seed, 3, ks))
    nsph_tot = int(round(frac * n))
    counts = _largest_remainder(nsph_tot, e / e.sum())
    sid = np.repeat(ks, counts)
    g = np.arange(nsph_tot, dtype=np.int64)

    # radius r = a / sqrt(X^(-2/3) - 1), resampled while r > rcut * a
    r = np.empty(nsph_tot)
    pending = g
    attempt = 0
    while pending.size:
        X = uniform_open(seed, 100 + attempt, pending)
        rr = a[sid[pending]] / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
        ok = rr <= rcut * a[sid[pending]]
        r[pending[ok]] = rr[ok]
        pending = pending[~ok]
        attempt += 1
    cth = 2.0 * uniform(seed, 4, g) - 1.0
    sth = np.sqrt(np.maximum(0.0, 1.0 - cth * cth))
    phi = 2.0 * math.pi * uniform(seed, 5, g)
    xs = centres[sid, 0] + r * sth * np.cos(phi)
    ys = centres[sid, 1] + r * sth * np.sin(phi)
    zs = centres[sid, 2] + r * cth
    nb = n - nsph_tot
    gb = np.arange(nb, dtype=np.int64)
    xb, yb, zb = uniform(seed, 10, gb), uniform(seed, 11, gb), uniform(seed, 12, gb)
    return np.concatenate([xs, xb]), np.concatenate([ys, yb]), np.concatenate([zs, zb])


def _pack(x, y, z, h, kernel, name, **extra):
    n = x.size
    d = {"name": name, "n": n, "x": np.ascontiguousarray(x), "y": np.ascontiguousarray(y),
         "z": np.ascontiguousarray(z), "h": np.ascontiguousarray(h), "pmass": 1.0 / n,
         "kernel": kernel, "quintic": kernel == "quintic",
         "hfact": 1.0 if kernel == "quintic" else 1.2}
    d.update(extra)
    return d


def config1(npartx=50, seed=1):
    """C1: lattice, 1% of h perturbed by U[-0.35, 0.35], cubic, density only, max 10 its."""
    x, y, z, _ = make_lattice(npartx)
    n = x.size
    idx = np.arange(n)
    h = np.full(n, 1.2 * n ** (-1.0 / 3.0))
    pick = uniform(seed, 0, idx) < 0.01
    delta = 0.7 * uniform(seed, 1, idx) - 0.35
    h = np.where(pick, h * (1.0 + delta), h)
    return _pack(x, y, z, h, "cubic", "C1 lattice npartx=%d" % npartx, max_iter=10)


def config2(npartx=100, seed=2):
    """C2: lattice with Gaussian displacement sigma = 0.05 dx, cubic."""
    x, y, z, dx = make_lattice(npartx)
    n = x.size
    idx = np.arange(n)
    s = 0.05 * dx
    x = x + s * gaussian(seed, 0, idx)
    y = y + s * gaussian(seed, 1, idx)
    z = z + s * gaussian(seed, 2, idx)
    h = np.full(n, 1.2 * n ** (-1.0 / 3.0))
    return _pack(x, y, z, h, "cubic", "C2 displaced lattice npartx=%d" % npartx, max_iter=100)


def config3(n=8_000_000, seed=3):
    """C3: Plummer mixture, quintic, h0 = N^-1/3."""
    x, y, z = plummer_mixture(n, seed)
    h = np.full(n, 1.0 * n ** (-1.0 / 3.0))
    return _pack(x, y, z, h, "quintic", "C3 Plummer mixture N=%d" % n, max_iter=100)


def _jittered_lattice(n, seed):
    npartx = 4
    while True:
        dx = 1.0 / npartx
        ny = int(1.0 / (dx * math.sqrt(3.0) / 2.0))
        nz = int(1.0 / (dx * math.sqrt(6.0) / 3.0))
        if npartx * ny * nz >= n:
            break
        npartx += 1
    x, y, z, dx = make_lattice(npartx)
    x, y, z = x[:n], y[:n], z[:n]
    idx = np.arange(n)
    x = x + dx * (0.5 * uniform(seed, 20, idx) - 0.25)
    y = y + dx * (0.5 * uniform(seed, 21, idx) - 0.25)
    z = z + dx * (0.5 * uniform(seed, 22, idx) - 0.25)
    return x, y, z, np.full(n, 1.2 * dx)


def config4_sizes(nsets=256, seed=4, nmin=10_000, nmax=500_000):
    u = uniform(seed, 0, np.arange(nsets))
    return np.floor(np.exp(math.log(nmin) + (math.log(nmax) - math.log(nmin)) * u)).astype(np.int64)


def config4(nsets=256, seed=4, nmin=10_000, nmax=500_000):
    """C4: list of independent sets (dicts), cubic kernel."""
    sets = []
    for b, nb in enumerate(config4_sizes(nsets, seed, nmin, nmax)):
        nb = int(nb)
        if b % 2 == 0:
            x, y, z, h = _jittered_lattice(nb, seed + b)
        else:
            x, y, z = plummer_mixture(nb, seed + b)
            h = np.full(nb, 1.2 * nb ** (-1.0 / 3.0))
        sets.append(_pack(x, y, z, h, "cubic", "C4 set %d N=%d" % (b, nb)))
    return sets


def config5(n=16_000_000, seed=5):
    """C5: Plummer mixture with velocities and thermodynamics for the force pass."""
    x, y, z = plummer_mixture(n, seed)
    idx = np.arange(n)
    vx = 0.1 * gaussian(seed, 30, idx) - (y - 0.5)
    vy = 0.1 * gaussian(seed, 31, idx) + (x - 0.5)
    vz = 0.1 * gaussian(seed, 32, idx)
    h = np.full(n, 1.0 * n ** (-1.0 / 3.0))
    gamma = 5.0 / 3.0
    return _pack(x, y, z, h, "quintic", "C5 Plummer mixture + force N=%d" % n,
                 vx=vx, vy=vy, vz=vz, u=np.ones(n), gamma=gamma,
                 cs=np.full(n, math.sqrt(gamma * (gamma - 1.0) * 1.0)),
                 alpha=np.ones(n), beta=2.0, alphau=1.0, dt=1e-4, nsteps=10)
