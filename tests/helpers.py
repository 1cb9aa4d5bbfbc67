"""Shared fixtures for the parity tests: small grids, random models and
fields in the canonical layout, comparison helpers."""
import numpy as np

from pyoracle import Grid


def small_grid(nx=28, ny=26, nz=30, fw=6, nt=60, fdcoeff=2, free_surf=0, dh=2.0, dt=2e-4):
    return Grid(nx=nx, ny=ny, nz=nz, fw=fw, dx=dh, dy=dh, dz=dh, dt=dt, nt=nt, fdcoeff=fdcoeff,
                free_surf=free_surf)


def random_model(g, rng, vp=4000.0, vs=2300.0, rho=2000.0, pert=0.1):
    shp = g.shape
    r = (rho * (1 + pert * rng.uniform(-1, 1, shp))).astype(np.float32)
    a = (vp * (1 + pert * rng.uniform(-1, 1, shp))).astype(np.float32)
    b = (vs * (1 + pert * rng.uniform(-1, 1, shp))).astype(np.float32)
    pi = (r * a * a).astype(np.float32)
    u = (r * b * b).astype(np.float32)
    return r, pi, u


def random_fields(g, rng, scale=1.0, n=9):
    return [(scale * rng.standard_normal(g.shape)).astype(np.float32) for _ in range(n)]


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def max_ulp(a, b):
    """largest distance in float32 units-in-the-last-place"""
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return int(np.abs(a - b).max()) if a.size else 0


# ---------------------------------------------------------------------------
# deterministic inputs independent of numpy's random streams: splitmix64 of
# (index, seed), top 24 bits -> [0, 1)
def det_uniform(shape, seed, lo=0.0, hi=1.0):
    n = int(np.prod(shape))
    with np.errstate(over="ignore"):
        z = np.arange(n, dtype=np.uint64) + np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(40)).astype(np.float64) / float(1 << 24)
    return (lo + (hi - lo) * u).reshape(shape).astype(np.float32)


def det_fields(g, seed, n=9, scale=1.0):
    return [det_uniform(g.shape, seed + 17 * m, -scale, scale) for m in range(n)]


def det_model(g, seed, vp=4000.0, vs=2300.0, rho=2000.0, pert=0.1):
    r = rho * (1 + det_uniform(g.shape, seed, -pert, pert).astype(np.float64))
    a = vp * (1 + det_uniform(g.shape, seed + 1, -pert, pert).astype(np.float64))
    b = vs * (1 + det_uniform(g.shape, seed + 2, -pert, pert).astype(np.float64))
    r = r.astype(np.float32)
    pi = (r * a * a).astype(np.float32)
    u = (r * b * b).astype(np.float32)
    return r, pi, u


def sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


def f32hex(a):
    return [format(int(x), "08x") for x in np.ascontiguousarray(a, np.float32).view(np.uint32).ravel()]


def from_f32hex(h):
    return np.array([int(x, 16) for x in h], np.uint32).view(np.float32)


# ---------------------------------------------------------------------------
# mirror symmetry of the staggered grid.  A canonical array index idx along
# an axis is the grid position idx - 2.  Mirrored about the cell centre c, a
# quantity centred on that axis maps position p to 2c - p and keeps its sign;
# one staggered half a cell up (vx and sxy, sxz along x, ...) maps p to
# 2c - 1 - p and changes sign (an odd vector or tensor component).
# Axes of the canonical arrays: 0 = y (j), 1 = x (i), 2 = z (k).
FIELD_STAGGER = {"vx": (1,), "vy": (0,), "vz": (2,), "sxx": (), "syy": (), "szz": (),
                 "sxy": (0, 1), "syz": (0, 2), "sxz": (1, 2)}
FIELD_NAMES = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "syz", "sxz")


def mirror_mismatch(a, axis, c, stag):
    """cells of canonical array a (or a stack of them, the grid in its last
    three axes) that break the mirror relation about centre c along grid
    axis `axis`; every pair of cells inside the array is compared"""
    ax = a.ndim - 3 + axis
    n = a.shape[ax]
    s = 2 * c + (3 if stag else 4)  # idx + mirror idx
    lo, hi = max(0, s - (n - 1)), min(n - 1, s)
    sl = [slice(None)] * a.ndim
    sl[ax] = slice(lo, hi + 1)
    part = a[tuple(sl)]
    mir = np.flip(part, ax)
    return int(np.count_nonzero(part != (-mir if stag else mir)))


def symmetric_table_model(shape, centre, seed, vp=4000.0, rho=2000.0, pert=0.1):
    """rho, pi, u of a heterogeneous model that is mirror-symmetric about the
    cell centre (cy, cx, cz): each value is looked up by (|y-cy|, |x-cx|,
    |z-cz|) in a seeded random table.  u takes the values 2^32 and 2^33
    (vs 1.4-2.2 km/s, bulk modulus pi - 4u/3 > 0): the harmonic averages sum
    four reciprocals in an
    order that the mirror reverses (av_mat.c:51-53), and reciprocals of
    powers of two sum exactly in any order, so the averages are symmetric
    bit for bit too"""
    rng = np.random.default_rng(seed)
    idx = [np.abs(np.arange(n) - 2 - c) for n, c in zip(shape, centre)]
    tshape = tuple(int(i.max()) + 1 for i in idx)
    r = (rho * (1 + pert * rng.uniform(-1, 1, tshape))).astype(np.float32)
    a = (vp * (1 + pert * rng.uniform(-1, 1, tshape))).astype(np.float32)
    u = np.ldexp(np.float32(1.0), rng.integers(32, 34, tshape)).astype(np.float32)
    r, a, u = [t[np.ix_(*idx)] for t in (r, a, u)]
    return r, (r * a * a).astype(np.float32), u


def symmetric_table_field(shape, centre, seed, lo, hi):
    """a positive field (e.g. tau) mirror-symmetric about the cell centre, in
    multiples of 2^-10, so that sums of four of them (the averages of
    av_mat.c:56-60) are exact in any order"""
    rng = np.random.default_rng(seed)
    idx = [np.abs(np.arange(n) - 2 - c) for n, c in zip(shape, centre)]
    t = rng.integers(int(lo * 1024), int(hi * 1024) + 1, tuple(int(i.max()) + 1 for i in idx))
    return (t.astype(np.float32) / np.float32(1024.0))[np.ix_(*idx)]


def homogeneous_model(shape, vp=4000.0, vs=2300.0, rho=2000.0):
    r = np.full(shape, rho, np.float32)
    return r, (r * np.float32(vp * vp)).astype(np.float32), (r * np.float32(vs * vs)).astype(np.float32)


def shifted_mismatch(a, b, shift, lo, hi):
    """cells of the box lo..hi (canonical indices, inclusive, per axis) where
    array b, shifted by `shift` cells, differs from a: a[p] vs b[p + shift]"""
    sa = tuple(slice(l, h + 1) for l, h in zip(lo, hi))
    sb = tuple(slice(l + s, h + 1 + s) for l, h, s in zip(lo, hi, shift))
    return int(np.count_nonzero(a[(Ellipsis,) + sa] != b[(Ellipsis,) + sb]))
