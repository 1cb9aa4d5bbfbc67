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
