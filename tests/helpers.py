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
