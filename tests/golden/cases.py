"""Inputs of the golden cases (deterministic, numpy-version independent).

make_golden.py runs these through the reference's own routines (oracle/_ref,
compiled from the reference sources) and stores the outputs as SHA-256 of
the exact float32 bytes plus a few sample values; test_oracle_golden.py runs
the same inputs through the oracle and compares.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))

from helpers import det_fields, det_model, det_uniform  # noqa: E402
from pyoracle import Grid  # noqa: E402

CPML_CASES = [  # fw, dh, vppml, fpml, dt
    (10, 2.0, 4000.0, 50.0, 2e-4),
    (8, 5.0, 5500.0, 15.0, 4e-4),
    (12, 0.8, 6200.0, 200.0, 5e-5),
    (3, 60.0, 200.0, 1.0, 1e-3),  # |d| < 1 at some points: integer-abs quirk
]


def step_grid(fdcoeff, free_surf):
    return Grid(nx=14, ny=13, nz=15, fw=4, dx=2.0, dy=2.5, dz=3.0, dt=2e-4, nt=30, fdcoeff=fdcoeff,
                free_surf=free_surf)


STEP_CASES = [(1, 0), (2, 0), (2, 1)]
STEP_NT = 5
# (i, j, k, tshift, fc, amp, type): forces of types 2-4 (two sharing a cell)
# and an explosive source
STEP_SOURCES = [(3, 4, 5, 0.0, 60.0, 1.0, 2), (3, 4, 5, 0.001, 45.0, 2.0, 3), (11, 9, 8, 0.0, 70.0, 1.5, 4),
                (7, 6, 7, 0.0, 55.0, 1.0, 1)]
STEP_RECEIVERS = np.array([[2, 3, 4], [2, 3, 4], [8, 10, 12], [13, 12, 14]], np.int32)


def step_inputs(g, seed):
    f = det_fields(g, seed, scale=1.0)
    rho, pi, u = det_model(g, seed + 100)
    psi = [det_uniform(s, seed + 200 + n, -1e-3, 1e-3) for n, s in enumerate(g.psi_shapes() * 2)]
    adj = [det_uniform((len(STEP_RECEIVERS), g.nt + 2), seed + 300 + c, -1e-3, 1e-3) for c in range(3)]
    return f, (rho, pi, u), psi, adj


def dft_grid():
    return Grid(nx=12, ny=11, nz=13, fw=3, dx=2.0, dy=2.0, dz=2.0, dt=2e-4, nt=50)


DFT_FINV = [40.0, 60.0, 80.0]
DFT_CALLS = [(4, 2, 0), (6, 2, 0), (10, 2, 1), (12, 2, 1)]  # (nt, ntast, back)


def dft_inputs(g, seed):
    v = det_fields(g, seed, n=3)
    nf = len(DFT_FINV)
    F = [det_uniform((nf,) + g.shape, seed + 50 + n, -0.5, 0.5) for n in range(12)]
    return v, F


def grad_inputs(g, seed):
    nf = len(DFT_FINV)
    F = [det_uniform((nf,) + g.shape, seed + 70 + n, -1.0, 1.0) for n in range(12)]
    for a in F:  # zero halos, as zero_invers leaves them (zero_invers.c:43-79)
        a[:, :3] = 0
        a[:, 3 + g.ny:] = 0
        a[:, :, :3] = 0
        a[:, :, 3 + g.nx:] = 0
        a[:, :, :, :3] = 0
        a[:, :, :, 3 + g.nz:] = 0
    grads = [det_uniform(g.shape, seed + 90 + n, -1.0, 1.0) for n in range(3)]
    rho, pi, u = det_model(g, seed + 95)
    return F, grads, (rho, pi, u)


HESS_NTR = 2  # receiver-side Green's-function sets of the hess_F case


def hess_inputs(g, seed, ntr=HESS_NTR):
    """forward DFT fields [nf], receiver-side fields [nf * (ntr + 1)] (entries
    of receiver n, frequency l at n*nf + l - 1; receiver 0 unused by hess_F),
    three Hessian accumulators and a model"""
    nf = len(DFT_FINV)
    Ff = [det_uniform((nf,) + g.shape, seed + 110 + n, -1.0, 1.0) for n in range(6)]
    Fb = [det_uniform((nf * (ntr + 1),) + g.shape, seed + 120 + n, -1.0, 1.0) for n in range(6)]
    for a in Ff + Fb:  # zero halos (zero_invers.c:43-79)
        a[:, :3] = 0
        a[:, 3 + g.ny:] = 0
        a[:, :, :3] = 0
        a[:, :, 3 + g.nx:] = 0
        a[:, :, :, :3] = 0
        a[:, :, :, 3 + g.nz:] = 0
    hess = [det_uniform(g.shape, seed + 130 + n, 0.0, 1.0) for n in range(3)]
    rho, pi, u = det_model(g, seed + 135)
    return Ff, Fb, hess, (rho, pi, u)


def residual_inputs(ntr, ns, seed):
    obs = [det_uniform((ntr, ns + 1), seed + c, -1.0, 1.0) for c in range(3)]
    syn = [det_uniform((ntr, ns + 1), seed + 10 + c, -1.0, 1.0) for c in range(3)]
    return obs, syn


def shot_grid(free_surf):
    return Grid(nx=24, ny=22, nz=26, fw=6, dx=2.0, dy=2.0, dz=2.0, dt=2e-4, nt=100, fdcoeff=2,
                free_surf=free_surf)


SHOT_SRC = (12, 9, 13, 0.0, 60.0, 1.0, 1)
SHOT_REC = np.array([[6 + 2 * n, 7, 13] for n in range(7)], np.int32)
SHOT_FINV = [50.0, 80.0]


def shot_models(g):
    shp = g.shape
    J, I, K = np.meshgrid(*(np.arange(s) - 2 for s in shp), indexing="ij")
    fac = 1.0 + 0.1 * ((((J - 11) ** 2 + (I - 12) ** 2 + (K - 13) ** 2) <= 9))
    out = []
    for f in (np.ones(shp), fac):
        vp, vs, rho = 4000.0 * f, 2300.0 * f, 2000.0 * f
        out.append([a.astype(np.float32) for a in (rho, rho * vp * vp, rho * vs * vs)])
    return out  # start, true


# g8: the viscoelastic stress update with one relaxation mechanism
VISCO_CASES = [(2, 0), (1, 1)]  # (fdcoeff, free_surf)
VISCO_ETA = 0.0125


def visco_inputs(g, seed):
    """wavefields, model, psi as step_inputs, plus memory variables r (6),
    taus and taup (relaxation-time ratios, positive)"""
    f, model, psi, _ = step_inputs(g, seed)
    r = [det_uniform(g.shape, seed + 400 + n, -1e-2, 1e-2) for n in range(6)]
    taus = det_uniform(g.shape, seed + 410, 0.01, 0.06)
    taup = det_uniform(g.shape, seed + 411, 0.01, 0.06)
    return f, model, psi, r, taus, taup


# SU seismogram files (outseis.c / readseis.c): (ntr, ns, receiver ids,
# source position or None for nsrc != 1, dx, dy, dz, dt, ndt)
SU_CASES = (
    (4, 9, (3, 7, 8, 12), (12.3, 4.56, 30.1), 0.8, 1.25, 0.7, 2.7e-4, 2),
    (3, 5, (1, 2, 3), None, 0.0005, 2.5, 0.125, 1.3e-3, 1),
)


def su_inputs(case, seed):
    """a section [ntr][ns+1] (index 0 zero) and receiver grid indices [ntr][3]"""
    ntr, ns = case[0], case[1]
    rng = np.random.default_rng(seed)
    sec = np.zeros((ntr, ns + 1), np.float32)
    sec[:, 1:] = rng.standard_normal((ntr, ns)).astype(np.float32)
    rec_xyz = rng.integers(1, 60, size=(ntr, 3)).astype(np.int32)
    return sec, rec_xyz
