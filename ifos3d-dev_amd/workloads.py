"""Synthetic workloads of the benchmark configurations (models, shots,
receivers, time step), in the canonical/device layouts of ifosgpu.

* toy84:   84^3 (64^3 + 10-cell CPML), dh 2 m, dt 2e-4 s, NT 600, homogeneous
           start model (vp 4000, vs 2300, rho 2000) and a true model with two
           +-10% spheres of radius 8 cells, 4 explosive Ricker shots (50 Hz),
           32 receivers, DFT at 40/60/80 Hz.
* grad276: 276^3 (256^3 + 10-cell CPML), dh 4 m, dt = 0.9 CFL, NT 1000,
           seeded Gaussian random field (correlation length 20 cells, vp in
           2500-5500 m/s, vp/vs in 1.7-2.0, Gardner density) as the true
           model and its 2-cell Gaussian smoothing as the start model,
           16 explosive Ricker shots (10 Hz) on a 4x4 grid, 1024 receivers
           (32x32) near the top, 6 DFT frequencies 4-14 Hz.

* fwd536:   536 x 280 x 536 (512 x 256 x 512 + 12-cell CPML), dh 5 m,
           dt = 0.9 CFL, NT 2000, the seeded random field (seed 2), one
           explosive 15 Hz Ricker source, 4096 receivers (64 x 64): the
           forward-only bandwidth case.
* visco404: 404 x 212 x 404 (384 x 192 x 384 + 10-cell CPML), dh 5 m, the
           seeded random field (seed 3) as the model, 3 relaxation mechanisms
           whose frequencies are fitted to constant Q over 5-40 Hz (the
           tau-method, the reference's mfiles/qapprox.m), Q uniform in
           20-100 per cell (seed 3) -> tau per cell, moduli relaxed so that
           the phase velocity at the first relaxation frequency is the
           model's (readmod.c:50-64,110-111), dt = 0.9 CFL of the unrelaxed
           P velocity (checkfd_ssg.c), NT 2000, one explosive 15 Hz source,
           4096 receivers, 8 DFT frequencies 5, 10, ..., 40 Hz.
* levels3:  three coarse-to-fine grid levels of one 768 m cube, interiors
           96^3 at dh 8 m, 128^3 at 6 m and 192^3 at 4 m, each + 10-cell
           CPML, dt = 0.9 CFL per level and NT = ceil(0.6 s / dt); the
           seeded random field (as grad276) built on the finest level and
           sampled at each level's cell positions; 32 explosive Ricker shots
           (seed 4, fc uniform in 5-20 Hz) at the same physical positions on
           every level; 64 receivers (8x8) near the top.  Forward solves.

The CFL limit is the reference's (checkfd_ssg.c:473-510):
dt <= dh / (gamma * sqrt(3) * vmax), gamma = 1.184614 at 4th order.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

GAMMA4 = 1.184614


def cfl_dt(dh: float, vmax: float, frac: float = 0.9) -> float:
    return frac * dh / (GAMMA4 * math.sqrt(3.0) * vmax)


@dataclass
class Workload:
    name: str
    nx: int
    ny: int
    nz: int
    fw: int
    dh: float
    dt: float
    nt: int
    finv: list
    tast: int
    vppml: float
    fpml: float
    shots: list            # (i, j, k, tshift, fc, amp, type)
    receivers: np.ndarray  # (ntr, 3) i, j, k
    fdcoeff: int = 2
    seed: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def cells(self):
        return self.nx * self.ny * self.nz


def toy84() -> Workload:
    fw, n = 10, 84
    shots = [(fw + i, fw + 12, fw + k, 0.0, 50.0, 1.0, 1) for i, k in ((21, 21), (21, 43), (43, 21), (43, 43))]
    rec = np.array([[fw + 1 + 2 * m, fw + 4, fw + 32] for m in range(32)], np.int32)
    return Workload("toy84", n, n, n, fw, 2.0, 2e-4, 600, [40.0, 60.0, 80.0], 100, 4000.0, 50.0, shots, rec)


def grad276() -> Workload:
    fw, n, dh = 10, 276, 4.0
    dt = cfl_dt(dh, 5500.0)
    pos = [fw + 32 + 64 * m for m in range(4)]
    shots = [(i, fw + 10, k, 0.0, 10.0, 1.0, 1) for i in pos for k in pos]
    rec = np.array([[fw + 1 + 8 * a, fw + 1, fw + 1 + 8 * b] for a in range(32) for b in range(32)], np.int32)
    finv = [4.0, 6.0, 8.0, 10.0, 12.0, 14.0]
    return Workload("grad276", n, n, n, fw, dh, float(np.float32(dt)), 1000, finv, 100, 5500.0, 10.0, shots, rec,
                    seed=5)


def fwd536() -> Workload:
    """the forward-only bandwidth case: 536 x 280 x 536 (512 x 256 x 512 + 12-cell
    CPML), dh 5 m, dt = 0.9 CFL, NT 2000, one explosive Ricker source (15 Hz)
    at the centre 10 cells below the top, 4096 receivers (64 x 64, every 8
    cells) at the top"""
    fw, dh = 12, 5.0
    nx, ny, nz = 536, 280, 536
    dt = cfl_dt(dh, 5500.0)
    shots = [(fw + 1 + 256, fw + 10, fw + 1 + 256, 0.0, 15.0, 1.0, 1)]
    rec = np.array([[fw + 1 + 8 * a, fw + 1, fw + 1 + 8 * b] for a in range(64) for b in range(64)], np.int32)
    return Workload("fwd536", nx, ny, nz, fw, dh, float(np.float32(dt)), 2000, [15.0], 100, 5500.0, 15.0, shots,
                    rec, seed=2)


def visco404() -> Workload:
    """the viscoelastic + DFT case (see the module docstring); dt is set by
    visco_models, which needs the model"""
    fw, dh = 10, 5.0
    nx, ny, nz = 404, 212, 404
    shots = [(fw + 1 + 192, fw + 10, fw + 1 + 192, 0.0, 15.0, 1.0, 1)]
    rec = np.array([[fw + 1 + 6 * a, fw + 1, fw + 1 + 6 * b] for a in range(64) for b in range(64)], np.int32)
    finv = [5.0 * (n + 1) for n in range(8)]
    return Workload("visco404", nx, ny, nz, fw, dh, 0.0, 2000, finv, 100, 5500.0, 15.0, shots, rec, seed=3,
                    extra={"L": 3, "qrange": (20.0, 100.0), "band": (5.0, 40.0)})


def q_of(ts, tau, f):
    """Q(f) of the L-fold standard linear solid of the tau-method
    (mfiles/qgsls.m: relaxation times ts, strain retardation ts (1 + tau))"""
    w = 2.0 * np.pi * np.asarray(f, np.float64)
    num = 1.0 - len(ts) + sum((1.0 + w * w * t * t * (1.0 + tau)) / (1.0 + w * w * t * t) for t in ts)
    den = sum(w * t * tau / (1.0 + w * w * t * t) for t in ts)
    return num / den


def fit_relaxation(L=3, band=(5.0, 40.0), q0=50.0):
    """relaxation frequencies and tau of L mechanisms for constant Q = q0
    over the band: least squares on 1/Q (the tau-method, mfiles/qapprox.m),
    the frequencies kept within a factor 3 of the band"""
    from scipy.optimize import least_squares
    f = np.linspace(band[0], band[1], 64)

    def res(x):
        ts = 1.0 / (2.0 * np.pi * np.exp(x[:L]))
        return 1.0 / q_of(ts, x[L], f) - 1.0 / q0

    x0 = np.concatenate([np.log(np.geomspace(band[0], band[1], L)), [2.0 / q0]])
    lo = np.concatenate([np.full(L, np.log(band[0] / 3.0)), [1e-4]])
    hi = np.concatenate([np.full(L, np.log(band[1] * 3.0)), [1.0]])
    x = least_squares(res, x0, bounds=(lo, hi), x_scale=np.concatenate([np.ones(L), [0.01]])).x
    return np.exp(x[:L]), float(x[L])


def visco_models(w: Workload, corr=20.0):
    """(rho, pi, u), (taus, taup), etas of the viscoelastic workload, and
    its dt (sets w.dt): the random field model, Q uniform in qrange per
    cell, tau = tau(q0 = 50) * 50 / Q (1/Q is proportional to tau in the
    tau-method), moduli relaxed as readmod.c:50-64,110-111 does with the
    first relaxation frequency as reference"""
    L = w.extra["L"]
    fl, tau50 = fit_relaxation(L, w.extra["band"])
    shape = (w.ny + 5, w.nx + 5, w.nz + 5)
    vp = _minmax(_grf(shape, corr, w.seed), 2500.0, 5500.0)
    vs = vp / _minmax(_grf(shape, corr, w.seed + 1000), 1.7, 2.0)
    rho = 310.0 * vp ** 0.25
    q = np.random.default_rng(w.seed).uniform(*w.extra["qrange"], size=shape)
    tau = tau50 * 50.0 / q
    ts = 1.0 / (2.0 * np.pi * fl)
    ws = 2.0 * np.pi * fl[0]
    fac = 1.0 + sum(ws * ws * t * t * tau / (1.0 + ws * ws * t * t) for t in ts)  # sumu = sumpi (ts = tp)
    pi = rho * vp * vp / fac
    u = rho * vs * vs / fac
    vmax = float(np.sqrt(pi * (1.0 + L * tau) / rho).max())
    w.dt = float(np.float32(cfl_dt(w.dh, vmax)))
    etas = np.array([w.dt / t for t in ts], np.float32)
    w.extra.update(fl=[float(x) for x in fl], tau50=tau50, vmax_unrelaxed=vmax)
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    return [f32(rho), f32(pi), f32(u)], [f32(tau), f32(tau)], etas


LEVELS3 = ((96, 8.0), (128, 6.0), (192, 4.0))  # (interior cells per axis, dh), one 768 m cube


def levels3(nshots: int = 32) -> list:
    """the three grid levels of one survey (see the module docstring)"""
    fw, side, tmax = 10, 768.0, 0.6
    rng = np.random.default_rng(4)
    # physical shot positions (x, depth, z) and frequencies, shared by the levels
    pos = np.column_stack([rng.uniform(0.1 * side, 0.9 * side, nshots), rng.uniform(40.0, 120.0, nshots),
                           rng.uniform(0.1 * side, 0.9 * side, nshots)])
    fcs = rng.uniform(5.0, 20.0, nshots)
    out = []
    for n, dh in LEVELS3:
        dt = float(np.float32(cfl_dt(dh, 5500.0)))
        nt = int(math.ceil(tmax / dt))
        idx = lambda x: fw + 1 + int(round(x / dh))
        shots = [(idx(p[0]), idx(p[1]), idx(p[2]), 0.0, float(f), 1.0, 1) for p, f in zip(pos, fcs)]
        rec = np.array([[idx(side * (a + 0.5) / 8), fw + 1, idx(side * (b + 0.5) / 8)]
                        for a in range(8) for b in range(8)], np.int32)
        out.append(Workload(f"levels3_{n}", n + 2 * fw, n + 2 * fw, n + 2 * fw, fw, dh, dt, nt, [10.0], 100,
                            5500.0, 10.0, shots, rec, seed=4, extra={"interior": n}))
    return out


def level_models(levels: list):
    """true models (rho, pi, u) of each level: the random field of the finest
    level sampled at every level's cell positions"""
    fine = levels[-1]
    shape = (fine.ny + 5, fine.nx + 5, fine.nz + 5)
    vp_f = _minmax(_grf(shape, 160.0 / fine.dh, fine.seed), 2500.0, 5500.0)
    ratio_f = _minmax(_grf(shape, 160.0 / fine.dh, fine.seed + 1000), 1.7, 2.0)
    out = []
    for w in levels:
        # canonical index c (grid index c - 2) of this level -> nearest canonical fine index
        def fidx(m, w=w):
            x = (np.arange(m) - 2 - w.fw - 1) * w.dh
            return np.clip(np.rint(x / fine.dh).astype(int) + fine.fw + 1 + 2, 0, shape[0] - 1)
        j, i, k = fidx(w.ny + 5), fidx(w.nx + 5), fidx(w.nz + 5)
        vp = vp_f[np.ix_(j, i, k)]
        vs = vp / ratio_f[np.ix_(j, i, k)]
        rho = 310.0 * vp ** 0.25
        out.append([np.ascontiguousarray(t, np.float32) for t in (rho, rho * vp * vp, rho * vs * vs)])
    return out


# ---------------------------------------------------------------------------
# models (host, numpy/scipy FFT)
def _kgrid(shape):
    ks = [np.fft.fftfreq(s) * 2 * np.pi for s in shape[:-1]]
    ks.append(np.fft.rfftfreq(shape[-1]) * 2 * np.pi)
    return ks[0][:, None, None] ** 2 + ks[1][None, :, None] ** 2 + ks[2][None, None, :] ** 2


def _grf(shape, corr_cells, seed):
    """zero-mean Gaussian random field with covariance exp(-r^2/l^2), via FFT"""
    import scipy.fft as sf
    w = np.random.default_rng(seed).standard_normal(shape).astype(np.float32)
    W = sf.rfftn(w, workers=-1)
    W *= np.exp(-_kgrid(shape) * corr_cells ** 2 / 8.0).astype(np.float32)
    return sf.irfftn(W, s=shape, workers=-1).astype(np.float64)


def _minmax(a, lo, hi):
    amin, amax = a.min(), a.max()
    return lo + (hi - lo) * (a - amin) / (amax - amin)


def _smooth(a, sigma_cells):
    import scipy.fft as sf
    A = sf.rfftn(a, workers=-1)
    A *= np.exp(-0.5 * _kgrid(a.shape) * sigma_cells ** 2)
    return sf.irfftn(A, s=a.shape, workers=-1)


def random_models(w: Workload, corr=20.0, smooth_cells=2.0):
    """(true, start) models as canonical-shape numpy arrays (rho, pi, u)"""
    shape = (w.ny + 5, w.nx + 5, w.nz + 5)
    vp = _minmax(_grf(shape, corr, w.seed), 2500.0, 5500.0)
    ratio = _minmax(_grf(shape, corr, w.seed + 1000), 1.7, 2.0)

    def moduli(vp_):
        vs_ = vp_ / ratio
        rho_ = 310.0 * vp_ ** 0.25
        return [np.ascontiguousarray(t, np.float32) for t in (rho_, rho_ * vp_ * vp_, rho_ * vs_ * vs_)]

    return moduli(vp), moduli(_smooth(vp, smooth_cells))


def sphere_models(w: Workload):
    """toy84: homogeneous start, two +-10% spheres of radius 8 cells (numpy)"""
    shape = (w.ny + 5, w.nx + 5, w.nz + 5)
    J, I, K = np.meshgrid(*(np.arange(s) - 2 for s in shape), indexing="ij")
    fac = np.ones(shape)
    for (ci, sgn) in ((22, 1.0), (42, -1.0)):
        c = (w.fw + 32, w.fw + ci, w.fw + 32)
        fac += sgn * 0.1 * (((J - c[0]) ** 2 + (I - c[1]) ** 2 + (K - c[2]) ** 2) <= 64)

    def moduli(f):
        vp, vs, rho = 4000.0 * f, 2300.0 * f, 2000.0 * f
        return [a.astype(np.float32) for a in (rho, rho * vp * vp, rho * vs * vs)]

    return moduli(fac), moduli(np.ones(shape))
