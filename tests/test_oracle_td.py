"""Pins of the oracle's time-domain gradient (og_td_gradient).

The time-domain gradient has no counterpart in the reference; it is defined
from the reference's own pieces, and these tests tie it to them:

* per time sample it is gradient_F's per-cell arithmetic (the restatement
  that the reference's golden vectors pin): og_td_gradient on the velocity
  histories equals, sample by sample, og_gradient_F applied to one "DFT
  field" holding the sample (real part only), with its frequency factor
  divided out and the lambda/mu terms taken from the displacements and the
  rho term from the velocities.  The displacements (sums from the end of
  the record) and the forward/adjoint pairing n <-> NT-1-n are formed here
  independently in numpy;
* summed over every DFT bin of the band, the reference's frequency-domain
  gradient is it: gradient_F over the frequencies k / (P DT), P = NT + 2
  (the period that holds both records' DFT times without overlap), equals
  (P DT / 2) times the time-domain gradient (Parseval over the positive
  bins; the 1 / (4 pi^2 f^2) of gradient_F turns velocity products into
  displacement products) once the wavefield has left the grid by the end
  of the record.  This pins the sign, the pairing and the weighting against
  the reference's own definition.
"""
import numpy as np
import pytest

from helpers import random_model, small_grid
from pyoracle import Oracle

SRC = (14, 11, 15, 0.0, 60.0, 1.0, 1)
REC = np.array([[8 + 3 * r, 9, 15] for r in range(5)] + [[20, 16, 10]], np.int32)


def _shot(o, g, seed=3):
    rng = np.random.default_rng(seed)
    true = random_model(g, rng)
    start = random_model(g, rng, pert=0.02)
    obs = o.run_shot(g, *[a.astype(o.dtype) for a in true], SRC, REC, [40.0], want_adjoint=False)["syn"]
    return start, obs


def _decompose(gr, rho, pi, u, interior):
    """(lambda, mu, rho) terms from gradient_F's vp, vs, rho outputs"""
    r, p, m = (interior(a).astype(np.float64) for a in (rho, pi, u))
    g1, g2, g3 = (interior(a).astype(np.float64) for a in gr)
    gl = g1 / (2.0 * np.sqrt(r * p))
    gm = (g2 + 4.0 * np.sqrt(r * m) * gl) / (2.0 * np.sqrt(r * m))
    grho = g3 - m / r * gm - (p - 2.0 * m) / r * gl
    return gl, gm, grho


def test_td_gradient_is_gradient_F_per_sample():
    o = Oracle("f64")
    g = small_grid(nx=22, ny=20, nz=24, fw=4, nt=24)
    (rho, pi, u), obs = _shot(o, g)
    out = o.run_shot(g, rho.astype(np.float64), pi.astype(np.float64), u.astype(np.float64), SRC, REC, [40.0],
                     obs=obs, want_history=True)
    vf, va = out["vfwd"], out["vadj"]
    N, dt = g.nt, np.float64(np.float32(g.dt))
    grads = [o.zeros(g) for _ in range(3)]
    o.td_gradient(g, vf, va, rho, pi, u, grads)
    # displacements from the end of the record; adjoint step m pairs with forward n = N - 1 - m
    uf = np.cumsum(vf[::-1], axis=0)[::-1] * dt      # uf[n-1] = sum_{n' >= n} vf[n'-1] dt
    ub = np.cumsum(va, axis=0) * dt                  # ub[m-1] = sum_{m' <= m} va[m'-1] dt
    f0 = np.float32(1.0 / (2 * np.pi))
    fd = 1.0 / (2.0 * 2.0 * np.float64(f0) * np.float64(f0) * np.pi * np.pi)
    zero = o.zeros(g)
    acc = [np.zeros_like(g.interior(zero), np.float64) for _ in range(3)]
    rho64, pi64, u64 = (a.astype(np.float64) for a in (rho, pi, u))
    for m in range(1, N - 1):
        n = N - 1 - m
        per = []
        for F, B in ((uf[n - 1], ub[m - 1]), (vf[n - 1], va[m - 1])):
            F, B = np.ascontiguousarray(F), np.ascontiguousarray(B)
            gr = [o.zeros(g) for _ in range(3)]
            o.gradient_F(g, [F[0], F[1], F[2], zero, zero, zero], [B[0], B[1], B[2], zero, zero, zero], gr,
                         rho64, pi64, u64, [f0])
            per.append(_decompose(gr, rho64, pi64, u64, g.interior))
        acc[0] += per[0][0] / fd * dt
        acc[1] += per[0][1] / fd * dt
        acc[2] += per[1][2] * dt
    r, p, mu = (g.interior(a).astype(np.float64) for a in (rho, pi, u))
    want = [2.0 * np.sqrt(r * p) * acc[0], -4.0 * np.sqrt(r * mu) * acc[0] + 2.0 * np.sqrt(r * mu) * acc[1],
            acc[2] + mu / r * acc[1] + (p - 2.0 * mu) / r * acc[0]]
    for n in range(3):
        got = g.interior(grads[n])
        scale = np.max(np.abs(want[n]))
        assert scale > 0
        assert np.max(np.abs(got - want[n])) <= 1e-9 * scale, f"gradient {n + 1}"


def test_td_gradient_is_the_full_band_frequency_gradient():
    """a record long enough for the wavefield to leave through the CPML, all
    DFT bins of the source band: the time-domain gradient times P DT / 2 is
    the reference's frequency-domain gradient (measured: cosine 1.0000, scale
    within 0.5 % of P DT / 2)"""
    o = Oracle("f64")
    g = small_grid(nx=30, ny=30, nz=30, fw=6, nt=200, dh=2.0, dt=2e-4)
    J, I, K = np.meshgrid(*(np.arange(s) - 2.0 for s in g.shape), indexing="ij")
    blob = np.exp(-((J - 18) ** 2 + (I - 15) ** 2 + (K - 15) ** 2) / 18.0)

    def model(f):
        vp, vs, rho = 4000.0 * f, 2300.0 * f, 2000.0 * np.ones(g.shape)
        return [np.ascontiguousarray(a, np.float64) for a in (rho, rho * vp * vp, rho * vs * vs)]

    src = (10, 10, 12, 0.0, 60.0, 1.0, 1)
    rec = np.array([[8 + 2 * n, 9, 8 + 2 * m] for n in range(8) for m in range(8)], np.int32)
    obs = o.run_shot(g, *model(1 + 0.05 * blob), src, rec, [40.0], want_adjoint=False)["syn"]
    rho, pi, u = model(np.ones(g.shape))
    out = o.run_shot(g, rho, pi, u, src, rec, [40.0], obs=obs, want_history=True)
    td = [o.zeros(g) for _ in range(3)]
    o.td_gradient(g, out["vfwd"], out["vadj"], rho, pi, u, td)
    P = g.nt + 2
    dt = float(np.float32(g.dt))
    finv = [k / (P * dt) for k in range(1, 31)]  # 24.75 .. 742 Hz: the whole band of a 60 Hz Ricker
    fq = o.run_shot(g, rho, pi, u, src, rec, finv, tast=2, obs=obs)["grads"]
    for n in range(3):
        a = g.interior(td[n]).ravel()
        b = g.interior(fq[n]).ravel()
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
        scale = float(a @ b / (a @ a))
        print(f"gradient {n + 1}: cosine {cos:.5f}, scale {scale:.6f} (P DT / 2 = {P * dt / 2:.6f})")
        assert cos > 0.999, f"gradient {n + 1}: cosine {cos}"
        assert abs(scale / (P * dt / 2) - 1) < 0.02, f"gradient {n + 1}: scale {scale}"


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_td_gradient_zero_without_residual(prec):
    """observed data equal to the synthetics: no adjoint field, no gradient"""
    o = Oracle(prec)
    g = small_grid(nx=20, ny=18, nz=22, fw=4, nt=16)
    rng = np.random.default_rng(1)
    rho, pi, u = (a.astype(o.dtype) for a in random_model(g, rng))
    syn = o.run_shot(g, rho, pi, u, (10, 9, 11, 0.0, 60.0, 1.0, 1), REC[:3] - 2, [40.0],
                     want_adjoint=False)["syn"]
    out = o.run_shot(g, rho, pi, u, (10, 9, 11, 0.0, 60.0, 1.0, 1), REC[:3] - 2, [40.0], obs=syn,
                     want_history=True)
    td = [o.zeros(g) for _ in range(3)]
    o.td_gradient(g, out["vfwd"], out["vadj"], rho, pi, u, td)
    assert all(np.all(t == 0) for t in td)
