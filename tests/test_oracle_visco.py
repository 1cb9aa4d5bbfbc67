"""The oracle's viscoelastic update with L relaxation mechanisms, pinned by
the closed form it discretises.

A cell of the stress update is driven by a uniform strain rate
vxx = eps0 w cos(w t) (vx linear in x, Taylor coefficients, so the 4th-order
difference is exact and every other derivative is zero).  After the
transients of the memory variables have decayed, sxx = Re(M) eps0 sin(wt) +
Im(M) eps0 cos(wt) with M(w) the complex P modulus of the L-fold standard
linear solid of the tau-method: relaxed modulus pi, unrelaxed pi (1 + L tau),
relaxation times tau_sigma,l = 1 / (2 pi f_l) (eta_l = DT / tau_sigma,l) and
tau_eps,l = tau_sigma,l (1 + tau):

    M(w) / pi = 1 - L + sum_l (1 + i w tau_eps,l) / (1 + i w tau_sigma,l),

whose ratio Re / Im is the Q of the reference's mfiles/qgsls.m:1-12 (and
whose real part at the first relaxation frequency is the relaxed-to-phase-
velocity factor of readmod.c:58-64).  The measured Q and Re(M) must match
the closed form to 1e-3 for L = 1, 2 and 3 and several frequencies; a
mechanism summed in the wrong place, a wrong eta or a dropped half step of
the memory term moves them by far more.  The discretisation error is
O((w DT)^2, (DT / tau_sigma)^2) ~ 1e-5 here.
"""
import math

import numpy as np
import pytest

from pyoracle import Grid, Oracle

DT = 2e-5
FL = (8.0, 30.0, 110.0)  # relaxation frequencies of the mechanisms (Hz)
TAU = 0.05


def qgsls(te, ts, L, w):
    """mfiles/qgsls.m"""
    sumz = sumn = 0.0
    for l in range(L):
        d = 1.0 + w * w * ts[l] * ts[l]
        sumz += (1.0 + w * w * te[l] * ts[l]) / d
        sumn += w * (te[l] - ts[l]) / d
    return (1.0 - L + sumz) / sumn


def modulus(te, ts, L, w):
    return 1.0 - L + sum((1.0 + 1j * w * te[l]) / (1.0 + 1j * w * ts[l]) for l in range(L))


def _drive(L, f, periods=2):
    o = Oracle("f64")
    g = Grid(nx=8, ny=8, nz=8, fw=2, dx=1.0, dy=1.0, dz=1.0, dt=DT, nt=10, fdcoeff=1)
    shp = g.shape
    one = np.ones(shp)
    pi, u = 2.0e10 * one, 0.7e10 * one
    f9 = [np.zeros(shp) for _ in range(9)]
    x = ((np.arange(shp[1]) - 2) * g.dx)[None, :, None] * np.ones(shp)
    taus = TAU * one
    taup = TAU * one
    r6 = [np.zeros((L,) + shp) for _ in range(6)]
    ts = [1.0 / (2.0 * math.pi * fl) for fl in FL[:L]]
    V = o.visco(r6, taus, taup, taus, taus, taus, [DT / t for t in ts])
    box = [4, 4, 4, 4, 4, 4]
    c = (4 + 2, 4 + 2, 4 + 2)
    w = 2.0 * math.pi * f
    eps0 = 1e-6
    # transients: the slowest mechanism relaxes with tau_sigma = 1 / (2 pi 8 Hz) = 20 ms
    nsettle = int(0.4 / DT)
    nper = int(round(periods / f / DT))
    sxx = np.empty(nsettle + nper)
    for n in range(1, nsettle + nper + 1):
        f9[0][...] = eps0 * w * math.cos(w * (n - 0.5) * DT) * x
        o.update_s_visco(g, box, f9, pi, u, u, u, u, V)
        sxx[n - 1] = f9[3][c]
    t = np.arange(nsettle + 1, nsettle + nper + 1) * DT
    A = np.column_stack([np.sin(w * t), np.cos(w * t)])
    a, b = np.linalg.lstsq(A, sxx[nsettle:], rcond=None)[0]
    te = [t_ * (1.0 + TAU) for t_ in ts]
    return a / eps0 / 2.0e10, b / eps0 / 2.0e10, te, ts


@pytest.mark.parametrize("L,f", [(1, 12.0), (2, 25.0), (3, 5.0), (3, 40.0)])
def test_q_of_the_mechanisms_matches_the_closed_form(L, f):
    re, im, te, ts = _drive(L, f)
    w = 2.0 * math.pi * f
    M = modulus(te, ts, L, w)
    q_meas, q_want = re / im, qgsls(te, ts, L, w)
    print(f"L={L} f={f} Hz: Q {q_meas:.6f} (closed form {q_want:.6f}), Re M/pi {re:.7f} ({M.real:.7f})", flush=True)
    assert q_meas == pytest.approx(q_want, rel=1e-3)
    assert re == pytest.approx(M.real, rel=1e-4)
    assert abs(q_want - M.real / M.imag) <= 1e-12 * abs(q_want)


def test_one_mechanism_in_the_general_form_is_the_reference_form():
    """L = 1 written as one mechanism of the L-fold code and the L = 1 form
    give the same bits (f32 build)"""
    o = Oracle("f32")
    g = Grid(nx=10, ny=9, nz=11, fw=2, dx=2.0, dy=2.0, dz=2.0, dt=2e-4, nt=10, fdcoeff=2)
    rng = np.random.default_rng(3)
    shp = g.shape
    f9 = [(1e-3 * rng.standard_normal(shp)).astype(np.float32) for _ in range(9)]
    pi = (2e10 * (1 + 0.1 * rng.random(shp))).astype(np.float32)
    u = (0.7e10 * (1 + 0.1 * rng.random(shp))).astype(np.float32)
    taus = (0.02 + 0.03 * rng.random(shp)).astype(np.float32)
    r1 = [(1e-2 * rng.standard_normal(shp)).astype(np.float32) for _ in range(6)]
    r3 = [np.concatenate([a[None], np.zeros((2,) + shp, np.float32)]) for a in r1]
    box = o.box(g)
    fa = [a.copy() for a in f9]
    fb = [a.copy() for a in f9]
    Va = o.visco([a.copy() for a in r1], taus, taus, taus, taus, taus, np.float32(0.017))
    Vb = o.visco(r3, taus, taus, taus, taus, taus, np.array([0.017, 0.0, 0.0], np.float32), L=1)
    o.update_s_visco(g, box, fa, pi, u, u, u, u, Va)
    o.update_s_visco(g, box, fb, pi, u, u, u, u, Vb)
    for n in range(9):
        assert np.array_equal(fa[n], fb[n])
