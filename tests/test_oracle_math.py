"""Closed-form checks of the oracle, independent of the reference build.

* staggered 4th-order differences with Taylor coefficients are exact for
  cubic polynomials, evaluated half a cell forward (sxx_x, syy_y, szz_z) or
  backward (sxy_x, sxy_y, ...) of the cell: pins b1 = 9/8, b2 = -1/24, the
  staggering direction of every term and the DT/DX scaling;
* the DFT of a constant field over n steps is a geometric series of the
  sampled cos/sin at t = nt*DT (forward) or TIME - nt*DT (adjoint);
* CPML profiles at the left-side grid points follow the published
  Komatitsch-Martin formulas (d0, cubic d, K, alpha, b, a);
* harmonic/arithmetic material averages of constants are the constants;
* the residual is the time-reversed running integral of (syn - obs).
"""
import numpy as np
import pytest

from helpers import det_uniform, small_grid
from pyoracle import Grid, Oracle


def _poly(x, c):
    return c[0] + c[1] * x + c[2] * x ** 2 + c[3] * x ** 3


def _dpoly(x, c):
    return c[1] + 2 * c[2] * x + 3 * c[3] * x ** 2


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_velocity_update_exact_on_cubics(prec):
    o = Oracle(prec)
    g = small_grid(nx=20, ny=18, nz=22, fw=4, fdcoeff=1, dh=1.0, dt=0.5)
    g.dx, g.dy, g.dz = 1.0, 1.25, 0.75  # exact in float32, as the oracle's grid holds them
    shp = g.shape
    J, I, K = np.meshgrid(*(np.arange(s) - 2.0 for s in shp), indexing="ij")
    x, y, z = I * g.dx, J * g.dy, K * g.dz
    cx, cy, cz = (0.3, -0.2, 0.05, 0.01), (0.1, 0.4, -0.03, 0.02), (-0.2, 0.1, 0.04, -0.015)
    dt = o.dtype
    f = [np.zeros(shp, dt) for _ in range(9)]
    f[3][:] = _poly(x, cx)                       # sxx(x)
    f[4][:] = _poly(y, cy)                       # syy(y)
    f[5][:] = _poly(z, cz)                       # szz(z)
    f[6][:] = _poly(x, cy) + _poly(y, cz)        # sxy(x, y)
    f[7][:] = _poly(y, cx) + _poly(z, cy)        # syz(y, z)
    f[8][:] = _poly(x, cz) + _poly(z, cx)        # sxz(x, z)
    rho = np.full(shp, 2.0, dt)
    box = o.box(g)
    o.update_v(g, box, 1, f, rho, rho, rho, [], None, None, None, 0)
    sl = (slice(3 + box[2] - 1, 3 + box[3]), slice(3 + box[0] - 1, 3 + box[1]), slice(3 + box[4] - 1, 3 + box[5]))
    xs, ys, zs = x[sl], y[sl], z[sl]
    h = 0.5
    # vx: d/dx sxx at x+dx/2, d/dy sxy at y-dy/2, d/dz sxz at z-dz/2
    ex = g.dt * (_dpoly(xs + h * g.dx, cx) + _dpoly(ys - h * g.dy, cz) + _dpoly(zs - h * g.dz, cx)) / 2.0
    ey = g.dt * (_dpoly(ys + h * g.dy, cy) + _dpoly(xs - h * g.dx, cy) + _dpoly(zs - h * g.dz, cy)) / 2.0
    ez = g.dt * (_dpoly(zs + h * g.dz, cz) + _dpoly(xs - h * g.dx, cz) + _dpoly(ys - h * g.dy, cx)) / 2.0
    tol = 2e-5 if prec == "f32" else 1e-11
    for got, want in ((f[0][sl], ex), (f[1][sl], ey), (f[2][sl], ez)):
        assert np.max(np.abs(got - want)) <= tol * np.max(np.abs(want))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_stress_update_exact_on_cubics(prec):
    o = Oracle(prec)
    g = small_grid(nx=20, ny=18, nz=22, fw=4, fdcoeff=1, dh=1.0, dt=0.25)
    g.dx, g.dy, g.dz = 0.875, 1.125, 1.25
    shp = g.shape
    J, I, K = np.meshgrid(*(np.arange(s) - 2.0 for s in shp), indexing="ij")
    x, y, z = I * g.dx, J * g.dy, K * g.dz
    c1, c2, c3 = (0.1, 0.3, -0.02, 0.01), (-0.2, 0.1, 0.05, -0.01), (0.05, -0.3, 0.01, 0.02)
    dt = o.dtype
    f = [np.zeros(shp, dt) for _ in range(9)]
    f[0][:] = _poly(x, c1) + _poly(y, c2)   # vx(x, y)
    f[1][:] = _poly(y, c3) + _poly(z, c1)   # vy(y, z)
    f[2][:] = _poly(z, c2) + _poly(x, c3)   # vz(z, x)
    lam2mu, mu = 3.0, 1.0
    pi = np.full(shp, lam2mu, dt)
    u = np.full(shp, mu, dt)
    box = o.box(g)
    o.update_s(g, box, f, pi, u, u, u, u)
    sl = (slice(3 + box[2] - 1, 3 + box[3]), slice(3 + box[0] - 1, 3 + box[1]), slice(3 + box[4] - 1, 3 + box[5]))
    xs, ys, zs = x[sl], y[sl], z[sl]
    h = 0.5
    vxx = _dpoly(xs - h * g.dx, c1)        # backward in x
    vyy = _dpoly(ys - h * g.dy, c3)        # backward in y
    vzz = _dpoly(zs - h * g.dz, c2)        # backward in z
    vxy = _dpoly(ys + h * g.dy, c2)        # forward in y
    vyz = _dpoly(zs + h * g.dz, c1)        # forward in z
    vzx = _dpoly(xs + h * g.dx, c3)        # forward in x
    div = vxx + vyy + vzz
    want = {3: g.dt * (lam2mu * div - 2 * mu * (vyy + vzz)), 4: g.dt * (lam2mu * div - 2 * mu * (vxx + vzz)),
            5: g.dt * (lam2mu * div - 2 * mu * (vxx + vyy)), 6: g.dt * mu * vxy,
            7: g.dt * mu * vyz, 8: g.dt * mu * vzx}
    tol = 2e-5 if prec == "f32" else 1e-11
    for n, w in want.items():
        assert np.max(np.abs(f[n][sl] - w)) <= tol * max(np.max(np.abs(w)), 1e-30), n


@pytest.mark.parametrize("back", [0, 1])
def test_dft_geometric_series(back):
    o = Oracle("f64")
    g = small_grid(nx=9, ny=9, nz=9, fw=2, nt=400, dt=2.5e-4)
    one = np.ones(g.shape)
    finv = [30.0, 55.0]
    Fr = [np.zeros((2,) + g.shape) for _ in range(3)]
    Fi = [np.zeros((2,) + g.shape) for _ in range(3)]
    ntast = 3
    steps = [n for n in range(1, g.nt + 1) if n % ntast == 0]
    for nt in steps:
        o.discfourier(g, nt, one, one, 2 * one, Fr, Fi, finv, ntast, back)
    dt = np.float32(g.dt)
    n = np.array(steps, np.float32)
    # sample times are float, as the reference computes them (disc_fourier.c:36-37)
    t = (n * dt).astype(np.float32) if back == 0 else (np.float32(g.time) - (n * dt).astype(np.float32))
    for l, fr in enumerate(finv):
        w = 2 * np.pi * float(fr)
        # sum of exp(i w t) over the sampled steps, times DT * ntast
        z = np.sum(np.exp(1j * w * t.astype(np.float64))) * float(dt) * ntast
        want_c, want_s = z.real, z.imag
        c = g.interior(Fr[0][l])
        s = g.interior(Fi[0][l])
        assert np.allclose(c, want_c, rtol=1e-6, atol=1e-9)
        assert np.allclose(s, want_s, rtol=1e-6, atol=1e-9)
        assert np.allclose(g.interior(Fr[2][l]), 2 * want_c, rtol=1e-6, atol=1e-9)


def test_cpml_profiles_follow_published_formulas():
    o = Oracle("f64")
    fw, dh, vppml, fpml, dt = 10, 3.0, 5000.0, 20.0, 3e-4
    g = small_grid(nx=30, ny=30, nz=30, fw=fw, dh=dh, dt=dt)
    prof, alpha = o.cpml_coeff(g, vppml, fpml)
    K, a, b, Kh, ah, bh = prof[0:6]
    L = fw * dh
    d0 = -(3 + 1) * vppml * np.log(0.0008) / (2 * L)
    amax = 2 * np.pi * fpml / 2
    for i in range(1, fw + 1):
        for pos, Kv, av, bv, al in (((fw - i + 1) * dh, K[i], a[i], b[i], alpha[0][i]),
                                    ((fw - i + 0.5) * dh, Kh[i], ah[i], bh[i], alpha[1][i])):
            p = pos / L
            d = d0 * (0.25 * p + 0.75 * p ** 3)
            k = 1 + (2 - 1) * p ** 3
            alp = amax * (1 - p)
            bb = np.exp(-(d / k + alp) * dt)
            aa = d * (bb - 1) / (k * (d + k * alp))
            assert np.isclose(Kv, k, rtol=1e-6) and np.isclose(al, alp, rtol=1e-6)
            assert np.isclose(bv, bb, rtol=1e-6) and np.isclose(av, aa, rtol=1e-6)
    # K is 1 and b is exp(-alpha_max dt) at the inner edge of the layer
    assert K[fw + 1] == 1.0 and np.isclose(b[fw + 1], np.exp(-amax * dt), rtol=1e-9)


def test_material_averages_of_constants():
    o = Oracle("f32")
    g = small_grid(nx=10, ny=11, nz=12, fw=2)
    rho = np.full(g.shape, 2100.0, np.float32)
    u = np.full(g.shape, 3.2e9, np.float32)
    avg = o.av_mat(g, rho, u)
    for n in range(3):
        assert np.allclose(g.interior(avg[n]), 3.2e9, rtol=1e-6)
    for n in range(3, 6):
        assert np.array_equal(g.interior(avg[n]), g.interior(rho))


def test_residual_is_reversed_running_integral():
    o = Oracle("f32")
    g = small_grid(nt=64)
    obs = det_uniform((4, g.nt + 1), 3, -1, 1)
    syn = det_uniform((4, g.nt + 1), 4, -1, 1)
    diff, l2 = o.residual(g, obs, syn)
    run = np.float32(g.dt) * np.cumsum((syn[:, 1:] - obs[:, 1:]).astype(np.float32), axis=1, dtype=np.float32)
    assert np.array_equal(diff[:, 1:], run[:, ::-1])
    assert np.isclose(l2, float(np.sum(run.astype(np.float64) ** 2)), rtol=1e-5)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_visco_without_relaxation_is_elastic(prec):
    """taus = taup = 0 and r = 0: update_s_ssg.c reduces to the elastic update
    (1 + L*tau = 1, the memory variables stay 0), bit for bit"""
    o = Oracle(prec)
    g = Grid(nx=12, ny=11, nz=13, fw=3, dx=2.0, dy=2.5, dz=3.0, dt=2e-4, nt=5)
    rng = np.random.default_rng(4)
    f = [rng.standard_normal(g.shape).astype(o.dtype) for _ in range(9)]
    rho = (2000 + 100 * rng.random(g.shape)).astype(o.dtype)
    pi = (rho * 1.6e7).astype(o.dtype)
    u = (rho * 5e6 * (1 + 0.1 * rng.random(g.shape))).astype(o.dtype)
    uipjp, ujpkp, uipkp, _, _, _ = o.av_mat(g, rho, u)
    zero = o.zeros(g)
    r6 = [o.zeros(g) for _ in range(6)]
    V = o.visco(r6, zero, zero, zero, zero, zero, 0.02)
    box = [1, g.nx, 1, g.ny, 1, g.nz]
    fe = [a.copy() for a in f]
    o.update_s(g, box, fe, pi, u, uipjp, ujpkp, uipkp)
    o.update_s_visco(g, box, f, pi, u, uipjp, ujpkp, uipkp, V)
    for n in range(9):
        assert np.array_equal(f[n], fe[n]), n
    assert all(not r.any() for r in r6)


def test_memory_variables_decay_geometrically():
    """with no strain the memory variables follow r <- b*c*r with
    b = 1/(1 + eta/2), c = 1 - eta/2 (update_s_ssg.c:211-224): after n steps
    r_n = r_0 * ((1 - eta/2)/(1 + eta/2))^n (FP64 oracle, to rounding)"""
    o = Oracle("f64")
    g = Grid(nx=10, ny=10, nz=10, fw=2, dx=2.0, dy=2.0, dz=2.0, dt=2e-4, nt=5)
    rng = np.random.default_rng(5)
    f = [o.zeros(g) for _ in range(9)]  # zero velocities: no strain
    rho = np.full(g.shape, 2000.0)
    pi = np.full(g.shape, 3.2e10)
    u = np.full(g.shape, 1.0e10)
    uipjp, ujpkp, uipkp, _, _, _ = o.av_mat(g, rho, u)
    taus = np.full(g.shape, 0.04)
    taup = np.full(g.shape, 0.03)
    tsa = o.av_tau(g, rho, u, taus)
    r0 = [rng.standard_normal(g.shape) for _ in range(6)]
    r6 = [a.copy() for a in r0]
    eta = 0.05
    V = o.visco(r6, taus, taup, *tsa, eta)
    box = [1, g.nx, 1, g.ny, 1, g.nz]
    n = 7
    for _ in range(n):
        o.update_s_visco(g, box, f, pi, u, uipjp, ujpkp, uipkp, V)
    q = ((1 - eta / 2) / (1 + eta / 2)) ** n
    for a, b in zip(r6, r0):
        np.testing.assert_allclose(g.interior(a), g.interior(b) * q, rtol=1e-13)
    # the stress picks up dt/2 * (r_old + r_new) per step: sxx = dt/2 * sum_k (r_k + r_{k+1})
    qq = (1 - eta / 2) / (1 + eta / 2)
    s = sum(qq ** m + qq ** (m + 1) for m in range(n)) * float(np.float32(g.dt)) / 2  # DT is a float global
    np.testing.assert_allclose(g.interior(f[3]), g.interior(r0[0]) * s, rtol=1e-12)
