"""Failure paths the reference has: the stability check of checkfd_ssg.c
(:77-158, :473-510, which aborts when DT exceeds the limit) and the NaN probe
of the time loop (ifos3d.c:761, which aborts when vy at the centre cell is
NaN).  The native ABI returns IFOS_EUNSTABLE and IFOS_ENAN instead of
aborting.  The unstable runs are ordinary float arithmetic growing to inf and
NaN: every kernel still runs to its end.
"""
import numpy as np
import pytest

from helpers import random_model, small_grid

pytestmark = pytest.mark.gpu

GAMMA = np.float32(1.184614)   # checkfd_ssg.c:477, FDORDER 4


def _geom(ig, g, dt=None, nt=None):
    return ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt if dt is None else dt,
                   g.nt if nt is None else nt, g.fdcoeff, g.free_surf)


def _dtstab(dh, cmax):
    return np.float32(np.float32(dh) / (np.float64(GAMMA) * np.sqrt(3.0) * np.float64(cmax)))


def test_stability_homogeneous_closed_form():
    """homogeneous vp 4000: cmax is vp exactly, the limit dh / (gamma sqrt(3)
    vp); DT just below passes, just above is IFOS_EUNSTABLE"""
    import ifosgpu as ig
    g = small_grid(dh=2.0)
    rho = np.full(g.shape, 2000.0, np.float32)
    pi = (rho * np.float32(4000.0 ** 2)).astype(np.float32)
    u = (rho * np.float32(2300.0 ** 2)).astype(np.float32)
    lim = _dtstab(2.0, 4000.0)
    for dt, want in ((float(lim) * 0.999, ig.OK), (float(lim) * 1.001, ig.EUNSTABLE)):
        G = _geom(ig, g, dt=dt)
        rc, cmax, ds = ig.check_stability(G, *[G.to_dev(a) for a in (rho, pi, u)])
        assert cmax == 4000.0 and ds == lim
        assert rc == want, (dt, rc)


@pytest.mark.parametrize("L", [0, 1, 3])
def test_stability_random_model(L):
    """cmax over a random model (and random tau with L mechanisms) equals the
    largest of the reference's per-cell phase velocities formed in numpy;
    halo cells (outside [1..N]^3) are left out even when they are faster"""
    import ifosgpu as ig
    g = small_grid()
    rng = np.random.default_rng(40 + L)
    rho, pi, u = random_model(g, rng)
    taup = rng.uniform(0.01, 0.08, g.shape).astype(np.float32)
    taus = rng.uniform(0.01, 0.08, g.shape).astype(np.float32)
    pi[0, 0, 0] = 1e12  # a halo cell: not part of the model
    I = (slice(2 + 1, 2 + g.ny + 1), slice(2 + 1, 2 + g.nx + 1), slice(2 + 1, 2 + g.nz + 1))
    r, p, s = [a[I].astype(np.float64) for a in (rho, pi, u)]
    if L:
        tp = (np.float32(L) * taup[I]).astype(np.float64)
        ts = (np.float32(L) * taus[I]).astype(np.float64)
        cp = np.sqrt(p * (1.0 + tp) / r).astype(np.float32)
        cs = np.sqrt(s * (1.0 + ts) / r).astype(np.float32)
    else:
        cp = np.sqrt((pi[I] / rho[I]).astype(np.float64)).astype(np.float32)
        cs = np.sqrt((u[I] / rho[I]).astype(np.float64)).astype(np.float32)
    want = max(cp.max(), cs.max())
    G = _geom(ig, g)
    d = [G.to_dev(a) for a in (rho, pi, u, taup, taus)]
    rc, cmax, ds = ig.check_stability(G, d[0], d[1], d[2], d[3] if L else None, d[4] if L else None, L)
    assert cmax == want
    assert ds == _dtstab(min(g.dx, g.dy, g.dz), want)
    assert rc == (ig.OK if g.dt <= ds else ig.EUNSTABLE)


def _unstable_setup(ig):
    g = small_grid(nx=30, ny=30, nz=30, fw=6, nt=600, dh=2.0)
    rho = np.full(g.shape, 2000.0, np.float32)
    pi = (rho * np.float32(4000.0 ** 2)).astype(np.float32)
    u = (rho * np.float32(2300.0 ** 2)).astype(np.float32)
    dt = 2.0 * float(_dtstab(2.0, 4000.0))   # twice the limit
    G = _geom(ig, g, dt=dt)
    src = (g.nx // 2, g.ny // 2, g.nz // 2, 0.0, 50.0, 1.0, 1)
    rec = np.array([[g.nx // 2 + 3, g.ny // 2, g.nz // 2]], np.int32)
    return g, G, (rho, pi, u), src, rec


def test_unstable_dt_is_reported_as_nan_plain():
    """DT at twice the limit: the stability check says so, and a shot run
    anyway stops at the NaN probe with IFOS_ENAN"""
    import ifosgpu as ig
    g, G, model, src, rec = _unstable_setup(ig)
    d = [G.to_dev(a) for a in model]
    assert ig.check_stability(G, *d)[0] == ig.EUNSTABLE
    ctx = ig.ShotContext(G, 1, len(rec))
    ctx.set_cpml(4000.0, 50.0)
    ctx.set_freqs([50.0], 100)
    ctx.set_model(*d)
    d_syn = ig.DBuf((3, len(rec), g.nt + 1))
    print("plain shot at twice the stability limit", flush=True)
    with pytest.raises(ig.IfosError) as e:
        ctx.run_shot(src, rec, d_syn, None, want_adjoint=False)
    assert e.value.code == ig.ENAN, str(e.value)
    ig.sync()
    ctx.close()


def test_unstable_dt_is_reported_as_nan_batched():
    import ifosgpu as ig
    g, G, model, src, rec = _unstable_setup(ig)
    b = ig.BatchContext(G, 2, 1, len(rec), ig.DFT_COMPAT, 4)
    b.set_cpml(4000.0, 50.0)
    b.set_freqs([50.0], 100)
    b.set_model(*[G.to_dev(a) for a in model])
    d_syn = ig.DBuf((2, 3, len(rec), g.nt + 1))
    print("batched shots at twice the stability limit", flush=True)
    with pytest.raises(ig.IfosError) as e:
        b.run([src, src], rec, d_syn=d_syn, want_adjoint=False)
    assert e.value.code == ig.ENAN, str(e.value)
    ig.sync()
    b.close()


def test_unstable_dt_stops_the_dropin_driver():
    """the drop-in behind an IFOS3D-shaped caller (tests/dropin/nr_driver.c),
    DT at twice the limit: the caller's own NaN check of vy at the centre cell
    (ifos3d.c:761) stops the forward loop, from the probe value the drop-in
    keeps in the caller's array (one step behind the device)"""
    import ctypes as C
    import os
    import ifosgpu as ig
    g, G, model, src, rec = _unstable_setup(ig)
    lib = C.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "dropin", "lib", "libnr_driver.so"))
    p = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
    rho, pi, u = [np.ascontiguousarray(a, np.float32) for a in model]
    si = np.array([src[0], src[1], src[2], src[6]], np.int32)
    sf = np.array([src[3], src[4], src[5]], np.float32)
    fv = np.array([50.0], np.float32)
    syn = np.zeros((3, len(rec), g.nt + 1), np.float32)
    grads = np.zeros((3,) + g.shape, np.float32)
    l2 = C.c_float(0.0)
    print("drop-in shot at twice the stability limit", flush=True)
    rc = lib.dd_run_shot(g.nx, g.ny, g.nz, g.fw, C.c_float(g.dx), C.c_float(g.dy), C.c_float(g.dz),
                         C.c_float(G.c.dt), g.nt, g.fdcoeff, g.free_surf, p(rho), p(pi), p(u), p(si), p(sf), 1,
                         len(rec), p(np.ascontiguousarray(rec)), None, p(fv), 1, 100, C.c_float(4000.0),
                         C.c_float(50.0), p(syn), p(grads), C.byref(l2))
    assert rc == -1
