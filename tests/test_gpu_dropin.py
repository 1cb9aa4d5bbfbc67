"""The drop-in boundary: the reference prototypes (update_v, update_v_CPML,
update_s_elastic, update_s_CPML_elastic, psource, seismo, surface_elastic,
discfourier, gradient_F, hess_F, av_mat, CPML_coeff, CPML_ini_elastic,
zero_wavefield, zero_invers, and the optimiser's lbfgs_savegrad, lbfgs,
conjugate, modelupdate) called on Numerical-Recipes arrays by an
IFOS3D-shaped driver (tests/dropin/nr_driver.c), in the order of
ifos3d.c:712-975 (and of its Hessian loop, ifos3d.c:880-983, for the
Hessian wavefields and hess_F).  Seismograms, misfit, gradients and
Hessians must equal the oracle's bit for bit."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from helpers import max_ulp
from pyoracle import Oracle
from golden import cases as cs

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dropin", "lib", "libnr_driver.so")


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def run_dropin(g, model, src, shape, rec, obs, finv, tast=100, vppml=4000.0, fpml=50.0):
    lib = C.CDLL(DRIVER)
    ntr = len(rec)
    syn = np.zeros((3, ntr, g.nt + 1), np.float32)
    grads = np.zeros((3,) + g.shape, np.float32)
    l2 = C.c_float(0.0)
    si = np.array([src[0], src[1], src[2], src[6]], np.int32)
    sf = np.array([src[3], src[4], src[5]], np.float32)
    recs = np.ascontiguousarray(rec, np.int32)
    fv = np.asarray(finv, np.float32)
    o = np.ascontiguousarray(np.stack(obs), np.float32) if obs is not None else None
    rho, pi, u = [np.ascontiguousarray(a, np.float32) for a in model]
    rc = lib.dd_run_shot(g.nx, g.ny, g.nz, g.fw, C.c_float(g.dx), C.c_float(g.dy), C.c_float(g.dz), C.c_float(g.dt),
                         g.nt, g.fdcoeff, g.free_surf, _p(rho), _p(pi), _p(u), _p(si), _p(sf), shape, ntr, _p(recs),
                         _p(o), _p(fv), len(fv), tast, C.c_float(vppml), C.c_float(fpml), _p(syn), _p(grads),
                         C.byref(l2))
    assert rc == 0
    return syn, grads, l2.value


@pytest.mark.parametrize("fs", [0, 1])
def test_dropin_gradient_shot_matches_oracle(fs):
    g = cs.shot_grid(fs)
    print(f"drop-in gradient shot free_surf={fs}", flush=True)
    o = Oracle("f32")
    start, true = cs.shot_models(g)
    obs = o.run_shot(g, *true, cs.SHOT_SRC, cs.SHOT_REC, cs.SHOT_FINV, want_adjoint=False)["syn"]
    ref = o.run_shot(g, *start, cs.SHOT_SRC, cs.SHOT_REC, cs.SHOT_FINV, obs=obs)
    syn_obs, _, _ = run_dropin(g, true, cs.SHOT_SRC, 1, cs.SHOT_REC, None, cs.SHOT_FINV)
    for c in range(3):
        assert max_ulp(syn_obs[c][:, 1:], obs[c][:, 1:]) == 0, f"observed comp {c}"
    syn, grads, l2 = run_dropin(g, start, cs.SHOT_SRC, 1, cs.SHOT_REC, obs, cs.SHOT_FINV)
    for c in range(3):
        assert max_ulp(syn[c][:, 1:], ref["syn"][c][:, 1:]) == 0, f"synthetic comp {c}"
    assert np.float32(l2) == np.float32(ref["l2"])
    for n in range(3):
        assert max_ulp(g.interior(grads[n]), g.interior(ref["grads"][n])) == 0, f"grad{n + 1}"


def test_dropin_force_source_forward():
    g = cs.shot_grid(0)
    print("drop-in force source forward", flush=True)
    o = Oracle("f32")
    start, _ = cs.shot_models(g)
    src = (10, 12, 14, 0.001, 45.0, 3.0, 3)
    ref = o.run_shot(g, *start, src, cs.SHOT_REC, [50.0], want_adjoint=False)
    syn, _, _ = run_dropin(g, start, src, 1, cs.SHOT_REC, None, [50.0])
    for c in range(3):
        assert max_ulp(syn[c][:, 1:], ref["syn"][c][:, 1:]) == 0


def test_dropin_viscoelastic_shot():
    """L = 1: the driver calls update_s / update_s_CPML (and av_mat forms the
    relaxation-ratio averages); the oracle's viscoelastic shot must match"""
    g = cs.shot_grid(0)
    print("drop-in viscoelastic gradient shot", flush=True)
    o = Oracle("f32")
    start, true = cs.shot_models(g)
    rng = np.random.default_rng(12)
    taus = (0.02 + 0.03 * rng.random(g.shape)).astype(np.float32)
    taup = (0.02 + 0.03 * rng.random(g.shape)).astype(np.float32)
    eta = np.float32(0.017)
    vis = (taus, taup, eta)
    obs = o.run_shot(g, *true, cs.SHOT_SRC, cs.SHOT_REC, cs.SHOT_FINV, want_adjoint=False, visco=vis)["syn"]
    ref = o.run_shot(g, *start, cs.SHOT_SRC, cs.SHOT_REC, cs.SHOT_FINV, obs=obs, visco=vis)
    lib = C.CDLL(DRIVER)
    lib.dd_set_visco(_p(taus), _p(taup), C.c_float(eta))
    try:
        syn, grads, l2 = run_dropin(g, start, cs.SHOT_SRC, 1, cs.SHOT_REC, obs, cs.SHOT_FINV)
    finally:
        lib.dd_set_visco(None, None, C.c_float(0.0))
    for c in range(3):
        assert max_ulp(syn[c][:, 1:], ref["syn"][c][:, 1:]) == 0, f"synthetic comp {c}"
    assert np.float32(l2) == np.float32(ref["l2"])
    for n in range(3):
        assert max_ulp(g.interior(grads[n]), g.interior(ref["grads"][n])) == 0, f"grad{n + 1}"


def test_dropin_hessian_matches_oracle():
    """discfourier with Hessian wavefields (pshot 1..2) and hess_F through the
    drop-in prototypes, against og_discfourier and og_hess_F"""
    g = cs.dft_grid()
    print("drop-in Hessian wavefields and hess_F", flush=True)
    lib = C.CDLL(DRIVER)
    o = Oracle("f32")
    rng = np.random.default_rng(17)
    nsteps, ntr_hess, finv = 4, 2, [40.0, 70.0]
    nf = len(finv)
    vf = (rng.standard_normal((nsteps, 3) + g.shape)).astype(np.float32)
    vb = (rng.standard_normal((ntr_hess, nsteps, 3) + g.shape)).astype(np.float32)
    for a in (vf, vb):  # the halos the velocity fields carry are 0
        a[..., :3, :, :] = 0; a[..., -2:, :, :] = 0
        a[..., :, :3, :] = 0; a[..., :, -2:, :] = 0
        a[..., :, :, :3] = 0; a[..., :, :, -2:] = 0
    rho, pi, u = (np.ascontiguousarray(x, np.float32) for x in cs.shot_models(g)[1])
    hess = np.zeros((3,) + g.shape, np.float32)
    fv = np.asarray(finv, np.float32)
    rc = lib.dd_hess(g.nx, g.ny, g.nz, g.fw, C.c_float(g.dx), C.c_float(g.dy), C.c_float(g.dz), C.c_float(g.dt),
                     g.nt, g.fdcoeff, _p(rho), _p(pi), _p(u), _p(vf), _p(vb), nsteps, ntr_hess, _p(fv), nf, _p(hess))
    assert rc == 0
    Ff = [np.zeros((nf,) + g.shape, np.float32) for _ in range(6)]
    Fb = [np.zeros((nf * (ntr_hess + 1),) + g.shape, np.float32) for _ in range(6)]
    for s in range(nsteps):
        o.discfourier(g, s + 1, *vf[s], Ff[:3], Ff[3:], finv, 1, 0)
    for p in range(1, ntr_hess + 1):
        sl = [np.ascontiguousarray(F[p * nf:(p + 1) * nf]) for F in Fb]
        for s in range(nsteps):
            o.discfourier(g, s + 1, *vb[p - 1][s], sl[:3], sl[3:], finv, 1, 1)
        for F, x in zip(Fb, sl):
            F[p * nf:(p + 1) * nf] = x
    want = [o.zeros(g) for _ in range(3)]
    o.hess_F(g, Ff, Fb, want, rho, pi, u, finv, ntr_hess)
    for n in range(3):
        w, got = g.interior(want[n]), g.interior(hess[n])
        assert np.abs(w).max() > 0
        assert max_ulp(got, w) == 0, f"hess{n + 1}"


@pytest.mark.parametrize("use_lbfgs,numpar", [(0, 2), (1, 1), (1, 2)])
def test_dropin_optimiser_matches_native(use_lbfgs, numpar):
    """lbfgs_savegrad / lbfgs or conjugate, then modelupdate, for four
    iterations through the drop-in prototypes on Numerical-Recipes arrays
    (the order of ifos3d.c:1005-1040): the model, the directions, beta and
    bfgsscale equal those of the same sequence through the device entry
    points (ifos_lbfgs_*, ifos_conjugate, ifos_modelupdate), which
    tests/test_gpu_optim.py pins to the oracle, bit for bit"""
    import ifosgpu as ig
    from helpers import small_grid
    g = small_grid(nx=18, ny=16, nz=20, fw=4)
    print(f"drop-in optimiser: lbfgs {use_lbfgs}, numpar {numpar}", flush=True)
    lib = C.CDLL(DRIVER)
    rng = np.random.default_rng(90 + 2 * use_lbfgs + numpar)
    niter, bfgsnum = 4, 3
    vp = 4000.0 * (1 + 0.05 * rng.random(g.shape))
    vs = vp / 1.8
    rho = 2000.0 * (1 + 0.05 * rng.random(g.shape))
    model = np.ascontiguousarray(np.stack([rho, rho * vp * vp, rho * vs * vs]), np.float32)
    grads = np.zeros((niter, 3) + g.shape, np.float32)
    inner = (slice(3, -2),) * 3  # cells 1..N of the canonical layout
    prev = rng.standard_normal((3,) + g.shape)
    for it in range(niter):
        prev = 0.7 * prev + 0.3 * rng.standard_normal((3,) + g.shape)
        for c in range(3):
            grads[it, c][inner] = prev[c][inner]
    steps = np.array([0.02, 0.015, 0.01, 0.012], np.float32)
    weight = np.array([1.0, 0.8, 0.5], np.float32)
    vp0, vs0, rho0 = 120.0, 70.0, 25.0
    m_out = np.zeros_like(model)
    d_out = np.zeros_like(grads)
    beta_out = np.zeros((niter, 3), np.float32)
    scale_out = np.zeros((niter, bfgsnum), np.float32)
    rc = lib.dd_optim(g.nx, g.ny, g.nz, g.fw, C.c_float(g.dx), C.c_float(g.dt), g.nt, _p(model), _p(grads), niter,
                      use_lbfgs, numpar, bfgsnum, _p(steps), _p(weight), C.c_float(vp0), C.c_float(vs0),
                      C.c_float(rho0), _p(m_out), _p(d_out), _p(beta_out), _p(scale_out))
    assert rc == 0

    G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dx, g.dx, g.dt, g.nt, 2, 0)
    dm = [G.to_dev(a) for a in model]
    dp = [G.to_dev(np.zeros(g.shape, np.float32)) for _ in range(6)]
    lb = ig.Lbfgs(G, numpar, bfgsnum) if use_lbfgs else None
    try:
        for it in range(niter):
            dg = [G.to_dev(a) for a in grads[it]]
            beta = np.zeros(3, np.float32)
            if lb is not None:
                g2 = dg[1] if numpar > 1 else None
                if it > 0:
                    lb.update(dg[0], g2, it + 1)
                else:
                    lb.savegrad(dg[0], g2)
            else:
                beta = np.asarray(ig.conjugate(G, dg, dp, 1 if it == 0 else 0)[0], np.float32)
            w = (it + 1) % bfgsnum or bfgsnum
            ig.modelupdate(G, dg, *dm, float(steps[it]), weight, vp0, vs0, rho0,
                           bfgsmod=lb.mod(w) if lb is not None else None, numpar=numpar)
            for c in range(3):
                want = g.interior(G.from_dev(dg[c]))
                assert np.abs(want).max() > 0
                assert max_ulp(g.interior(d_out[it, c]), want) == 0, f"direction {c + 1}, iteration {it + 1}"
            assert np.array_equal(beta_out[it], beta), f"beta, iteration {it + 1}"
            if lb is not None:
                assert np.array_equal(scale_out[it], lb.scale().astype(np.float32)), f"bfgsscale, iteration {it + 1}"
        if lb is not None:
            assert np.all(scale_out[-1] != 0)
        for c in range(3):
            want = g.interior(G.from_dev(dm[c]))
            assert not np.array_equal(want, g.interior(model[c]))
            assert max_ulp(g.interior(m_out[c]), want) == 0, f"model {c}"
    finally:
        if lb is not None:
            lb.close()
