"""The drop-in boundary: the reference prototypes (update_v, update_v_CPML,
update_s_elastic, update_s_CPML_elastic, psource, seismo, surface_elastic,
discfourier, gradient_F, hess_F, av_mat, CPML_coeff, CPML_ini_elastic,
zero_wavefield, zero_invers) called on Numerical-Recipes arrays by an
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
