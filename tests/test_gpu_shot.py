"""GPU parity of whole gradient shots (the plain path) against the oracle.

A shot is the call order of ifos3d.c:712-975: forward NT steps with the DFT
every ntast steps and seismograms every step, residual against observed
data, adjoint NT steps, gradient_F.  The plain path (one kernel per
reference routine per step) must reproduce the compat oracle bit for bit;
the path's stated tolerances (seismograms rel. L2 <= 1e-5, gradients
<= 1e-4) are asserted as well and reported.
"""
import sys

import numpy as np
import pytest

from helpers import max_ulp, rel_l2, small_grid
from pyoracle import Oracle

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _models(g, seed):
    """homogeneous start model and a true model with two +-10% spheres"""
    rng = np.random.default_rng(seed)
    shp = g.shape
    J, I, K = np.meshgrid(np.arange(shp[0]) - 2, np.arange(shp[1]) - 2, np.arange(shp[2]) - 2, indexing="ij")
    vp = np.full(shp, 4000.0)
    vs = np.full(shp, 2300.0)
    rho = np.full(shp, 2000.0)
    start = [rho.astype(np.float32), (rho * vp * vp).astype(np.float32), (rho * vs * vs).astype(np.float32)]
    r = max(2, g.nx // 8)
    c1 = (g.ny // 2, g.nx // 3, g.nz // 2)
    c2 = (g.ny // 2, 2 * g.nx // 3, g.nz // 2)
    s1 = (J - c1[0]) ** 2 + (I - c1[1]) ** 2 + (K - c1[2]) ** 2 <= r * r
    s2 = (J - c2[0]) ** 2 + (I - c2[1]) ** 2 + (K - c2[2]) ** 2 <= r * r
    fac = 1.0 + 0.1 * s1 - 0.1 * s2 + 0.01 * rng.uniform(-1, 1, shp) * 0
    vp2, vs2, rho2 = vp * fac, vs * fac, rho * fac
    true = [rho2.astype(np.float32), (rho2 * vp2 * vp2).astype(np.float32), (rho2 * vs2 * vs2).astype(np.float32)]
    return start, true


def _run_case(g, src, rec, finv, tast, vppml, fpml, precision="compat"):
    import ifosgpu as ig
    o = Oracle("f32")
    start, true = _models(g, 1)
    obs = o.run_shot(g, *true, src, rec, finv, tast=tast, vppml=vppml, fpml=fpml, want_adjoint=False)["syn"]
    ref = o.run_shot(g, *start, src, rec, finv, tast=tast, vppml=vppml, fpml=fpml, obs=obs)
    G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt, g.fdcoeff, g.free_surf)
    prec = ig.DFT_COMPAT if precision == "compat" else ig.DFT_KAHAN
    ctx = ig.ShotContext(G, len(finv), len(rec), prec)
    ctx.set_cpml(vppml, fpml)
    ctx.set_freqs(finv, tast)
    ntr = len(rec)
    d_syn = ig.DBuf((3, ntr, g.nt + 1))
    grads = [G.empty() for _ in range(3)]
    # observed data through the device path too (forward only on the true model)
    ctx.set_model(*[G.to_dev(a) for a in true])
    ctx.run_shot(src, rec, d_syn, None, want_adjoint=False)
    d_obs = d_syn.clone()
    ctx.set_model(*[G.to_dev(a) for a in start])
    l2 = ctx.run_shot(src, rec, d_syn, grads, d_obs=d_obs)
    ig.sync()
    out = dict(obs=d_obs.numpy(), syn=d_syn.numpy(), l2=l2,
               grads=[g.interior(G.from_dev(t)) for t in grads])
    ctx.close()
    return ref, obs, out


# nz 90: the interior box is 74 cells deep in k, so its launches take the
# LDS-staged marches (a full 64-lane k tile and a 10-lane one)
@pytest.mark.parametrize("fdcoeff,free_surf,nz", [(2, 0, 38), (1, 0, 38), (2, 1, 38), (2, 1, 90)])
def test_gradient_shot_matches_oracle(fdcoeff, free_surf, nz):
    g = small_grid(nx=36, ny=34, nz=nz, fw=8, nt=160 if nz < 64 else 90, fdcoeff=fdcoeff, free_surf=free_surf)
    src = (18, 12, 19, 0.0, 50.0, 1.0, 1)
    rec = np.array([[9 + 2 * n, 10, 19] for n in range(9)], np.int32)
    finv = [40.0, 60.0, 80.0]
    _note(f"gradient shot {g.nx}x{g.ny}x{g.nz} NT={g.nt} fdcoeff={fdcoeff} free_surf={free_surf}")
    ref, obs, out = _run_case(g, src, rec, finv, tast=100, vppml=4000.0, fpml=50.0)
    for c in range(3):
        assert max_ulp(out["obs"][c][:, 1:], obs[c][:, 1:]) == 0, f"observed comp {c}"
        e = rel_l2(out["syn"][c][:, 1:], ref["syn"][c][:, 1:])
        assert e <= 1e-5
        assert max_ulp(out["syn"][c][:, 1:], ref["syn"][c][:, 1:]) == 0, f"synthetic comp {c}"
    for n in range(3):
        e = rel_l2(out["grads"][n], g.interior(ref["grads"][n]))
        _note(f"  grad{n + 1}: rel-L2 {e:.2e}, max ulp {max_ulp(out['grads'][n], g.interior(ref['grads'][n]))}")
        assert e <= 1e-4
        assert max_ulp(out["grads"][n], g.interior(ref["grads"][n])) == 0
    assert abs(out["l2"] - ref["l2"]) <= 1e-5 * abs(ref["l2"])


def test_force_source_shot():
    """single force (type 4, vertical) and non-cubic grid, forward only"""
    import ifosgpu as ig
    g = small_grid(nx=40, ny=30, nz=44, fw=7, nt=120)
    _note("force-source forward shot")
    o = Oracle("f32")
    start, _ = _models(g, 2)
    src = (20, 15, 22, 0.002, 45.0, 1e6, 4)
    rec = np.array([[8 + 3 * n, 9, 22] for n in range(8)], np.int32)
    ref = o.run_shot(g, *start, src, rec, [50.0], want_adjoint=False)
    G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt, g.fdcoeff, g.free_surf)
    ctx = ig.ShotContext(G, 1, len(rec))
    ctx.set_cpml(4000.0, 50.0)
    ctx.set_freqs([50.0])
    ctx.set_model(*[G.to_dev(a) for a in start])
    d_syn = ig.DBuf((3, len(rec), g.nt + 1))
    ctx.run_shot(src, rec, d_syn, None, want_adjoint=False)
    syn = d_syn.numpy()
    ctx.close()
    for c in range(3):
        assert max_ulp(syn[c][:, 1:], ref["syn"][c][:, 1:]) == 0


@pytest.mark.parametrize("free_surf,nz,L", [(0, 34, 1), (1, 34, 1), (0, 84, 1), (0, 34, 3), (1, 84, 3)])
def test_viscoelastic_gradient_shot(free_surf, nz, L):
    """the plain path with the viscoelastic update (L relaxation mechanisms;
    the reference's L = 1) against the oracle's viscoelastic shot, bit for
    bit; at nz 84 the interior launches take the LDS-staged marches (70
    cells deep in k)"""
    import ifosgpu as ig
    g = small_grid(nx=32, ny=30, nz=nz, fw=7, nt=120 if nz < 64 else 80, free_surf=free_surf)
    src = (16, 11, 17, 0.0, 50.0, 1.0, 1)
    rec = np.array([[8 + 2 * n, 9, 17] for n in range(8)], np.int32)
    finv = [40.0, 70.0]
    _note(f"viscoelastic gradient shot free_surf={free_surf} nz={nz} L={L}")
    o = Oracle("f32")
    start, true = _models(g, 3)
    rng = np.random.default_rng(7)
    taus = (0.02 + 0.03 * rng.random(g.shape)).astype(np.float32)
    taup = (0.02 + 0.03 * rng.random(g.shape)).astype(np.float32)
    eta = np.array([0.021, 0.0063, 0.047][:L], np.float32)
    vis = (taus, taup, eta)
    obs = o.run_shot(g, *true, src, rec, finv, want_adjoint=False, visco=vis)["syn"]
    ref = o.run_shot(g, *start, src, rec, finv, obs=obs, visco=vis)
    G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt, g.fdcoeff, g.free_surf)
    ctx = ig.ShotContext(G, len(finv), len(rec))
    ctx.set_cpml(4000.0, 50.0)
    ctx.set_freqs(finv)
    ctx.set_visco(G.to_dev(taus), G.to_dev(taup), eta)
    ctx.set_model(*[G.to_dev(a) for a in start])
    d_obs = ig.DBuf.from_numpy(np.stack(obs))
    d_syn = ig.DBuf((3, len(rec), g.nt + 1))
    grads = [G.empty() for _ in range(3)]
    l2 = ctx.run_shot(src, rec, d_syn, grads, d_obs=d_obs)
    ig.sync()
    syn = d_syn.numpy()
    for c in range(3):
        assert max_ulp(syn[c][:, 1:], ref["syn"][c][:, 1:]) == 0, f"synthetic comp {c}"
    for n in range(3):
        got = g.interior(G.from_dev(grads[n]))
        assert max_ulp(got, g.interior(ref["grads"][n])) == 0, f"gradient {n}"
    assert abs(l2 - ref["l2"]) <= 1e-5 * abs(ref["l2"])
    ctx.close()


@pytest.mark.parametrize("nt,nf", [(160, 3), (700, 8)])
def test_errors_against_fp64_oracle(nt, nf):
    """the FP32 device shot (compat and Kahan DFT) against the FP64 oracle
    run of the same shot: the error the FP32 method carries, stated and held
    to the path's tolerances (seismograms rel. L2 <= 1e-5, gradients <= 1e-4)"""
    import ifosgpu as ig
    g = small_grid(nx=36, ny=34, nz=38, fw=8, nt=nt)
    src = (18, 12, 19, 0.0, 50.0, 1.0, 1)
    rec = np.array([[9 + 2 * n, 10, 19] for n in range(9)], np.int32)
    finv = [40.0 + 10.0 * n for n in range(nf)]
    _note(f"FP32 device vs FP64 oracle: NT={nt}, {nf} frequencies")
    o32, o64 = Oracle("f32"), Oracle("f64")
    start, true = _models(g, 1)
    obs = o32.run_shot(g, *true, src, rec, finv, want_adjoint=False)["syn"]
    ref = o64.run_shot(g, *[a.astype(np.float64) for a in start], src, rec, finv, obs=obs)
    G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt, g.fdcoeff, g.free_surf)
    d_obs = ig.DBuf.from_numpy(np.stack(obs))
    errs = {}
    for name, prec in (("compat", ig.DFT_COMPAT), ("kahan", ig.DFT_KAHAN)):
        ctx = ig.ShotContext(G, nf, len(rec), prec)
        ctx.set_cpml(4000.0, 50.0)
        ctx.set_freqs(finv, 100)
        ctx.set_model(*[G.to_dev(a) for a in start])
        d_syn = ig.DBuf((3, len(rec), g.nt + 1))
        grads = [G.empty() for _ in range(3)]
        ctx.run_shot(src, rec, d_syn, grads, d_obs=d_obs)
        ig.sync()
        syn = d_syn.numpy()
        es = max(rel_l2(syn[c][:, 1:], ref["syn"][c][:, 1:]) for c in range(3))
        eg = [rel_l2(g.interior(G.from_dev(grads[n])), g.interior(ref["grads"][n])) for n in range(3)]
        ctx.close()
        errs[name] = (es, eg)
        _note(f"  {name}: seismograms rel-L2 {es:.2e}; gradients vp {eg[0]:.2e} vs {eg[1]:.2e} rho {eg[2]:.2e}")
        assert es <= 1e-5
        assert max(eg) <= 1e-4
