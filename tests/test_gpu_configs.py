"""Parity at the benchmark workloads' own shapes.

* toy84 -- the whole CPU-runnable toy gradient: 84^3 (64^3 + 10-cell CPML),
  NT 600, 4 explosive shots, 32 receivers, DFT at 40/60/80 Hz, two +-10 %
  spheres as the true model.  Observed data, synthetics, misfits and the
  gradient accumulated over the 4 shots, through the plain one-shot path and
  through the batched path (one batch of 4, tuned tiles, hipGraph replay),
  against the oracle's outputs stored in tests/golden/configs.npz (written
  by tests/golden/make_configs.py): bit for bit.
* the uneven grid levels -- 116^3 at 8 m, 148^3 at 6 m, 212^3 at 4 m, each
  with its own 0.9-CFL time step and model, 3 shots per level run together
  by ifos_levels_run (one graph, one branch per level): the first shot of
  every level against the oracle, bit for bit.
* the viscoelastic DFT case's shape -- 404 x 212 x 404 (384 x 192 x 384 +
  10-cell CPML), relaxation mechanisms and per-cell Q, 8 DFT frequencies
  (5-40 Hz): a short gradient shot (forward + DFT, residual, adjoint + DFT,
  gradient_F) through the plain and the batched path against the oracle run
  here, bit for bit in compat precision.
"""
import os
import sys

import numpy as np
import pytest

from helpers import max_ulp, rel_l2, sha

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "configs.npz")


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fix():
    return np.load(FIX, allow_pickle=False)


def _geom(ig, w):
    return ig.Geom(w.nx, w.ny, w.nz, w.fw, w.dh, w.dh, w.dh, w.dt, w.nt, w.fdcoeff, 0)


def _check_toy(F, G, w, obs, syns, l2s, grads, tag):
    from pyoracle import Grid
    g = Grid(w.nx, w.ny, w.nz, w.fw, w.dh, w.dh, w.dh, w.dt, w.nt, w.fdcoeff)
    for n in range(len(w.shots)):
        if sha(obs[n][:, :, 1:]) != str(F[f"toy84_obs_sha_{n}"]):
            detail = f" (max ulp {max_ulp(obs[n], F['toy84_obs_0'])})" if n == 0 else ""
            raise AssertionError(f"{tag}: observed data of shot {n} differ from the oracle{detail}")
        if sha(syns[n][:, :, 1:]) != str(F[f"toy84_syn_sha_{n}"]):
            detail = f" (max ulp {max_ulp(syns[n], F['toy84_syn_0'])})" if n == 0 else ""
            raise AssertionError(f"{tag}: synthetics of shot {n} differ from the oracle{detail}")
        # the misfit: the library sums the squares of the integrated residual
        # (residual.c:43-53) in FP64, the reference in float, one term after
        # the other (57 600 terms here: the float sum is off by ~3e-5)
        want = float(F[f"toy84_l2_{n}"])
        assert abs(l2s[n] - want) <= 2e-4 * abs(want), f"{tag}: misfit of shot {n}: {l2s[n]} vs {want}"
        R = np.cumsum(syns[n][:, :, 1:] - obs[n][:, :, 1:], axis=2, dtype=np.float32) * np.float32(w.dt)
        l2_64 = float((R.astype(np.float64) ** 2).sum())
        assert abs(l2s[n] - l2_64) <= 1e-6 * l2_64, f"{tag}: misfit of shot {n}: {l2s[n]} vs FP64 sum {l2_64}"
    jc = g.shape[0] // 2
    for q in range(3):
        a = G.from_dev(grads[q])
        if sha(g.interior(a)) != str(F[f"toy84_grad_sha_{q}"]):
            raise AssertionError(f"{tag}: gradient {q + 1} differs from the oracle (centre slice: rel-L2 "
                                 f"{rel_l2(a[jc], F[f'toy84_grad_slice_{q}']):.3e}, max ulp "
                                 f"{max_ulp(a[jc], F[f'toy84_grad_slice_{q}'])})")
        assert float(F[f"toy84_grad_l2_{q}"]) > 0


@pytest.mark.timeout(600)
def test_toy84_plain_matches_oracle():
    import ifosgpu as ig
    import workloads as wl
    F = _fix()
    w = wl.toy84()
    G = _geom(ig, w)
    true, start = wl.sphere_models(w)
    ntr = len(w.receivers)
    _note("toy84, plain path: observed data and 4 gradient shots")
    ctx = ig.ShotContext(G, len(w.finv), ntr)
    ctx.set_cpml(w.vppml, w.fpml)
    ctx.set_freqs(w.finv, w.tast)
    d_syn = ig.DBuf((3, ntr, w.nt + 1))
    ctx.set_model(*[G.to_dev(a) for a in true])
    obs, d_obs = [], []
    for src in w.shots:
        ctx.run_shot(src, w.receivers, d_syn, None, want_adjoint=False)
        d_obs.append(d_syn.clone())
        obs.append(d_obs[-1].numpy())
    ctx.set_model(*[G.to_dev(a) for a in start])
    grads = [G.empty() for _ in range(3)]
    syns, l2s = [], []
    for n, src in enumerate(w.shots):
        l2s.append(ctx.run_shot(src, w.receivers, d_syn, grads, d_obs=d_obs[n]))
        syns.append(d_syn.numpy())
    ig.sync()
    ctx.close()
    _check_toy(F, G, w, obs, syns, l2s, grads, "plain")


@pytest.mark.timeout(600)
def test_toy84_batched_matches_oracle():
    import ifosgpu as ig
    import workloads as wl
    F = _fix()
    w = wl.toy84()
    G = _geom(ig, w)
    true, start = wl.sphere_models(w)
    ns, ntr = len(w.shots), len(w.receivers)
    _note("toy84, batched path: one batch of 4 shots, tuned tiles, graph replay")
    b = ig.BatchContext(G, ns, len(w.finv), ntr, ig.DFT_COMPAT, 16)
    b.set_cpml(w.vppml, w.fpml)
    b.set_freqs(w.finv, w.tast)
    b.set_model(*[G.to_dev(a) for a in true])
    b.autotune(reps=1)
    d_syn = ig.DBuf((ns, 3, ntr, w.nt + 1))
    b.run(w.shots, w.receivers, d_syn=d_syn, want_adjoint=False, use_graph=True)
    hb = d_syn.numpy().copy()
    obs = [hb[n] for n in range(ns)]
    d_obs = [ig.DBuf.from_numpy(o) for o in obs]
    b.set_model(*[G.to_dev(a) for a in start])
    grads = [G.empty() for _ in range(3)]
    l2s = b.run(w.shots, w.receivers, d_syn=d_syn, grads=grads, d_obs=d_obs, use_graph=True)
    hs = d_syn.numpy()
    ig.sync()
    b.close()
    _check_toy(F, G, w, obs, [hs[n] for n in range(ns)], l2s, grads, "batched")


@pytest.mark.timeout(600)
def test_level_shapes_match_oracle():
    import ifosgpu as ig
    import workloads as wl
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_configs as mc
    F = _fix()
    levels = mc.level_workloads()
    models = wl.level_models(levels)
    ctxs, srcs, recs, d_syn = [], [], [], []
    for w, m in zip(levels, models):
        G = _geom(ig, w)
        rr = [mc.level_receivers(w, s) for s in w.shots]
        assert len({len(r) for r in rr}) == 1
        _note(f"level {w.extra['interior']}^3 (+{w.fw}), dh {w.dh}, NT {w.nt}: {len(w.shots)} shots, "
              f"{len(rr[0])} receivers")
        b = ig.BatchContext(G, len(w.shots), 1, len(rr[0]), ig.DFT_COMPAT, 1)
        b.set_cpml(w.vppml, w.fpml)
        b.set_model(*[G.to_dev(a) for a in m])
        b.autotune(reps=1)
        ctxs.append(b)
        srcs.append(w.shots)
        recs.append(rr)
        d_syn.append(ig.DBuf((len(w.shots), 3, len(rr[0]), w.nt + 1)))
    _note("ifos_levels_run: the three levels in one graph")
    ig.levels_run(ctxs, srcs, recs, d_syn=d_syn, use_graph=True)
    ig.sync()
    for w, d in zip(levels, d_syn):
        n = w.extra["interior"]
        syn = d.numpy()
        if sha(syn[0][:, :, 1:]) != str(F[f"level_{n}_syn_sha"]):
            raise AssertionError(f"level {n}^3: first shot differs from the oracle "
                                 f"(max ulp {max_ulp(syn[0][:, :, 1:], F[f'level_{n}_syn'][:, :, 1:])})")
        assert np.abs(syn).max() > 0
    for b in ctxs:
        b.close()


@pytest.mark.timeout(900)
def test_visco404_shape_matches_oracle():
    """the viscoelastic + DFT case's own grid, model, relaxation mechanisms
    (3, per-cell Q) and 8 frequencies: a 4-step gradient shot (a source next
    to the receivers so the short record is not empty) through the plain path
    and the batched path (compat DFT) against the oracle bit for bit; the
    Kahan DFT within 1e-4 of it"""
    import ifosgpu as ig
    import workloads as wl
    from pyoracle import Grid, Oracle
    w = wl.visco404()
    model, (taus, taup), etas = wl.visco_models(w)
    w.nt = 4
    src = (w.fw + 1 + 192, w.fw + 2, w.fw + 1 + 192, 0.0, 15.0, 1e6, 1)
    rec = np.array([[w.fw + 1 + 192 + d, w.fw + 1, w.fw + 1 + 192 + e] for d in (-2, 0, 2) for e in (-2, 0, 2)],
                   np.int32)
    g = Grid(w.nx, w.ny, w.nz, w.fw, w.dh, w.dh, w.dh, w.dt, w.nt, w.fdcoeff)
    obs = [np.zeros((len(rec), w.nt + 1), np.float32) for _ in range(3)]
    _note(f"oracle: visco404 ({w.nx}x{w.ny}x{w.nz}), L=3, {len(w.finv)} frequencies, NT={w.nt}")
    ref = Oracle("f32").run_shot(g, *model, src, rec, w.finv, tast=w.tast, vppml=w.vppml, fpml=w.fpml, obs=obs,
                                 visco=(taus, taup, etas))
    G = _geom(ig, w)
    d_model = [G.to_dev(a) for a in model]
    d_tau = [G.to_dev(a) for a in (taus, taup)]
    d_obs = ig.DBuf.from_numpy(np.stack(obs))
    for prec_name, prec in (("compat", ig.DFT_COMPAT), ("kahan", ig.DFT_KAHAN)):
        _note(f"plain path, {prec_name} DFT")
        ctx = ig.ShotContext(G, len(w.finv), len(rec), prec)
        ctx.set_cpml(w.vppml, w.fpml)
        ctx.set_freqs(w.finv, w.tast)
        ctx.set_visco(*d_tau, etas)
        ctx.set_model(*d_model)
        d_syn = ig.DBuf((3, len(rec), w.nt + 1))
        grads = [G.empty() for _ in range(3)]
        ctx.run_shot(src, rec, d_syn, grads, d_obs=d_obs)
        ig.sync()
        syn = d_syn.numpy()
        for c in range(3):
            assert max_ulp(syn[c][:, 1:], ref["syn"][c][:, 1:]) == 0, f"{prec_name}: seismogram {c}"
        assert np.abs(syn).max() > 0
        for q in range(3):
            a, want = g.interior(G.from_dev(grads[q])), g.interior(ref["grads"][q])
            assert np.abs(want).max() > 0
            if prec_name == "compat":
                assert max_ulp(a, want) == 0, f"plain compat: gradient {q + 1}"
            else:
                assert rel_l2(a, want) <= 1e-4, f"plain kahan: gradient {q + 1}"
        ctx.close()
        del grads
    _note("batched path, one shot, compat DFT")
    b = ig.BatchContext(G, 1, len(w.finv), len(rec), ig.DFT_COMPAT, 4)
    b.set_cpml(w.vppml, w.fpml)
    b.set_freqs(w.finv, w.tast)
    b.set_visco(*d_tau, etas)
    b.set_model(*d_model)
    grads = [G.empty() for _ in range(3)]
    d_b = ig.DBuf((1, 3, len(rec), w.nt + 1))
    b.run([src], rec, d_syn=d_b, grads=grads, d_obs=[d_obs])
    ig.sync()
    for q in range(3):
        assert max_ulp(g.interior(G.from_dev(grads[q])), g.interior(ref["grads"][q])) == 0, f"batched: gradient {q + 1}"
    b.close()
