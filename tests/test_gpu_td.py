"""The checkpointed time-domain gradient (ifos_ctx_run_shot_td) against the
oracle's restatement (og_td_gradient over the full velocity histories).

The device stores the forward state every K steps and recomputes each
segment during the adjoint pass; the oracle keeps every step.  The
wavefields are the plain path's (bit-exact to the oracle), the per-sample
products are gradient_F's arithmetic and the sums are FP64 in the same
order, so the gradients agree bit for bit, for every checkpoint interval
(1, a divisor of NT, one that leaves a short last segment, NT itself and the
automatic round(sqrt(NT))), elastic and viscoelastic, with and without the
free surface.  The seismograms and the misfit equal those of the plain
gradient shot.  tests/test_oracle_td.py pins the oracle's definition to
gradient_F and to the frequency-domain gradient over the full band.
"""
import sys

import numpy as np
import pytest

from helpers import max_ulp, small_grid
from pyoracle import Oracle

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _models(g, seed):
    rng = np.random.default_rng(seed)
    J, I, K = np.meshgrid(*(np.arange(s) - 2.0 for s in g.shape), indexing="ij")
    blob = np.exp(-((J - g.ny / 2) ** 2 + (I - g.nx / 3) ** 2 + (K - g.nz / 2) ** 2) / 12.0)

    def mod(f):
        vp, vs, rho = 4000.0 * f, 2300.0 * f, 2000.0 * (1 + 0.02 * rng.uniform(-1, 1, g.shape))
        return [a.astype(np.float32) for a in (rho, rho * vp * vp, rho * vs * vs)]

    return mod(np.ones(g.shape)), mod(1 + 0.08 * blob)


CASES = {
    "elastic": dict(free_surf=0, visco=False, every=(0, 1, 6, 7, 48)),
    "free_surface": dict(free_surf=1, visco=False, every=(5,)),
    "viscoelastic": dict(free_surf=0, visco=1, every=(0, 11)),
    "viscoelastic_3": dict(free_surf=0, visco=3, every=(0, 7)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_td_gradient_matches_oracle(case):
    import ifosgpu as ig
    c = CASES[case]
    g = small_grid(nx=28, ny=26, nz=30, fw=6, nt=48, free_surf=c["free_surf"])
    src = (14, 10 if not c["free_surf"] else 5, 15, 0.0, 60.0, 1.0, 1)
    rec = np.array([[8 + 2 * n, 8, 15] for n in range(6)] + [[20, 4, 12]], np.int32)
    o = Oracle("f32")
    start, true = _models(g, 2)
    vis = None
    if c["visco"]:
        rng = np.random.default_rng(9)
        vis = ((0.02 + 0.03 * rng.random(g.shape)).astype(np.float32),
               (0.02 + 0.03 * rng.random(g.shape)).astype(np.float32),
               np.array([0.021, 0.0063, 0.047][:c["visco"]], np.float32))
    obs = o.run_shot(g, *true, src, rec, [40.0], want_adjoint=False, visco=vis)["syn"]
    ref = o.run_shot(g, *start, src, rec, [40.0], obs=obs, visco=vis, want_history=True)
    want = [o.zeros(g) for _ in range(3)]
    o.td_gradient(g, ref["vfwd"], ref["vadj"], *start, want)
    G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt, g.fdcoeff, g.free_surf)
    ctx = ig.ShotContext(G, 1, len(rec))
    ctx.set_cpml(4000.0, 50.0)
    if vis is not None:
        ctx.set_visco(G.to_dev(vis[0]), G.to_dev(vis[1]), vis[2])
    ctx.set_model(*[G.to_dev(a) for a in start])
    d_obs = ig.DBuf.from_numpy(np.stack(obs))
    d_syn = ig.DBuf((3, len(rec), g.nt + 1))
    try:
        for every in c["every"]:
            _note(f"td gradient {case}: checkpoint every {every or 'sqrt(NT)'} steps, {g.nx}x{g.ny}x{g.nz} NT={g.nt}")
            grads = [G.empty() for _ in range(3)]
            l2 = ctx.run_shot_td(src, rec, d_syn, grads, d_obs, every=every)
            ig.sync()
            syn = d_syn.numpy()
            for comp in range(3):
                assert max_ulp(syn[comp][:, 1:], ref["syn"][comp][:, 1:]) == 0, f"seismogram {comp}"
            assert abs(l2 - ref["l2"]) <= 1e-5 * abs(ref["l2"])
            for n in range(3):
                got = g.interior(G.from_dev(grads[n]))
                w = g.interior(want[n])
                assert np.max(np.abs(w)) > 0
                assert max_ulp(got, w) == 0, f"gradient {n + 1}, every {every}: max ulp {max_ulp(got, w)}"
    finally:
        ctx.close()


def test_td_accumulates_and_rejects():
    """gradients accumulate over calls (shot after shot); missing inputs are errors"""
    import ifosgpu as ig
    g = small_grid(nx=24, ny=22, nz=26, fw=5, nt=30)
    src = (12, 9, 13, 0.0, 60.0, 1.0, 1)
    rec = np.array([[7 + 2 * n, 7, 13] for n in range(5)], np.int32)
    _note("td accumulation and errors")
    start, true = _models(g, 4)
    G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt, g.fdcoeff, g.free_surf)
    ctx = ig.ShotContext(G, 1, len(rec))
    ctx.set_cpml(4000.0, 50.0)
    try:
        d_syn = ig.DBuf((3, len(rec), g.nt + 1))
        ctx.set_model(*[G.to_dev(a) for a in true])
        ctx.set_freqs([40.0], 100)
        ctx.run_shot(src, rec, d_syn, None, want_adjoint=False)
        d_obs = d_syn.clone()
        ctx.set_model(*[G.to_dev(a) for a in start])
        one = [G.empty() for _ in range(3)]
        ctx.run_shot_td(src, rec, d_syn, one, d_obs, every=4)
        two = [G.empty() for _ in range(3)]
        ctx.run_shot_td(src, rec, d_syn, two, d_obs, every=4)
        ctx.run_shot_td(src, rec, d_syn, two, d_obs, every=9)
        for n in range(3):
            a, b = G.from_dev(one[n]).astype(np.float64), G.from_dev(two[n]).astype(np.float64)
            assert np.allclose(b, 2 * a, rtol=1e-6, atol=1e-6 * np.max(np.abs(a)))
        with pytest.raises(ig.IfosError):
            ctx.run_shot_td(src, rec, None, one, d_obs)
        with pytest.raises(ig.IfosError):
            ctx.run_shot_td((0, 9, 13, 0.0, 60.0, 1.0, 1), rec, d_syn, one, d_obs)
    finally:
        ctx.close()
