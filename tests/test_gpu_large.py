"""Parity at the full shape of the forward-only bandwidth case: 536 x 280 x
536 cells (512 x 256 x 512 + 12-cell CPML), 4096 receivers on a 64 x 64
grid near the top, a heterogeneous model.

The oracle is single-threaded C, so the check runs the first steps of the
time loop at the real shape (83 M cells per field, 355 MB per field in the
device layout) through the plain path and the batched path and compares
every velocity of the grid after the last step with the oracle's, bit for
bit; the seismograms of those steps are compared too.  The full 2000-step
run is measured by `bench.py --config fwd536`.

The gradient workload's 276^3 grid is checked the same way for the batched
path: two gradient shots of a 200-step record through the batched kernels
(the tiles the autotuner picks on that grid) against the plain path, which
the oracle pins, bit for bit.
"""
import sys

import numpy as np
import pytest

from helpers import max_ulp
from pyoracle import Grid, Oracle

pytestmark = pytest.mark.gpu

NT = 3


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.timeout(900)
def test_bandwidth_case_shape_matches_oracle():
    import ifosgpu as ig
    fw = 12
    g = Grid(nx=536, ny=280, nz=536, fw=fw, dx=5.0, dy=5.0, dz=5.0, dt=3.99e-4, nt=NT, fdcoeff=2, free_surf=0)
    rng = np.random.default_rng(2)
    shp = g.shape
    vp = (2500.0 + 3000.0 * rng.random(shp)).astype(np.float32)
    vs = (vp / (1.7 + 0.3 * rng.random(shp))).astype(np.float32)
    rho = (310.0 * vp ** 0.25).astype(np.float32)
    model = [rho, (rho * vp * vp).astype(np.float32), (rho * vs * vs).astype(np.float32)]
    del vp, vs
    # a source next to the receivers so that the seismograms of the first steps are not zero
    src = (fw + 1 + 256, fw + 2, fw + 1 + 256, 0.0, 15.0, 1e6, 1)
    rec = np.array([[fw + 1 + 8 * a, fw + 1, fw + 1 + 8 * b] for a in range(64) for b in range(64)], np.int32)
    _note(f"oracle: {NT} steps at {g.nx}x{g.ny}x{g.nz}")
    o = Oracle("f32")
    ref = o.run_shot(g, *model, src, rec, [15.0], want_adjoint=False, want_history=True)
    last = ref["vfwd"][NT - 1]
    G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt, g.fdcoeff, g.free_surf)
    d_model = [G.to_dev(a) for a in model]
    ntr = len(rec)

    _note("plain path")
    ctx = ig.ShotContext(G, 1, ntr)
    ctx.set_cpml(4000.0, 50.0)
    ctx.set_freqs([15.0], 100)
    ctx.set_model(*d_model)
    d_syn = ig.DBuf((3, ntr, g.nt + 1))
    ctx.run_shot(src, rec, d_syn, None, want_adjoint=False)
    ig.sync()
    f = ctx.fields()
    for c, p in enumerate((f.vx, f.vy, f.vz)):
        got = G.from_dev(ig.DBuf.wrap(p, G.dshape))
        assert max_ulp(g.interior(got), g.interior(last[c])) == 0, f"plain path, velocity {c}"
    syn = d_syn.numpy()
    for c in range(3):
        assert max_ulp(syn[c][:, 1:], ref["syn"][c][:, 1:]) == 0, f"plain path, seismogram {c}"
    assert np.abs(syn).max() > 0
    ctx.close()

    _note("batched path, one shot")
    b = ig.BatchContext(G, 1, 1, ntr, ig.DFT_COMPAT, 1)
    b.set_cpml(4000.0, 50.0)
    b.set_model(*d_model)
    d_b = ig.DBuf((1, 3, ntr, g.nt + 1))
    b.run([src], rec, d_syn=d_b, want_adjoint=False)
    sb = d_b.numpy()[0]
    for c in range(3):
        assert max_ulp(sb[c][:, 1:], ref["syn"][c][:, 1:]) == 0, f"batched path, seismogram {c}"
    b.close()


@pytest.mark.timeout(900)
def test_gradient_workload_shape_batched_matches_plain():
    """the 276^3 gradient workload's grid, model, receivers and frequencies (a
    200-step record): two shots through the batched path (autotuned tiles)
    give the plain path's seismograms, misfits and accumulated gradient bit
    for bit"""
    import ifosgpu as ig
    import workloads as wl
    w = wl.grad276()
    w.nt = 200
    true, start = wl.random_models(w)
    G = ig.Geom(w.nx, w.ny, w.nz, w.fw, w.dh, w.dh, w.dh, w.dt, w.nt, w.fdcoeff, 0)
    d_true = [G.to_dev(a) for a in true]
    d_start = [G.to_dev(a) for a in start]
    del true, start
    shots = [w.shots[0], w.shots[5]]
    ntr = len(w.receivers)
    _note(f"grad276 grid, NT={w.nt}: observed data")
    ctx = ig.ShotContext(G, len(w.finv), ntr)
    ctx.set_cpml(w.vppml, w.fpml)
    ctx.set_freqs(w.finv, w.tast)
    d_syn = ig.DBuf((3, ntr, w.nt + 1))
    ctx.set_model(*d_true)
    obs = []
    for s in shots:
        ctx.run_shot(s, w.receivers, d_syn, None, want_adjoint=False)
        obs.append(d_syn.clone())
    _note("plain path, two gradient shots")
    ctx.set_model(*d_start)
    gp = [G.empty() for _ in range(3)]
    l2p, synp = [], []
    for n, s in enumerate(shots):
        l2p.append(ctx.run_shot(s, w.receivers, d_syn, gp, d_obs=obs[n]))
        synp.append(d_syn.numpy())
    ctx.close()
    _note("batched path, the same two shots")
    b = ig.BatchContext(G, 2, len(w.finv), ntr, ig.DFT_COMPAT, 8)
    b.set_cpml(w.vppml, w.fpml)
    b.set_freqs(w.finv, w.tast)
    b.set_model(*d_start)
    b.autotune(reps=1)
    gb = [G.empty() for _ in range(3)]
    d_b = ig.DBuf((2, 3, ntr, w.nt + 1))
    l2b = b.run(shots, w.receivers, d_syn=d_b, grads=gb, d_obs=obs)
    sb = d_b.numpy()
    b.close()
    for n in range(2):
        assert max_ulp(sb[n][:, :, 1:], synp[n][:, :, 1:]) == 0, f"seismograms, shot {n}"
        assert l2b[n] == l2p[n], f"misfit, shot {n}"
    for q in range(3):
        a, c = G.from_dev(gb[q]), G.from_dev(gp[q])
        assert np.abs(c).max() > 0
        assert max_ulp(a, c) == 0, f"gradient {q + 1}"
