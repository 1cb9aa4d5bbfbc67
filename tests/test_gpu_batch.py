"""The batched path against the plain path, shot for shot.

ifos_batch_run propagates several shots of one model together (fused
interior+CPML updates over all shots, chunked DFT from a snapshot ring,
optionally replayed from a hipGraph).  Every number it produces must be
bit-identical to running the same shots one by one through
ifos_ctx_run_shot, which is itself bit-identical to the oracle
(test_gpu_shot.py): seismograms, per-shot misfits, DFT accumulators and the
gradient accumulated over the shots in order.

The cases put sources and receivers both inside the interior box and inside
CPML slabs, with two receivers sharing a cell, because the batched path
reorders the point injections around the fused update (slab points before,
interior points after) to keep the reference's rounding order.
"""
import sys

import numpy as np
import pytest

from helpers import max_ulp, rel_l2, small_grid

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _models(g):
    shp = g.shape
    J, I, K = np.meshgrid(*(np.arange(s) - 2 for s in shp), indexing="ij")
    fac = 1.0 + 0.08 * np.exp(-((J - g.ny / 2) ** 2 + (I - g.nx / 3) ** 2 + (K - g.nz / 2) ** 2) / 20.0)
    vp, vs, rho = 4000.0 * fac, 2300.0 * fac, 2000.0 * fac
    true = [a.astype(np.float32) for a in (rho, rho * vp * vp, rho * vs * vs)]
    one = np.ones(shp)
    start = [a.astype(np.float32) for a in (2000.0 * one, 2000.0 * 4000.0 ** 2 * one, 2000.0 * 2300.0 ** 2 * one)]
    return true, start


def _setup(ig, g, finv, tast, prec):
    G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt, g.fdcoeff, g.free_surf)
    return G


def _plain(ig, G, g, true, start, srcs, rec, finv, tast, prec, visco=None):
    """observed data, seismograms, misfits and accumulated gradient, one shot at a time"""
    ctx = ig.ShotContext(G, len(finv), len(rec), prec)
    ctx.set_cpml(4000.0, 50.0)
    ctx.set_freqs(finv, tast)
    if visco is not None:
        ctx.set_visco(G.to_dev(visco[0]), G.to_dev(visco[1]), visco[2])
    ntr = len(rec)
    obs, syn, l2 = [], [], []
    d_syn = ig.DBuf((3, ntr, g.nt + 1))
    ctx.set_model(*[G.to_dev(a) for a in true])
    for src in srcs:
        ctx.run_shot(src, rec, d_syn, None, want_adjoint=False)
        obs.append(d_syn.clone())
    ctx.set_model(*[G.to_dev(a) for a in start])
    grads = [G.empty() for _ in range(3)]
    dft_last = None
    for n, src in enumerate(srcs):
        l2.append(ctx.run_shot(src, rec, d_syn, grads, d_obs=obs[n]))
        syn.append(d_syn.numpy())
    ig.sync()
    fe = G.field_elems()
    dft_last = [ig.DBuf.wrap(p, (len(finv) * fe,)).numpy().copy() for p in ctx.dft_ptrs(1)]
    out = dict(obs=obs, syn=syn, l2=l2, grads=[G.from_dev(t) for t in grads], dft_last=dft_last)
    ctx.close()
    return out


def _batched(ig, G, g, true, start, srcs, rec, finv, tast, prec, chunk, use_graph, obs=None, tiles=None,
             visco=None):
    ns, ntr = len(srcs), len(rec)
    b = ig.BatchContext(G, ns, len(finv), ntr, prec, chunk)
    if tiles is not None:
        b.set_tiles(*tiles)
    if visco is not None:
        b.set_visco(G.to_dev(visco[0]), G.to_dev(visco[1]), visco[2])
    b.set_cpml(4000.0, 50.0)
    b.set_freqs(finv, tast)
    d_syn = ig.DBuf((ns, 3, ntr, g.nt + 1))
    b.set_model(*[G.to_dev(a) for a in true])
    b.run(srcs, rec, d_syn=d_syn, want_adjoint=False, use_graph=use_graph)
    obs_b = d_syn.numpy().copy()
    if obs is None:
        obs = [ig.DBuf.from_numpy(obs_b[n]) for n in range(ns)]
    b.set_model(*[G.to_dev(a) for a in start])
    grads = [G.empty() for _ in range(3)]
    l2 = b.run(srcs, rec, d_syn=d_syn, grads=grads, d_obs=obs, use_graph=use_graph)
    fe = G.field_elems()
    dft_last = [ig.DBuf.wrap(p, (len(finv) * fe,)).numpy().copy() for p in b.dft_ptrs(ns - 1, 1)]
    out = dict(obs=obs_b, syn=d_syn.numpy(), l2=l2, grads=[G.from_dev(t) for t in grads], dft_last=dft_last, ctx=b)
    return out


def _oracle(g, true, start, srcs, rec, finv, tast):
    """the oracle's observed data, synthetics, misfits and the gradient
    accumulated over the shots in order (ifos3d.c:712-975)"""
    from pyoracle import Oracle
    o = Oracle("f32")
    obs = [o.run_shot(g, *true, s, rec, finv, tast=tast, want_adjoint=False)["syn"] for s in srcs]
    grads = [o.zeros(g) for _ in range(3)]
    syn, l2 = [], []
    for n, s in enumerate(srcs):
        r = o.run_shot(g, *start, s, rec, finv, tast=tast, obs=obs[n], grads=grads)
        syn.append(r["syn"])
        l2.append(r["l2"])
    return dict(obs=obs, syn=syn, l2=l2, grads=grads)


CASES = {
    # explosive sources in the interior, receivers in the interior and in the y slab, two sharing a cell
    "interior": dict(free_surf=0, srcs=[(18, 12, 19, 0.0, 50.0, 1.0, 1), (12, 14, 22, 0.0, 45.0, 1.0, 1),
                                        (24, 13, 15, 0.001, 55.0, 1.0, 1)]),
    # a vertical force inside the x slab, an explosive source in the z slab
    "slab_sources": dict(free_surf=0, srcs=[(4, 14, 20, 0.0, 50.0, 1e6, 4), (18, 12, 35, 0.0, 50.0, 1.0, 1)]),
    "free_surface": dict(free_surf=1, srcs=[(18, 6, 19, 0.0, 50.0, 1.0, 1), (14, 8, 16, 0.0, 40.0, 1.0, 1)]),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("prec_name,chunk,use_graph", [("compat", 4, False), ("compat", 3, True),
                                                       ("kahan", 5, False), ("compat", 24, False),
                                                       ("kahan", 29, False)])
def test_batch_matches_plain(case, prec_name, chunk, use_graph):
    import ifosgpu as ig
    c = CASES[case]
    g = small_grid(nx=36, ny=34, nz=38, fw=8, nt=150, free_surf=c["free_surf"])
    rec = np.array([[9 + 2 * n, 10, 19] for n in range(8)] + [[13, 10, 19], [20, 5, 21], [30, 4, 33]], np.int32)
    finv, tast = [40.0, 60.0, 80.0], 100
    prec = ig.DFT_KAHAN if prec_name == "kahan" else ig.DFT_COMPAT
    _note(f"batch vs plain: case={case} prec={prec_name} chunk={chunk} graph={use_graph}")
    true, start = _models(g)
    G = _setup(ig, g, finv, tast, prec)
    ref = _plain(ig, G, g, true, start, c["srcs"], rec, finv, tast, prec)
    out = _batched(ig, G, g, true, start, c["srcs"], rec, finv, tast, prec, chunk, use_graph)
    for n in range(len(c["srcs"])):
        assert max_ulp(out["obs"][n], ref["obs"][n].numpy()) == 0, f"observed data, shot {n}"
        assert max_ulp(out["syn"][n], ref["syn"][n]) == 0, f"synthetics, shot {n}"
        assert out["l2"][n] == ref["l2"][n], f"misfit, shot {n}"
    for n in range(6):
        assert max_ulp(out["dft_last"][n], ref["dft_last"][n]) == 0, f"adjoint DFT component {n}, last shot"
    for n in range(3):
        assert max_ulp(out["grads"][n], ref["grads"][n]) == 0, f"gradient {n}"
    assert np.abs(ref["grads"][0]).max() > 0
    # the batched outputs against the oracle directly (not only through the
    # plain path): compat bit for bit; Kahan changes the DFT's rounding only
    orc = _oracle(g, true, start, c["srcs"], rec, finv, tast)
    kahan = prec_name == "kahan"
    for n in range(len(c["srcs"])):
        for q in range(3):
            assert max_ulp(out["obs"][n][q][:, 1:], orc["obs"][n][q][:, 1:]) == 0, f"oracle: observed, shot {n}"
            assert max_ulp(out["syn"][n][q][:, 1:], orc["syn"][n][q][:, 1:]) == 0, f"oracle: synthetics, shot {n}"
        assert abs(out["l2"][n] - orc["l2"][n]) <= 1e-5 * abs(orc["l2"][n]), f"oracle: misfit, shot {n}"
    for n in range(3):
        a, want = g.interior(out["grads"][n]), g.interior(orc["grads"][n])
        if kahan:
            assert rel_l2(a, want) <= 1e-4, f"oracle: gradient {n} (Kahan DFT)"
        else:
            assert max_ulp(a, want) == 0, f"oracle: gradient {n}"
    if use_graph:
        # replays: the captured loops run again with eager work in between
        b = out["ctx"]
        for rep in range(2):
            ig.DBuf((1 << 20,)).zero_()
            grads = [G.empty() for _ in range(3)]
            obs = [ig.DBuf.from_numpy(out["obs"][n]) for n in range(len(c["srcs"]))]
            l2 = b.run(c["srcs"], rec, grads=grads, d_obs=obs, use_graph=True)
            assert l2 == out["l2"], f"replay {rep}: misfits"
            for n in range(3):
                assert max_ulp(G.from_dev(grads[n]), ref["grads"][n]) == 0, f"replay {rep}: gradient {n}"
    out["ctx"].close()


def test_batch_partial_and_rejects():
    """a batch may run fewer shots than it was sized for; bad input is refused"""
    import ifosgpu as ig
    g = small_grid(nx=30, ny=28, nz=32, fw=6, nt=60)
    G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt, g.fdcoeff, g.free_surf)
    rec = np.array([[10, 9, 12], [12, 9, 14]], np.int32)
    b = ig.BatchContext(G, 4, 2, 2, ig.DFT_COMPAT, 4)
    b.set_cpml(4000.0, 50.0)
    b.set_freqs([40.0, 60.0])
    _, start = _models(g)
    b.set_model(*[G.to_dev(a) for a in start])
    d_syn = ig.DBuf((2, 3, 2, g.nt + 1))
    b.run([(15, 10, 16, 0.0, 50.0, 1.0, 1), (12, 11, 14, 0.0, 50.0, 1.0, 1)], rec, d_syn=d_syn, want_adjoint=False)
    assert np.abs(d_syn.numpy()).max() > 0
    with pytest.raises(ig.IfosError):
        b.run([(15, 10, 16, 0.0, 50.0, 1.0, 1)] * 5, rec, want_adjoint=False)
    with pytest.raises(ig.IfosError):
        b.run([(0, 10, 16, 0.0, 50.0, 1.0, 1)], rec, want_adjoint=False)
    b.close()


@pytest.mark.parametrize("tiles", [((4, 0, 1, 0, 0), (2, 0, 4, 1, 1)), ((8, 5, 0, 1, 1), (1, 0, 16, 0, 0)),
                                   ((1, 3, 2, 1, 0), (8, 0, 1, 1, 1)), ((2, 4, 3, 0, 1), (4, 7, 0, 0, 1))])
def test_batch_tile_shapes(tiles):
    """every tile shape the autotuner may pick gives the same bits"""
    import ifosgpu as ig
    c = CASES["slab_sources"]
    g = small_grid(nx=36, ny=34, nz=38, fw=8, nt=90)
    rec = np.array([[9 + 2 * n, 10, 19] for n in range(8)] + [[20, 5, 21], [30, 4, 33]], np.int32)
    finv, tast = [40.0, 60.0], 100
    _note(f"batch tile shapes v={tiles[0]} s={tiles[1]}")
    true, start = _models(g)
    G = _setup(ig, g, finv, tast, ig.DFT_COMPAT)
    ref = _plain(ig, G, g, true, start, c["srcs"], rec, finv, tast, ig.DFT_COMPAT)
    out = _batched(ig, G, g, true, start, c["srcs"], rec, finv, tast, ig.DFT_COMPAT, 4, False, tiles=tiles)
    for n in range(len(c["srcs"])):
        assert max_ulp(out["syn"][n], ref["syn"][n]) == 0, f"synthetics, shot {n}"
    for n in range(3):
        assert max_ulp(out["grads"][n], ref["grads"][n]) == 0, f"gradient {n}"
    out["ctx"].close()


@pytest.mark.parametrize("tiles", [((8, 6, 0, 0, 1, 4), (8, 0, 2, 0, 0, 2)), ((4, 0, 1, 1, 0, 4), (8, 5, 0, 0, 1, 8)),
                                   ((8, 0, 0, 0, 0, 8), (4, 3, 0, 1, 1, 2)),
                                   ((8, 6, 0, 0, 1, 1, 2), (8, 0, 2, 0, 0, 2, 2)),
                                   ((4, 0, 1, 1, 0, 1, 4), (8, 5, 0, 0, 1, 1, 4)), ((2, 7, 0, 0, 1, 1, 2), (8, 0, 0, 0, 0, 4, 2))])
def test_batch_k_columns(tiles):
    """workgroups of several wavefront columns along k (kb > 1) on a grid
    three columns deep (nz = 150), and of several shots (sb > 1, three
    shots: the last shot group half empty), give the same bits as the plain
    path"""
    import ifosgpu as ig
    g = small_grid(nx=22, ny=20, nz=150, fw=6, nt=60)
    srcs = [(11, 10, 75, 0.0, 50.0, 1.0, 1), (4, 9, 140, 0.0, 50.0, 1.0, 1), (15, 12, 64, 0.0, 50.0, 1.0, 1)]
    rec = np.array([[3 + 2 * n, 10, 60 + 13 * n] for n in range(7)] + [[20, 5, 145], [12, 18, 3]], np.int32)
    finv, tast = [40.0, 60.0], 100
    _note(f"batch k columns v={tiles[0]} s={tiles[1]}")
    true, start = _models(g)
    G = _setup(ig, g, finv, tast, ig.DFT_COMPAT)
    ref = _plain(ig, G, g, true, start, srcs, rec, finv, tast, ig.DFT_COMPAT)
    out = _batched(ig, G, g, true, start, srcs, rec, finv, tast, ig.DFT_COMPAT, 4, False, tiles=tiles)
    for n in range(len(srcs)):
        assert max_ulp(out["syn"][n], ref["syn"][n]) == 0, f"synthetics, shot {n}"
    for n in range(3):
        assert max_ulp(out["grads"][n], ref["grads"][n]) == 0, f"gradient {n}"
    out["ctx"].close()


def test_batch_autotune_keeps_results():
    import ifosgpu as ig
    c = CASES["interior"]
    g = small_grid(nx=36, ny=34, nz=38, fw=8, nt=60)
    rec = np.array([[9 + 2 * n, 10, 19] for n in range(6)], np.int32)
    finv, tast = [40.0, 60.0], 100
    true, start = _models(g)
    G = _setup(ig, g, finv, tast, ig.DFT_COMPAT)
    ref = _plain(ig, G, g, true, start, c["srcs"], rec, finv, tast, ig.DFT_COMPAT)
    ns = len(c["srcs"])
    b = ig.BatchContext(G, ns, len(finv), len(rec), ig.DFT_COMPAT, 4)
    b.set_cpml(4000.0, 50.0)
    b.set_freqs(finv, tast)
    b.set_model(*[G.to_dev(a) for a in start])
    bv, bs, table = b.autotune(reps=1)
    _note(f"autotune picked v={bv} s={bs}")
    assert len(table) == len(ig.BatchContext.tune_candidates()) and all(t["v_ms"] > 0 for t in table)
    grads = [G.empty() for _ in range(3)]
    b.run(c["srcs"], rec, grads=grads, d_obs=ref["obs"])
    for n in range(3):
        assert max_ulp(G.from_dev(grads[n]), ref["grads"][n]) == 0, f"gradient {n}"
    b.close()


@pytest.mark.parametrize("tiles,L", [(None, 1), (((4, 0, 1, 1, 1), (2, 0, 4, 1, 1)), 1), (None, 3),
                                     (((4, 0, 1, 1, 1), (2, 0, 4, 1, 1)), 3)])
def test_batch_viscoelastic(tiles, L):
    """the batched path with the viscoelastic update (L mechanisms) against
    the plain one"""
    import ifosgpu as ig
    c = CASES["slab_sources"]
    g = small_grid(nx=36, ny=34, nz=38, fw=8, nt=100)
    rec = np.array([[9 + 2 * n, 10, 19] for n in range(8)] + [[20, 5, 21]], np.int32)
    finv, tast = [40.0, 60.0], 100
    rng = np.random.default_rng(21)
    visco = ((0.02 + 0.03 * rng.random(g.shape)).astype(np.float32),
             (0.02 + 0.03 * rng.random(g.shape)).astype(np.float32),
             np.array([0.019, 0.0071, 0.043][:L], np.float32))
    _note(f"batched viscoelastic, tiles {tiles}, L={L}")
    true, start = _models(g)
    G = _setup(ig, g, finv, tast, ig.DFT_COMPAT)
    ref = _plain(ig, G, g, true, start, c["srcs"], rec, finv, tast, ig.DFT_COMPAT, visco=visco)
    out = _batched(ig, G, g, true, start, c["srcs"], rec, finv, tast, ig.DFT_COMPAT, 4, False, tiles=tiles,
                   visco=visco)
    for n in range(len(c["srcs"])):
        assert max_ulp(out["syn"][n], ref["syn"][n]) == 0, f"synthetics, shot {n}"
    for n in range(3):
        assert max_ulp(out["grads"][n], ref["grads"][n]) == 0, f"gradient {n}"
    out["ctx"].close()


@pytest.mark.parametrize("nz,free_surf,visco,tiles", [
    (70, 0, False, ((4, 32, 0, 0, 0, 1, 1, 1), (4, 32, 0, 0, 0, 1, 1, 1))),
    (70, 1, False, ((2, 5, 0, 0, 0, 1, 1, 1), (8, 0, 1, 0, 0, 1, 1, 1))),
    (70, 0, True, ((4, 0, 0, 0, 1, 1, 1, 1), (2, 7, 0, 0, 0, 1, 1, 1))),
    (140, 0, False, ((4, 0, 0, 1, 0, 1, 1, 1), (4, 0, 0, 1, 0, 1, 1, 1))),
    (140, 1, True, ((8, 9, 3, 0, 0, 1, 1, 1), (8, 9, 3, 0, 0, 1, 1, 1))),
    (70, 0, 3, ((4, 0, 0, 0, 1, 1, 1, 1), (4, 32, 0, 0, 0, 1, 1, 1))),
    (140, 0, 3, ((4, 0, 0, 1, 0, 1, 1, 1), (8, 9, 3, 0, 0, 1, 1, 1)))])
def test_batch_lds_march(nz, free_surf, visco, tiles):
    """the LDS-staged march (64-lane rows: nz >= 64; partial k and i tiles,
    CPML slabs in the same launch or split off, free surface, viscoelastic)
    gives the same bits as the plain path"""
    import ifosgpu as ig
    g = small_grid(nx=30, ny=28, nz=nz, fw=6, nt=80, free_surf=free_surf)
    srcs = [(15, 10, 35, 0.0, 50.0, 1.0, 1), (4, 12, nz - 4, 0.0, 50.0, 1.0, 1), (20, 14, 64, 0.0, 45.0, 1.0, 1)]
    rec = np.array([[3 + 3 * n, 9, 60 + n] for n in range(8)] + [[29, 27, nz], [12, 2, 1]], np.int32)
    finv, tast = [40.0, 60.0], 100
    vis = None
    if visco:
        rng = np.random.default_rng(5)
        # visco: the number of relaxation mechanisms (True: one)
        vis = ((0.02 + 0.03 * rng.random(g.shape)).astype(np.float32),
               (0.02 + 0.03 * rng.random(g.shape)).astype(np.float32),
               np.array([0.019, 0.0071, 0.043][:int(visco)], np.float32))
    _note(f"lds march nz={nz} free_surf={free_surf} visco={visco} tiles {tiles}")
    true, start = _models(g)
    G = _setup(ig, g, finv, tast, ig.DFT_COMPAT)
    ref = _plain(ig, G, g, true, start, srcs, rec, finv, tast, ig.DFT_COMPAT, visco=vis)
    out = _batched(ig, G, g, true, start, srcs, rec, finv, tast, ig.DFT_COMPAT, 4, False, tiles=tiles, visco=vis)
    for n in range(len(srcs)):
        assert max_ulp(out["syn"][n], ref["syn"][n]) == 0, f"synthetics, shot {n}"
    for n in range(3):
        assert max_ulp(out["grads"][n], ref["grads"][n]) == 0, f"gradient {n}"
    out["ctx"].close()


@pytest.mark.parametrize("fz,prec_name,chunk,use_graph", [((16, 0), "compat", 4, False), ((8, 5), "compat", 3, True),
                                                          ((16, 7), "kahan", 5, False), ((8, 0), "compat", 8, False)])
def test_batch_fused_step(fz, prec_name, chunk, use_graph):
    """the fused stress(n) + velocity(n+1) step (all points inside the
    interior box, no free surface) against the plain path, bit for bit"""
    import ifosgpu as ig
    g = small_grid(nx=40, ny=37, nz=70, fw=8, nt=130)
    srcs = [(18, 12, 19, 0.0, 50.0, 1.0, 1), (25, 14, 40, 0.0, 45.0, 1.0, 1), (12, 20, 55, 0.001, 55.0, 1e6, 4)]
    rec = np.array([[10 + 2 * n, 10, 19 + 3 * n] for n in range(9)] + [[14, 10, 22], [14, 10, 22]], np.int32)
    finv, tast = [40.0, 60.0, 80.0], 100
    prec = ig.DFT_KAHAN if prec_name == "kahan" else ig.DFT_COMPAT
    _note(f"fused step ti={fz[0]} jc={fz[1]} prec={prec_name} chunk={chunk} graph={use_graph}")
    true, start = _models(g)
    G = _setup(ig, g, finv, tast, prec)
    ref = _plain(ig, G, g, true, start, srcs, rec, finv, tast, prec)
    ns, ntr = len(srcs), len(rec)
    b = ig.BatchContext(G, ns, len(finv), ntr, prec, chunk)
    b.set_fused(True, *fz)
    b.set_cpml(4000.0, 50.0)
    b.set_freqs(finv, tast)
    d_syn = ig.DBuf((ns, 3, ntr, g.nt + 1))
    b.set_model(*[G.to_dev(a) for a in true])
    b.run(srcs, rec, d_syn=d_syn, want_adjoint=False, use_graph=use_graph)
    assert b.last_fused()
    obs_b = d_syn.numpy().copy()
    for n in range(ns):
        assert max_ulp(obs_b[n], ref["obs"][n].numpy()) == 0, f"observed data, shot {n}"
    b.set_model(*[G.to_dev(a) for a in start])
    for rep in range(2 if use_graph else 1):
        grads = [G.empty() for _ in range(3)]
        obs = [ig.DBuf.from_numpy(obs_b[n]) for n in range(ns)]
        l2 = b.run(srcs, rec, d_syn=d_syn, grads=grads, d_obs=obs, use_graph=use_graph)
        assert b.last_fused()
        syn = d_syn.numpy()
        for n in range(ns):
            assert max_ulp(syn[n], ref["syn"][n]) == 0, f"rep {rep}: synthetics, shot {n}"
            assert l2[n] == ref["l2"][n], f"rep {rep}: misfit, shot {n}"
        fe = G.field_elems()
        dft_last = [ig.DBuf.wrap(p, (len(finv) * fe,)).numpy() for p in b.dft_ptrs(ns - 1, 1)]
        for n in range(6):
            assert max_ulp(dft_last[n], ref["dft_last"][n]) == 0, f"rep {rep}: adjoint DFT component {n}"
        for n in range(3):
            assert max_ulp(G.from_dev(grads[n]), ref["grads"][n]) == 0, f"rep {rep}: gradient {n}"
        ig.DBuf((1 << 20,)).zero_()
    b.close()
