"""Grid levels of uneven size run together (ifos_levels_run).

Three batch contexts on three grids (different sizes, spacings, time steps
and step counts, as the coarse-to-fine levels of an inversion) run their
forward solves in one call: enqueued on the contexts' own streams, or
captured into one graph with a branch per level and replayed.  Each level's
seismograms must be bit-identical to ifos_batch_run on that context alone,
and the first shot of each level to the plain one-shot path.  The graph is
replayed with new sources, and recaptured when a level's shot count or tile
shape changes.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import max_ulp, random_model, small_grid

pytestmark = pytest.mark.gpu

# (nx, ny, nz, fw, dh, dt, nt, shots)
LEVELS = ((28, 26, 30, 6, 8.0, 6e-4, 60, 3), (36, 34, 38, 6, 6.0, 4.5e-4, 80, 2),
          (44, 42, 46, 8, 4.0, 3e-4, 110, 3))


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _level(ig, n, lv, rng):
    nx, ny, nz, fw, dh, dt, nt, ns = lv
    g = small_grid(nx=nx, ny=ny, nz=nz, fw=fw, nt=nt, dh=dh, dt=dt)
    G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt, g.fdcoeff, g.free_surf)
    model = random_model(g, rng)
    rec = np.array([[fw + 2 + 3 * r, fw + 3, nz // 2] for r in range(6)] + [[nx // 2, 3, nz // 2]], np.int32)
    return dict(g=g, G=G, model=model, rec=rec, ns=ns)


def _srcs(L, rng, ns):
    g = L["g"]
    out = []
    for _ in range(ns):
        i, j, k = (int(rng.integers(g.fw + 2, n - g.fw - 1)) for n in (g.nx, g.ny, g.nz))
        out.append((i, j, k, 0.0, float(rng.uniform(30.0, 60.0)), 1.0, 1))
    return out


def _ctx(ig, L, ns):
    b = ig.BatchContext(L["G"], ns, 1, len(L["rec"]), ig.DFT_COMPAT, 4)
    b.set_cpml(4000.0, 50.0)
    b.set_model(*[L["G"].to_dev(a) for a in L["model"]])
    return b


def _alone(ig, L, srcs):
    """the level's shots through ifos_batch_run on a context of its own"""
    b = _ctx(ig, L, len(srcs))
    d = ig.DBuf((len(srcs), 3, len(L["rec"]), L["g"].nt + 1))
    b.run(srcs, L["rec"], d_syn=d, want_adjoint=False)
    out = d.numpy()
    b.close()
    return out


def _plain_first(ig, L, src):
    ctx = ig.ShotContext(L["G"], 1, len(L["rec"]), ig.DFT_COMPAT)
    ctx.set_cpml(4000.0, 50.0)
    ctx.set_freqs([40.0], 100)
    ctx.set_model(*[L["G"].to_dev(a) for a in L["model"]])
    d = ig.DBuf((3, len(L["rec"]), L["g"].nt + 1))
    ctx.run_shot(src, L["rec"], d, None, want_adjoint=False)
    out = d.numpy()
    ctx.close()
    return out


def _check(ig, levels, srcs, got, what):
    for n, L in enumerate(levels):
        want = _alone(ig, L, srcs[n])
        assert max_ulp(got[n], want[:len(srcs[n])]) == 0, f"{what}: level {n} against its batch alone"


@pytest.mark.parametrize("use_graph", [False, True])
def test_levels_match_alone(use_graph):
    import ifosgpu as ig
    rng = np.random.default_rng(4)
    levels = [_level(ig, n, lv, rng) for n, lv in enumerate(LEVELS)]
    ctxs = [_ctx(ig, L, L["ns"]) for L in levels]
    try:
        for rep in range(2):
            _note(f"levels graph={use_graph} call {rep}")
            srcs = [_srcs(L, rng, L["ns"]) for L in levels]
            d = [ig.DBuf((L["ns"], 3, len(L["rec"]), L["g"].nt + 1)) for L in levels]
            ig.levels_run(ctxs, srcs, [L["rec"] for L in levels], d_syn=d, use_graph=use_graph)
            got = [t.numpy() for t in d]
            for n, L in enumerate(levels):
                assert np.abs(got[n]).max() > 0, f"level {n} recorded nothing"
            _check(ig, levels, srcs, got, f"call {rep}")
            if rep == 0:
                for n, L in enumerate(levels):
                    p = _plain_first(ig, L, srcs[n][0])
                    assert max_ulp(got[n][0], p) == 0, f"level {n}, shot 0 against the plain path"
        if use_graph:
            # fewer shots on one level, then another tile shape on another: both recapture
            _note("levels graph: fewer shots on level 1")
            srcs = [_srcs(L, rng, L["ns"]) for L in levels]
            srcs[1] = srcs[1][:1]
            d = [ig.DBuf((len(s), 3, len(L["rec"]), L["g"].nt + 1)) for s, L in zip(srcs, levels)]
            ig.levels_run(ctxs, srcs, [L["rec"] for L in levels], d_syn=d, use_graph=True)
            _check(ig, levels, srcs, [t.numpy() for t in d], "fewer shots")
            _note("levels graph: new tiles on level 2")
            ctxs[2].set_tiles((2, 0, 0, 0, 1), (8, 32, 0, 0, 0))
            srcs = [_srcs(L, rng, L["ns"]) for L in levels]
            d = [ig.DBuf((L["ns"], 3, len(L["rec"]), L["g"].nt + 1)) for L in levels]
            ig.levels_run(ctxs, srcs, [L["rec"] for L in levels], d_syn=d, use_graph=True)
            _check(ig, levels, srcs, [t.numpy() for t in d], "new tiles")
    finally:
        for b in ctxs:
            b.close()


def test_levels_errors():
    import ifosgpu as ig
    _note("levels errors")
    rng = np.random.default_rng(5)
    L = _level(ig, 0, LEVELS[0], rng)
    b = _ctx(ig, L, 2)
    try:
        srcs = _srcs(L, rng, 2)
        with pytest.raises(ig.IfosError, match="share one context"):
            ig.levels_run([b, b], [srcs, srcs], [L["rec"], L["rec"]])
        with pytest.raises(ig.IfosError, match="need 1..2 shots"):
            ig.levels_run([b], [_srcs(L, rng, 3)], [L["rec"]])
        # an adjoint request is refused before anything runs
        shots, _keep = b._shots(srcs, L["rec"], None, 1, True)
        arr = (C.POINTER(ig.IfosShot) * 1)(C.cast(shots, C.POINTER(ig.IfosShot)))
        rc = ig.lib().ifos_levels_run((C.c_void_p * 1)(b.h.value), 1, arr, (C.c_int * 1)(2), None, 0)
        assert rc != 0 and b"forward solves only" in ig.lib().ifos_last_error()
    finally:
        b.close()
