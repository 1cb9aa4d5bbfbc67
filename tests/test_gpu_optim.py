"""The device optimiser steps against the oracle (oracle/optim_oracle.py).

ifos_lbfgs_update runs lbfgs.c's two-loop recursion with FP64 dot products;
the oracle in its FP64-dot mode forms the same sums in another order, so the
FP32 results agree to a few ulp, not bit for bit (the reference itself sums
in FP32 and all-reduces, which orders the sums differently again).  The
history slots, bfgsscale and the saved gradients are compared too, over more
iterations than slots so that the window wraps.  ifos_conjugate is checked
the same way through a cdf 1 call and two cdf 0 calls.
"""
import os
import sys

import numpy as np
import pytest

from helpers import small_grid

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
from optim_oracle import LbfgsOracle, conjugate  # noqa: E402

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)


def _geom(ig, g):
    return ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt, g.fdcoeff, g.free_surf)


def _field(g, rng, scale=1.0):
    a = np.zeros(g.shape, np.float32)
    g.interior(a)[...] = (scale * rng.standard_normal((g.ny, g.nx, g.nz))).astype(np.float32)
    return a


def _close(a, b, rtol=2e-6):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b).max() <= rtol * np.abs(b).max()


@pytest.mark.parametrize("numpar,bfgsnum", [(1, 3), (2, 3), (2, 1)])
def test_lbfgs_matches_oracle(numpar, bfgsnum):
    import ifosgpu as ig
    g = small_grid(nx=18, ny=16, nz=20, fw=4)
    _note(f"lbfgs numpar={numpar} bfgsnum={bfgsnum}")
    G = _geom(ig, g)
    rng = np.random.default_rng(40 + numpar + bfgsnum)
    n1 = g.nx * g.ny * g.nz
    o = LbfgsOracle(numpar * n1, bfgsnum, dot="f64", vec=np.float32)
    lb = ig.Lbfgs(G, numpar, bfgsnum)
    try:
        grads = [_field(g, rng) for _ in range(2)]
        d = [G.to_dev(a) for a in grads]
        compact = lambda gs: np.concatenate([g.interior(a).ravel() for a in gs[:numpar]])
        lb.savegrad(d[0], d[1] if numpar == 2 else None)
        o.savegrad(compact(grads))
        for it in range(1, 7):
            # the model step of iteration it (modelupdate.c:146-149), written by the caller
            w = it % bfgsnum or bfgsnum
            s = (0.01 * rng.standard_normal(numpar * n1)).astype(np.float32)
            o.mod[w - 1] = s
            lb.mod(w).copy_from(ig.DBuf.from_numpy(s))
            # a new gradient correlated with the old one so that y.s stays away from 0
            grads = [(0.7 * a + 0.3 * _field(g, rng)).astype(np.float32) for a in grads]
            d = [G.to_dev(a) for a in grads]
            h0 = lb.update(d[0], d[1] if numpar == 2 else None, it + 1)
            q, h0_o = o.update(compact(grads), it + 1)
            assert h0 == pytest.approx(h0_o, rel=1e-12)
            got = np.concatenate([g.interior(G.from_dev(t)).ravel() for t in d[:numpar]])
            assert _close(got, q), f"direction, iteration {it + 1}"
            np.testing.assert_allclose(lb.scale(), o.scale, rtol=1e-12)
            for v in range(1, bfgsnum + 1):
                assert _close(lb.grad(v).numpy(), o.grad[v - 1]), f"bfgsgrad slot {v}"
            # the next iteration starts from the oracle's direction so both agree bitwise on inputs
            grads = [a.copy() for a in grads]
    finally:
        lb.close()


def test_conjugate_matches_oracle():
    import ifosgpu as ig
    g = small_grid(nx=18, ny=16, nz=20, fw=4)
    _note("conjugate gradient")
    G = _geom(ig, g)
    rng = np.random.default_rng(77)
    grads = [_field(g, rng) for _ in range(3)]
    priors = [np.zeros(g.shape, np.float32) for _ in range(6)]
    dp = [G.to_dev(a) for a in priors]
    for step, cdf in enumerate((1, 0, 0)):
        dg = [G.to_dev(a) for a in grads]
        beta, sums = ig.conjugate(G, dg, dp, cdf)
        og = [g.interior(a).copy() for a in grads]
        op = [g.interior(a).copy() for a in priors]
        beta_o, sums_o = conjugate(og, op, cdf)
        if cdf == 0:
            np.testing.assert_allclose(sums, sums_o, rtol=1e-12)
        np.testing.assert_allclose(beta, beta_o, rtol=1e-6)
        for n in range(3):
            assert _close(g.interior(G.from_dev(dg[n])), og[n]), f"grad{n + 1}, call {step}"
        for n in range(6):
            assert _close(g.interior(G.from_dev(dp[n])), op[n]), f"prior{n + 1}, call {step}"
        # next call: the oracle's state, and a new gradient
        for n in range(6):
            g.interior(priors[n])[...] = op[n]
        dp = [G.to_dev(a) for a in priors]
        grads = [(0.6 * _field(g, rng) + 0.4 * a).astype(np.float32) for a in grads]


# ---- precongrad.c and modelupdate.c on the device -------------------------
from optim_oracle import modelupdate as mu_oracle, precon_grad as pg_oracle  # noqa: E402


def _max_ulp(a, b):
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    return int(np.abs(a - b).max()) if a.size else 0


@pytest.mark.parametrize("damptype,free_surf", [(0, 0), (1, 0), (2, 0), (3, 0), (2, 1)])
def test_precon_grad_matches_oracle(damptype, free_surf):
    """double exp on the device against numpy's: the tapers are stored to
    float, so a last-bit difference of exp rarely moves them; a cell's factor
    chain is a product of (sources + receivers + 6) of them, so agreement is
    to a few ulp"""
    import ifosgpu as ig
    g = small_grid(nx=22, ny=18, nz=26, fw=4, free_surf=free_surf)
    _note(f"precon_grad DAMPTYPE={damptype} FREE_SURF={free_surf}")
    G = _geom(ig, g)
    rng = np.random.default_rng(60 + damptype)
    grads = [_field(g, rng) for _ in range(3)]
    src = np.array([[11, 5, 13], [3, 9, 20]], np.int32)
    rec = np.array([[2 + 2 * n, 4, 10 + n] for n in range(9)], np.int32)
    d = [G.to_dev(a) for a in grads]
    ig.precon_grad(G, d, damptype, src, rec)
    og = [g.interior(a).copy() for a in grads]
    pg_oracle(og, damptype, src, rec, (g.nx, g.ny, g.nz), g.fw, free_surf)
    for n in range(3):
        got = g.interior(G.from_dev(d[n]))
        assert _max_ulp(got, og[n]) <= 4, f"grad{n + 1}"


@pytest.mark.parametrize("lbfgs,numpar", [(False, 2), (True, 2), (True, 1)])
def test_modelupdate_matches_oracle(lbfgs, numpar):
    import ifosgpu as ig
    g = small_grid(nx=20, ny=16, nz=24, fw=4)
    _note(f"modelupdate lbfgs={lbfgs} numpar={numpar}")
    G = _geom(ig, g)
    rng = np.random.default_rng(70 + numpar)
    shp = g.shape
    rho = (2000.0 * (1 + 0.1 * rng.random(shp))).astype(np.float32)
    vp = (4000.0 * (1 + 0.1 * rng.random(shp))).astype(np.float32)
    vs = (2300.0 * (1 + 0.1 * rng.random(shp))).astype(np.float32)
    pi = (rho * vp * vp).astype(np.float32)
    u = (rho * vs * vs).astype(np.float32)
    grads = [_field(g, rng) for _ in range(3)]
    d = [G.to_dev(a) for a in grads]
    dm = [G.to_dev(a) for a in (rho, pi, u)]
    n1 = g.nx * g.ny * g.nz
    mod = ig.DBuf((numpar * n1,)) if lbfgs else None
    mx = ig.modelupdate(G, d, *dm, 0.02, (1.0, 0.8, 0.5), 4000.0, 2300.0, 2000.0, bfgsmod=mod, numpar=numpar)
    o = [g.interior(a).copy() for a in (rho, pi, u)]
    og = [g.interior(a).copy() for a in grads]
    mx_o, mod_o = mu_oracle(og, o[0], o[1], o[2], 0.02, (1.0, 0.8, 0.5), 4000.0, 2300.0, 2000.0, lbfgs, numpar)
    np.testing.assert_array_equal(mx, mx_o)
    for n, name in enumerate(("rho", "pi", "u")):
        assert _max_ulp(g.interior(G.from_dev(dm[n])), o[n]) == 0, name
    for n in range(3):
        assert _max_ulp(g.interior(G.from_dev(d[n])), og[n]) == 0, f"grad{n + 1}"
    if lbfgs:
        assert _max_ulp(mod.numpy(), mod_o) == 0, "bfgsmod"


def test_steplength_f64_matches_oracle():
    """ifos_steplength_f64 (one device thread, FP64) against the FP64 oracle
    bit for bit, and the float entry point against the reference's own
    steplength.c outputs (tests/golden/optim_golden.npz)"""
    import ifosgpu as ig
    from optim_oracle import steplength
    cases = [([3.0, 1.0, 2.0], [0, 0.5, 1.0, 0, 1.0], 0, 1),
             ([1.0, 3.0, 2.0], [0, 0.5, 1.0, 0, 1.0], 0, 1),
             ([2.0, 1.0, 3.5], [0, 0.5, 1.0, 0, 0.1], 0, 1),
             ([5.0, 2.0, 1.0], [0, 0.5, 1.0, 0, 4.0], 1, 3),
             ([2.0, 3.0, 1.0], [0, 0.5, 1.0, 0, 1.0], 0, 1),
             ([1.2345678901, 0.98765432, 1.1111111], [0, 0.013, 0.026, 0, 0.02], 0, 1)]
    for L2, st, lb, it in cases:
        want = steplength(L2, st, lb, it, 0.8, dtype=np.float64)
        got = ig.steplength_f64(L2, st, lb, it, 0.8)
        np.testing.assert_array_equal(got[1], want[1])
        np.testing.assert_array_equal(got[0][3:], want[0][3:])
    F = np.load(os.path.join(os.path.dirname(__file__), "golden", "optim_golden.npz"), allow_pickle=False)
    for c in range(5):
        x = F[f"step_{c}_in"]
        l2, s = ig.steplength(x[:3], x[3:8], int(x[8]), int(x[9]), float(x[10]))
        np.testing.assert_array_equal(s, F[f"step_{c}_step"])
        np.testing.assert_array_equal(l2[3:], F[f"step_{c}_L2"][3:])
