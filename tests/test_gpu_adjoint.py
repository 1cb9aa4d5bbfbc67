"""The consistent (exact discrete) adjoint on the device,
ifos_ctx_run_shot_adj.

* Its seismograms are the plain path's forward, bit for bit, and its
  misfit J = 1/2 sum (syn - obs)^2.
* Its gradient agrees with the oracle's consistent adjoint
  (og_run_shot_adj, f32 build; pinned to 1e-6 by FP64 finite differences in
  test_oracle_adjoint.py) to 1e-4 relative L2 per parameter: the transposed
  steps round in another order (the FP32 adjoint fields) and the model sums
  are FP64 on both sides.
* Every checkpoint interval gives the same bits (the recomputed segments
  repeat the forward exactly).
* It is the derivative of the misfit: central finite differences of J
  computed by the FP64 oracle along a seeded smooth perturbation (seed 6,
  1 % rms of vp, vs and rho) agree with <gradient, delta> to better than
  1e-3 relative, as do the single-parameter directions.
"""
import sys

import numpy as np
import pytest

from helpers import det_uniform, rel_l2
from pyoracle import Grid, Oracle

pytestmark = pytest.mark.gpu

SRC = (12, 6, 13, 0.0, 150.0, 1.0, 1)
REC = np.array([[4 + 2 * n, 17, 13] for n in range(9)] + [[12, 18, 6], [12, 18, 20], [2, 3, 24]], np.int32)


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _grid(nt=120):
    return Grid(nx=24, ny=22, nz=26, fw=4, dx=2.0, dy=2.0, dz=2.0, dt=2e-4, nt=nt, fdcoeff=2)


def _bump(g, cj, ci, ck, r):
    J, I, K = np.meshgrid(*(np.arange(s) - 2.0 for s in g.shape), indexing="ij")
    return np.exp(-((J - cj) ** 2 + (I - ci) ** 2 + (K - ck) ** 2) / r ** 2)


def _model(vp, vs, rho):
    return [rho, rho * vp * vp, rho * vs * vs]


def _smooth_dir(g, seed):
    import scipy.ndimage as nd
    d = nd.gaussian_filter(det_uniform(g.shape, seed, -1.0, 1.0).astype(np.float64), 2.0)
    return d / np.sqrt((d ** 2).mean())


class _Case:
    def __init__(self):
        import ifosgpu as ig
        self.ig = ig
        g = _grid()
        self.g = g
        one = np.ones(g.shape)
        self.base = (4000.0 * one, 2300.0 * one, 2000.0 * one)
        b = _bump(g, 11, 12, 13, 4)
        true = (self.base[0] * (1 + 0.05 * b), self.base[1] * (1 + 0.05 * b), self.base[2] * (1 + 0.03 * b))
        self.obs = Oracle("f32").run_shot(g, *[a.astype(np.float32) for a in _model(*true)], SRC, REC, [100.0],
                                          want_adjoint=False)["syn"]
        self.G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt, g.fdcoeff, 0)
        self.ctx = ig.ShotContext(self.G, 1, len(REC))
        self.ctx.set_cpml(4000.0, 50.0)
        self.d_obs = ig.DBuf.from_numpy(np.stack(self.obs).astype(np.float32))
        self.d_syn = ig.DBuf((3, len(REC), g.nt + 1))

    def gradient(self, model32, every=0):
        G = self.G
        self.ctx.set_model(*[G.to_dev(a) for a in model32])
        grads = [G.empty() for _ in range(3)]
        J = self.ctx.run_shot_adj(SRC, REC, self.d_syn, grads, self.d_obs, every=every)
        self.ig.sync()
        return J, [G.from_dev(t) for t in grads], self.d_syn.numpy()


def test_consistent_gradient_matches_oracle():
    c = _Case()
    g = c.g
    m32 = [a.astype(np.float32) for a in _model(*c.base)]
    _note("consistent adjoint on the device and in the oracle (f32)")
    J, grads, syn = c.gradient(m32)
    ref = Oracle("f32").run_shot_adj(g, *m32, SRC, REC, c.obs)
    fwd = Oracle("f32").run_shot(g, *m32, SRC, REC, [100.0], want_adjoint=False)["syn"]
    for q in range(3):
        assert np.array_equal(syn[q][:, 1:], fwd[q][:, 1:]), f"forward seismograms, component {q}"
    assert J == pytest.approx(ref["J"], rel=1e-6)
    for q in range(3):
        a, b = g.interior(grads[q]), g.interior(ref["grads"][q])
        err = rel_l2(a, b)
        _note(f"gradient {q + 1}: rel-L2 vs oracle {err:.2e}")
        assert np.abs(b).max() > 0
        assert err <= 1e-4, f"gradient {q + 1}"
    c.ctx.close()


def test_checkpoint_intervals_give_the_same_bits():
    c = _Case()
    m32 = [a.astype(np.float32) for a in _model(*c.base)]
    J0, g0, _ = c.gradient(m32, every=0)
    for every in (1, 7, 19, c.g.nt):
        _note(f"checkpoint every {every} steps")
        J, gr, _ = c.gradient(m32, every=every)
        assert J == J0
        for q in range(3):
            assert np.array_equal(gr[q], g0[q]), f"every={every}, gradient {q + 1}"
    c.ctx.close()


@pytest.mark.parametrize("name", ["seeded", "vp", "vs", "rho"])
def test_gradient_is_the_derivative_of_the_misfit(name):
    """FD of J (FP64 oracle) against <device gradient, delta>, relative error < 1e-3"""
    c = _Case()
    g = c.g
    zero = np.zeros(g.shape)
    if name == "seeded":
        dirs = [0.01 * _smooth_dir(g, 6 + 1000 * q) for q in range(3)]
    else:
        dirs = [zero] * 3
        dirs[("vp", "vs", "rho").index(name)] = 0.01 * _bump(g, 10, 13, 12, 3)
    _, grads, _ = c.gradient([a.astype(np.float32) for a in _model(*c.base)])
    ad = sum(float((g.interior(grads[q]).astype(np.float64) * g.interior(c.base[q] * dirs[q])).sum())
             for q in range(3))
    o = Oracle("f64")
    eps = 1e-3
    Jp = o.run_shot_adj(g, *_model(*[b * (1 + eps * d) for b, d in zip(c.base, dirs)]), SRC, REC, c.obs)["J"]
    Jm = o.run_shot_adj(g, *_model(*[b * (1 - eps * d) for b, d in zip(c.base, dirs)]), SRC, REC, c.obs)["J"]
    fd = (Jp - Jm) / (2 * eps)
    err = abs(fd - ad) / abs(fd)
    _note(f"{name}: FD {fd:.8e}  <grad, delta> {ad:.8e}  rel {err:.2e}")
    assert err < 1e-3
    c.ctx.close()


def test_rejects_what_it_does_not_cover():
    import ifosgpu as ig
    c = _Case()
    m32 = [a.astype(np.float32) for a in _model(*c.base)]
    c.ctx.set_model(*[c.G.to_dev(a) for a in m32])
    grads = [c.G.empty() for _ in range(3)]
    force = (12, 6, 13, 0.0, 150.0, 1.0, 2)
    with pytest.raises(ig.IfosError):
        c.ctx.run_shot_adj(force, REC, c.d_syn, grads, c.d_obs)
    c.ctx.close()
