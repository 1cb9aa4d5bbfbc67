"""The oracle's consistent (exact discrete) adjoint, og_run_shot_adj, pinned
by what the mathematics fixes.

* Its forward is og_run_shot's, bit for bit (the same seismograms).
* J = 1/2 sum (syn - obs)^2 is what the seismograms give.
* Its gradient is the exact derivative of J: central finite differences of
  J in FP64 along model perturbations agree with <gradient, delta> to 1e-6
  (the FD error falls as the square of the step, ~2e-8 at the step used).  The directions are
  chosen so that a plausible mistake fails one of them: vp only, vs only,
  rho only (a wrong term or sign in one parameter's chain), one inside a
  CPML layer (the transposed memory-variable recursion), and one on the
  first and last layer of cells of each axis (av_mat's averages and
  matcopy's halo copies, which wrap cell 1 into the N+1 halo).
* With obs equal to the synthetics, J = 0 and the gradient vanishes.
"""
import numpy as np
import pytest

from helpers import det_uniform
from pyoracle import Grid, Oracle


def _grid():
    return Grid(nx=24, ny=22, nz=26, fw=4, dx=2.0, dy=2.0, dz=2.0, dt=2e-4, nt=120, fdcoeff=2)


def _coords(g):
    return np.meshgrid(*(np.arange(s) - 2.0 for s in g.shape), indexing="ij")


def _bump(g, cj, ci, ck, r):
    J, I, K = _coords(g)
    return np.exp(-((J - cj) ** 2 + (I - ci) ** 2 + (K - ck) ** 2) / r ** 2)


def _model(vp, vs, rho):
    return [rho, rho * vp * vp, rho * vs * vs]


SRC = (12, 6, 13, 0.0, 150.0, 1.0, 1)
REC = np.array([[4 + 2 * n, 17, 13] for n in range(9)] + [[12, 18, 6], [12, 18, 20], [2, 3, 24]], np.int32)


@pytest.fixture(scope="module")
def setup():
    g = _grid()
    one = np.ones(g.shape)
    base = (4000.0 * one, 2300.0 * one, 2000.0 * one)
    b = _bump(g, 11, 12, 13, 4)
    true = (base[0] * (1 + 0.05 * b), base[1] * (1 + 0.05 * b), base[2] * (1 + 0.03 * b))
    o = Oracle("f64")
    obs = o.run_shot(g, *_model(*true), SRC, REC, [100.0], want_adjoint=False)["syn"]
    r = o.run_shot_adj(g, *_model(*base), SRC, REC, obs)
    return g, o, base, obs, r


def _fd(g, o, base, obs, dirs, eps=1e-3):
    Jp = o.run_shot_adj(g, *_model(*[b * (1 + eps * d) for b, d in zip(base, dirs)]), SRC, REC, obs)["J"]
    Jm = o.run_shot_adj(g, *_model(*[b * (1 - eps * d) for b, d in zip(base, dirs)]), SRC, REC, obs)["J"]
    return (Jp - Jm) / (2 * eps)


def _ad(g, grads, base, dirs):
    return sum(float((g.interior(grads[q]) * g.interior(base[q] * dirs[q])).sum()) for q in range(3))


def test_forward_and_misfit(setup):
    g, o, base, obs, r = setup
    fwd = o.run_shot(g, *_model(*base), SRC, REC, [100.0], want_adjoint=False)["syn"]
    for q in range(3):
        assert np.array_equal(fwd[q][:, 1:], r["syn"][q][:, 1:])
    J = 0.5 * sum(float(((r["syn"][q][:, 1:] - obs[q][:, 1:]) ** 2).sum()) for q in range(3))
    assert r["J"] == pytest.approx(J, rel=1e-12)
    assert r["J"] > 0


def _smooth_dir(g, seed):
    import scipy.ndimage as nd
    d = nd.gaussian_filter(det_uniform(g.shape, seed, -1.0, 1.0).astype(np.float64), 2.0)
    return d / np.sqrt((d ** 2).mean())


@pytest.mark.parametrize("name", ["vp", "vs", "rho", "cpml", "edges", "all"])
def test_gradient_is_the_derivative(setup, name):
    g, o, base, obs, r = setup
    zero = np.zeros(g.shape)
    if name in ("vp", "vs", "rho"):
        q = ("vp", "vs", "rho").index(name)
        dirs = [zero] * 3
        dirs[q] = 0.01 * _bump(g, 10, 13, 12, 3)
    elif name == "cpml":
        # inside the low-x and high-z layers, where the memory variables act
        b = _bump(g, 11, 2, 13, 2) + _bump(g, 11, 12, 24, 2)
        dirs = [0.01 * b, 0.01 * b, 0.01 * b]
    elif name == "edges":
        # the first and the last layer of every axis (cell 1 feeds the N+1 halo)
        J, I, K = _coords(g)
        m = ((J == 1) | (J == g.ny) | (I == 1) | (I == g.nx) | (K == 1) | (K == g.nz)).astype(np.float64)
        dirs = [0.01 * m, 0.02 * m, 0.03 * m]
    else:
        dirs = [0.01 * _smooth_dir(g, 100 + q) for q in range(3)]
    fd = _fd(g, o, base, obs, dirs)
    ad = _ad(g, r["grads"], base, dirs)
    print(f"{name}: FD {fd:.10e}  <grad, delta> {ad:.10e}", flush=True)
    assert fd != 0.0
    assert abs(fd - ad) <= 1e-6 * abs(fd)


def test_zero_residual_zero_gradient(setup):
    """(the f32 build: its seismograms are the float32 samples it returns)"""
    g, _, base, _, _ = setup
    o = Oracle("f32")
    m = [a.astype(np.float32) for a in _model(*base)]
    r = o.run_shot_adj(g, *m, SRC, REC, [np.zeros((len(REC), g.nt + 1), np.float32)] * 3)
    r0 = o.run_shot_adj(g, *m, SRC, REC, r["syn"])
    assert r0["J"] == 0.0
    for q in range(3):
        assert not np.any(r0["grads"][q])


def test_rejects_what_it_does_not_cover():
    g = _grid()
    g.free_surf = 1
    o = Oracle("f64")
    one = np.ones(g.shape)
    obs = [np.zeros((len(REC), g.nt + 1), np.float32) for _ in range(3)]
    with pytest.raises(ValueError):
        o.run_shot_adj(g, *_model(4000.0 * one, 2300.0 * one, 2000.0 * one), SRC, REC, obs)
