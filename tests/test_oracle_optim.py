"""The optimiser oracle (oracle/optim_oracle.py) pinned without the device.

* the L-BFGS two-loop recursion (lbfgs.c:26-181) equals the closed-form
  BFGS inverse-Hessian recursion H <- V^T H V + rho s s^T from H0 = h0 I
  (Nocedal & Wright eq. 6.17 / Algorithm 9.1) applied to the gradient, over
  a window that wraps round the BFGSNUM history slots, on a quadratic; in
  FP64 vectors to 1e-11, and with FP32 vectors with both FP32 (the
  reference's) and FP64 (the device's) sums to 1e-4;
* conjugategrad.c's beta is the Polak-Ribiere ratio g.(g - g_prev) /
  g_prev.g_prev per parameter, and its clamps as written leave beta[1] and
  beta[2] unclamped (a hand-worked case).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
from optim_oracle import LbfgsOracle, conjugate, dense_bfgs_direction  # noqa: E402


def _quadratic_run(n, bfgsnum, vec, dot, iters, seed=1):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    A = A @ A.T + n * np.eye(n)

    def grad(x):  # the reference's grad is minus the misfit gradient
        return -(A @ x)

    o = LbfgsOracle(n, bfgsnum, dot=dot, vec=vec)
    x = rng.standard_normal(n)
    g = grad(x)
    o.savegrad(g)
    pairs, errs, q = [], [], None
    for it in range(1, iters + 1):
        s = 0.1 * np.asarray(g if it == 1 else q, np.float64)  # step along the last direction
        o.mod[o.slot(it)] = s                                 # modelupdate.c:146-149 writes slot it
        x_new = x + s
        g_new = grad(x_new)
        pairs.append((s, np.asarray(g, np.float64) - np.asarray(g_new, np.float64)))
        q, h0 = o.update(g_new, it + 1)
        s_, y_ = pairs[-1]
        assert h0 == pytest.approx(np.dot(y_, s_) / np.dot(y_, y_), rel=1e-5)
        ref = dense_bfgs_direction(pairs[-bfgsnum:], np.dot(y_, s_) / np.dot(y_, y_), g_new)
        errs.append(np.abs(np.asarray(q, np.float64) - ref).max() / np.abs(ref).max())
        x, g = x_new, g_new
    return errs


@pytest.mark.parametrize("bfgsnum", [1, 2, 3, 7])
def test_two_loop_equals_dense_bfgs(bfgsnum):
    assert max(_quadratic_run(9, bfgsnum, np.float64, "f64", 8)) < 1e-11


@pytest.mark.parametrize("dot", ["f32", "f64"])
def test_two_loop_fp32_vectors(dot):
    assert max(_quadratic_run(9, 3, np.float32, dot, 6, seed=4)) < 1e-4


def test_lbfgs_slots_and_scale():
    """y lands in slot (iteration-1) % BFGSNUM (0 -> BFGSNUM), the raw
    gradient in slot iteration % BFGSNUM, rho = 1 / y.s"""
    o = LbfgsOracle(3, 2, dot="f64", vec=np.float64)
    g1 = np.array([1.0, 2.0, 3.0])
    o.savegrad(g1)
    o.mod[0] = [0.5, 0.0, 0.0]
    g2 = np.array([0.0, 1.0, 1.0])
    o.update(g2, 2)
    np.testing.assert_array_equal(o.grad[0], g1 - g2)    # y in slot 1
    np.testing.assert_array_equal(o.grad[1], g2)         # g2 kept in slot 2
    assert o.scale[0] == 1.0 / np.dot(g1 - g2, o.mod[0])


def test_conjugate_polak_ribiere_and_clamps():
    g = [np.array([1.0, 2.0], np.float32), np.array([1.0, 0.0], np.float32), np.array([3.0, 0.0], np.float32)]
    p = [np.array([1.0, 1.0], np.float32), np.array([2.0, 1.0], np.float32), np.array([1.0, 1.0], np.float32),
         np.array([0.5, -1.0], np.float32), np.array([1.0, 2.0], np.float32), np.array([-1.0, 0.25], np.float32)]
    g0 = [a.copy() for a in g]
    beta, sums = conjugate(g, p, 0)
    # sums: g.(g - p) and p.p per parameter
    assert list(sums) == [2.0, 2.0, -1.0, 5.0, 6.0, 2.0]
    # beta[0] = 1 clamped to 0.5; beta[1] = -0.2 and beta[2] = 3 left alone
    assert beta[0] == np.float32(0.5) and beta[1] == np.float32(-0.2) and beta[2] == np.float32(3.0)
    np.testing.assert_array_equal(g[0], np.float32(0.5) * np.array([0.5, -1.0], np.float32) + g0[0])
    np.testing.assert_array_equal(g[1], np.float32(-0.2) * np.array([1.0, 2.0], np.float32) + g0[1])
    np.testing.assert_array_equal(p[3], g[0])
    np.testing.assert_array_equal(p[0], g0[0])   # prior1..3 = the gradient before the update


def test_conjugate_first_call_copies():
    g = [np.full(3, n + 1.0, np.float32) for n in range(3)]
    p = [np.zeros(3, np.float32) for _ in range(6)]
    beta, sums = conjugate(g, p, 1)
    assert sums is None and not beta.any()
    for n in range(3):
        np.testing.assert_array_equal(p[n], g[n])
        np.testing.assert_array_equal(p[n + 3], g[n])
