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


# ---- steplength.c, precongrad.c, modelupdate.c ----------------------------
from optim_oracle import modelupdate, precon_grad, steplength  # noqa: E402


@pytest.mark.parametrize("alpha,smin", [(2.0, 0.3), (0.5, 0.9), (7.0, 0.05)])
def test_steplength_parabola_minimum(alpha, smin):
    """three points of L2 = alpha (s - smin)^2 + 3: the fitted minimum is smin
    (before the reference's clamps) and L2[3] is 3"""
    st = [0.0, 0.5, 1.0, 0.0, 1.0]
    L2 = [alpha * (s - smin) ** 2 + 3.0 for s in st[:3]]
    l2, s = steplength(L2, st, 0, 1, 1.0)
    expect = max(min(smin, 2.5 * st[4]), 0.1 * st[1])
    assert s[3] == pytest.approx(expect, rel=1e-5)
    if expect == smin:
        assert l2[3] == pytest.approx(3.0, rel=1e-5)
    assert s[4] == pytest.approx(min(max(s[3] / 2, 0.2), 1.0), rel=1e-6)


def test_steplength_branches_match_the_library():
    """the a < 0 branches and the clamps, oracle against the C ABI bit for bit"""
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "ifos3d-dev_amd"))
    import ifosgpu as ig
    cases = [([3.0, 1.0, 2.0], [0, 0.5, 1.0, 0, 1.0], 0, 1),    # a > 0
             ([1.0, 3.0, 2.0], [0, 0.5, 1.0, 0, 1.0], 0, 1),    # a < 0, L2[0] smallest
             ([2.0, 1.0, 3.5], [0, 0.5, 1.0, 0, 0.1], 0, 1),    # clamp 2.5 step[4]
             ([5.0, 2.0, 1.0], [0, 0.5, 1.0, 0, 4.0], 1, 3),    # L-BFGS: step <= 1, step[4] = 0.5
             ([2.0, 3.0, 1.0], [0, 0.5, 1.0, 0, 1.0], 0, 1)]    # a < 0, L2[2] smallest
    for L2, st, lb, it in cases:
        a = steplength(L2, st, lb, it, 0.8)
        b = ig.steplength(L2, st, lb, it, 0.8)
        np.testing.assert_array_equal(a[0][3:], b[0][3:])
        np.testing.assert_array_equal(a[1], b[1])


@pytest.mark.parametrize("alpha,smin", [(2.0, 0.3), (0.5, 0.9), (7.0, 0.05), (1e-3, 0.7)])
def test_steplength_fp64_parabola_minimum(alpha, smin):
    """the FP64 fit: three points of L2 = alpha (s - smin)^2 + 3 (exact in
    double at these steps) give the minimum smin and L2[3] = 3 to the double
    rounding of the fit, where the float fit is off by float rounding"""
    st = [0.0, 0.5, 1.0, 0.0, 1.0]
    L2 = [alpha * (s - smin) ** 2 + 3.0 for s in st[:3]]
    l2, s = steplength(L2, st, 0, 1, 1.0, dtype=np.float64)
    assert s[3] == pytest.approx(smin, rel=1e-12, abs=1e-15)
    assert l2[3] == pytest.approx(3.0, rel=1e-13)
    l2f, sf = steplength(L2, st, 0, 1, 1.0)
    assert abs(float(sf[3]) - smin) <= 1e-5 * max(smin, 1.0) * (1.0 + 1.0 / alpha)


def test_steplength_fp64_branches_follow_the_float_fit():
    """the branch structure is the reference's: on the clamp cases the FP64
    fit lands where the float fit does, to float rounding"""
    cases = [([1.0, 3.0, 2.0], [0, 0.5, 1.0, 0, 1.0], 0, 1),
             ([2.0, 1.0, 3.5], [0, 0.5, 1.0, 0, 0.1], 0, 1),
             ([5.0, 2.0, 1.0], [0, 0.5, 1.0, 0, 4.0], 1, 3),
             ([2.0, 3.0, 1.0], [0, 0.5, 1.0, 0, 1.0], 0, 1)]
    for L2, st, lb, it in cases:
        a = steplength(L2, st, lb, it, 0.8)
        b = steplength(L2, st, lb, it, 0.8, dtype=np.float64)
        np.testing.assert_allclose(b[1], a[1], rtol=1e-6)


def _grads(shape, seed=3):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape).astype(np.float32) for _ in range(3)]


def test_precon_damptype0_is_a_constant_factor():
    """with the global indices left at 0 (DAMPTYPE 0), every cell is scaled by
    the same six edge tapers 1/(1 + 2000 exp(-0.6 r)) at r = 0 and r = N*G"""
    shape = (6, 5, 7)
    g = _grads(shape)
    g0 = [a.copy() for a in g]
    precon_grad(g, 0, [], [], (5, 6, 7), 4, 0)
    t = lambda r: np.float32(1.0 / (1.0 + 2000.0 * np.exp(-0.6 * r)))
    for n in range(3):
        expect = g0[n]
        for G0, G1 in ((t(0.0), t(6.0)), (t(0.0), t(5.0)), (t(0.0), t(7.0))):
            expect = (expect * G0).astype(np.float32) * G1
        np.testing.assert_array_equal(g[n], expect.astype(np.float32))


def test_precon_receiver_taper_and_quirks():
    shape = (12, 10, 14)
    g = _grads(shape)
    g1 = [a.copy() for a in g]
    precon_grad(g1, 1, [], [(4, 6, 7)], (10, 12, 14), 4, 0)
    # at the receiver cell r = 0: factor 1 / 1001 times the edge tapers
    t = lambda r: np.float32(1.0 / (1.0 + 2000.0 * np.exp(-0.6 * r)))
    j, i, k = 6, 4, 7
    expect = np.float32(g[0][j - 1, i - 1, k - 1] * np.float32(1.0 / 1001.0))
    for lo, hi in ((j, 12 - j), (i, 10 - i), (k, 14 - k)):
        expect = np.float32(np.float32(expect * t(lo)) * t(hi))
    assert g1[0][j - 1, i - 1, k - 1] == expect
    # three cells from the receiver the receiver factor is 1 / (1 + 1000 exp(-0.7 * 9))
    g_norec = [a.copy() for a in g]
    precon_grad(g_norec, 1, [], [], (10, 12, 14), 4, 0)
    ratio = g1[0][j - 1, i - 1, k + 2] / g_norec[0][j - 1, i - 1, k + 2]
    assert ratio == pytest.approx(1.0 / (1.0 + 1000.0 * np.exp(-0.7 * 9.0)), rel=1e-6)
    # FREE_SURF zeroes the y factor of DAMPTYPE 0-2; DAMPTYPE 3 with j = 0 zeroes G[0]
    for dt, fs in ((1, 1), (0, 1), (3, 0)):
        gz = [a.copy() for a in g]
        precon_grad(gz, dt, [], [(4, 6, 7)], (10, 12, 14), 4, fs)
        assert all(not a.any() for a in gz)


def test_modelupdate_by_hand():
    """one cell: vp 4000 -> 4000 + 0.1*1*gvp*VP0 etc., pi/u/rho rebuilt;
    without L-BFGS the gradients are first divided by their maxima"""
    sh = (1, 1, 2)
    rho = np.full(sh, 2000.0, np.float32)
    pi = (rho * np.float32(4000.0) ** 2).astype(np.float32)
    u = (rho * np.float32(2300.0) ** 2).astype(np.float32)
    g = [np.array([[[0.5, -1.0]]], np.float32), np.array([[[2.0, 1.0]]], np.float32),
         np.array([[[0.0, 0.0]]], np.float32)]
    mx, mod = modelupdate(g, rho, pi, u, 0.1, (1.0, 1.0, 1.0), 4000.0, 2300.0, 2000.0, lbfgs=False)
    assert list(mx) == [4000.0, 2300.0, 2000.0]
    np.testing.assert_array_equal(g[0], [[[0.5, -1.0]]])           # divided by max |g| = 1
    np.testing.assert_array_equal(g[1], [[[1.0, 0.5]]])            # divided by 2
    vp = np.sqrt(pi / rho)
    assert vp[0, 0, 0] == pytest.approx(4000.0 + 0.1 * 0.5 * 4000.0, rel=1e-6)
    assert np.sqrt(u / rho)[0, 0, 1] == pytest.approx(2300.0 + 0.1 * 0.5 * 2300.0, rel=1e-6)
    assert mod is None
    rho2 = np.full(sh, 2000.0, np.float32)
    pi2 = (rho2 * np.float32(4000.0) ** 2).astype(np.float32)
    u2 = (rho2 * np.float32(2300.0) ** 2).astype(np.float32)
    g2 = [np.array([[[0.5, -1.0]]], np.float32), np.array([[[2.0, 1.0]]], np.float32),
          np.array([[[1.0, 0.0]]], np.float32)]
    _, mod = modelupdate(g2, rho2, pi2, u2, 0.1, (1.0, 1.0, 1.0), 4000.0, 2300.0, 2000.0, lbfgs=True, numpar=2)
    np.testing.assert_allclose(mod, [0.05, -0.1, 0.2, 0.1], rtol=1e-5)   # (vpnew - vp) / VP0, then vs
    assert rho2[0, 0, 0] == np.float32(2000.0 + 0.1 * 1.0 * 2000.0)
