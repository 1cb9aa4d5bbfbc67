"""CPU restatement of the reference's optimiser steps (TEST INFRASTRUCTURE).

Only the test suite imports this module, as the checker of the device
optimiser (ifos_lbfgs_*, ifos_conjugate).

* LbfgsOracle follows lbfgs.c:26-181 and lbfgs_save.c:25-54: the history
  slots bfgsgrad / bfgsmod [1..BFGSNUM], bfgsscale, the two-loop recursion
  of Nocedal & Wright Algorithm 9.1 with H0 = (y.s / y.y) I from the newest
  pair.  `dot` selects how the sums are formed: "f32" accumulates in FP32 in
  index order as the reference does (one rank), "f64" forms them in FP64 (the
  device's choice); `vec` is the dtype of the vector updates (FP32 as in the
  reference; FP64 only for the closed-form check against the dense BFGS
  inverse-Hessian recursion).
* conjugate follows conjugategrad.c:26-110, including its clamps, which test
  the already-clamped beta[0] so that only beta[0] ends in [0, 0.5].

Vectors are compact 1-D arrays: numpar blocks of NX*NY*NZ cells in the
reference's l order (j outer, i, k inner).
"""
from __future__ import annotations

import numpy as np


def _dot(a, b, mode, scale=None):
    """sum of (scale*)a*b: "f32" sequential FP32 (the reference's loops), "f64"
    FP64"""
    if mode == "f32":
        t = a.astype(np.float32) * b.astype(np.float32) if scale is None else \
            (np.float32(scale) * a.astype(np.float32)) * b.astype(np.float32)
        return float(np.add.accumulate(t, dtype=np.float32)[-1]) if t.size else 0.0
    s = float(np.dot(a.astype(np.float64), b.astype(np.float64)))
    return s if scale is None else scale * s


class LbfgsOracle:
    def __init__(self, n, bfgsnum, dot="f64", vec=np.float32):
        self.n, self.bfgsnum, self.dot, self.vec = n, bfgsnum, dot, vec
        self.grad = np.zeros((bfgsnum, n), vec)   # bfgsgrad[w] at w-1
        self.mod = np.zeros((bfgsnum, n), vec)    # bfgsmod[w]
        self.scale = np.zeros(bfgsnum, np.float64)

    def slot(self, v):
        w = v % self.bfgsnum
        return (self.bfgsnum if w == 0 else w) - 1

    def savegrad(self, g):
        """lbfgs_save.c:31-53"""
        self.grad[0] = np.asarray(g, self.vec)

    def update(self, g, iteration):
        """lbfgs.c:26-181; returns (direction, h0)"""
        g = np.asarray(g, self.vec)
        m = max(iteration - self.bfgsnum, 1)
        w = self.slot(iteration - 1)
        self.grad[w] = self.grad[w] + (-g)          # lbfgs.c:52-53
        q = g.copy()
        ys = _dot(self.grad[w], self.mod[w], self.dot)
        yy = _dot(self.grad[w], self.grad[w], self.dot)
        if self.dot == "f32":                        # lbfgs.c:85, :87 in FP32
            self.scale[w] = float(np.float32(1.0) / np.float32(ys))
            h0 = float(np.float32(ys) / np.float32(yy))
        else:
            self.scale[w] = 1.0 / ys
            h0 = ys / yy
        alpha = np.zeros(self.bfgsnum, np.float64)
        for v in range(iteration - 1, m - 1, -1):    # lbfgs.c:94-116
            w = self.slot(v)
            if self.dot == "f32":
                alpha[w] = _dot(self.mod[w], q, "f32", scale=np.float32(self.scale[w]))
            else:
                alpha[w] = self.scale[w] * _dot(self.mod[w], q, "f64")
            q = q + (self.vec(-alpha[w]) * self.grad[w])
        q = q * self.vec(h0)                         # lbfgs.c:120-124
        for v in range(m, iteration):                # lbfgs.c:136-151
            w = self.slot(v)
            if self.dot == "f32":
                beta = _dot(self.grad[w], q, "f32", scale=np.float32(self.scale[w]))
                coef = np.float32(np.float32(alpha[w]) - np.float32(beta))
            else:
                beta = self.scale[w] * _dot(self.grad[w], q, "f64")
                coef = self.vec(alpha[w] - beta)
            q = q + (self.mod[w] * coef)
        self.grad[self.slot(iteration)] = g          # lbfgs.c:154-165
        return q, h0


def dense_bfgs_direction(pairs, h0, g):
    """closed form the two-loop recursion equals: H = h0 I updated by
    H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T for the pairs
    (s, y) oldest first; returns H g (FP64)"""
    n = len(g)
    H = h0 * np.eye(n)
    for s, y in pairs:
        s = np.asarray(s, np.float64)
        y = np.asarray(y, np.float64)
        rho = 1.0 / np.dot(y, s)
        V = np.eye(n) - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    return H @ np.asarray(g, np.float64)


def conjugate(grads, priors, cdf, dot="f64"):
    """conjugategrad.c:26-110 on interior arrays; grads[3] and priors[6] are
    updated in place; returns (beta[3] as float32, the six sums)"""
    beta = np.zeros(3, np.float32)
    if cdf == 1:
        for n in range(3):
            priors[n][...] = grads[n]
            priors[n + 3][...] = grads[n]
        return beta, None
    sums = np.zeros(6, np.float64)
    for n in range(3):
        g = grads[n].astype(np.float32).ravel()
        p = priors[n].astype(np.float32).ravel()
        d = (g - p).astype(np.float32)
        sums[2 * n] = _dot(g, d, dot)
        sums[2 * n + 1] = _dot(p, p, dot)
        priors[n][...] = grads[n]
    if cdf == 0:
        for n in range(3):
            if abs(sums[2 * n + 1]) > 0.0:
                beta[n] = np.float32(sums[2 * n] / sums[2 * n + 1])
    # conjugategrad.c:84-89 as written
    if beta[0] < 0: beta[0] = 0.0
    if beta[0] < 0: beta[1] = 0.0
    if beta[0] < 0: beta[2] = 0.0
    if beta[0] > 0.5: beta[0] = 0.5
    if beta[0] > 0.5: beta[1] = 0.5
    if beta[0] > 0.5: beta[2] = 0.5
    for n in range(3):
        v = (priors[n + 3].astype(np.float32) * beta[n]) + grads[n].astype(np.float32)
        grads[n][...] = v
        priors[n + 3][...] = v
    return beta, sums
