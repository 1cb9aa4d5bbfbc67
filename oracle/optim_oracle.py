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


# ---------------------------------------------------------------------------
# precongrad.c:28-336, modelupdate.c:26-160, steplength.c:26-91

def _taper(c, e):
    """(float)(1 / (1 + c * exp(e))) with e in double, as the reference forms it"""
    return np.float32(1.0 / (1.0 + c * np.exp(e)))


def precon_grad(grads, damptype, src_xyz, rec_xyz, nglob, fw, free_surf, pos=(0, 0, 0)):
    """grads: three interior arrays [NY][NX][NZ] (float32), updated in place.
    src_xyz / rec_xyz: integer grid positions (x, y, z) per source / receiver."""
    ny, nx, nz = grads[0].shape
    nxg, nyg, nzg = nglob
    sh = rh = fw + 40
    J, I, K = np.meshgrid(np.arange(1, ny + 1), np.arange(1, nx + 1), np.arange(1, nz + 1), indexing="ij")
    if damptype in (1, 2):
        ii, jj, kk = I + pos[0] * nx, J + pos[1] * ny, K + pos[2] * nz
    else:  # the global indices stay 0 outside DAMPTYPE 1 and 2 (precongrad.c:35)
        ii = jj = kk = np.zeros_like(I)
    g = [a.astype(np.float32) for a in grads]

    def mul(G):
        for n in range(3):
            g[n] = (g[n] * G).astype(np.float32)

    if damptype == 2:
        for sx, sy, sz in src_xyz:
            r = np.sqrt(((jj - sy) ** 2 + (ii - sx) ** 2 + (kk - sz) ** 2).astype(np.float64)).astype(np.float32)
            mul(_taper(1000.0, -0.7 * r.astype(np.float64)))
    if damptype in (1, 2):
        for rx, ry, rz in rec_xyz:
            r = np.sqrt(((jj - ry) ** 2 + (ii - rx) ** 2 + (kk - rz) ** 2).astype(np.float64)).astype(np.float32)
            rd = r.astype(np.float64)
            mul(_taper(1000.0, -0.7 * rd * rd))
    f = lambda v: np.asarray(v, np.float64).astype(np.float32).astype(np.float64)
    if damptype != 3:
        G0 = np.zeros_like(g[0]) if free_surf else _taper(2000.0, -0.6 * f(jj * 1.0))
        G1 = _taper(2000.0, -0.6 * f((nyg - jj) * 1.0))
    else:
        r = f((jj - sh) * 1.0)
        G0 = np.where(r < 0, np.float32(0), _taper(1000.0, -0.7 * r * r)).astype(np.float32)
        r = f((nyg - jj - rh) * 1.0)
        G1 = np.where(r < 0, np.float32(0), _taper(1000.0, -0.7 * r)).astype(np.float32)
    for n in range(3):
        g[n] = ((g[n] * G0).astype(np.float32) * G1).astype(np.float32)
    for lo, hi in ((ii, nxg - ii), (kk, nzg - kk)):
        G0 = _taper(2000.0, -0.6 * f(lo * 1.0))
        G1 = _taper(2000.0, -0.6 * f(hi * 1.0))
        for n in range(3):
            g[n] = ((g[n] * G0).astype(np.float32) * G1).astype(np.float32)
    for n in range(3):
        grads[n][...] = g[n]


def modelupdate(grads, rho, pi, u, step, weight, vp0, vs0, rho0, lbfgs, numpar=2):
    """interior arrays updated in place; returns (maxima[3], bfgsmod or None)"""
    f32 = np.float32
    vp = np.sqrt((pi / rho).astype(f32).astype(np.float64))
    vs = np.sqrt((u / rho).astype(f32).astype(np.float64))
    maxima = np.array([max(0.0, vp.astype(f32).max()), max(0.0, vs.astype(f32).max()), max(0.0, rho.max())], f32)
    g = [a.astype(f32) for a in grads]
    if not lbfgs:
        m = [np.abs(a).max() for a in g]
        if m[2] == 0.0:
            m[2] = f32(1.0)
        if m[0] > 0.0 and m[1] > 0.0:
            g = [(g[n] / f32(m[n])).astype(f32) for n in range(3)]
            for n in range(3):
                grads[n][...] = g[n]
    s1, s2, s3 = (f32(w) for w in weight)
    step, vp0, vs0, rho0 = f32(step), f32(vp0), f32(vs0), f32(rho0)
    vpnew = (vp + (((step * s1) * g[0]) * vp0).astype(f32).astype(np.float64)).astype(f32)
    vsnew = (vs + (((step * s2) * g[1]) * vs0).astype(f32).astype(np.float64)).astype(f32)
    rhonew = (rho + (((s3 * step) * g[2]) * rho0).astype(f32)).astype(f32)
    mod = None
    if lbfgs:
        parts = [((vpnew.astype(np.float64) - vp) / np.float64(vp0)).astype(f32).ravel()]
        if numpar > 1:
            parts.append(((vsnew.astype(np.float64) - vs) / np.float64(vs0)).astype(f32).ravel())
        mod = np.concatenate(parts)
    u[...] = ((rhonew * vsnew).astype(f32) * vsnew).astype(f32)
    pi[...] = ((rhonew * vpnew).astype(f32) * vpnew).astype(f32)
    rho[...] = rhonew
    return maxima, mod


def steplength(L2, step, lbfgs, it_group, teststep, dtype=np.float32):
    """steplength.c:26-91; with float32 (the reference's arithmetic: the
    double literals promote, the results are stored back to float) or
    float64 throughout (the FP64 fit); returns (L2[4], step[5])"""
    f = dtype
    L2 = [f(x) for x in L2[:3]] + [f(0)]
    s = [f(x) for x in step[:5]]
    a = ((s[0] - s[1]) * (L2[0] - L2[2]) - (s[0] - s[2]) * (L2[0] - L2[1])) / \
        (s[0] * s[2] * (s[0] - s[2]) + s[0] * s[1] * (s[1] - s[0]) + s[1] * s[2] * (s[2] - s[1]))
    b = (L2[0] - L2[1]) / (s[0] - s[1])
    b = b - a * (s[0] + s[1])
    c = L2[0] - a * s[0] * s[0] - b * s[0]
    s[3] = f(0)
    if a > 0:
        s[3] = -b / (f(2) * a)
    if a < 0:
        if L2[0] < L2[1] and L2[0] < L2[2]:
            s[3] = f(0.1 * float(s[1]))
        if L2[1] < L2[0] and L2[1] < L2[2]:
            s[3] = s[1]
        if L2[2] < L2[0] and L2[2] < L2[1]:
            s[3] = s[2]
    if float(s[3]) > 2.5 * float(s[4]):
        s[3] = f(2.5 * float(s[4]))
    if lbfgs and it_group > 1 and float(s[3]) > 1.0:
        s[3] = f(1.0)
    if s[3] <= 0:
        s[3] = f(0.1 * float(s[1]))
    if float(s[3]) < 0.1 * float(s[1]):
        s[3] = f(0.1 * float(s[1]))
    L2[3] = a * s[3] * s[3] + b * s[3] + c
    s[4] = s[3] / f(2)
    if float(s[4]) < 0.2 * float(teststep):
        s[4] = f(0.2 * float(teststep))
    if float(s[4]) > float(teststep):
        s[4] = f(teststep)
    if lbfgs and it_group > 1:
        s[4] = f(0.5)
    return np.array(L2, dtype), np.array(s, dtype)
