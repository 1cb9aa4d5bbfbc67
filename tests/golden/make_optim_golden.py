#!/usr/bin/env python3
"""Write tests/golden/optim_golden.npz from the reference's optimiser routines.

The optimiser sources of the reference (lbfgs.c, lbfgs_save.c,
conjugategrad.c, modelupdate.c, precongrad.c, steplength.c) are compiled in
place with the stencil routines by oracle/Makefile.ref into
oracle/_ref/libifos_ref.so; this script runs seeded inputs through them in
the order ifos3d.c:1005-1138 calls them and stores inputs and outputs (small
float32 arrays, so a mismatch can be located):

* lbfgs_np<P>: five iterations (it_group 1..5) of BFGSNUM 3 with NUMPAR P:
  lbfgs_savegrad at it_group 1, lbfgs after, each followed by the L-BFGS
  modelupdate that writes the history slot bfgsmod[it_group % BFGSNUM];
* cg: four iterations of conjugate (cdf 1, then 0) on gradients whose
  Polak-Ribiere ratios fall inside, below and above the [0, 0.5] clamp;
* mu_<n>: modelupdate without L-BFGS (normalised to the maxima);
* pre_<damptype>_<free_surf>: precon_grad for DAMPTYPE 0-3;
* step_<n>: steplength on parabolas with a minimum, with a maximum, and
  the clamps.

    make -C oracle -f Makefile.ref && python tests/golden/make_optim_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from pyoracle import Ref, RefOptim  # noqa: E402

OUT = os.path.join(HERE, "optim_golden.npz")
SHAPE = (6, 5, 7)          # ny, nx, nz
BFGSNUM = 3
MODEL = (2000.0, 4000.0, 2300.0)   # rho, vp, vs
WEIGHT = (1.0, 1.0, 0.5)
SCALE0 = (4000.0, 2300.0, 2000.0)  # VP0, VS0, RHO0
STEPS = ((0.0, 0.01, 0.02, 0.0, 0.02),          # a > 0: parabola minimum
         (0.0, 0.01, 0.02, 0.0, 0.02),          # a < 0, L2[0] smallest
         (0.0, 0.01, 0.02, 0.0, 0.02),          # a < 0, L2[2] smallest
         (0.0, 0.01, 0.02, 0.0, 0.004),         # minimum beyond 2.5 step[4]
         (0.0, 0.5, 1.0, 0.0, 0.5))             # L-BFGS, it_group 2: capped at 1
MISFITS = ((10.0, 7.0, 8.0), (5.0, 9.0, 8.0), (10.0, 9.9, 6.0), (10.0, 9.0, 8.1), (10.0, 6.0, 3.0))


def grads_of(rng, n):
    """smooth + noisy gradients, like preconditioned gradients"""
    ny, nx, nz = SHAPE
    J, I, K = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    base = np.sin(0.7 * J + 0.3 * n) * np.cos(0.5 * I - 0.2 * n) + 0.3 * np.sin(0.9 * K)
    return [(1e-3 * (base * (q + 1) + 0.3 * rng.standard_normal(SHAPE))).astype(np.float32) for q in range(3)]


def model_of(rng):
    rho = (MODEL[0] * (1 + 0.05 * rng.standard_normal(SHAPE))).astype(np.float32)
    vp = (MODEL[1] * (1 + 0.05 * rng.standard_normal(SHAPE))).astype(np.float32)
    vs = (MODEL[2] * (1 + 0.05 * rng.standard_normal(SHAPE))).astype(np.float32)
    return rho, (rho * vp * vp).astype(np.float32), (rho * vs * vs).astype(np.float32)


def lbfgs_case(R, out, numpar):
    rng = np.random.default_rng(100 + numpar)
    n = int(np.prod(SHAPE))
    rho, pi, u = model_of(rng)
    out[f"lbfgs_np{numpar}_model_in"] = np.stack([rho, pi, u])
    scale = np.zeros(BFGSNUM, np.float32)
    mod = np.zeros((BFGSNUM, numpar * n), np.float32)
    hist = np.zeros((BFGSNUM, numpar * n), np.float32)
    for it in range(1, 6):
        g = grads_of(rng, it)
        out[f"lbfgs_np{numpar}_g_{it}"] = np.stack(g)
        if it == 1:
            hist[0] = R.savegrad(g[0], g[1], numpar)
            d = g
        else:
            (d0, d1), scale, hist = R.lbfgs(g[0], g[1], scale, mod, hist, it, numpar)
            d = [d0, d1 if numpar > 1 else g[1], g[2]]
        out[f"lbfgs_np{numpar}_dir_{it}"] = np.stack(d)
        out[f"lbfgs_np{numpar}_scale_{it}"] = scale.copy()
        out[f"lbfgs_np{numpar}_hist_{it}"] = hist.copy()
        step = np.float32(0.02 / it)
        d, (rho, pi, u), mod = R.modelupdate(d, rho, pi, u, mod, step, it, True, numpar, WEIGHT, *SCALE0)
        out[f"lbfgs_np{numpar}_step_{it}"] = step
        out[f"lbfgs_np{numpar}_mod_{it}"] = mod.copy()
        out[f"lbfgs_np{numpar}_model_{it}"] = np.stack([rho, pi, u])


def cg_case(R, out):
    rng = np.random.default_rng(200)
    priors = [np.zeros(SHAPE, np.float32) for _ in range(6)]
    for it in range(1, 5):
        g = grads_of(rng, it)
        if it == 3:      # beta[0] below 0
            g[0] = (0.5 * priors[0]).astype(np.float32)
        if it == 4:      # beta[0] above 0.5
            g[0] = (3.0 * priors[0]).astype(np.float32)
        out[f"cg_g_{it}"] = np.stack(g)
        out[f"cg_prior_in_{it}"] = np.stack(priors)
        g, priors, beta = R.conjugate(g, priors, it, 1 if it == 1 else 0)
        out[f"cg_dir_{it}"] = np.stack(g)
        out[f"cg_prior_{it}"] = np.stack(priors)
        out[f"cg_beta_{it}"] = beta


def mu_case(R, out):
    for c, zero_rho in enumerate((False, True)):
        rng = np.random.default_rng(300 + c)
        rho, pi, u = model_of(rng)
        g = grads_of(rng, c)
        if zero_rho:     # max |grad rho| = 0 is replaced by 1 (modelupdate.c:112)
            g[2][...] = 0.0
        out[f"mu_{c}_in"] = np.stack(g + [rho, pi, u])
        d, (rho, pi, u), _ = R.modelupdate(g, rho, pi, u, None, 0.03, 1, False, 2, WEIGHT, *SCALE0)
        out[f"mu_{c}_out"] = np.stack(d + [rho, pi, u])


PRE_SRC = np.array([[8.0, 6.0, 10.0], [4.0, 12.0, 2.0]], np.float32)   # metres, dh 2 m
PRE_REC = np.array([[2, 1, 3], [4, 2, 5], [5, 6, 7]], np.int32)        # grid (x, y, z)
PRE_DH = (2.0, 2.0, 2.0)
PRE_FW = 2


def pre_case(R, out):
    rng = np.random.default_rng(400)
    g = grads_of(rng, 0)
    out["pre_in"] = np.stack(g)
    for damptype in (0, 1, 2, 3):
        for fs in (0, 1):
            r = R.precon_grad(g, PRE_SRC, PRE_REC, PRE_DH, PRE_FW, damptype, fs)
            out[f"pre_{damptype}_{fs}"] = np.stack(r)


def step_case(R, out):
    for c, (st, l2) in enumerate(zip(STEPS, MISFITS)):
        lbfgs, it_group = (1, 2) if c == 4 else (0, 1)
        L2, s = R.steplength(l2, st, 1, it_group, lbfgs, 0.02)
        out[f"step_{c}_in"] = np.array(list(l2) + list(st) + [lbfgs, it_group, 0.02], np.float32)
        out[f"step_{c}_L2"] = L2
        out[f"step_{c}_step"] = s


def main():
    if not Ref.available():
        sys.exit("oracle/_ref/libifos_ref.so is not built (make -C oracle -f Makefile.ref)")
    R = RefOptim(*SHAPE)
    out = {"generator": np.array("reference lbfgs.c, lbfgs_save.c, conjugategrad.c, modelupdate.c, "
                                 "precongrad.c, steplength.c compiled from source (oracle/_ref)")}
    for numpar in (1, 2):
        lbfgs_case(R, out, numpar)
    cg_case(R, out)
    mu_case(R, out)
    pre_case(R, out)
    step_case(R, out)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
