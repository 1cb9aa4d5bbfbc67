#!/usr/bin/env python3
"""Write tests/golden/configs.npz: oracle outputs at the benchmark workloads'
own shapes, for the GPU parity tests of test_gpu_configs.py.

The oracle (oracle/ifos_oracle.c, pinned to the reference by golden.json)
takes minutes on these grids, so its results are computed once here and kept
as small fixtures: the SHA-256 of every float32 array the GPU must reproduce
bit for bit, plus enough of the data (one shot's traces, the centre slice of
each gradient, per-shot misfits) to say how far a mismatch is.

* toy84: the CPU-runnable toy -- 84^3 (64^3 + 10-cell CPML), dh 2 m, dt 2e-4 s,
  NT 600, two +-10 % spheres in a homogeneous model, 4 explosive 50 Hz
  Ricker shots, 32 receivers, DFT at 40/60/80 Hz: observed data, synthetics,
  misfits and the gradient accumulated over the 4 shots in order.
* level_<n>: the three coarse-to-fine grid shapes of the uneven-levels case
  (116^3 at 8 m, 148^3 at 6 m, 212^3 at 4 m, each with its 0.9-CFL time step
  and model), the first of the level's shots, LEVEL_NT steps, receivers on
  a line through the source and on the top layer (level_receivers()).

    make -C oracle && OMP_NUM_THREADS=8 python tests/golden/make_configs.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "ifos3d-dev_amd")):
    sys.path.insert(0, p)

from helpers import sha  # noqa: E402
from pyoracle import Grid, Oracle  # noqa: E402
import workloads as wl  # noqa: E402

OUT = os.path.join(HERE, "configs.npz")
LEVEL_NT = 80


def level_receivers(w, src):
    """receivers of the level parity case: every 3 cells along i through the
    source (offsets -12..12, 0 left out) and 4 on the top interior layer"""
    i, j, k = src[0], src[1], src[2]
    rec = [[i + d, j, k] for d in range(-12, 13, 3) if d != 0 and 1 <= i + d <= w.nx]
    rec += [[i + d, w.fw + 1, k] for d in (-16, -4, 4, 16) if 1 <= i + d <= w.nx]
    return np.array(rec, np.int32)


def level_workloads():
    levels = wl.levels3(3)
    for w in levels:
        w.nt = LEVEL_NT
    return levels


def grid_of(w):
    return Grid(w.nx, w.ny, w.nz, w.fw, w.dh, w.dh, w.dh, w.dt, w.nt, w.fdcoeff)


def toy84(o, out):
    w = wl.toy84()
    g = grid_of(w)
    true, start = wl.sphere_models(w)
    obs, grads = [], [o.zeros(g) for _ in range(3)]
    for n, src in enumerate(w.shots):
        t = time.time()
        r = o.run_shot(g, *true, src, w.receivers, w.finv, tast=w.tast, vppml=w.vppml, fpml=w.fpml,
                       want_adjoint=False)
        obs.append(np.stack(r["syn"]))
        out[f"toy84_obs_sha_{n}"] = np.array(sha(obs[-1][:, :, 1:]))
        print(f"toy84 observed data, shot {n}: {time.time() - t:.1f} s", flush=True)
    for n, src in enumerate(w.shots):
        t = time.time()
        r = o.run_shot(g, *start, src, w.receivers, w.finv, tast=w.tast, vppml=w.vppml, fpml=w.fpml,
                       obs=list(obs[n]), grads=grads)
        syn = np.stack(r["syn"])
        out[f"toy84_syn_sha_{n}"] = np.array(sha(syn[:, :, 1:]))
        out[f"toy84_l2_{n}"] = np.float32(r["l2"])
        if n == 0:
            out["toy84_obs_0"] = obs[0]
            out["toy84_syn_0"] = syn
        print(f"toy84 gradient shot {n}: {time.time() - t:.1f} s", flush=True)
    jc = g.shape[0] // 2
    for q in range(3):
        out[f"toy84_grad_sha_{q}"] = np.array(sha(g.interior(grads[q])))
        out[f"toy84_grad_slice_{q}"] = np.ascontiguousarray(grads[q][jc])
        out[f"toy84_grad_l2_{q}"] = np.float64(np.linalg.norm(g.interior(grads[q]).astype(np.float64)))


def levels(o, out):
    lv = level_workloads()
    models = wl.level_models(lv)
    for w, m in zip(lv, models):
        g = grid_of(w)
        n = w.extra["interior"]
        src = w.shots[0]
        rec = level_receivers(w, src)
        t = time.time()
        r = o.run_shot(g, *m, src, rec, w.finv, tast=w.tast, vppml=w.vppml, fpml=w.fpml, want_adjoint=False)
        syn = np.stack(r["syn"])
        assert np.abs(syn).max() > 0
        out[f"level_{n}_syn"] = syn
        out[f"level_{n}_syn_sha"] = np.array(sha(syn[:, :, 1:]))
        print(f"level {n}^3: {time.time() - t:.1f} s", flush=True)


def main():
    o = Oracle("f32")
    out = {"generator": np.array("oracle/ifos_oracle.c (f32 build), tests/golden/make_configs.py")}
    levels(o, out)
    toy84(o, out)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
