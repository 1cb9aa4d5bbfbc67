#!/usr/bin/env python3
"""Benchmark: gradient shots per second on the 276^3 (256^3 + CPML) elastic
FWI workload, one MI355X.

A gradient shot is: forward NT=1000 steps with the on-the-fly DFT (6
frequencies), residual against observed data, adjoint NT steps with the
DFT, gradient_F.  A step of this benchmark is one batch: all 16 shot
positions propagated together by the batched path (ifos_batch_run).  The
plain one-shot path (ifos_ctx_run_shot, the drop-in's shape), the
checkpointed time-domain gradient (ifos_ctx_run_shot_td) and the drop-in
library itself (the reference prototypes called on Numerical-Recipes host
arrays by an IFOS3D-shaped driver, host copies included) are timed on one
shot each beside it, and the batched path once more with compensated (Kahan)
DFT sums and hipGraph replay of the time loops (two batches of 8 shots: the
second set of accumulators does not fit 16).  Inputs (models, observed data) are resident in HBM
before the timed region.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config grad276|toy84]
                  [--path batched|plain] [--shots-per-batch S] [--chunk C] [--graph]

Two other workloads print a line of their own:
  --config fwd536   forward-only bandwidth case, 536 x 280 x 536, NT 2000 (Gcell-updates/s)
  --config levels3  three uneven grid levels x 32 shots in one graph launch (forward solves/s)
  --config aux276   the optimiser, Hessian, residual and SU steps at 276^3 (ms per call)
  --config visco404 the viscoelastic (3 mechanisms) + 8-frequency DFT case at 404 x 212 x 404: Gcell-updates/s
                    and the compat and Kahan DFT accumulation errors against FP64
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ifos3d-dev_amd"))

METRIC = "gradient shots/sec at 256³ (batched vs plain); Gcell-updates/s; % HBM peak"
HBM_PEAK_GBS = 8000.0
# kernel (as rocprofv3 names it) behind each launch-timer category; the
# march kernels' template arguments follow the tile shape in use


def kernel_names(path, nf, chunk, tiles=None):
    b = lambda x: "true" if x else "false"
    if path == "plain":
        # interior boxes at least 64 cells deep in k take the LDS marches
        return {"v_interior": "ifos::k_v_lds<4, false, false>",
                "s_interior": "ifos::k_s_lds<4, false, 0>",
                "v_cpml": "ifos::k_v_slabs", "s_cpml": "ifos::k_s_slabs<0>",
                "dft": f"ifos::k_dft_nf<{nf}>", "dft_kahan": f"ifos::k_dft_kahan_nf<{nf}>",
                "gradient": "ifos::k_grad_march<4, 0>"}
    tv, ts = tiles
    # the LDS-staged marches run when asked for with TI >= 2, one k column
    # and one shot per workgroup
    lds = lambda t: len(t) > 7 and t[7] and t[0] >= 2 and t[5] <= 1 and t[6] == 1
    return {"v_fused": f"ifos::k_v_lds<{tv[0]}, true, false>" if lds(tv)
            else f"ifos::k_v_march<{tv[0]}, true, false, {b(tv[4])}>",
            "s_fused": f"ifos::k_s_lds<{ts[0]}, true, 0>" if lds(ts)
            else f"ifos::k_s_march<{ts[0]}, true, 0, {b(ts[4])}>",
            "vs_fused": "ifos::k_vs_fused<16, false>",
            "dft": f"ifos::k_dft_chunk<{nf}, {chunk}, false>", "dft_kahan": f"ifos::k_dft_chunk<{nf}, {chunk}, true>",
            "gradient": "ifos::k_grad_march<4, 0>"}


# tile shapes {TI, jc, group, split, pipe} used without the autotuner (its
# usual choice on grad276)
DEFAULT_TILES = ((4, 32, 0, 0, 0, 1, 1, 1), (4, 32, 0, 0, 0, 1, 1, 1))
CATS = ["v_interior", "v_cpml", "v_fused", "s_interior", "s_cpml", "s_fused", "dft", "gradient", "sources", "other",
        "vs_fused"]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def slab_stats(w, free_surf=0):
    """cells of the interior box, of the CPML slabs, and the number of
    (cell, axis) pairs with an active CPML memory variable"""
    fw = w.fw
    bi = (w.nx - 2 * fw) * (w.ny - (fw if free_surf else 2 * fw)) * (w.nz - 2 * fw)
    slab = w.cells - bi
    axes = 2 * fw * w.ny * w.nz + (fw if free_surf else 2 * fw) * w.nx * w.nz + 2 * fw * w.nx * w.ny
    return bi, slab, axes


def algorithmic_bytes(w, nf, kahan=False):
    """bytes each kernel category must move per launch (FP32, each array
    counted once per pass): v 60 B/cell, sigma 80 B/cell, +24 B per active
    CPML axis, DFT 12 + 48*nf B/cell (x2 accumulators with Kahan),
    gradient 48*nf + 36 B/cell"""
    bi, slab, axes = slab_stats(w)
    acc = 96 if kahan else 48
    return {
        "v_interior": 60 * bi, "v_cpml": 60 * slab + 24 * axes, "v_fused": 60 * w.cells + 24 * axes,
        "s_interior": 80 * bi, "s_cpml": 80 * slab + 24 * axes, "s_fused": 80 * w.cells + 24 * axes,
        "dft": (12 + acc * nf) * w.cells, "gradient": (48 * nf + 36) * w.cells,
    }


def algorithmic_bytes_batched(w, nf, S, chunk, ntast, kahan=False):
    """average bytes per launch of the batched path's kernels for a batch of
    S shots: the fused velocity march moves 48 B per cell and shot plus the
    shared model (12 B per cell) plus 24 B per active CPML axis and shot, and
    12 B per cell and shot more at the DFT sampling steps (the snapshot); the
    fused stress march 60 B per cell and shot plus 20 B of shared model plus
    the CPML term; a DFT chunk of c snapshots 12*c + 48*nf B per cell and shot
    (x2 accumulators with Kahan); the gradient 48*nf*S + 36 B per cell"""
    _, _, axes = slab_stats(w)
    cells = w.cells
    nsamp = w.nt // ntast
    nchunks = (nsamp + chunk - 1) // chunk
    acc = 96 if kahan else 48
    v_base = (48 * cells + 24 * axes) * S + 12 * cells
    v_avg = v_base + 12 * cells * S * nsamp / w.nt
    s_avg = (60 * cells + 24 * axes) * S + 20 * cells
    dft_avg = (12 * nsamp + acc * nf * nchunks) * cells * S / nchunks
    # fused step: v and sigma read and written once (72 B), the whole model
    # once per tile (32 B shared), 48 B per active CPML axis (the stress
    # memory read and rewritten, the velocity memory updated in place), the
    # snapshot at sampling steps
    vs_avg = (72 * cells + 48 * axes) * S + 32 * cells + 12 * cells * S * nsamp / w.nt
    return {"v_fused": v_avg, "s_fused": s_avg, "dft": dft_avg, "gradient": (48 * nf * S + 36) * cells,
            "vs_fused": vs_avg}


def _prof_start(ig):
    lib = ig.lib()
    lib.ifos_prof_reset()
    lib.ifos_prof_enable(1)


def _prof_stop(ig):
    import ctypes as C
    lib = ig.lib()
    lib.ifos_prof_enable(0)
    ms = (C.c_double * len(CATS))()
    cnt = (C.c_long * len(CATS))()
    ig.check(lib.ifos_prof_read(ms, cnt, len(CATS)), "ifos_prof_read")
    lib.ifos_prof_reset()
    return {CATS[n]: (ms[n], cnt[n]) for n in range(len(CATS))}


def setup(w):
    import ifosgpu as ig
    import workloads as wl
    G = ig.Geom(w.nx, w.ny, w.nz, w.fw, w.dh, w.dh, w.dh, w.dt, w.nt, w.fdcoeff, 0)
    log(f"grid {w.nx}x{w.ny}x{w.nz} pitch {G.pitch} NT {w.nt} dt {w.dt:.4e} nf {len(w.finv)}")
    true, start = wl.sphere_models(w) if w.name == "toy84" else wl.random_models(w)
    return G, [G.to_dev(a) for a in true], [G.to_dev(a) for a in start], start


def observed_data(w, G, d_true, S, chunk, fused=False):
    """forward runs of every shot position on the true model (untimed setup),
    through the batched path; one device buffer [3][ntr][nt+1] per shot"""
    import ifosgpu as ig
    ntr = len(w.receivers)
    b = ig.BatchContext(G, S, len(w.finv), ntr, ig.DFT_COMPAT, chunk)
    b.set_cpml(w.vppml, w.fpml)
    b.set_freqs(w.finv, w.tast)
    b.set_model(*d_true)
    b.set_fused(fused)
    b.set_tiles(DEFAULT_TILES[0], DEFAULT_TILES[1])
    obs = []
    for b0 in range(0, len(w.shots), S):
        srcs = w.shots[b0:b0 + S]
        d_syn = ig.DBuf((len(srcs), 3, ntr, w.nt + 1))
        b.run(srcs, w.receivers, d_syn=d_syn, want_adjoint=False)
        h = d_syn.numpy()
        obs += [ig.DBuf.from_numpy(h[n]) for n in range(len(srcs))]
        d_syn.free()
        log(f"observed data: shots {b0 + 1}-{b0 + len(srcs)} of {len(w.shots)}")
    b.close()
    return obs


def run_batched(w, G, d_start, obs, steps, warmup, precision, S, chunk, use_graph, autotune=True, fused=False):
    import ifosgpu as ig
    ntr = len(w.receivers)
    log(f"batched path: {S} shots per batch, DFT chunk {chunk}, graph {use_graph}, "
        f"{ig.BatchContext.bytes_needed(G, S, len(w.finv), ntr, precision, chunk) / 2**30:.1f} GiB")
    b = ig.BatchContext(G, S, len(w.finv), ntr, precision, chunk)
    b.set_cpml(w.vppml, w.fpml)
    b.set_freqs(w.finv, w.tast)
    b.set_model(*d_start)
    b.set_fused(fused)
    b.set_tiles(DEFAULT_TILES[0], DEFAULT_TILES[1])
    tune = None
    if autotune:
        bv, bs, table = b.autotune(reps=2)
        tune = {"v": bv, "s": bs, "table": table, "fused": b.fused_table}
        log(f"autotune: velocity march {bv}, stress march {bs}, fused step {b.fused_table}")
    grads = [G.empty() for _ in range(3)]
    batches = [list(range(b0, min(b0 + S, len(w.shots)))) for b0 in range(0, len(w.shots), S)]

    def one(n):
        ids = batches[n % len(batches)]
        return b.run([w.shots[i] for i in ids], w.receivers, grads=grads, d_obs=[obs[i] for i in ids],
                     use_graph=use_graph), len(ids)

    for n in range(warmup):
        one(n)
        log(f"warmup batch {n + 1}/{warmup} done")
    ig.sync()
    if not use_graph:
        _prof_start(ig)
    t0 = time.perf_counter()
    l2, nshots = 0.0, 0
    for n in range(warmup, warmup + steps):
        l2s, k = one(n)
        l2 += sum(l2s)
        nshots += k
        log(f"timed batch {n - warmup + 1}/{steps}: {k} shots")
    ig.sync()
    elapsed = time.perf_counter() - t0
    prof = _prof_stop(ig) if not use_graph else None
    gnorm = [float(np.linalg.norm(G.from_dev(t).astype(np.float64))) for t in grads]
    fused_used = b.last_fused()
    b.close()
    return dict(elapsed=elapsed, shots=nshots, prof=prof, l2=l2, gnorm=gnorm, tune=tune, fused=fused_used)


def run_levels(args):
    """the uneven grid-level workload: forward solves of three levels (32
    shots each) in one ifos_levels_run call, against the same solves level by
    level through ifos_batch_run and against plain one-shot launches (a
    bounded sample of shots per level, extrapolated to all of them)"""
    import ifosgpu as ig
    import workloads as wl
    levels = wl.levels3(args.level_shots)
    models = wl.level_models(levels)
    ctxs, geoms = [], []
    for w, m in zip(levels, models):
        G = ig.Geom(w.nx, w.ny, w.nz, w.fw, w.dh, w.dh, w.dh, w.dt, w.nt, w.fdcoeff, 0)
        b = ig.BatchContext(G, len(w.shots), 1, len(w.receivers), ig.DFT_COMPAT, 1)
        b.set_cpml(w.vppml, w.fpml)
        b.set_model(*[G.to_dev(a) for a in m])
        tune = None
        if not args.no_autotune:
            bv, bs, _ = b.autotune(reps=2)
            tune = {"v": bv, "s": bs}
        log(f"{w.name}: {w.nx}^3, dh {w.dh}, dt {w.dt:.4e}, NT {w.nt}, {len(w.shots)} shots, tiles {tune}")
        ctxs.append(b)
        geoms.append((G, tune))
    srcs = [w.shots for w in levels]
    recs = [w.receivers for w in levels]
    d_syn = [ig.DBuf((len(w.shots), 3, len(w.receivers), w.nt + 1)) for w in levels]
    solves = sum(len(w.shots) for w in levels)
    cell_steps = sum(w.cells * w.nt * len(w.shots) for w in levels)
    for n in range(args.warmup):
        ig.levels_run(ctxs, srcs, recs, d_syn=d_syn, use_graph=True)
        log(f"warmup {n + 1}/{args.warmup} done")
    ig.sync()
    t0 = time.perf_counter()
    for n in range(args.steps):
        ig.levels_run(ctxs, srcs, recs, d_syn=d_syn, use_graph=True)
        log(f"timed levels run {n + 1}/{args.steps}")
    ig.sync()
    el = time.perf_counter() - t0
    together = solves * args.steps / el
    log(f"levels together (one graph): {together:.3f} solves/s")
    # the same solves level after level, each level's batch alone (graph replay)
    for b, w, d in zip(ctxs, levels, d_syn):
        b.run(w.shots, w.receivers, d_syn=d, want_adjoint=False, use_graph=True)
    ig.sync()
    t0 = time.perf_counter()
    for b, w, d in zip(ctxs, levels, d_syn):
        b.run(w.shots, w.receivers, d_syn=d, want_adjoint=False, use_graph=True)
    ig.sync()
    seq = solves / (time.perf_counter() - t0)
    log(f"levels one after another: {seq:.3f} solves/s")
    for b in ctxs:
        b.close()
    # plain one-shot launches: the first shots of every level
    plain_t, plain_n = 0.0, 0
    for w, (G, _), m in zip(levels, geoms, models):
        ctx = ig.ShotContext(G, 1, len(w.receivers), ig.DFT_COMPAT)
        ctx.set_cpml(w.vppml, w.fpml)
        ctx.set_freqs(w.finv, w.tast)
        ctx.set_model(*[G.to_dev(a) for a in m])
        d = ig.DBuf((3, len(w.receivers), w.nt + 1))
        ctx.run_shot(w.shots[0], w.receivers, d, None, want_adjoint=False)
        ig.sync()
        t0 = time.perf_counter()
        for src in w.shots[:args.plain_shots]:
            ctx.run_shot(src, w.receivers, d, None, want_adjoint=False)
        ig.sync()
        dt_ = time.perf_counter() - t0
        plain_t += dt_ * len(w.shots) / args.plain_shots
        plain_n += len(w.shots)
        log(f"plain {w.name}: {args.plain_shots / dt_:.3f} solves/s")
        ctx.close()
    plain = plain_n / plain_t
    # stencil bytes of the whole run: 140 B per cell and shot-step (v 60, sigma 80)
    # less the model share (32 B) the shots of a batch read together, + CPML
    abytes = 0.0
    for w in levels:
        _, _, axes = slab_stats(w)
        S = len(w.shots)
        abytes += ((108 * w.cells + 48 * axes) * S + 32 * w.cells) * w.nt
    achieved = abytes / (el / args.steps) / 1e9
    desc = "; ".join(f"{w.extra['interior']}^3+{w.fw} CPML at dh {w.dh} m, NT {w.nt}" for w in levels)
    out = {
        "metric": "forward solves/s, 3 uneven grid levels x 32 shots in one graph launch",
        "value": round(together, 4),
        "unit": "forward solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * el / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"levels3: {desc}; {args.level_shots} explosive Ricker shots per level "
                               f"(fc 5-20 Hz), 64 receivers; elastic FP32 4th order, forward only",
                   "path": "ifos_levels_run: one hipGraph, one branch per level, autotuned tiles per level"},
        "gcell_updates_per_s": round(cell_steps * args.steps / el / 1e9, 3),
        "levels_sequential_solves_per_s": round(seq, 4),
        "plain_solves_per_s": round(plain, 4),
        "plain_sample": f"{args.plain_shots} shots per level, extrapolated to {plain_n}",
        "together_over_plain": round(together / plain, 3),
        "roofline": {"kernel": "all (concurrent levels)", "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "autotune": [t for _, t in geoms],
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_fwd(args):
    """the forward-only bandwidth case: one shot of NT steps with seismograms,
    through the plain path and the batched path (one shot, autotuned)"""
    import ifosgpu as ig
    import workloads as wl
    w = wl.fwd536()
    true, _ = wl.random_models(w)
    G = ig.Geom(w.nx, w.ny, w.nz, w.fw, w.dh, w.dh, w.dh, w.dt, w.nt, w.fdcoeff, 0)
    d_model = [G.to_dev(a) for a in true]
    del true
    ntr = len(w.receivers)
    src = w.shots[0]
    res = {}
    ctx = ig.ShotContext(G, 1, ntr)
    ctx.set_cpml(w.vppml, w.fpml)
    ctx.set_freqs(w.finv, w.tast)
    ctx.set_model(*d_model)
    d_syn = ig.DBuf((3, ntr, w.nt + 1))
    ctx.run_shot(src, w.receivers, d_syn, None, want_adjoint=False)
    ig.sync()
    _prof_start(ig)
    t0 = time.perf_counter()
    ctx.run_shot(src, w.receivers, d_syn, None, want_adjoint=False)
    ig.sync()
    res["plain"] = (time.perf_counter() - t0, _prof_stop(ig))
    ctx.close()
    log(f"plain forward: {res['plain'][0]:.2f} s")
    b = ig.BatchContext(G, 1, 1, ntr, ig.DFT_COMPAT, 1)
    b.set_cpml(w.vppml, w.fpml)
    b.set_model(*d_model)
    tune = None
    if not args.no_autotune:
        bv, bs, _ = b.autotune(reps=2)
        tune = {"v": bv, "s": bs}
    d_b = ig.DBuf((1, 3, ntr, w.nt + 1))
    b.run([src], w.receivers, d_syn=d_b, want_adjoint=False)
    ig.sync()
    _prof_start(ig)
    t0 = time.perf_counter()
    b.run([src], w.receivers, d_syn=d_b, want_adjoint=False)
    ig.sync()
    res["batched"] = (time.perf_counter() - t0, _prof_stop(ig))
    b.close()
    log(f"batched forward (one shot): {res['batched'][0]:.2f} s, tiles {tune}")
    cells = w.cells
    ab_p = algorithmic_bytes(w, 1)
    _, _, axes = slab_stats(w)
    ab_b = {"v_fused": 48 * cells + 24 * axes + 12 * cells, "s_fused": 60 * cells + 24 * axes + 20 * cells,
            "vs_fused": 72 * cells + 48 * axes + 32 * cells}
    best = min(res, key=lambda k: res[k][0])
    el, prof = res[best]
    ab = ab_p if best == "plain" else ab_b
    dom, avg_ms, achieved = roofline_of(prof, ab)
    st_cats = [k for k in ("v_interior", "v_cpml", "v_fused", "s_interior", "s_cpml", "s_fused") if k in ab]
    st_bytes = sum(ab[k] * prof[k][1] for k in st_cats)
    st_ms = sum(prof[k][0] for k in st_cats)
    out = {
        "metric": "forward Gcell-updates/s, 512x256x512 + CPML bandwidth case",
        "value": round(cells * w.nt / el / 1e9, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": round(1000.0 * el, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"fwd536: {w.nx}x{w.ny}x{w.nz} (512x256x512 + {w.fw}-cell CPML), dh {w.dh} m, "
                               f"dt {w.dt:.4e} s, NT {w.nt}, seeded random model, one explosive 15 Hz source, "
                               f"{ntr} receivers, elastic FP32 4th order, forward only",
                   "path": f"{best} (fastest of plain and one-shot batched)"},
        "plain_gcell_updates_per_s": round(cells * w.nt / res["plain"][0] / 1e9, 3),
        "batched_gcell_updates_per_s": round(cells * w.nt / res["batched"][0] / 1e9, 3),
        "hbm_frac_stencils": round(st_bytes / (st_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if st_ms else None,
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": int(ab[dom]), "avg_launch_ms": round(avg_ms, 5)},
        "kernels": kernel_table(prof, ab),
        "autotune": tune,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_visco(args):
    """the viscoelastic + DFT case: 404 x 212 x 404 (384 x 192 x 384 + 10-cell
    CPML), three relaxation mechanisms fitted over 5-40 Hz, per-cell Q ~
    U[20, 100], 8 DFT frequencies 5-40 Hz sampled every step, NT 2000.  One
    gradient shot through the plain path with the reference's DFT rounding
    (compat) and with Kahan sums, and through the batched path (one shot,
    Kahan); Gcell-updates/s and the stress update against HBM peak.  The
    DFT's accumulation error: the forward DFT fields at 4096 cells of the
    mid-depth plane (receivers there record the velocities every step)
    against the FP64 sum of the same terms, disc_fourier.c:36-54"""
    import ctypes as C
    import ifosgpu as ig
    import workloads as wl
    w = wl.visco404()
    model, (taus, taup), etas = wl.visco_models(w)
    L = len(etas)
    G = ig.Geom(w.nx, w.ny, w.nz, w.fw, w.dh, w.dh, w.dh, w.dt, w.nt, w.fdcoeff, 0)
    d_model = [G.to_dev(a) for a in model]
    d_tau = [G.to_dev(a) for a in (taus, taup)]
    del model, taus, taup
    nf = len(w.finv)
    jm = w.fw + 1 + (w.ny - 2 * w.fw) // 2
    rec = np.array([[w.fw + 1 + 6 * a, jm, w.fw + 1 + 6 * b] for a in range(64) for b in range(64)], np.int32)
    ntr = len(rec)
    src = w.shots[0]
    ntast = ig.ntast(w.finv[-1], w.tast, w.dt)
    d_obs = ig.DBuf((3, ntr, w.nt + 1))
    log(f"visco404: {w.nx}x{w.ny}x{w.nz}, L={L}, {nf} frequencies, NT {w.nt}, dt {w.dt:.4e}, ntast {ntast}")
    # the device offsets of the sample cells within plane jm
    lib = ig.lib()
    base = lib.ifos_offset(G.ref, jm, -2, -2)
    cell = np.array([lib.ifos_offset(G.ref, jm, int(r[0]), int(r[2])) - base for r in rec], np.int64)
    plane = (w.nx + 5) * G.pitch
    fe = G.field_elems()
    res, fields, syn = {}, {}, None
    for name, prec in (("compat", ig.DFT_COMPAT), ("kahan", ig.DFT_KAHAN)):
        ctx = ig.ShotContext(G, nf, ntr, prec)
        ctx.set_cpml(w.vppml, w.fpml)
        ctx.set_freqs(w.finv, w.tast)
        ctx.set_visco(*d_tau, etas)
        ctx.set_model(*d_model)
        d_syn = ig.DBuf((3, ntr, w.nt + 1))
        grads = [G.empty() for _ in range(3)]
        ig.sync()
        _prof_start(ig)
        t0 = time.perf_counter()
        ctx.run_shot(src, rec, d_syn, grads, d_obs=d_obs)
        ig.sync()
        res[name] = (time.perf_counter() - t0, _prof_stop(ig))
        log(f"plain gradient shot, {name} DFT: {res[name][0]:.2f} s")
        buf = np.empty(plane, np.float32)
        F = np.empty((6, nf, ntr), np.float64)
        for n, ptr in enumerate(ctx.dft_ptrs(0)):
            for l in range(nf):
                src_p = C.c_void_p(ptr + 4 * (l * fe + base))
                ig.check(lib.ifos_memcpy_d2h(buf.ctypes.data_as(C.c_void_p), src_p, C.c_size_t(4 * plane)), "d2h")
                F[n, l] = buf[cell]
        fields[name] = F
        if syn is None:
            syn = d_syn.numpy().astype(np.float64)
        ctx.close()
        del grads, d_syn
    # FP64 DFT of the same FP32 velocities, the terms of disc_fourier.c:42-54
    nts = np.arange(ntast, w.nt + 1, ntast)
    t = (nts.astype(np.float32) * np.float32(w.dt)).astype(np.float64)
    F64 = np.empty((6, nf, ntr), np.float64)
    for l, f in enumerate(w.finv):
        c = np.cos(2.0 * t * np.float64(np.float32(f)) * np.pi) * np.float64(np.float32(w.dt)) * ntast
        sn = np.sin(2.0 * t * np.float64(np.float32(f)) * np.pi) * np.float64(np.float32(w.dt)) * ntast
        for comp in range(3):
            F64[comp, l] = syn[comp][:, nts] @ c
            F64[3 + comp, l] = syn[comp][:, nts] @ sn
    err = {k: float(np.linalg.norm(v - F64) / np.linalg.norm(F64)) for k, v in fields.items()}
    log(f"DFT accumulation error against FP64: {err}")
    # the batched path, one shot, Kahan sums
    b = ig.BatchContext(G, 1, nf, ntr, ig.DFT_KAHAN, args.chunk)
    b.set_cpml(w.vppml, w.fpml)
    b.set_freqs(w.finv, w.tast)
    b.set_visco(*d_tau, etas)
    b.set_model(*d_model)
    grads = [G.empty() for _ in range(3)]
    d_b = ig.DBuf((1, 3, ntr, w.nt + 1))
    ig.sync()
    _prof_start(ig)
    t0 = time.perf_counter()
    b.run([src], rec, d_syn=d_b, grads=grads, d_obs=[d_obs])
    ig.sync()
    res["batched"] = (time.perf_counter() - t0, _prof_stop(ig))
    b.close()
    log(f"batched gradient shot (one shot, Kahan): {res['batched'][0]:.2f} s")
    cells = w.cells
    bi, slab, axes = slab_stats(w)
    vis = 68 + 48 * (L - 1)   # per cell of the stress update: tau arrays, the memory variables
    ab_p = {"v_interior": 60 * bi, "v_cpml": 60 * slab + 24 * axes,
            "s_interior": (80 + vis) * bi, "s_cpml": (80 + vis) * slab + 24 * axes,
            "dft": (12 + 48 * nf) * cells, "gradient": (48 * nf + 36) * cells}
    ab_k = dict(ab_p, dft=(12 + 96 * nf) * cells)
    # the batched DFT: a chunk of c snapshots, 12 B per cell each, folded
    # into 2 x 48 B per cell and frequency of Kahan accumulators
    nchunks = -(-(w.nt // ntast) // args.chunk)
    ab_b = {"v_fused": 60 * cells + 24 * axes, "s_fused": (80 + vis) * cells + 24 * axes,
            "dft": (12 * (w.nt // ntast) + 96 * nf * nchunks) * cells / nchunks,
            "gradient": (48 * nf + 36) * cells}
    out_k = {}
    for name, ab in (("compat", ab_p), ("kahan", ab_k), ("batched", ab_b)):
        el, prof = res[name]
        dom = "s_interior" if name != "batched" else "s_fused"
        avg = prof[dom][0] / prof[dom][1]
        out_k[name] = {"gradient_shot_s": round(el, 3), "gcell_updates_per_s": round(cells * 2 * w.nt / el / 1e9, 3),
                       "stress_kernel_ms": round(avg, 4),
                       "stress_kernel_hbm_frac": round(ab[dom] / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "kernels": kernel_table(prof, ab)}
    el, prof = res["batched"]
    avg = prof["s_fused"][0] / prof["s_fused"][1]
    out = {
        "metric": "viscoelastic gradient shot: Gcell-updates/s and DFT error against FP64, 384x192x384 + CPML",
        "value": round(cells * 2 * w.nt / el / 1e9, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": round(1000.0 * el, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"visco404: {w.nx}x{w.ny}x{w.nz} (384x192x384 + {w.fw}-cell CPML), dh {w.dh} m, "
                               f"dt {w.dt:.4e} s, NT {w.nt}, seeded random model, {L} relaxation mechanisms "
                               f"(5-40 Hz), Q ~ U[20, 100] per cell, {nf} DFT frequencies {w.finv[0]}-{w.finv[-1]} Hz "
                               f"every {ntast} step(s), {ntr} receivers on the mid-depth plane",
                   "path": "batched (one shot, Kahan DFT sums); the plain path in compat and Kahan precision beside it"},
        "dft_error_vs_fp64": {"compat": err["compat"], "kahan": err["kahan"],
                              "what": f"rel. L2 of the forward DFT fields at {ntr} cells (6 components x {nf} "
                                      f"frequencies) against the FP64 sum of the same FP32 velocity terms"},
        "paths": out_k,
        "roofline": {"kernel": "s_fused", "bound": "hbm", "achieved": round(ab_b["s_fused"] / (avg * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ab_b["s_fused"] / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": int(ab_b["s_fused"]), "avg_launch_ms": round(avg, 5)},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_aux(args):
    """the steps on either side of the propagation at the 276^3 gradient
    workload's size, each timed over a few calls: the L-BFGS update (two
    parameters, five stored pairs, FP64 dots), conjugate, precon_grad
    (DAMPTYPE 2: tapers round the 16 sources and 1024 receivers), modelupdate,
    hess_F (one receiver-side set, 6 frequencies), the residual of one shot's
    traces, and an SU write and read of one trace component"""
    import tempfile
    import ifosgpu as ig
    import workloads as wl
    w = wl.grad276()
    G = ig.Geom(w.nx, w.ny, w.nz, w.fw, w.dh, w.dh, w.dh, w.dt, w.nt, w.fdcoeff, 0)
    n1 = (w.nx * w.ny * w.nz)
    cells = n1
    rng = np.random.default_rng(3)
    fe = G.field_elems()

    def field(scale=1.0):
        return ig.DBuf.from_numpy((scale * rng.standard_normal(G.dshape)).astype(np.float32))

    def timed(fn, reps):
        fn()
        ig.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ig.sync()
        return (time.perf_counter() - t0) / reps * 1e3

    res = {}
    grads = [field() for _ in range(3)]
    # L-BFGS: fill the five slots, then time updates that use all of them
    lb = ig.Lbfgs(G, 2, 5)
    lb.savegrad(grads[0], grads[1])
    step = ig.DBuf.from_numpy((0.01 * rng.standard_normal(2 * n1)).astype(np.float32))
    it = [1]

    def lb_update():
        it[0] += 1
        ww = (it[0] - 1) % 5 or 5
        lb.mod(ww).copy_from(step)
        lb.update(grads[0], grads[1], it[0])

    for _ in range(6):
        lb_update()
    ms = timed(lb_update, 3)
    k, n = 5, 2 * n1
    res["lbfgs_update"] = {"ms": round(ms, 3), "bytes": 4 * (10 + 10 * k) * n,
                           "GBps": round(4 * (10 + 10 * k) * n / (ms * 1e-3) / 1e9, 1),
                           "bytes_rule": "(10 + 10 k) n floats, k = 5 pairs, n = 2 N^3"}
    lb.close()
    priors = [field() for _ in range(6)]
    ig.conjugate(G, grads, priors, 1)
    ms = timed(lambda: ig.conjugate(G, grads, priors, 0), 3)
    res["conjugate"] = {"ms": round(ms, 3)}
    src = np.array([[s[0], s[1], s[2]] for s in w.shots], np.int32)
    ms = timed(lambda: ig.precon_grad(G, grads, 2, src, w.receivers), 2)
    res["precon_grad"] = {"ms": round(ms, 3), "bytes": 24 * cells,
                          "note": f"{len(src)} source and {len(w.receivers)} receiver tapers per cell (FP64 exp)"}
    true, _ = wl.random_models(w)
    dm = [G.to_dev(a) for a in true]
    mod = ig.DBuf((2 * n1,))
    small = [field(1e-6) for _ in range(3)]
    ms = timed(lambda: ig.modelupdate(G, small, *dm, 1e-6, (1.0, 1.0, 1.0), 4000.0, 2300.0, 2000.0, bfgsmod=mod,
                                      numpar=2), 3)
    res["modelupdate"] = {"ms": round(ms, 3), "bytes": 44 * cells,
                          "GBps": round(44 * cells / (ms * 1e-3) / 1e9, 1),
                          "bytes_rule": "3 gradients + 3 model fields read, 3 model fields + 2 step slots written"}
    nf = len(w.finv)
    Ff = [ig.DBuf((nf * fe,)) for _ in range(6)]  # zero-filled: the timing does not depend on the values
    Fb = [ig.DBuf((2 * nf * fe,)) for _ in range(6)]
    hess = [G.empty() for _ in range(3)]
    ms = timed(lambda: ig.hess_F(G, Ff, Fb, 1, dm, w.finv, hess), 2)
    hb = (12 * 4 * nf + 36) * cells
    res["hess_F"] = {"ms": round(ms, 3), "bytes": hb, "GBps": round(hb / (ms * 1e-3) / 1e9, 1),
                     "bytes_rule": "48 nf + 36 B per cell (12 DFT fields per frequency, 3 Hessians and the model)"}
    ntr, ns = len(w.receivers), w.nt
    obs = ig.DBuf.from_numpy(rng.standard_normal((3, ntr, ns + 1)).astype(np.float32))
    syn = ig.DBuf.from_numpy(rng.standard_normal((3, ntr, ns + 1)).astype(np.float32))
    adj = ig.DBuf((3, ntr, ns + 2))
    ms = timed(lambda: ig.residual(G, obs, syn, adj, ntr, ns), 5)
    res["residual"] = {"ms": round(ms, 3), "traces": 3 * ntr, "samples": ns}
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "vy.su")
        sec = ig.DBuf.from_numpy(rng.standard_normal((ntr, ns + 1)).astype(np.float32))
        ms_w = timed(lambda: ig.su_write(path, sec, ntr, ns, w.receivers, w.dh, w.dh, w.dh, w.dt, 1,
                                         src_xyz=(1.0, 2.0, 3.0)), 3)
        ms_r = timed(lambda: ig.su_read(path, ntr, ns, sec), 3)
        mb = ntr * (240 + 4 * ns) / 1e6
    res["su_write"] = {"ms": round(ms_w, 3), "MB": round(mb, 2)}
    res["su_read"] = {"ms": round(ms_r, 3), "MB": round(mb, 2)}
    out = {
        "metric": "auxiliary steps at 276^3 (ms per call)",
        "value": res["lbfgs_update"]["ms"],
        "unit": "ms per L-BFGS update",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": res["lbfgs_update"]["ms"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (FP64 sums)",
        "data": "synthetic",
        "config": {"workload": f"grad276 grid {w.nx}^3, 2 inversion parameters, 5 L-BFGS pairs, {nf} frequencies, "
                               f"{ntr} receivers x {ns} samples"},
        "steps_ms": res,
        "roofline": {"kernel": "lbfgs_update", "bound": "hbm", "achieved": res["lbfgs_update"]["GBps"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(res["lbfgs_update"]["GBps"] / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_plain(w, G, d_start, obs, shots, precision):
    """the one-shot-per-launch path on the first `shots` shot positions"""
    import ifosgpu as ig
    ntr = len(w.receivers)
    ctx = ig.ShotContext(G, len(w.finv), ntr, precision)
    ctx.set_cpml(w.vppml, w.fpml)
    ctx.set_freqs(w.finv, w.tast)
    ctx.set_model(*d_start)
    grads = [G.empty() for _ in range(3)]
    d_syn = ig.DBuf((3, ntr, w.nt + 1))
    ig.sync()
    _prof_start(ig)
    t0 = time.perf_counter()
    for n in range(shots):
        ctx.run_shot(w.shots[n], w.receivers, d_syn, grads, d_obs=obs[n])
        log(f"plain shot {n + 1}/{shots}")
    ig.sync()
    elapsed = time.perf_counter() - t0
    prof = _prof_stop(ig)
    ctx.close()
    return dict(elapsed=elapsed, shots=shots, prof=prof)


def run_td(w, G, d_start, obs, shots):
    """the checkpointed time-domain (full-band) gradient, one shot at a time,
    checkpoints every round(sqrt(NT)) steps"""
    import ifosgpu as ig
    ntr = len(w.receivers)
    log(f"time-domain gradient: {ig.ShotContext.td_bytes(G) / 2**30:.1f} GiB of checkpoints and buffers")
    ctx = ig.ShotContext(G, 1, ntr)
    ctx.set_cpml(w.vppml, w.fpml)
    ctx.set_model(*d_start)
    grads = [G.empty() for _ in range(3)]
    d_syn = ig.DBuf((3, ntr, w.nt + 1))
    ctx.run_shot_td(w.shots[0], w.receivers, d_syn, grads, obs[0])  # allocates the checkpoints
    ig.sync()
    _prof_start(ig)
    t0 = time.perf_counter()
    for n in range(shots):
        ctx.run_shot_td(w.shots[n], w.receivers, d_syn, grads, obs[n])
        log(f"time-domain gradient shot {n + 1}/{shots}")
    ig.sync()
    elapsed = time.perf_counter() - t0
    prof = _prof_stop(ig)
    ctx.close()
    return dict(elapsed=elapsed, shots=shots, prof=prof)


def run_dropin(w, host_start, obs0):
    """one gradient shot through the reference prototypes of
    libifos3d_dropin.so, called on Numerical-Recipes host arrays by an
    IFOS3D-shaped driver (tests/dropin/nr_driver.c) in the order of
    ifos3d.c:712-975; the driver's host set-up (NR tensors, matcopy, CPML
    profiles), the per-step NaN probe and seismogram copies, the host
    residual and the gradient download are inside the time"""
    import ctypes as C
    lib = C.CDLL(os.path.join(ROOT, "tests", "dropin", "lib", "libnr_driver.so"))
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    src = w.shots[0]
    si = np.array([src[0], src[1], src[2], src[6]], np.int32)
    sf = np.array([src[3], src[4], src[5]], np.float32)
    recs = np.ascontiguousarray(w.receivers, np.int32)
    ntr = len(recs)
    fv = np.asarray(w.finv, np.float32)
    rho, pi, u = [np.ascontiguousarray(a, np.float32) for a in host_start]
    obs = np.ascontiguousarray(obs0, np.float32)
    syn = np.zeros((3, ntr, w.nt + 1), np.float32)
    grads = np.zeros((3,) + rho.shape, np.float32)
    l2 = C.c_float(0.0)
    t0 = time.perf_counter()
    rc = lib.dd_run_shot(w.nx, w.ny, w.nz, w.fw, C.c_float(w.dh), C.c_float(w.dh), C.c_float(w.dh),
                         C.c_float(w.dt), w.nt, w.fdcoeff, 0, p(rho), p(pi), p(u), p(si), p(sf), 1, ntr, p(recs),
                         p(obs), p(fv), len(fv), w.tast, C.c_float(w.vppml), C.c_float(w.fpml), p(syn), p(grads),
                         C.byref(l2))
    elapsed = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError(f"dd_run_shot returned {rc}")
    t = (C.c_double * 3).in_dll(lib, "dd_last_times")
    return dict(elapsed=elapsed, shots=1, l2=l2.value, shot_calls_s=t[0], setup_s=t[1])


def copy_probe(mib=1024, reps=5):
    """plain device copy of a known size: the practical HBM ceiling, and the
    reference kernel the rocprof FETCH_SIZE/WRITE_SIZE counters are scaled by"""
    import ifosgpu as ig
    n = mib * (1 << 20) // 4
    a = ig.DBuf((n,))
    b = ig.DBuf((n,))
    ig.copy(b, a, n)
    _prof_start(ig)
    for _ in range(reps):
        ig.copy(b, a, n)
    prof = _prof_stop(ig)
    ms = prof["other"][0] / max(prof["other"][1], 1)
    a.free()
    b.free()
    return {"bytes_per_copy": 2 * n * 4, "avg_ms": round(ms, 4), "GBps": round(2 * n * 4 / (ms * 1e-3) / 1e9, 1)}


def cpu_baseline(w, host_model, steps=8):
    """The oracle's C restatement (oracle/ifos_oracle.c, bit-identical to the
    reference's routines on the golden vectors) with its cell loops split
    over every host core by OpenMP: a bounded sample of one gradient shot of
    the same workload -- `steps` forward steps (update_v, update_v_CPML,
    update_s, update_s_CPML, psource, seismo, the DFT at its sampling steps),
    `steps` adjoint steps (with the adjoint injection and the DFT) and one
    gradient_F over every frequency -- extrapolated to the NT + NT steps of a
    shot.  The steps start from seeded random wavefields of normal magnitude,
    so every cell does the full arithmetic (a shot's early steps are mostly
    zeros)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    o = po.Oracle("f32")
    g = po.Grid(nx=w.nx, ny=w.ny, nz=w.nz, fw=w.fw, dx=w.dh, dy=w.dh, dz=w.dh, dt=w.dt, nt=w.nt,
                fdcoeff=w.fdcoeff)
    ntast = o.ntast(max(w.finv), w.tast, w.dt)
    nf = len(w.finv)
    rho, pi, u = [np.ascontiguousarray(a, np.float32) for a in host_model]
    for m in (rho, pi, u):
        o.matcopy_wrap(g, m)
    uipjp, ujpkp, uipkp, rjp, rkp, rip = o.av_mat(g, rho, u)
    prof, _ = o.cpml_coeff(g, w.vppml, w.fpml)
    box = o.box(g)
    rng = np.random.default_rng(7)
    f = [(1e-3 * rng.standard_normal(g.shape)).astype(np.float32) for _ in range(9)]
    psi_s = [np.zeros(sh, np.float32) for sh in g.psi_shapes()]
    psi_v = [np.zeros(sh, np.float32) for sh in g.psi_shapes()]
    src = w.shots[0]
    sig = o.wavelet(g, [src], 1)
    rec = np.ascontiguousarray(w.receivers, np.int32)
    ntr = len(rec)
    radj = [(1e-6 * rng.standard_normal((ntr, w.nt + 2))).astype(np.float32) for _ in range(3)]
    rsrc = [(int(r[0]), int(r[1]), int(r[2]), 0.0, 0.0, 0.0, 0) for r in rec]
    Ff = [np.zeros((nf,) + g.shape, np.float32) for _ in range(6)]
    Fb = [np.zeros((nf,) + g.shape, np.float32) for _ in range(6)]
    sec = [np.zeros((ntr, w.nt + 1), np.float32) for _ in range(3)]

    def step(nt, back):
        if back:
            o.update_v(g, box, nt, f, rjp, rkp, rip, rsrc, radj[0], radj[1], radj[2], 1)
            o.update_v_cpml(g, box, nt, f, rjp, rkp, rip, rsrc, radj[0], radj[1], radj[2], 1, prof, psi_s)
        else:
            o.update_v(g, box, nt, f, rjp, rkp, rip, [src], sig, sig, sig, 0)
            o.update_v_cpml(g, box, nt, f, rjp, rkp, rip, [src], sig, sig, sig, 0, prof, psi_s)
        o.update_s(g, box, f, pi, u, uipjp, ujpkp, uipkp)
        o.update_s_cpml(g, box, f, pi, u, uipjp, ujpkp, uipkp, prof, psi_v)
        if not back:
            o.psource(g, nt, f[3], f[4], f[5], [src], sig)
        if nt % ntast == 0:
            F = Fb if back else Ff
            o.discfourier(g, nt, f[0], f[1], f[2], F[:3], F[3:], w.finv, ntast, back)
        if not back:
            o.seismo(g, nt, rec, sec, f)

    t0 = time.perf_counter()
    for nt in range(1, steps + 1):
        step(nt, 0)
    t1 = time.perf_counter()
    for nt in range(1, steps + 1):
        step(nt, 1)
    t2 = time.perf_counter()
    grads = [o.zeros(g) for _ in range(3)]
    o.gradient_F(g, Ff, Fb, grads, rho, pi, u, w.finv)
    t3 = time.perf_counter()
    tf, ta, tg = (t1 - t0) / steps, (t2 - t1) / steps, t3 - t2
    shot_s = w.nt * (tf + ta) + tg
    cores = o.threads()
    return {"value": 1.0 / shot_s, "unit": "gradient shots/s", "cores": cores, "kind": "port",
            "sample": (f"oracle/ifos_oracle.c (OpenMP, {cores} threads of {os.cpu_count()} host CPUs): "
                       f"{steps} forward + {steps} adjoint time steps of one gradient shot on the full "
                       f"{w.nx}x{w.ny}x{w.nz} grid (update_v, update_v_CPML, update_s_elastic, "
                       f"update_s_CPML_elastic, psource/adjoint injection, seismo, discfourier of {nf} "
                       f"frequencies every {ntast} steps) from seeded random wavefields, plus one gradient_F: "
                       f"{tf:.3f} s/forward step, {ta:.3f} s/adjoint step, {tg:.2f} s gradient_F, "
                       f"extrapolated to {w.nt}+{w.nt} steps per shot"),
            "s_per_forward_step": round(tf, 4), "s_per_adjoint_step": round(ta, 4), "s_gradient_F": round(tg, 3)}


def roofline_of(prof, ab):
    """the dominant kernel category by total time, its average launch and
    achieved algorithmic GB/s"""
    dom = max((k for k in ab if prof[k][1] > 0), key=lambda k: prof[k][0])
    avg_ms = prof[dom][0] / prof[dom][1]
    return dom, avg_ms, ab[dom] / (avg_ms * 1e-3) / 1e9


def kernel_table(prof, ab):
    return {k: {"ms_total": round(prof[k][0], 3), "launches": prof[k][1],
                "avg_ms": round(prof[k][0] / prof[k][1], 5) if prof[k][1] else None,
                "GBps": round(ab[k] / (prof[k][0] / prof[k][1] * 1e-3) / 1e9, 1)
                if prof[k][1] and k in ab else None}
            for k in CATS if prof[k][1]}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="timed batches (batched) or shots (plain)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="grad276",
                    choices=["grad276", "toy84", "levels3", "fwd536", "aux276", "visco404"])
    ap.add_argument("--level-shots", type=int, default=32, help="shots per grid level (levels3)")
    ap.add_argument("--path", default="batched", choices=["batched", "plain"])
    ap.add_argument("--shots-per-batch", type=int, default=16)
    ap.add_argument("--chunk", type=int, default=16, help="DFT samples per accumulator pass")
    ap.add_argument("--graph", action="store_true", help="replay the time loops from hipGraphs")
    ap.add_argument("--no-autotune", action="store_true", help="default tile shapes instead of the autotuner's")
    ap.add_argument("--fused", action="store_true",
                    help="fused stress+velocity step without the autotuner's verdict (the autotuner "
                         "turns it on by itself only where it is faster)")
    ap.add_argument("--tiles", default="",
                    help="tile shapes used without the autotuner, velocity/stress march, each "
                         "TI,jc,group,split,pipe[,kb[,sb[,lds]]] (e.g. 4,32,0,0,0/4,32,0,0,1)")
    ap.add_argument("--nt", type=int, default=0,
                    help="shorter time loop (profiling runs only: the line then reports a smaller workload)")
    ap.add_argument("--plain-shots", type=int, default=1, help="plain-path shots timed beside the batch")
    ap.add_argument("--no-kahan-graph", action="store_true",
                    help="skip the batched run with Kahan DFT sums and hipGraph replay")
    ap.add_argument("--dropin-shots", type=int, default=1,
                    help="gradient shots timed through the drop-in library on NR host arrays (grad276)")
    ap.add_argument("--td-shots", type=int, default=1,
                    help="time-domain (full-band, checkpointed) gradient shots timed beside the batch")
    ap.add_argument("--precision", default="compat", choices=["compat", "kahan"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.tiles:
        global DEFAULT_TILES
        DEFAULT_TILES = tuple(tuple(int(x) for x in part.split(",")) for part in args.tiles.split("/"))
        if len(DEFAULT_TILES) != 2 or any(len(t) not in (5, 6, 7, 8) for t in DEFAULT_TILES):
            ap.error("--tiles wants two groups of five to eight integers separated by '/'")
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(shots would be sharded one rank per GPU). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import ifosgpu as ig
    import workloads as wl
    ig.lib()  # fail loudly if the HIP library is not built
    if ig.device_count() < 1:
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    log(f"device: {ig.device_name()}")
    copy = copy_probe()
    log(f"copy probe: {copy}")
    if args.config == "levels3":
        run_levels(args)
        return
    if args.config == "fwd536":
        run_fwd(args)
        return
    if args.config == "aux276":
        run_aux(args)
        return
    if args.config == "visco404":
        run_visco(args)
        return
    w = wl.grad276() if args.config == "grad276" else wl.toy84()
    if args.nt > 0:
        w.nt = args.nt
        w.name += f"_nt{args.nt}"
    prec = ig.DFT_KAHAN if args.precision == "kahan" else ig.DFT_COMPAT
    kahan = prec == ig.DFT_KAHAN
    nf = len(w.finv)
    S = min(args.shots_per_batch, len(w.shots))
    ntast = ig.ntast(w.finv[-1], w.tast, w.dt)
    G, d_true, d_start, host_start = setup(w)
    obs = observed_data(w, G, d_true, S, args.chunk, args.fused)
    del d_true

    plain = None
    if args.path == "plain" or args.plain_shots > 0:
        n_plain = args.steps if args.path == "plain" else args.plain_shots
        plain = run_plain(w, G, d_start, obs, n_plain, prec)
        plain["value"] = plain["shots"] / plain["elapsed"]
        log(f"plain path: {plain['value']:.4f} shots/s")
    batched = None
    if args.path == "batched":
        batched = run_batched(w, G, d_start, obs, args.steps, args.warmup, prec, S, args.chunk, args.graph,
                              not args.no_autotune, args.fused)
        batched["value"] = batched["shots"] / batched["elapsed"]
        log(f"batched path: {batched['value']:.4f} shots/s")

    # the north star's own configuration of the batched path: compensated
    # (Kahan) DFT accumulators and the time loops replayed from hipGraphs;
    # Kahan doubles the accumulators, so 16 shots run as two batches of 8
    # the legs beside the headline figure are reported, never required: one
    # that fails is logged and left out of the line
    def leg(name, fn):
        try:
            r = fn()
            r["value"] = r["shots"] / r["elapsed"]
            log(f"{name}: {r['value']:.4f} shots/s")
            return r
        except Exception as e:
            log(f"{name} failed: {e}")
            return None

    kg = None
    if args.path == "batched" and not args.no_kahan_graph and args.config == "grad276":
        Sk = min(8, len(w.shots))
        kg = leg(f"batched path, Kahan DFT, hipGraph replay, {Sk} shots per batch",
                 lambda: run_batched(w, G, d_start, obs, max(1, len(w.shots) // Sk), 1, ig.DFT_KAHAN, Sk,
                                     args.chunk, True, autotune=False))
    td = leg("time-domain gradient", lambda: run_td(w, G, d_start, obs, args.td_shots)) if args.td_shots > 0 else None
    dropin = None
    if args.dropin_shots > 0 and args.config == "grad276":
        dropin = leg("drop-in (reference prototypes on NR host arrays)",
                     lambda: run_dropin(w, host_start, obs[0].numpy()))

    res = batched if args.path == "batched" else plain
    el, value = res["elapsed"], res["value"]
    if args.path == "batched":
        ab = algorithmic_bytes_batched(w, nf, S, args.chunk, ntast, kahan)
    else:
        ab = algorithmic_bytes(w, nf, kahan)
    roof = None
    kernels = None
    if res["prof"] is not None:
        dom, avg_ms, achieved = roofline_of(res["prof"], ab)
        kernels = kernel_table(res["prof"], ab)
        path_tag = "batched" if args.path == "batched" else "plain"
        traffic = None
        tpath = os.path.join(ROOT, "profiles", f"{w.name}_{path_tag}_traffic.txt")
        tiles = None
        if args.path == "batched":
            tune = batched.get("tune")
            tiles = (tuple(tune["v"]), tuple(tune["s"])) if tune else DEFAULT_TILES
        kname = kernel_names(path_tag, nf, args.chunk, tiles).get("dft_kahan" if dom == "dft" and kahan else dom, "")
        if os.path.exists(tpath) and kname:
            import profsum
            traffic = profsum.read_traffic(tpath).get(kname)
        roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": os.path.relpath(tpath, ROOT) if traffic else None,
                "algorithmic_bytes_per_launch": int(ab[dom]), "avg_launch_ms": round(avg_ms, 5),
                "rocprof_kernel": kname or None}
        st_cats = [k for k in ("v_interior", "v_cpml", "v_fused", "s_interior", "s_cpml", "s_fused", "vs_fused")
                   if k in ab]
        st_bytes = sum(ab[k] * res["prof"][k][1] for k in st_cats)
        st_ms = sum(res["prof"][k][0] for k in st_cats)
        hbm_frac_stencils = (st_bytes / (st_ms * 1e-3) / 1e9) / HBM_PEAK_GBS if st_ms > 0 else None
    else:
        hbm_frac_stencils = None
    gcells = w.cells * 2 * w.nt * res["shots"] / el / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (the oracle restatement, OpenMP on every host core, bounded sample)")
        try:
            cpu = cpu_baseline(w, host_start)
        except Exception as e:  # the baseline is reported, never the product
            log(f"CPU baseline failed: {e}")
    path_desc = (f"batched: {S} shots per launch, "
                 f"{'fused stress+velocity step (one launch per time step)' if batched and batched['fused'] else 'separate velocity and stress marches'}, "
                 f"DFT in chunks of {args.chunk} samples, {'hipGraph replay' if args.graph else 'direct launches'}"
                 if args.path == "batched" else
                 "plain (one kernel per reference routine per step, one shot per launch)")
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "gradient shots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * el / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"{w.name}: {w.nx}x{w.ny}x{w.nz} ({w.nx - 2 * w.fw}^3 + {w.fw}-cell CPML), "
                                f"elastic FP32 4th order, dh {w.dh} m, dt {w.dt:.4e} s, NT {w.nt}, "
                                f"{len(w.shots)} shot positions, {len(w.receivers)} receivers, "
                                f"{nf} DFT frequencies {w.finv[0]}-{w.finv[-1]} Hz, TAST {w.tast}"),
                   "path": path_desc, "dft_precision": args.precision,
                   "shots_per_step": res["shots"] // args.steps},
        "gcell_updates_per_s": round(gcells, 3),
        "hbm_frac_stencils": round(hbm_frac_stencils, 4) if hbm_frac_stencils else None,
        "plain_shots_per_s": round(plain["value"], 4) if plain else None,
        "batched_over_plain": round(batched["value"] / plain["value"], 3) if (batched and plain) else None,
        "td_shots_per_s": round(td["value"], 4) if td else None,
        "dropin_shots_per_s": round(dropin["value"], 4) if dropin else None,
        "dropin_s": {"shot_calls": round(dropin["shot_calls_s"], 3), "host_setup": round(dropin["setup_s"], 3),
                     "whole_call": round(dropin["elapsed"], 3)} if dropin else None,
        "kahan_graph_shots_per_s": round(kg["value"], 4) if kg else None,
        "td_kernels_ms": {k: round(v[0], 1) for k, v in td["prof"].items() if v[1]} if td else None,
        "roofline": roof,
        "kernels": kernels,
        "plain_kernels": kernel_table(plain["prof"], algorithmic_bytes(w, nf, kahan))
        if (plain and args.path == "batched") else None,
        "autotune": batched.get("tune") if batched else None,
        "copy_probe": copy,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
