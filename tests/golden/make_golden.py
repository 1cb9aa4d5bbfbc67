#!/usr/bin/env python3
"""Write tests/golden/golden.json from the reference's own routines.

The reference ships no test vectors, so they are generated here: the hot-path
sources of the reference are compiled in place by oracle/Makefile.ref into
oracle/_ref/libifos_ref.so, and this script runs the inputs of cases.py
through them.  Each output is stored as the SHA-256 of its exact float32
bytes (the oracle must match bit for bit), its L2 norm and a few samples.
Run in the development container (the reference sources are not shipped):

    make -C oracle -f Makefile.ref && python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import cases as cs  # noqa: E402
from helpers import f32hex, sha  # noqa: E402
from pyoracle import Ref  # noqa: E402


def rec(a):
    a = np.ascontiguousarray(a, np.float32)
    flat = a.ravel()
    idx = np.linspace(0, flat.size - 1, 6).astype(int)
    return {"sha256": sha(a), "shape": list(a.shape), "l2": float(np.linalg.norm(flat.astype(np.float64))),
            "samples": {str(int(i)): f32hex(flat[i:i + 1])[0] for i in idx}}


def main():
    if not Ref.available():
        sys.exit("oracle/_ref/libifos_ref.so is not built (make -C oracle -f Makefile.ref)")
    out = {"generator": "reference routines compiled from source (oracle/_ref)", "cases": {}}
    C = out["cases"]

    # g1: CPML profiles (CPML_coeff.c)
    for n, (fw, dh, vppml, fpml, dt) in enumerate(cs.CPML_CASES):
        g = cs.Grid(nx=2 * fw + 3, ny=2 * fw + 4, nz=2 * fw + 5, fw=fw, dx=dh, dy=dh * 1.25, dz=dh * 1.5, dt=dt,
                    nt=10)
        r = Ref(g, vppml=vppml, fpml=fpml)
        C[f"cpml_{n}"] = {"prof": [f32hex(p) for p in r.prof()]}

    # g2: one step of every stencil routine, sources, adjoint injection
    for fdc, fs in cs.STEP_CASES:
        g = cs.step_grid(fdc, fs)
        key = f"step_fd{fdc}_fs{fs}"
        f, (rho, pi, u), psi, adj = cs.step_inputs(g, 11 * fdc + fs)
        r = Ref(g, vppml=4500.0, fpml=40.0, nfmax=1, nsrc=len(cs.STEP_SOURCES), ntr=len(cs.STEP_RECEIVERS))
        for name, a in zip(Ref.FIELDS[:9], f):
            r.set(name, a)
        for name, a in zip(("rho", "pi", "u"), (rho, pi, u)):
            r.set(name, a)
        for n in range(18):
            r.psi_set(n, psi[n])
        for s, src in enumerate(cs.STEP_SOURCES):
            r.set_source(s, *src)
        sig = r.wavelet(1)
        r.set_receivers(cs.STEP_RECEIVERS)
        r.set_adjoint(*adj)
        r.call("matcopy")
        r.call("av_mat")
        res = {"signals": rec(sig[:, 1:g.nt + 1])}
        for name in ("rho", "pi", "u"):
            res["wrapped_" + name] = rec(r.get(name)[2:-1, 2:-1, 2:-1])
        for name in ("uipjp", "ujpkp", "uipkp", "rjp", "rkp", "rip"):
            res[name] = rec(g.interior(r.get(name)))
        r.call("update_v", cs.STEP_NT, 0)
        r.call("update_v_cpml", cs.STEP_NT, 0)
        for name in ("vx", "vy", "vz"):
            res["v0_" + name] = rec(g.interior(r.get(name)))
        for n in range(9):
            res[f"v0_psi{n}"] = rec(r.psi_get(n))
        r.call("update_s")
        r.call("update_s_cpml")
        r.call("psource", cs.STEP_NT)
        if fs:
            r.call("surface")
        for name in ("sxx", "syy", "szz", "sxy", "syz", "sxz"):
            res["s_" + name] = rec(r.get(name)[1:-1, 1:-1, 1:-1])
        for n in range(9):
            res[f"s_psi{n}"] = rec(r.psi_get(9 + n))
        r.call("update_v", cs.STEP_NT + 1, 1)
        r.call("update_v_cpml", cs.STEP_NT + 1, 1)
        for name in ("vx", "vy", "vz"):
            res["v1_" + name] = rec(g.interior(r.get(name)))
        C[key] = res

    # g3: discfourier (forward and adjoint time conventions)
    g = cs.dft_grid()
    v, F = cs.dft_inputs(g, 5)
    r = Ref(g, nfmax=len(cs.DFT_FINV))
    for name, a in zip(("vx", "vy", "vz"), v):
        r.set(name, a)
    r.set_freqs(cs.DFT_FINV)
    for n in range(12):
        for m in range(len(cs.DFT_FINV)):
            r.F_set(n, m + 1, F[n][m])
    for nt, ntast, back in cs.DFT_CALLS:
        r.call("discfourier", nt, ntast, back)
    C["dft"] = {f"F{n}": rec(np.stack([g.interior(r.F_get(n, m + 1)) for m in range(len(cs.DFT_FINV))]))
                for n in range(12)}

    # g4: gradient_F
    F, grads, (rho, pi, u) = cs.grad_inputs(g, 7)
    r = Ref(g, nfmax=len(cs.DFT_FINV))
    for n in range(12):
        for m in range(len(cs.DFT_FINV)):
            r.F_set(n, m + 1, F[n][m])
    for n in range(3):
        r.set(f"grad{n + 1}", grads[n])
    for name, a in zip(("rho", "pi", "u"), (rho, pi, u)):
        r.set(name, a)
    r.set_freqs(cs.DFT_FINV)
    r.call("gradient_F")
    C["gradient"] = {f"grad{n + 1}": rec(g.interior(r.get(f"grad{n + 1}"))) for n in range(3)}

    # g9: hess_F (hess_F.c:28-192), two receiver-side sets
    Ff, Fb, hess, (rho, pi, u) = cs.hess_inputs(g, 9)
    nf = len(cs.DFT_FINV)
    r = Ref(g, nfmax=nf * (cs.HESS_NTR + 1))
    for n in range(6):
        for m in range(nf):
            r.F_set(n, m + 1, Ff[n][m])
        for m in range(nf * (cs.HESS_NTR + 1)):
            r.F_set(6 + n, m + 1, Fb[n][m])
    for n in range(3):
        r.set(f"grad{n + 1}", hess[n])
    for name, a in zip(("rho", "pi", "u"), (rho, pi, u)):
        r.set(name, a)
    r.set_freqs(cs.DFT_FINV)
    r.call("hess_F", cs.HESS_NTR)
    C["hess"] = {f"hess{n + 1}": rec(g.interior(r.get(f"grad{n + 1}"))) for n in range(3)}

    # g10: SU seismogram files (outseis.c SEIS_FORMAT 1, readseis.c); the
    # header words outseis.c leaves unassigned are zeroed before storing
    import tempfile
    from suio_oracle import mask_unassigned
    for n, case in enumerate(cs.SU_CASES):
        ntr, ns, rid, src, dx, dy, dz, dt, ndt = case
        sec, rxyz = cs.su_inputs(case, 100 + n)
        with tempfile.TemporaryDirectory() as td:
            prefix = os.path.join(td, "obs")
            path = f"{prefix}_vy.su.shot1.0"
            Ref.outseis(path, 2, sec, rxyz, rid, src, dx, dy, dz, dt, ndt)
            data = open(path, "rb").read()
            full = Ref.readseis(prefix, 1, 2, ntr, ns)
            cut = 240 + 4 * ns + 240 + 4 * (ns // 2)  # second trace ends mid-way
            with open(path, "r+b") as fh:
                fh.truncate(cut)
            short = Ref.readseis(prefix, 1, 2, ntr, ns)
        C[f"su_{n}"] = {"bytes": mask_unassigned(data, ns).hex(), "read": rec(full), "read_short": rec(short),
                        "short_len": cut}

    # g5: residual
    g = cs.dft_grid()
    ntr = 5
    obs, syn = cs.residual_inputs(ntr, g.nt, 9)
    r = Ref(g, ntr=ntr)
    r.set_sections(syn)
    l2 = r.residual(obs)
    adj = r.get_adjoint()
    C["residual"] = {"adj": rec(np.stack([a[:, 1:g.nt + 1] for a in adj])), "l2": f32hex([l2])[0]}

    # g6: whole gradient shots (ifos3d.c:712-975 order)
    for fs in (0, 1):
        g = cs.shot_grid(fs)
        start, true = cs.shot_models(g)
        r = Ref(g, vppml=4000.0, fpml=50.0, nfmax=len(cs.SHOT_FINV), nsrc=1, ntr=len(cs.SHOT_REC))
        r.set_freqs(cs.SHOT_FINV)
        r.set_source(0, *cs.SHOT_SRC)
        r.set_receivers(cs.SHOT_REC)
        for name, a in zip(("rho", "pi", "u"), true):
            r.set(name, a)
        r.run_shot(1, 100, None, want_adjoint=False)
        obs = r.sections()
        for name, a in zip(("rho", "pi", "u"), start):
            r.set(name, a)
        r.call("zero_grad")
        l2 = r.run_shot(1, 100, obs, want_adjoint=True)
        res = {"obs": rec(np.stack([o[:, 1:] for o in obs])), "l2": f32hex([l2])[0]}
        for n in range(3):
            res[f"grad{n + 1}"] = rec(g.interior(r.get(f"grad{n + 1}")))
        C[f"shot_fs{fs}"] = res

    # g8: viscoelastic stress update, interior box and CPML slabs (update_s_ssg.c,
    # update_s_ssg_CPML.c), two calls so the memory variables carry over
    for fdc, fs in cs.VISCO_CASES:
        g = cs.step_grid(fdc, fs)
        f, (rho, pi, u), psi, r6, taus, taup = cs.visco_inputs(g, 50 + 7 * fdc + fs)
        r = Ref(g, vppml=4500.0, fpml=40.0)
        for name, a in zip(Ref.FIELDS[:9], f):
            r.set(name, a)
        for name, a in zip(("rho", "pi", "u", "taus", "taup"), (rho, pi, u, taus, taup)):
            r.set(name, a)
        for name, a in zip(("rxx", "ryy", "rzz", "rxy", "ryz", "rxz"), r6):
            r.set(name, a)
        for n in range(18):
            r.psi_set(n, psi[n])
        r.set_visco(1, cs.VISCO_ETA)
        r.call("matcopy")
        r.call("av_mat")
        res = {name: rec(g.interior(r.get(name))) for name in ("tausipjp", "tausjpkp", "tausipkp")}
        for _ in range(2):
            r.call("update_s_visco")
            r.call("update_s_cpml_visco")
        for name in ("sxx", "syy", "szz", "sxy", "syz", "sxz", "rxx", "ryy", "rzz", "rxy", "ryz", "rxz"):
            res[name] = rec(g.interior(r.get(name)))
        for n in range(9):
            res[f"psi{n}"] = rec(r.psi_get(9 + n))
        C[f"visco_fd{fdc}_fs{fs}"] = res

    path = os.path.join(HERE, "golden.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(C)} cases)")


if __name__ == "__main__":
    main()
