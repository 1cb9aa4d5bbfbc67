"""The oracle against the golden vectors of the reference's own routines.

tests/golden/golden.json holds, for the inputs of tests/golden/cases.py, the
outputs of the reference compiled from its sources (make_golden.py): CPML
profiles, one step of every stencil routine with sources and adjoint
injection, material averaging and halo wrap, the DFT, gradient_F, the
residual and whole gradient shots.  The oracle (compat precision) must
reproduce every array bit for bit (SHA-256 of the float32 bytes); on a
mismatch the stored norms and samples show how far off it is.
"""
import json
import os

import numpy as np
import pytest

from golden import cases as cs
from helpers import f32hex, from_f32hex, sha
from pyoracle import Oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.json")


@pytest.fixture(scope="module")
def gold():
    with open(GOLDEN) as fh:
        return json.load(fh)["cases"]


def check(want, got, what):
    got = np.ascontiguousarray(got, np.float32)
    assert list(got.shape) == want["shape"], f"{what}: shape {got.shape} != {want['shape']}"
    if sha(got) != want["sha256"]:
        flat = got.ravel()
        bad = {i: (f32hex(flat[int(i):int(i) + 1])[0], h) for i, h in want["samples"].items()
               if f32hex(flat[int(i):int(i) + 1])[0] != h}
        l2 = float(np.linalg.norm(flat.astype(np.float64)))
        pytest.fail(f"{what}: differs from the reference (L2 {l2:.9e} vs {want['l2']:.9e}; samples {bad})")


@pytest.mark.parametrize("n", range(len(cs.CPML_CASES)))
def test_cpml_profiles(gold, n):
    fw, dh, vppml, fpml, dt = cs.CPML_CASES[n]
    g = cs.Grid(nx=2 * fw + 3, ny=2 * fw + 4, nz=2 * fw + 5, fw=fw, dx=dh, dy=dh * 1.25, dz=dh * 1.5, dt=dt, nt=10)
    prof, alpha = Oracle("f32").cpml_coeff(g, vppml, fpml)
    want = gold[f"cpml_{n}"]["prof"]
    # reference order per axis: K a b K_half a_half b_half alpha alpha_half
    for ax in range(3):
        ours = prof[6 * ax:6 * ax + 6] + [alpha[2 * ax], alpha[2 * ax + 1]]
        for m in range(8):
            w = from_f32hex(want[8 * ax + m])
            assert np.array_equal(ours[m][1:].view(np.uint32), w[1:].view(np.uint32)), (ax, m)


@pytest.mark.parametrize("fdc,fs", cs.STEP_CASES)
def test_one_step(gold, fdc, fs):
    g = cs.step_grid(fdc, fs)
    want = gold[f"step_fd{fdc}_fs{fs}"]
    o = Oracle("f32")
    f, (rho, pi, u), psi, adj = cs.step_inputs(g, 11 * fdc + fs)
    sig = o.wavelet(g, cs.STEP_SOURCES, 1)
    check(want["signals"], sig[:, 1:g.nt + 1], "wavelet")
    for a in (rho, pi, u):
        o.matcopy_wrap(g, a)
    for name, a in zip(("rho", "pi", "u"), (rho, pi, u)):
        check(want["wrapped_" + name], a[2:-1, 2:-1, 2:-1], "matcopy " + name)
    avg = o.av_mat(g, rho, u)
    for name, a in zip(("uipjp", "ujpkp", "uipkp", "rjp", "rkp", "rip"), avg):
        check(want[name], g.interior(a), "av_mat " + name)
    prof, _ = o.cpml_coeff(g, 4500.0, 40.0)
    box = o.box(g)
    nt = cs.STEP_NT
    ps, pv = psi[:9], psi[9:]
    o.update_v(g, box, nt, f, avg[3], avg[4], avg[5], cs.STEP_SOURCES, sig, sig, sig, 0)
    o.update_v_cpml(g, box, nt, f, avg[3], avg[4], avg[5], cs.STEP_SOURCES, sig, sig, sig, 0, prof, ps)
    for n, name in enumerate(("vx", "vy", "vz")):
        check(want["v0_" + name], g.interior(f[n]), "update_v " + name)
    for n in range(9):
        check(want[f"v0_psi{n}"], ps[n], f"update_v_CPML psi {n}")
    o.update_s(g, box, f, pi, u, avg[0], avg[1], avg[2])
    o.update_s_cpml(g, box, f, pi, u, avg[0], avg[1], avg[2], prof, pv)
    o.psource(g, nt, f[3], f[4], f[5], cs.STEP_SOURCES, sig)
    if fs:
        o.surface(g, u, pi, f)
    for n, name in enumerate(("sxx", "syy", "szz", "sxy", "syz", "sxz")):
        check(want["s_" + name], f[3 + n][1:-1, 1:-1, 1:-1], "update_s " + name)
    for n in range(9):
        check(want[f"s_psi{n}"], pv[n], f"update_s_CPML psi {n}")
    recsrc = [(int(r[0]), int(r[1]), int(r[2]), 0.0, 0.0, 0.0, 0) for r in cs.STEP_RECEIVERS]
    o.update_v(g, box, nt + 1, f, avg[3], avg[4], avg[5], recsrc, adj[0], adj[1], adj[2], 1)
    o.update_v_cpml(g, box, nt + 1, f, avg[3], avg[4], avg[5], recsrc, adj[0], adj[1], adj[2], 1, prof, ps)
    for n, name in enumerate(("vx", "vy", "vz")):
        check(want["v1_" + name], g.interior(f[n]), "adjoint update_v " + name)


def test_discfourier(gold):
    g = cs.dft_grid()
    o = Oracle("f32")
    v, F = cs.dft_inputs(g, 5)
    for nt, ntast, back in cs.DFT_CALLS:
        b = 6 * back  # forward fields accumulate in F[0:6], adjoint ones in F[6:12]
        o.discfourier(g, nt, v[0], v[1], v[2], F[b:b + 3], F[b + 3:b + 6], cs.DFT_FINV, ntast, back)
    for n in range(12):
        check(gold["dft"][f"F{n}"], F[n][:, 3:3 + g.ny, 3:3 + g.nx, 3:3 + g.nz], f"F{n}")


def test_gradient_F(gold):
    g = cs.dft_grid()
    o = Oracle("f32")
    F, grads, (rho, pi, u) = cs.grad_inputs(g, 7)
    o.gradient_F(g, F[:6], F[6:], grads, rho, pi, u, cs.DFT_FINV)
    for n in range(3):
        check(gold["gradient"][f"grad{n + 1}"], g.interior(grads[n]), f"grad{n + 1}")


def test_hess_F(gold):
    """hess_F.c on seeded DFT fields, two receiver-side sets"""
    g = cs.dft_grid()
    o = Oracle("f32")
    Ff, Fb, hess, (rho, pi, u) = cs.hess_inputs(g, 9)
    o.hess_F(g, Ff, Fb, hess, rho, pi, u, cs.DFT_FINV, cs.HESS_NTR)
    for n in range(3):
        check(gold["hess"][f"hess{n + 1}"], g.interior(hess[n]), f"hess{n + 1}")


def test_residual(gold):
    g = cs.dft_grid()
    o = Oracle("f32")
    obs, syn = cs.residual_inputs(5, g.nt, 9)
    adj = []
    # the reference accumulates L2 over the three components in one float
    import ctypes as C
    l2c = C.c_float(0.0)
    og = g.og()
    for c in range(3):
        d = np.zeros_like(obs[c])
        o.lib.og_residual(C.byref(og), obs[c].ctypes.data_as(C.c_void_p), syn[c].ctypes.data_as(C.c_void_p),
                          d.ctypes.data_as(C.c_void_p), 5, g.nt, C.byref(l2c))
        adj.append(d[:, 1:g.nt + 1])
    check(gold["residual"]["adj"], np.stack(adj), "residual traces")
    assert f32hex([l2c.value])[0] == gold["residual"]["l2"]


@pytest.mark.parametrize("fs", [0, 1])
def test_gradient_shot(gold, fs):
    g = cs.shot_grid(fs)
    want = gold[f"shot_fs{fs}"]
    o = Oracle("f32")
    start, true = cs.shot_models(g)
    obs = o.run_shot(g, *true, cs.SHOT_SRC, cs.SHOT_REC, cs.SHOT_FINV, want_adjoint=False)["syn"]
    check(want["obs"], np.stack([a[:, 1:] for a in obs]), "observed seismograms")
    res = o.run_shot(g, *start, cs.SHOT_SRC, cs.SHOT_REC, cs.SHOT_FINV, obs=obs)
    assert f32hex([res["l2"]])[0] == want["l2"]
    for n in range(3):
        check(want[f"grad{n + 1}"], g.interior(res["grads"][n]), f"grad{n + 1}")


@pytest.mark.parametrize("fdc,fs", cs.VISCO_CASES)
def test_visco_stress(gold, fdc, fs):
    """update_s + update_s_CPML with one relaxation mechanism, twice"""
    g = cs.step_grid(fdc, fs)
    want = gold[f"visco_fd{fdc}_fs{fs}"]
    o = Oracle("f32")
    f, (rho, pi, u), psi, r6, taus, taup = cs.visco_inputs(g, 50 + 7 * fdc + fs)
    for a in (rho, pi, u, taus, taup):
        o.matcopy_wrap(g, a)
    uipjp, ujpkp, uipkp, _, _, _ = o.av_mat(g, rho, u)
    tsa = o.av_tau(g, rho, u, taus)
    for name, a in zip(("tausipjp", "tausjpkp", "tausipkp"), tsa):
        check(want[name], g.interior(a), name)
    prof, _ = o.cpml_coeff(g, 4500.0, 40.0)
    psv = [p.copy() for p in psi[9:]]
    V = o.visco(r6, taus, taup, *tsa, np.float32(cs.VISCO_ETA))
    box = o.box(g)
    for _ in range(2):
        o.update_s_visco(g, box, f, pi, u, uipjp, ujpkp, uipkp, V)
        o.update_s_cpml_visco(g, box, f, pi, u, uipjp, ujpkp, uipkp, prof, psv, V)
    for n, name in enumerate(("sxx", "syy", "szz", "sxy", "syz", "sxz")):
        check(want[name], g.interior(f[3 + n]), name)
    for n, name in enumerate(("rxx", "ryy", "rzz", "rxy", "ryz", "rxz")):
        check(want[name], g.interior(r6[n]), name)
    for n in range(9):
        check(want[f"psi{n}"], psv[n], f"psi{n}")


@pytest.mark.parametrize("n", range(len(cs.SU_CASES)))
def test_su_io(gold, n):
    """outseis.c (SU) bytes and readseis.c samples, full and short file"""
    from suio_oracle import su_read, su_write
    ntr, ns, rid, src, dx, dy, dz, dt, ndt = cs.SU_CASES[n]
    sec, rxyz = cs.su_inputs(cs.SU_CASES[n], 100 + n)
    want = bytes.fromhex(gold[f"su_{n}"]["bytes"])
    got = su_write(sec, rxyz, dx, dy, dz, dt, ndt, rec_id=rid, src_xyz=src)
    assert len(got) == len(want) == ntr * (240 + 4 * ns)
    bad = [i for i in range(len(want)) if got[i] != want[i]]
    assert not bad, f"SU bytes differ at {bad[:8]}"
    check(gold[f"su_{n}"]["read"], su_read(want, ntr, ns), "readseis")
    check(gold[f"su_{n}"]["read_short"], su_read(want[:gold[f"su_{n}"]["short_len"]], ntr, ns), "short readseis")
