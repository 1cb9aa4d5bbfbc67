"""GPU parity of every kernel against the oracle (compat precision).

Each kernel runs once on seeded random inputs of a small grid, on the device
through the C ABI and on the CPU in the oracle; FP32 results must be
bit-identical (the kernels keep the reference's operation order, no FMA
contraction, IEEE division).  The tolerance stated by the path for
seismograms (rel. L2 1e-5) is reported alongside.
"""
import sys

import numpy as np
import pytest

from helpers import max_ulp, random_fields, random_model, rel_l2, small_grid
from pyoracle import Oracle

pytestmark = pytest.mark.gpu

NAMES = ["vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "syz", "sxz"]


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    import ifosgpu as ig
    assert ig.device_count() > 0, "GPU tests need a device"
    return ig, ig


def _geom(ig, g):
    return ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt, g.fdcoeff, g.free_surf)


def _setup(g, seed, vppml=4000.0, fpml=50.0):
    o = Oracle("f32")
    rng = np.random.default_rng(seed)
    rho, pi, u = random_model(g, rng)
    for a in (rho, pi, u):
        o.matcopy_wrap(g, a)
    avg = o.av_mat(g, rho, u)  # uipjp ujpkp uipkp rjp rkp rip
    prof, _ = o.cpml_coeff(g, vppml, fpml)
    f = random_fields(g, rng)
    psi = [(1e-3 * rng.standard_normal(s)).astype(np.float32) for s in g.psi_shapes()]
    return o, rng, rho, pi, u, avg, prof, f, psi


def _upload(ig, G, rho, pi, u, avg, f, psi, prof):
    dmodel = [G.to_dev(a) for a in (rho, pi, u, avg[0], avg[1], avg[2], avg[3], avg[4], avg[5])]
    df = [G.to_dev(a) for a in f]
    dpsi = [ig.DBuf.from_numpy(p) for p in psi]
    dprof = ig.DBuf.from_numpy(np.concatenate(prof).astype(np.float32))
    return dmodel, df, dpsi, dprof


# nz 150: k rows of 64 cells fill whole wavefronts (k-neighbours through lane shifts) next to a partial row
@pytest.mark.parametrize("nz", [30, 150])
@pytest.mark.parametrize("free_surf", [0, 1])
@pytest.mark.parametrize("fdcoeff", [1, 2])
def test_velocity_update_regions(env, free_surf, fdcoeff, nz):
    _, ig = env
    g = small_grid(fdcoeff=fdcoeff, free_surf=free_surf, nz=nz)
    _note(f"velocity update parity fdcoeff={fdcoeff} free_surf={free_surf} grid={g.nx}x{g.ny}x{g.nz}")
    o, rng, rho, pi, u, avg, prof, f, psi = _setup(g, 11 + fdcoeff + 2 * free_surf)
    G = _geom(ig, g)
    dmodel, df, dpsi, dprof = _upload(ig, G, rho, pi, u, avg, f, psi, prof)
    box = o.box(g)
    # oracle: interior then CPML, no sources
    fo = [a.copy() for a in f]
    po = [p.copy() for p in psi]
    o.update_v(g, box, 1, fo, avg[3], avg[4], avg[5], [], None, None, None, 0)
    o.update_v_cpml(g, box, 1, fo, avg[3], avg[4], avg[5], [], None, None, None, 0, prof, po)
    # device, reference-shaped (two launches)
    ig.update_v(G, ig.REGION_INTERIOR, df, dmodel, dprof, dpsi)
    ig.update_v(G, ig.REGION_CPML, df, dmodel, dprof, dpsi)
    ig.sync()
    for n in range(3):
        got = G.from_dev(df[n])
        assert max_ulp(g.interior(got), g.interior(fo[n])) == 0, NAMES[n]
    for n in range(9):
        assert np.array_equal(dpsi[n].numpy(), po[n]), f"psi_s[{n}]"
    # device, fused single pass
    dmodel2, df2, dpsi2, _ = _upload(ig, G, rho, pi, u, avg, f, psi, prof)
    ig.update_v(G, ig.REGION_ALL, df2, dmodel2, dprof, dpsi2)
    ig.sync()
    for n in range(3):
        assert max_ulp(g.interior(G.from_dev(df2[n])), g.interior(fo[n])) == 0, "fused " + NAMES[n]
    for n in range(9):
        assert np.array_equal(dpsi2[n].numpy(), po[n]), f"fused psi_s[{n}]"


@pytest.mark.parametrize("nz", [29, 150])
@pytest.mark.parametrize("dh", [2.0, 2.5])  # power-of-two spacing: multiply by 1/dh; else IEEE division
@pytest.mark.parametrize("free_surf", [0, 1])
@pytest.mark.parametrize("fdcoeff", [1, 2])
def test_stress_update_regions(env, free_surf, fdcoeff, dh, nz):
    _, ig = env
    g = small_grid(fdcoeff=fdcoeff, free_surf=free_surf, nx=30, ny=27, nz=nz, dh=dh)
    _note(f"stress update parity fdcoeff={fdcoeff} free_surf={free_surf} dh={dh}")
    o, rng, rho, pi, u, avg, prof, f, psi = _setup(g, 21 + fdcoeff + 2 * free_surf)
    G = _geom(ig, g)
    dmodel, df, dpsi, dprof = _upload(ig, G, rho, pi, u, avg, f, psi, prof)
    box = o.box(g)
    fo = [a.copy() for a in f]
    po = [p.copy() for p in psi]
    o.update_s(g, box, fo, pi, u, avg[0], avg[1], avg[2])
    o.update_s_cpml(g, box, fo, pi, u, avg[0], avg[1], avg[2], prof, po)
    ig.update_s(G, ig.REGION_INTERIOR, df, dmodel, dprof, dpsi)
    ig.update_s(G, ig.REGION_CPML, df, dmodel, dprof, dpsi)
    ig.sync()
    for n in range(3, 9):
        assert max_ulp(g.interior(G.from_dev(df[n])), g.interior(fo[n])) == 0, NAMES[n]
    for n in range(9):
        assert np.array_equal(dpsi[n].numpy(), po[n]), f"psi_v[{n}]"
    dmodel2, df2, dpsi2, _ = _upload(ig, G, rho, pi, u, avg, f, psi, prof)
    ig.update_s(G, ig.REGION_ALL, df2, dmodel2, dprof, dpsi2)
    ig.sync()
    for n in range(3, 9):
        assert max_ulp(g.interior(G.from_dev(df2[n])), g.interior(fo[n])) == 0, "fused " + NAMES[n]
    for n in range(9):
        assert np.array_equal(dpsi2[n].numpy(), po[n]), f"fused psi_v[{n}]"


@pytest.mark.parametrize("tiles", [(8, 0, 0, 0, 0, 1, 1, 1), (2, 7, 0, 0, 1, 1, 1, 0), (8, 16, 0, 0, 0, 2, 1, 0)])
def test_plain_tile_shapes_match_oracle(env, tiles):
    """the one-shot updates with other tile shapes (ifos_set_plain_tiles:
    LDS and register marches, several rows or k columns per workgroup)
    against the oracle, as the default shape above"""
    _, ig = env
    old = ig.get_plain_tiles()
    _note(f"plain tile shape {tiles}")
    ig.set_plain_tiles(tiles)
    try:
        for nz in (30, 150):
            test_velocity_update_regions(env, 1, 2, nz)
            test_stress_update_regions(env, 0, 2, 2.0, nz - 1)
    finally:
        ig.set_plain_tiles(old)


@pytest.mark.parametrize("free_surf,fdcoeff,nz,L", [(0, 2, 29, 1), (1, 1, 29, 1), (0, 2, 150, 1), (0, 2, 29, 3),
                                                     (1, 1, 29, 2), (0, 2, 150, 3), (0, 1, 70, 4)])
def test_visco_stress_update_regions(env, free_surf, fdcoeff, nz, L):
    """update_s / update_s_CPML with L relaxation mechanisms (L = 1: the
    reference's), two steps so the memory variables carry over; interior+CPML
    launches and the fused one"""
    _, ig = env
    g = small_grid(fdcoeff=fdcoeff, free_surf=free_surf, nx=30, ny=27, nz=nz)
    _note(f"viscoelastic stress update parity fdcoeff={fdcoeff} free_surf={free_surf} L={L}")
    o, rng, rho, pi, u, avg, prof, f, psi = _setup(g, 31 + fdcoeff + 2 * free_surf)
    taus = (0.01 + 0.05 * rng.random(g.shape)).astype(np.float32)
    taup = (0.01 + 0.05 * rng.random(g.shape)).astype(np.float32)
    for a in (taus, taup):
        o.matcopy_wrap(g, a)
    tsa = o.av_tau(g, rho, u, taus)
    r6 = [(1e-2 * rng.standard_normal((L,) + g.shape)).astype(np.float32) for _ in range(6)]
    eta = np.array([0.0137, 0.0051, 0.031, 0.0089][:L], np.float32)
    G = _geom(ig, g)
    # averages on the device
    dtaus = G.to_dev(taus)
    dtsa = [G.empty() for _ in range(3)]
    ig.av_tau(G, dtaus, dtsa)
    ig.sync()
    for n in range(3):
        assert max_ulp(g.interior(G.from_dev(dtsa[n])), g.interior(tsa[n])) == 0, f"tau average {n}"
    box = o.box(g)
    fo = [a.copy() for a in f]
    po = [p.copy() for p in psi]
    ro = [a.copy() for a in r6]
    V = o.visco(ro, taus, taup, *tsa, eta)
    for _ in range(2):
        o.update_s_visco(g, box, fo, pi, u, avg[0], avg[1], avg[2], V)
        o.update_s_cpml_visco(g, box, fo, pi, u, avg[0], avg[1], avg[2], prof, po, V)
    dtaup = G.to_dev(taup)
    for fused in (False, True):
        dmodel, df, dpsi, dprof = _upload(ig, G, rho, pi, u, avg, f, psi, prof)
        dr = [G.to_dev(a) for a in r6]
        vs = ig.visco_struct(dr, dtaus, dtaup, *dtsa, eta)
        for _ in range(2):
            if fused:
                ig.update_s_visco(G, ig.REGION_ALL, df, dmodel, vs, dprof, dpsi)
            else:
                ig.update_s_visco(G, ig.REGION_INTERIOR, df, dmodel, vs, dprof, dpsi)
                ig.update_s_visco(G, ig.REGION_CPML, df, dmodel, vs, dprof, dpsi)
        ig.sync()
        tag = "fused " if fused else ""
        for n in range(3, 9):
            assert max_ulp(g.interior(G.from_dev(df[n])), g.interior(fo[n])) == 0, tag + NAMES[n]
        for n in range(6):
            assert max_ulp(g.interior(G.from_dev(dr[n])), g.interior(ro[n])) == 0, f"{tag}r[{n}]"
        for n in range(9):
            assert np.array_equal(dpsi[n].numpy(), po[n]), f"{tag}psi_v[{n}]"


def test_sources_and_receivers(env):
    _, ig = env
    g = small_grid()
    _note("sources, adjoint injection, seismo")
    o, rng, rho, pi, u, avg, prof, f, psi = _setup(g, 5)
    G = _geom(ig, g)
    dmodel, df, dpsi, dprof = _upload(ig, G, rho, pi, u, avg, f, psi, prof)
    # force sources of every type, two sharing a cell, and an explosive one
    srcs = [(5, 7, 9, 0.0, 50.0, 1.0, 2), (5, 7, 9, 0.001, 40.0, 2.0, 3), (20, 8, 11, 0.0, 60.0, 1.5, 4),
            (12, 14, 15, 0.0, 45.0, 1.0, 1)]
    sig = o.wavelet(g, srcs, 1)
    nt = 9
    fo = [a.copy() for a in f]
    box = o.box(g)
    # forces alone: use a zero-region update by comparing only the injection delta
    # oracle force injection happens inside og_update_v after the interior loop; emulate on a
    # copy whose stresses are zero so the interior loop adds exactly 0
    z = [np.zeros_like(a) for a in fo]
    zero_f = fo[:3] + z[3:]
    o.update_v(g, box, nt, zero_f, avg[3], avg[4], avg[5], srcs, sig, sig, sig, 0)
    o.psource(g, nt, fo[3], fo[4], fo[5], srcs, sig)
    dsrc = ig.DBuf.from_numpy(np.array([[s[0], s[1], s[2], s[6]] for s in srcs], np.int32))
    dsig = ig.DBuf.from_numpy(sig)
    ig.add_forces(G, df, dmodel, dsrc, dsig, len(srcs), nt)
    ig.psource(G, df, dsrc, dsig, len(srcs), nt)
    ig.sync()
    for n in range(6):
        ref = zero_f[n] if n < 3 else fo[n]
        assert max_ulp(g.interior(G.from_dev(df[n])), g.interior(ref)) == 0, NAMES[n]
    # adjoint injection: receivers with traces, two on one cell
    rec = np.array([[4, 5, 6], [4, 5, 6], [22, 20, 25]], np.int32)
    adj = [(1e-3 * rng.standard_normal((3, g.nt + 2))).astype(np.float32) for _ in range(3)]
    fo2 = [a.copy() for a in f]
    recsrc = [(int(r[0]), int(r[1]), int(r[2]), 0.0, 0.0, 0.0, 0) for r in rec]
    zero2 = fo2[:3] + [np.zeros_like(a) for a in fo2[3:]]
    o.update_v(g, box, nt, zero2, avg[3], avg[4], avg[5], recsrc, adj[0], adj[1], adj[2], 1)
    dmodel, df, dpsi, dprof = _upload(ig, G, rho, pi, u, avg, f, psi, prof)
    drec = ig.DBuf.from_numpy(rec)
    dadj = ig.DBuf.from_numpy(np.stack(adj))
    ig.add_adjoint(G, df, drec, dadj, len(rec), nt)
    ig.sync()
    for n in range(3):
        assert max_ulp(g.interior(G.from_dev(df[n])), g.interior(zero2[n])) == 0, "adjoint " + NAMES[n]
    # same with the precomputed same-cell chain
    dmodel_c, df_c, _, _ = _upload(ig, G, rho, pi, u, avg, f, psi, prof)
    dchain = ig.DBuf.from_numpy(ig.rec_chain(rec))
    ig.add_adjoint(G, df_c, drec, dadj, len(rec), nt, d_chain=dchain)
    ig.sync()
    for n in range(3):
        assert max_ulp(g.interior(G.from_dev(df_c[n])), g.interior(zero2[n])) == 0, "chained adjoint " + NAMES[n]
    # seismo
    ns = g.nt
    dsec = ig.DBuf((3, len(rec), ns + 1))
    ig.seismo(G, df, drec, len(rec), 7, ns, dsec)
    ig.sync()
    sec = dsec.numpy()
    sx = np.zeros((len(rec), ns + 1), np.float32)
    sy, sz = sx.copy(), sx.copy()
    o.lib.og_seismo  # noqa: B018 (presence)
    import ctypes as C
    og = g.og()
    got = [G.from_dev(df[n]) for n in range(3)]
    o.lib.og_seismo(C.byref(og), 7, len(rec), rec.ctypes.data_as(C.c_void_p), sx.ctypes.data_as(C.c_void_p),
                    sy.ctypes.data_as(C.c_void_p), sz.ctypes.data_as(C.c_void_p), ns,
                    got[0].ctypes.data_as(C.c_void_p), got[1].ctypes.data_as(C.c_void_p),
                    got[2].ctypes.data_as(C.c_void_p))
    assert np.array_equal(sec[0], sx) and np.array_equal(sec[1], sy) and np.array_equal(sec[2], sz)


@pytest.mark.parametrize("nf", [3, 9])
def test_dft_compat_and_kahan(env, nf):
    _, ig = env
    g = small_grid()
    _note(f"discfourier compat and Kahan nf={nf}")
    o = Oracle("f32")
    rng = np.random.default_rng(8)
    f = random_fields(g, rng, n=3)
    finv = [40.0, 60.0, 80.0, 20.0, 30.0, 50.0, 70.0, 90.0, 100.0][:nf]
    nf = len(finv)
    Fr = [(rng.standard_normal((nf,) + g.shape)).astype(np.float32) for _ in range(3)]
    Fi = [(rng.standard_normal((nf,) + g.shape)).astype(np.float32) for _ in range(3)]
    G = _geom(ig, g)
    dv = [G.to_dev(a) for a in f] + [G.empty() for _ in range(6)]
    dF = [G.to_dev(a) for a in Fr + Fi]
    for back in (0, 1):
        nt = 17 + back
        o.discfourier(g, nt, f[0], f[1], f[2], Fr, Fi, finv, 2, back)
        ig.discfourier(G, dv, dF, None, finv, nt, 2, back)
        ig.sync()
        for n, ref in enumerate(Fr + Fi):
            got = G.from_dev(dF[n])
            assert max_ulp(got[:, 3:3 + g.ny, 3:3 + g.nx, 3:3 + g.nz],
                           ref[:, 3:3 + g.ny, 3:3 + g.nx, 3:3 + g.nz]) == 0, f"back={back} F[{n}]"
    # Kahan: 400 accumulation steps of a smooth signal against the FP64 oracle
    o64 = Oracle("f64")
    f64 = [a.astype(np.float64) for a in f]
    R64 = [np.zeros((nf,) + g.shape) for _ in range(3)]
    I64 = [np.zeros((nf,) + g.shape) for _ in range(3)]
    R32 = [np.zeros((nf,) + g.shape, np.float32) for _ in range(3)]
    I32 = [np.zeros((nf,) + g.shape, np.float32) for _ in range(3)]
    dS = [G.empty(nf) for _ in range(6)]
    dC = [G.empty(nf) for _ in range(6)]
    for nt in range(1, 401):
        o64.discfourier(g, nt, f64[0], f64[1], f64[2], R64, I64, finv, 1, 0)
        o.discfourier(g, nt, f[0], f[1], f[2], R32, I32, finv, 1, 0)
        ig.discfourier(G, dv, dS, dC, finv, nt, 1, 0, precision=ig.DFT_KAHAN)
    ig.sync()
    ek, ec = [], []
    for n, (ref, c32) in enumerate(zip(R64 + I64, R32 + I32)):
        got = G.from_dev(dS[n])
        sl = (slice(None), slice(3, 3 + g.ny), slice(3, 3 + g.nx), slice(3, 3 + g.nz))
        ek.append(rel_l2(got[sl], ref[sl]))
        ec.append(rel_l2(c32[sl], ref[sl]))
    _note(f"DFT rel-L2 vs FP64 after 400 steps: kahan {max(ek):.2e}  compat {max(ec):.2e}")
    assert max(ek) < 1e-6
    assert max(ek) <= max(ec)


@pytest.mark.parametrize("dh,ntr", [(2.0, 2), (3.0, 3)])
def test_hess_F(env, dh, ntr):
    """hess_F.c: forward fields [nf], receiver-side fields [nf*(ntr+1)]"""
    _, ig = env
    g = small_grid(nx=22, ny=20, nz=24, dh=dh)
    _note(f"hess_F parity dh={dh} ntr_hess={ntr}")
    o = Oracle("f32")
    rng = np.random.default_rng(19)
    rho, pi, u = random_model(g, rng)
    finv = [40.0, 60.0]
    nf = len(finv)
    Ff = [rng.standard_normal((nf,) + g.shape).astype(np.float32) for _ in range(6)]
    Fb = [rng.standard_normal((nf * (ntr + 1),) + g.shape).astype(np.float32) for _ in range(6)]
    for a in Ff + Fb:  # zero halos (zero_invers.c:43-79)
        a[:, :3] = 0; a[:, 3 + g.ny:] = 0; a[:, :, :3] = 0; a[:, :, 3 + g.nx:] = 0
        a[:, :, :, :3] = 0; a[:, :, :, 3 + g.nz:] = 0
    hess = [rng.random(g.shape).astype(np.float32) for _ in range(3)]
    ho = [a.copy() for a in hess]
    o.hess_F(g, Ff, Fb, ho, rho, pi, u, finv, ntr)
    G = _geom(ig, g)
    dFf = [G.to_dev(a) for a in Ff]
    dFb = [G.to_dev(a) for a in Fb]
    dm = [G.to_dev(a) for a in (rho, pi, u)] + [G.empty() for _ in range(6)]
    dh_ = [G.to_dev(a) for a in hess]
    ig.hess_F(G, dFf, dFb, ntr, dm, finv, dh_)
    ig.sync()
    for n in range(3):
        got = g.interior(G.from_dev(dh_[n]))
        assert max_ulp(got, g.interior(ho[n])) == 0, f"hess{n + 1}"


@pytest.mark.parametrize("dh", [2.0, 3.0])
def test_gradient_F(env, dh):
    _, ig = env
    g = small_grid(nx=26, ny=24, nz=28, dh=dh)
    _note(f"gradient_F parity dh={dh}")
    o = Oracle("f32")
    rng = np.random.default_rng(9)
    rho, pi, u = random_model(g, rng)
    finv = [40.0, 60.0, 80.0]
    nf = len(finv)
    Ff = [rng.standard_normal((nf,) + g.shape).astype(np.float32) for _ in range(6)]
    Fb = [rng.standard_normal((nf,) + g.shape).astype(np.float32) for _ in range(6)]
    # the reference's DFT fields have zero halos (zero_invers.c:43-79)
    for a in Ff + Fb:
        a[:, :3] = 0; a[:, 3 + g.ny:] = 0; a[:, :, :3] = 0; a[:, :, 3 + g.nx:] = 0
        a[:, :, :, :3] = 0; a[:, :, :, 3 + g.nz:] = 0
    grads = [rng.standard_normal(g.shape).astype(np.float32) for _ in range(3)]
    go = [a.copy() for a in grads]
    o.gradient_F(g, Ff, Fb, go, rho, pi, u, finv)
    G = _geom(ig, g)
    dFf = [G.to_dev(a) for a in Ff]
    dFb = [G.to_dev(a) for a in Fb]
    dm = [G.to_dev(a) for a in (rho, pi, u)] + [G.empty() for _ in range(6)]
    dg = [G.to_dev(a) for a in grads]
    ig.gradient_F(G, dFf, dFb, dm, finv, dg)
    ig.sync()
    for n in range(3):
        got = g.interior(G.from_dev(dg[n]))
        ref = g.interior(go[n])
        _note(f"grad{n + 1}: max ulp {max_ulp(got, ref)} rel-L2 {rel_l2(got, ref):.2e}")
        assert max_ulp(got, ref) == 0


def test_model_preparation(env):
    _, ig = env
    g = small_grid(nx=25, ny=23, nz=27)
    _note("matcopy wrap and av_mat")
    o = Oracle("f32")
    rng = np.random.default_rng(4)
    rho, pi, u = random_model(g, rng)
    G = _geom(ig, g)
    drho, du = G.to_dev(rho), G.to_dev(u)
    o.matcopy_wrap(g, rho)
    o.matcopy_wrap(g, u)
    ig.matcopy_wrap(G, drho)
    ig.matcopy_wrap(G, du)
    ig.sync()
    inner = (slice(2, 3 + g.ny + 1), slice(2, 3 + g.nx + 1), slice(2, 3 + g.nz + 1))
    assert np.array_equal(G.from_dev(drho)[inner], rho[inner])
    avg = o.av_mat(g, rho, u)
    out = [G.empty() for _ in range(6)]
    ig.av_mat(G, drho, du, out)
    ig.sync()
    for n in range(6):
        assert np.array_equal(g.interior(G.from_dev(out[n])), g.interior(avg[n])), n


def test_residual(env):
    _, ig = env
    g = small_grid(nt=80)
    _note("residual")
    o = Oracle("f32")
    rng = np.random.default_rng(6)
    ntr, ns = 5, g.nt
    obs = [rng.standard_normal((ntr, ns + 1)).astype(np.float32) for _ in range(3)]
    syn = [rng.standard_normal((ntr, ns + 1)).astype(np.float32) for _ in range(3)]
    l2o = 0.0
    diffs = []
    for c in range(3):
        d, l2 = o.residual(g, obs[c], syn[c])
        diffs.append(d)
        l2o += l2
    G = _geom(ig, g)
    dobs = ig.DBuf.from_numpy(np.stack(obs))
    dsyn = ig.DBuf.from_numpy(np.stack(syn))
    dadj = ig.DBuf((3, ntr, ns + 2))
    l2 = ig.residual(G, dobs, dsyn, dadj, ntr, ns)
    adj = dadj.numpy()
    for c in range(3):
        assert np.array_equal(adj[c][:, 1:ns + 1], diffs[c][:, 1:ns + 1])
    # the kernel sums each float square in double
    l2_64 = sum(float(np.sum((diffs[c][:, 1:] * diffs[c][:, 1:]).astype(np.float64))) for c in range(3))
    assert abs(l2 - l2_64) <= 1e-12 * abs(l2_64) + 1e-300
    assert abs(l2 - l2o) <= 1e-5 * abs(l2o)
