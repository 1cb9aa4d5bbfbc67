"""Mathematical invariants of the oracle's time step, independent of the
golden vectors (the reference ships no tests; these fix what any correct
4th-order staggered-grid scheme must do).

* Mirror symmetry.  With a model that is mirror-symmetric about a cell
  centre c and an explosive source at c, the exact update keeps every field
  symmetric (centred components) or antisymmetric (the component staggered
  along the mirror axis) about c, and in IEEE arithmetic it does so bit for
  bit: a mirrored difference is the exact negative of the original one, and
  sums, products and quotients of mirrored operands are the same operations
  on the same values (update_v_ssg.c:106-125, update_s_ssg_elastic.c:106-138
  use only such operations).  A wrong index, a forward difference where a
  backward one belongs, a dropped or sign-flipped term breaks it.  It holds
  until the fields' support (3 cells per step) reaches the CPML, whose
  profiles the reference mirrors asymmetrically (CPML_coeff.c:150-190).  The
  model's averages sum four neighbours in an order the mirror reverses
  (av_mat.c:51-60), so the test models use values whose sums are exact in
  any order (helpers.symmetric_table_model).
* Energy.  With zero halos and no CPML the velocity step uses the
  difference operator D and the stress step E = -D^T (forward and backward
  4th-order differences are negative adjoints under zero extension), so the
  leapfrog conserves H^n = 1/2 <S sigma^n, sigma^n> + 1/2 <M v^n, v^(n+1)>
  exactly (S the compliance, M the averaged densities); in FP64 to
  round-off.  A non-adjoint pair (a coefficient, a sign, a stagger or the
  update box wrong) makes H drift.
"""
import numpy as np
import pytest

from helpers import (FIELD_STAGGER, homogeneous_model, mirror_mismatch, random_model, shifted_mismatch,
                     symmetric_table_field, symmetric_table_model)
from pyoracle import Grid, Oracle


def _mirror_failures(hist, centre):
    """(step, component, axis, cells) of every broken mirror relation in a
    velocity history [nt][3][ny+5][nx+5][nz+5]"""
    bad = []
    for c, name in enumerate(("vx", "vy", "vz")):
        for ax in range(3):
            n = mirror_mismatch(hist[:, c], ax, centre[ax], ax in FIELD_STAGGER[name])
            if n:
                bad.append((name, ax, n))
    return bad


def _shot(o, g, centre, src_cell, seed=7, L=0):
    rho, pi, u = [a.astype(o.dtype) for a in symmetric_table_model(g.shape, centre, seed)]
    visco = None
    if L:
        taus, taup = [symmetric_table_field(g.shape, centre, seed + n, 0.01, 0.08).astype(o.dtype)
                      for n in (1, 2)]
        visco = (taus, taup, [g.dt / t for t in (0.02, 0.005, 0.001)[:L]])
    fc = 50.0
    src = (src_cell[1], src_cell[0], src_cell[2], -1.5 / fc, fc, 1.0e6, 1)  # (i, j, k, ..): peak at t = 0
    rec = np.array([[centre[1], centre[0], centre[2]]], np.int32)
    out = o.run_shot(g, rho, pi, u, src, rec, [fc], want_adjoint=False, want_history=True, visco=visco)
    return out["vfwd"]


@pytest.mark.parametrize("prec,L", [("f32", 0), ("f64", 0), ("f32", 3)])
def test_mirror_symmetry_of_a_centred_shot(prec, L):
    """61^3 with a 6-cell CPML, source at the centre cell 31: 9 steps (the
    velocities' support grows by 3 cells a step and reaches 8..53 at step 9;
    at step 10 it enters the CPML, 25 cells out, and the symmetry ends there),
    every velocity of every step mirror-(anti)symmetric in x, y and z;
    elastic, and viscoelastic with three relaxation mechanisms"""
    g = Grid(nx=61, ny=61, nz=61, fw=6, dx=2.0, dy=2.0, dz=2.0, dt=2e-4, nt=9, fdcoeff=2)
    c = (31, 31, 31)
    hist = _shot(Oracle(prec), g, c, c, L=L)
    assert np.count_nonzero(hist[-1]) > 30000
    assert _mirror_failures(hist, c) == []


def test_mirror_symmetry_detects_an_offset_source():
    """the same check with the source one cell off centre along x: the x
    mirror breaks, y and z still hold"""
    g = Grid(nx=61, ny=61, nz=61, fw=6, dx=2.0, dy=2.0, dz=2.0, dt=2e-4, nt=9, fdcoeff=2)
    c = (31, 31, 31)
    hist = _shot(Oracle("f32"), g, c, (31, 32, 31))
    bad = _mirror_failures(hist, c)
    assert bad and all(ax == 1 for _, ax, _ in bad)


def test_translation_invariance_in_a_homogeneous_model():
    """the same shot moved by (5, 7, 3) cells gives the same velocities
    moved by as much, bit for bit, at every step until the support nears the
    CPML (61^3, 6-cell CPML, 7 steps)"""
    g = Grid(nx=61, ny=61, nz=61, fw=6, dx=2.0, dy=2.0, dz=2.0, dt=2e-4, nt=7, fdcoeff=2)
    o = Oracle("f32")
    model = homogeneous_model(g.shape)
    fc = 50.0
    rec = np.array([[31, 31, 31]], np.int32)
    h = [o.run_shot(g, *model, (i, j, k, -1.5 / fc, fc, 1.0e6, 1), rec, [fc], want_adjoint=False,
                    want_history=True)["vfwd"] for (i, j, k) in ((28, 27, 30), (33, 34, 33))]
    # the support after 7 steps spans 21 cells round each source; the box
    # below covers it with room for the shift (canonical index = cell + 2)
    lo, hi = (27 + 2 - 12,) * 3, (27 + 2 + 12,) * 3
    assert np.count_nonzero(h[0][-1]) > 10000
    assert shifted_mismatch(h[0], h[1], (7, 5, 3), lo, hi) == 0
    assert shifted_mismatch(h[0], h[1], (7, 5, 2), lo, hi) > 0  # the check sees a wrong shift


def _energy(f, v_next, m):
    """H = 1/2 <S sigma, sigma> + 1/2 <M v, v_next> in FP64 over the grid"""
    vx, vy, vz, sxx, syy, szz, sxy, syz, sxz = f
    pi, u, uipjp, ujpkp, uipkp, rjp, rkp, rip = m
    inner = (slice(3, -2),) * 3  # [1..N]^3 (canonical index pos + 2, N + 5 per axis)
    pi, u, uipjp, ujpkp, uipkp, rjp, rkp, rip = [a[inner] for a in (pi, u, uipjp, ujpkp, uipkp, rjp, rkp, rip)]
    lam = pi - 2.0 * u
    s1, s2, s3 = sxx[inner], syy[inner], szz[inner]
    # compliance of the normal block 2u I + lam J: (I - lam/(2u + 3 lam) J) / (2u)
    normal = (s1 * s1 + s2 * s2 + s3 * s3 - lam / (2.0 * u + 3.0 * lam) * (s1 + s2 + s3) ** 2) / (2.0 * u)
    shear = sxy[inner] ** 2 / uipjp + syz[inner] ** 2 / ujpkp + sxz[inner] ** 2 / uipkp
    kin = rip * vx[inner] * v_next[0][inner] + rjp * vy[inner] * v_next[1][inner] + rkp * vz[inner] * v_next[2][inner]
    return 0.5 * float(np.sum(normal + shear)) + 0.5 * float(np.sum(kin))


def _energy_run(vbox_trim=0, steps=200):
    """the FP64 oracle's interior updates over [1..N]^3 (zero halos, no CPML,
    no source) from random fields; H at every step.  vbox_trim > 0 plants a
    mistake: the velocity update leaves out the last vbox_trim x-rows"""
    g = Grid(nx=20, ny=22, nz=24, fw=4, dx=2.0, dy=2.5, dz=3.0, dt=2e-4, nt=steps, fdcoeff=2)
    o = Oracle("f64")
    rng = np.random.default_rng(11)
    rho, pi, u = [a.astype(np.float64) for a in random_model(g, rng)]
    uipjp, ujpkp, uipkp, rjp, rkp, rip = o.av_mat(g, rho, u)
    m = (pi, u, uipjp, ujpkp, uipkp, rjp, rkp, rip)
    f = [o.zeros(g) for _ in range(9)]
    inner = (slice(3, -2),) * 3
    for n, a in enumerate(f):
        a[inner] = rng.uniform(-1.0, 1.0, a[inner].shape) * (1.0 if n < 3 else 1.0e7)
    box = [1, g.nx, 1, g.ny, 1, g.nz]  # i0 i1 j0 j1 k0 k1
    vbox = [1, g.nx - vbox_trim, 1, g.ny, 1, g.nz]
    H = []
    for nt in range(1, steps + 1):
        v_old = [a.copy() for a in f[:3]]
        o.update_v(g, vbox, nt, f, rjp, rkp, rip, [], None, None, None, 0)
        # sigma^(n-1) (not yet updated this step) with v^(n-1) and v^n
        H.append(_energy(v_old + f[3:], f[:3], m))
        o.update_s(g, box, f, pi, u, uipjp, ujpkp, uipkp)
    return np.array(H), f


def test_energy_is_conserved_by_the_fp64_step():
    """200 steps of a heterogeneous model with unequal spacings: H constant to
    1e-12 relative, while the fields themselves change at O(1)"""
    H, f = _energy_run()
    assert H[0] > 0
    drift = np.abs(H - H[0]).max() / abs(H[0])
    assert drift < 1e-12, drift


def test_energy_drift_detects_a_non_adjoint_update():
    """planted mistake: the velocity update skips the last x-row, so the
    velocity and stress operators are no longer negative adjoints and H
    drifts by orders of magnitude more than round-off"""
    H, _ = _energy_run(vbox_trim=1)
    drift = np.abs(H - H[0]).max() / abs(H[0])
    assert drift > 1e-6, drift
