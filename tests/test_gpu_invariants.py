"""Mirror symmetry of the HIP path at the workloads' full shapes: a property
that does not depend on the grid size, so it checks every cell of grids the
oracle is too slow for.

With a model mirror-symmetric about a cell centre c (helpers.
symmetric_table_model: averages exact in any order) and explosive sources
at c, every field of the exact update is symmetric about c, or
antisymmetric for the component staggered along the mirror axis, bit for
bit (the argument is in test_oracle_invariants.py), until the fields'
support, growing 3 cells a step, reaches the CPML.  Each case runs just
short of that:

* 276^3 (the gradient workload's grid, 10-cell CPML), elastic, 40 steps;
* 404 x 212 x 404 (the viscoelastic case), three relaxation mechanisms, 32
  steps (the support reaches 14..197 of the 212 planes in y);
* 536 x 280 x 536 (the forward-only case, 12-cell CPML), elastic, 42 steps.

The plain path is checked on all nine fields of the whole grid after the
last step, the batched path (two shots, different wavelets, several tile
shapes, direct launches and hipGraph replay) on the seismograms of a cube of
receivers round the source at every step.

Translation: in a homogeneous model the same shot moved by (5, 7, 3) cells
gives the same fields moved by as much, bit for bit, until the support
nears the CPML.  The offset is no multiple of any tile size, so the two
wavefields cross the tiles', LDS rows' and k-halos' edges at different
places; the batched case runs both positions as two shots of one launch.
"""
import sys

import numpy as np
import pytest

from helpers import (FIELD_NAMES, FIELD_STAGGER, homogeneous_model, mirror_mismatch, symmetric_table_field,
                     symmetric_table_model)

pytestmark = pytest.mark.gpu

CASES = {
    # name: (nx, ny, nz, fw, dh, dt, nt, L)
    "grad276": (276, 276, 276, 10, 4.0, 3.19e-4, 40, 0),
    "visco404": (404, 212, 404, 10, 5.0, 3.72e-4, 32, 3),
    "fwd536": (536, 280, 536, 12, 5.0, 3.99e-4, 42, 0),
}
ETAS = (0.02, 0.005, 0.001)  # relaxation times of the three mechanisms (s)


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup(ig, name):
    nx, ny, nz, fw, dh, dt, nt, L = CASES[name]
    G = ig.Geom(nx, ny, nz, fw, dh, dh, dh, dt, nt, 2, 0)
    shape = (ny + 5, nx + 5, nz + 5)
    c = (ny // 2, nx // 2, nz // 2)  # (y, x, z) centre cell
    model = symmetric_table_model(shape, c, 21)
    d_model = [G.to_dev(a) for a in model]
    del model
    visco = None
    if L:
        tau = [G.to_dev(symmetric_table_field(shape, c, 22 + n, 0.01, 0.08)) for n in range(2)]
        visco = (tau[0], tau[1], [dt / t for t in ETAS[:L]])
    return G, c, d_model, visco


def _cube(c, R):
    """receivers on the cells c-R-1 .. c+R of each axis (every mirror partner
    of a staggered or centred component inside), as (i, j, k) rows"""
    ax = [np.arange(cc - R - 1, cc + R + 1) for cc in c]
    jj, ii, kk = np.meshgrid(*ax, indexing="ij")
    return np.stack([ii.ravel(), jj.ravel(), kk.ravel()], 1).astype(np.int32)


def _seismogram_mismatches(syn, rec, c):
    """broken mirror relations among the receivers' traces syn[3][ntr][ns]"""
    where = {tuple(r): n for n, r in enumerate(rec.tolist())}
    bad = []
    for comp, name in enumerate(("vx", "vy", "vz")):
        for ax in range(3):  # 0 y, 1 x, 2 z as in the canonical arrays
            stag = ax in FIELD_STAGGER[name]
            col = (1, 0, 2)[ax]  # column of that axis in an (i, j, k) row
            s = 2 * c[ax] - (1 if stag else 0)
            pairs = []
            for n, r in enumerate(rec.tolist()):
                m = list(r)
                m[col] = s - r[col]
                p = where.get(tuple(m))
                if p is not None:
                    pairs.append((n, p))
            a, b = np.array(pairs).T
            ok = syn[comp][a] == (-syn[comp][b] if stag else syn[comp][b])
            if not ok.all():
                bad.append((name, ax, int((~ok).sum())))
    return bad


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name", list(CASES))
def test_plain_path_fields_are_mirror_symmetric(name):
    import ifosgpu as ig
    G, c, d_model, visco = _setup(ig, name)
    nt = CASES[name][6]
    rec = _cube(c, 4)
    fc = 15.0
    src = (c[1], c[0], c[2], -1.5 / fc, fc, 1.0e6, 1)  # (i, j, k, tshift, fc, amp, type): peak at t = 0
    _note(f"plain path, {name}, {nt} steps, source at (i, j, k) = ({c[1]}, {c[0]}, {c[2]})")
    ctx = ig.ShotContext(G, 1, len(rec))
    ctx.set_cpml(5000.0, fc)
    ctx.set_freqs([fc], 100)
    if visco:
        ctx.set_visco(*visco)
    ctx.set_model(*d_model)
    d_syn = ig.DBuf((3, len(rec), nt + 1))
    ctx.run_shot(src, rec, d_syn, None, want_adjoint=False)
    ig.sync()
    f = ctx.fields()
    bad = []
    for fname in FIELD_NAMES:
        a = G.from_dev(ig.DBuf.wrap(getattr(f, fname), G.dshape))
        assert np.count_nonzero(a) > 100000, fname
        for ax in range(3):
            n = mirror_mismatch(a, ax, c[ax], ax in FIELD_STAGGER[fname])
            if n:
                bad.append((fname, ax, n))
        del a
    assert bad == [], bad
    assert _seismogram_mismatches(d_syn.numpy(), rec, c) == []
    ctx.close()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name,tiles,graph", [
    ("grad276", (4, 32, 0, 0, 0, 1, 1, 1), False),   # LDS marches, shots slowest
    ("grad276", (4, 0, 1, 0, 1, 1, 1, 0), True),     # register marches, shots fastest, pipelined, graph
    ("visco404", (4, 32, 0, 0, 0, 1, 1, 1), False),
])
def test_batched_path_seismograms_are_mirror_symmetric(name, tiles, graph):
    import ifosgpu as ig
    G, c, d_model, visco = _setup(ig, name)
    nt = CASES[name][6]
    rec = _cube(c, 6)
    srcs = [(c[1], c[0], c[2], -1.5 / fc, fc, amp, 1) for fc, amp in ((15.0, 1.0e6), (22.0, 3.0e5))]
    _note(f"batched path, {name}, tiles {tiles}, graph {graph}, {nt} steps, {len(rec)} receivers")
    b = ig.BatchContext(G, 2, 1, len(rec), ig.DFT_COMPAT, 4)
    b.set_cpml(5000.0, 15.0)
    b.set_freqs([15.0], 100)
    if visco:
        b.set_visco(*visco)
    b.set_model(*d_model)
    b.set_tiles(tiles, tiles)
    d_b = ig.DBuf((2, 3, len(rec), nt + 1))
    b.run(srcs, rec, d_syn=d_b, want_adjoint=False, use_graph=graph)
    syn = d_b.numpy()
    b.close()
    for s in range(2):
        assert np.count_nonzero(syn[s]) > 0.5 * syn[s][:, :, 2:].size, s
        assert _seismogram_mismatches(syn[s], rec, c) == [], s


SHIFT = (5, 7, 3)  # (i, j, k)
TCASES = {
    # name: (nx, ny, nz, fw, dh, dt, nt, L, first source (i, j, k))
    "grad276": (276, 276, 276, 10, 4.0, 3.19e-4, 35, 0, (130, 128, 131)),
    "visco404": (404, 212, 404, 10, 5.0, 3.72e-4, 28, 3, (200, 100, 199)),
}


def _setup_homogeneous(ig, name):
    nx, ny, nz, fw, dh, dt, nt, L, p0 = TCASES[name]
    G = ig.Geom(nx, ny, nz, fw, dh, dh, dh, dt, nt, 2, 0)
    shape = (ny + 5, nx + 5, nz + 5)
    d_model = [G.to_dev(a) for a in homogeneous_model(shape)]
    visco = None
    if L:
        tau = [G.to_dev(np.full(shape, t, np.float32)) for t in (0.05, 0.04)]
        visco = (tau[0], tau[1], [dt / t for t in ETAS[:L]])
    p1 = tuple(a + b for a, b in zip(p0, SHIFT))
    return G, d_model, visco, nt, p0, p1


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name", list(TCASES))
def test_plain_path_translation_invariance(name):
    import ifosgpu as ig
    G, d_model, visco, nt, p0, p1 = _setup_homogeneous(ig, name)
    fc = 15.0
    rec = np.array([p0], np.int32)
    ctxs = []
    for p in (p0, p1):
        _note(f"plain path, {name}, {nt} steps, source at (i, j, k) = {p}")
        ctx = ig.ShotContext(G, 1, 1)
        ctx.set_cpml(5000.0, fc)
        ctx.set_freqs([fc], 100)
        if visco:
            ctx.set_visco(*visco)
        ctx.set_model(*d_model)
        ctx.run_shot((p[0], p[1], p[2], -1.5 / fc, fc, 1.0e6, 1), rec, ig.DBuf((3, 1, nt + 1)), None,
                     want_adjoint=False)
        ctxs.append(ctx)
    ig.sync()
    f0, f1 = ctxs[0].fields(), ctxs[1].fields()
    sh = (SHIFT[1], SHIFT[0], SHIFT[2])  # canonical axes y, x, z
    bad = []
    for fname in FIELD_NAMES:
        a = G.from_dev(ig.DBuf.wrap(getattr(f0, fname), G.dshape))
        b = G.from_dev(ig.DBuf.wrap(getattr(f1, fname), G.dshape))
        assert np.count_nonzero(a) > 100000, fname
        part = tuple(slice(0, n - s) for n, s in zip(a.shape, sh))
        moved = tuple(slice(s, n) for n, s in zip(a.shape, sh))
        n = int(np.count_nonzero(a[part] != b[moved]))
        if n:
            bad.append((fname, n))
    for ctx in ctxs:
        ctx.close()
    assert bad == [], bad


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name,tiles,graph", [
    ("grad276", (4, 32, 0, 0, 0, 1, 1, 1), False),
    ("grad276", (4, 0, 1, 0, 1, 1, 1, 0), True),
    ("visco404", (4, 32, 0, 0, 0, 1, 1, 1), False),
])
def test_batched_path_translation_invariance(name, tiles, graph):
    import ifosgpu as ig
    G, d_model, visco, nt, p0, p1 = _setup_homogeneous(ig, name)
    fc = 15.0
    cube = _cube((p0[1], p0[0], p0[2]), 6)  # (i, j, k) rows round the first source
    rec = np.concatenate([cube, cube + np.array(SHIFT, np.int32)])
    _note(f"batched path, {name}, tiles {tiles}, graph {graph}, sources {p0} and {p1}")
    b = ig.BatchContext(G, 2, 1, len(rec), ig.DFT_COMPAT, 4)
    b.set_cpml(5000.0, fc)
    b.set_freqs([fc], 100)
    if visco:
        b.set_visco(*visco)
    b.set_model(*d_model)
    b.set_tiles(tiles, tiles)
    d_b = ig.DBuf((2, 3, len(rec), nt + 1))
    b.run([(p[0], p[1], p[2], -1.5 / fc, fc, 1.0e6, 1) for p in (p0, p1)], rec, d_syn=d_b, want_adjoint=False,
          use_graph=graph)
    syn = d_b.numpy()
    b.close()
    m = len(cube)
    first, second = syn[0][:, :m], syn[1][:, m:]
    assert np.count_nonzero(first) > 0.5 * first[:, :, 2:].size
    assert np.array_equal(first, second)
