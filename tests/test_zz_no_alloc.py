"""The allocation-free residual (ifos_residual_noalloc) against ifos_residual.

Both are called through the C ABI on the same seeded traces.  The adjoint
traces and the accumulated FP64 misfit must be the same bits, on two calls in
a row, the second with fewer traces and samples, so that it runs over what the
first call left in the static partial-sum array.  A trace count above the
array's maximum is refused before any device work.
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _residual(ig, fn, G, dobs, dsyn, dadj, ntr, ns, l2_0):
    acc = C.c_double(l2_0)
    rc = getattr(ig.lib(), fn)(G.ref, ig._ptr(dobs), ig._ptr(dsyn), ig._ptr(dadj), ntr, ns, C.byref(acc), None)
    ig.check(rc, fn)
    return acc.value


def test_noalloc_residual_matches_residual():
    import ifosgpu as ig

    assert ig.device_count() > 0, "GPU tests need a device"
    G = ig.Geom(28, 26, 30, 6, 2.0, 2.0, 2.0, 3.1e-4, 700)
    # the second call is smaller than the first in both traces and samples
    for ntr, ns, l2_0, seed in [(3000, 700, 1.25, 11), (37, 90, 0.0, 12)]:
        _note(f"residual_noalloc: ntr {ntr} ns {ns} seed {seed}")
        rng = np.random.default_rng(seed)
        obs = rng.standard_normal((3, ntr, ns + 1)).astype(np.float32)
        syn = rng.standard_normal((3, ntr, ns + 1)).astype(np.float32)
        junk = rng.standard_normal((3, ntr, ns + 2)).astype(np.float32)  # slots 0 and ns+1 stay as they were
        dobs, dsyn = ig.DBuf.from_numpy(obs), ig.DBuf.from_numpy(syn)
        dadj_a, dadj_b = ig.DBuf.from_numpy(junk), ig.DBuf.from_numpy(junk)
        l2_a = _residual(ig, "ifos_residual", G, dobs, dsyn, dadj_a, ntr, ns, l2_0)
        l2_b = _residual(ig, "ifos_residual_noalloc", G, dobs, dsyn, dadj_b, ntr, ns, l2_0)
        ig.sync()
        a, b = dadj_a.numpy(), dadj_b.numpy()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), "adjoint traces differ"
        assert np.float64(l2_a).tobytes() == np.float64(l2_b).tobytes(), (l2_a, l2_b)
        # and the misfit is the sum of the traces' float squares, added in FP64
        d = b[:, :, 1:ns + 1]
        want = l2_0 + float(np.sum((d * d).astype(np.float64)))
        assert abs(l2_b - want) <= 1e-10 * abs(want)


def test_noalloc_residual_rejects_too_many_traces():
    import ifosgpu as ig

    assert ig.device_count() > 0, "GPU tests need a device"
    G = ig.Geom(28, 26, 30, 6, 2.0, 2.0, 2.0, 3.1e-4, 10)
    d = ig.DBuf((3, 1, 12))
    acc = C.c_double(0.0)
    n = ig.RESIDUAL_NOALLOC_MAX_TRACES + 1
    rc = ig.lib().ifos_residual_noalloc(G.ref, ig._ptr(d), ig._ptr(d), ig._ptr(d), n, 10, C.byref(acc), None)
    assert rc == ig.EINVAL
    assert b"scratch maximum" in ig.lib().ifos_last_error()
    assert acc.value == 0.0
