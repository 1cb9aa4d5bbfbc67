"""CPU tests of the HIP library's host side and ABI (no GPU calls).

* the library loads and exports every function declared in include/*.h;
* the host-computed CPML profiles and source wavelets are bit-identical to
  the oracle (same float/double promotions as CPML_coeff.c and wavelet.c).
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from helpers import small_grid
from pyoracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b([a-zA-Z_][a-zA-Z0-9_]*)\s*\([^;{]*\)\s*;", txt)
    return sorted(set(n for n in names if n not in ("if", "for", "while", "return", "sizeof")))


def test_native_library_exports_header():
    import ifosgpu as ig
    lib = ig.lib()
    names = _declared("ifos3d_hip.h")
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_dropin_library_exports_header():
    import ctypes as C
    drv = os.path.join(ROOT, "tests", "dropin", "lib", "libnr_driver.so")
    lib = C.CDLL(drv)  # defines the globals, pulls in libifos3d_dropin.so
    dropin = C.CDLL(os.path.join(ROOT, "ifos3d-dev_amd", "lib", "libifos3d_dropin.so"))
    names = _declared("ifos3d_dropin.h")
    assert {"update_v", "update_v_CPML", "update_s_elastic", "update_s_CPML_elastic", "discfourier",
            "gradient_F"} <= set(names)
    missing = [n for n in names if not hasattr(dropin, n)]
    assert not missing, missing
    assert hasattr(lib, "dd_run_shot")


@pytest.mark.parametrize("fw,dh,vppml,fpml,dt", [(10, 2.0, 4000.0, 50.0, 2e-4), (8, 5.0, 5500.0, 15.0, 4e-4),
                                                 (12, 0.8, 6200.0, 200.0, 5e-5), (3, 50.0, 300.0, 1.0, 1e-3)])
def test_cpml_profiles_match_oracle(fw, dh, vppml, fpml, dt):
    import ifosgpu as ig
    g = small_grid(nx=2 * fw + 5, ny=2 * fw + 6, nz=2 * fw + 7, fw=fw, dh=dh, dt=dt)
    G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt)
    got = ig.cpml_profiles(G, vppml, fpml).reshape(18, 2 * fw + 1)
    ref, _ = Oracle("f32").cpml_coeff(g, vppml, fpml)
    for n in range(18):
        assert np.array_equal(got[n], ref[n]), n


@pytest.mark.parametrize("shape", [1, 2, 4, 5])
def test_wavelet_matches_oracle(shape):
    import ifosgpu as ig
    g = small_grid(nt=700)
    G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt)
    for tshift, fc, amp in [(0.0, 50.0, 1.0), (0.003, 17.5, 2.5e3)]:
        got = ig.wavelet(G, shape, tshift, fc, amp)
        ref = Oracle("f32").wavelet(g, [(1, 1, 1, tshift, fc, amp, 1)], shape)[0]
        assert np.array_equal(got[1:g.nt + 1], ref[1:g.nt + 1])


def test_ntast_and_geometry():
    import ifosgpu as ig
    o = Oracle("f32")
    for fmax, tast, dt in [(80.0, 100, 2e-4), (14.0, 100, 3.19e-4), (40.0, 10, 4e-4)]:
        assert ig.ntast(fmax, tast, dt) == o.ntast(fmax, tast, dt)
    G = ig.Geom(84, 84, 84, 10, 2.0, 2.0, 2.0, 2e-4, 600)
    assert G.pitch % 32 == 0 and G.pitch >= 84 + 34
    # k = 1 starts a 128-byte line, k in [-2, nz+2] fits the pitch
    off = ig.lib().ifos_offset(G.ref, 1, 1, 1)
    assert off % 32 == 0
    assert ig.lib().ifos_offset(G.ref, 0, 0, 85) - ig.lib().ifos_offset(G.ref, 0, 0, -2) < G.pitch


def test_invalid_configuration_is_rejected():
    import ifosgpu as ig
    with pytest.raises(ig.IfosError):
        ig.Geom(10, 10, 10, 10, 2.0, 2.0, 2.0, 2e-4, 100)  # grid smaller than 2*FW+1
    with pytest.raises(ig.IfosError):
        ig.Geom(40, 40, 40, 5, 2.0, 2.0, 2.0, 2e-4, 100, fdcoeff=3)


def test_receiver_chain():
    import ifosgpu as ig
    rec = np.array([[1, 2, 3], [4, 5, 6], [1, 2, 3], [7, 7, 7], [1, 2, 3], [4, 5, 6]], np.int32)
    ch = ig.rec_chain(rec)
    assert ch[:, 0].tolist() == [1, 1, 0, 1, 0, 0]
    assert ch[:, 1].tolist() == [2, 5, 4, -1, -1, -1]
    assert ig.rec_chain(np.zeros((0, 3), np.int32)).shape == (0, 2)


def test_entry_points_reject_bad_arguments_before_any_device_call():
    """argument checks of the SU I/O, grid-level and time-domain entry points
    return IFOS_EINVAL with a message and touch no device"""
    import ifosgpu as ig
    lib = ig.lib()
    EINVAL = -1
    last = lambda: lib.ifos_last_error().decode()
    # SU I/O: ns out of range, a missing section, nothing to read
    assert lib.ifos_su_write(b"/tmp/x.su", None, 2, 0, None, None, None, C.c_float(1), C.c_float(1), C.c_float(1),
                             C.c_float(1e-3), 1) == EINVAL
    assert "ifos_su_write" in last()
    assert lib.ifos_su_write(b"/tmp/x.su", None, 2, 70000, None, None, None, C.c_float(1), C.c_float(1),
                             C.c_float(1), C.c_float(1e-3), 1) == EINVAL
    assert lib.ifos_su_read(b"/nonexistent/x.su", 3, 10, None) == EINVAL
    assert lib.ifos_su_read(b"/nonexistent/x.su", 0, 10, None) == 0  # readseis.c:53: no traces, no file
    # grid levels: none, too many
    assert lib.ifos_levels_run(None, 0, None, None, None, 0) == EINVAL
    assert "ifos_levels_run" in last()
    # time-domain gradient: no context / no buffers
    assert lib.ifos_ctx_run_shot_td(None, None, 0, None, None, None, None) == EINVAL
    assert "ifos_ctx_run_shot_td" in last()
    # batched path: DFT chunks of 1..32 snapshots, 1..1024 shots, 1..8 frequencies
    G = ig.Geom(20, 20, 20, 4, 2.0, 2.0, 2.0, 2e-4, 10)
    h = C.c_void_p()
    for shots, nf, chunk in ((1, 1, 33), (1, 1, 0), (0, 1, 8), (1025, 1, 8), (1, 9, 8)):
        assert lib.ifos_batch_create(C.byref(h), G.ref, shots, nf, 1, 0, chunk) == EINVAL, (shots, nf, chunk)
        assert "chunk in [1,32]" in last()


def test_plain_tile_shapes():
    """the one-shot march shape: set and read back; bad shapes rejected with
    the previous shape kept"""
    import ifosgpu as ig
    lib = ig.lib()
    old = ig.get_plain_tiles()
    assert old == (4, 32, 0, 0, 0, 1, 1, 1)
    try:
        for cfg in [(8, 32, 0, 0, 0, 1, 1, 1), (2, 0, 0, 0, 0, 1, 1, 1), (8, 16, 0, 0, 1, 2, 1, 0)]:
            ig.set_plain_tiles(cfg)
            assert ig.get_plain_tiles() == cfg
        for bad in [(10, 32, 0, 0, 0, 1, 1, 1), (5, 32, 0, 0, 0, 1, 1, 1), (3, 32, 0, 0, 0, 1, 1, 1),
                    (4, 32, 0, 0, 0, 3, 1, 0), (4, 32, 0, 0, 0, 2, 4, 0)]:
            with pytest.raises(ig.IfosError):
                ig.set_plain_tiles(bad)
            assert "ifos_set_plain_tiles" in lib.ifos_last_error().decode()
            assert ig.get_plain_tiles() == (8, 16, 0, 0, 1, 2, 1, 0)
    finally:
        ig.set_plain_tiles(old)


def test_td_bytes_formula():
    """ifos_td_bytes: (ceil(NT/K) + 1) states + K velocity triples + 6 running
    displacement fields + 3 FP64 sums"""
    import ifosgpu as ig
    g = small_grid(nx=30, ny=28, nz=32, fw=6, nt=100)
    G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt)
    fe = G.field_elems()
    psi = sum(G.psi_elems(n) for n in range(9))
    for every, L in ((10, 0), (7, 1), (0, 0)):
        K = every if every > 0 else round(np.sqrt(g.nt))
        nck = -(-g.nt // K)
        state = (9 + (6 if L else 0)) * fe * 4 + 2 * psi * 4
        want = (nck + 1) * state + 3 * K * fe * 4 + 6 * fe * 4 + 3 * fe * 8
        assert ig.ShotContext.td_bytes(G, every, L) == want
