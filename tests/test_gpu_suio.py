"""SU seismogram files through the C ABI (ifos_su_write / ifos_su_read).

The writer's bytes are compared with the golden files that the reference's
outseis.c wrote for the same sections (its unassigned header words zeroed),
and with the oracle's restatement on seeded random geometries; the reader
with readseis.c's samples of the golden files, full and cut short.  Sections
come from device memory and go back to it, as the shot context holds them.
"""
import json
import os
import sys

import numpy as np
import pytest

from golden import cases as cs

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
from suio_oracle import su_read, su_write  # noqa: E402

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.json")


def _note(msg):
    print(msg, flush=True)


def _gold(n):
    with open(GOLDEN) as fh:
        return json.load(fh)["cases"][f"su_{n}"]


def _sha_ok(want, a):
    from helpers import sha
    return sha(np.ascontiguousarray(a, np.float32)) == want["sha256"]


@pytest.mark.parametrize("n", range(len(cs.SU_CASES)))
def test_su_golden(tmp_path, n):
    import ifosgpu as ig
    ntr, ns, rid, src, dx, dy, dz, dt, ndt = cs.SU_CASES[n]
    _note(f"su golden case {n}: ntr={ntr} ns={ns}")
    sec, rxyz = cs.su_inputs(cs.SU_CASES[n], 100 + n)
    gold = _gold(n)
    d = ig.DBuf.from_numpy(sec)
    path = tmp_path / "obs_vy.su"
    ig.su_write(path, d, ntr, ns, rxyz, dx, dy, dz, dt, ndt, rec_id=rid, src_xyz=src)
    got = path.read_bytes()
    want = bytes.fromhex(gold["bytes"])
    bad = [i for i in range(max(len(got), len(want))) if i >= len(got) or i >= len(want) or got[i] != want[i]]
    assert not bad, f"SU bytes differ at {bad[:8]}"
    back = ig.DBuf((ntr, ns + 1))
    ig.su_read(path, ntr, ns, back)
    assert _sha_ok(gold["read"], back.numpy())
    path.write_bytes(want[:gold["short_len"]])
    ig.su_read(path, ntr, ns, back)
    assert _sha_ok(gold["read_short"], back.numpy())


@pytest.mark.parametrize("seed", range(6))
def test_su_write_matches_oracle(tmp_path, seed):
    import ifosgpu as ig
    rng = np.random.default_rng(700 + seed)
    ntr, ns = int(rng.integers(1, 40)), int(rng.integers(1, 300))
    _note(f"su random seed={seed} ntr={ntr} ns={ns}")
    sec = np.zeros((ntr, ns + 1), np.float32)
    sec[:, 1:] = rng.standard_normal((ntr, ns)).astype(np.float32)
    rxyz = rng.integers(1, 400, size=(ntr, 3)).astype(np.int32)
    rid = rng.permutation(5 * ntr)[:ntr].astype(np.int32) + 1
    dx, dy, dz = (float(np.float32(v)) for v in rng.uniform(0.05, 25.0, 3))
    dt = float(np.float32(rng.uniform(1e-5, 5e-3)))
    ndt = int(rng.integers(1, 5))
    src = None if seed == 0 else tuple(float(np.float32(v)) for v in rng.uniform(0.0, 4000.0, 3))
    path = tmp_path / "s.su"
    ig.su_write(path, ig.DBuf.from_numpy(sec), ntr, ns, rxyz, dx, dy, dz, dt, ndt,
                rec_id=None if seed == 1 else rid, src_xyz=src)
    want = su_write(sec, rxyz, dx, dy, dz, dt, ndt, rec_id=None if seed == 1 else rid, src_xyz=src)
    got = path.read_bytes()
    assert got == want
    assert np.array_equal(su_read(got, ntr, ns), sec)


def test_su_round_trip_full_section(tmp_path):
    """a section the size of one component of a 256^3 gradient shot's
    (1024 receivers, 1000 steps)"""
    import ifosgpu as ig
    ntr, ns = 1024, 1000
    _note(f"su round trip ntr={ntr} ns={ns}")
    rng = np.random.default_rng(5)
    sec = np.zeros((ntr, ns + 1), np.float32)
    sec[:, 1:] = rng.standard_normal((ntr, ns)).astype(np.float32)
    rxyz = rng.integers(1, 256, size=(ntr, 3)).astype(np.int32)
    path = tmp_path / "big.su"
    ig.su_write(path, ig.DBuf.from_numpy(sec), ntr, ns, rxyz, 10.0, 10.0, 10.0, 8e-4, 1, src_xyz=(1280.0, 50.0, 1280.0))
    assert path.stat().st_size == ntr * (240 + 4 * ns)
    back = ig.DBuf((ntr, ns + 1))
    ig.su_read(path, ntr, ns, back)
    assert np.array_equal(back.numpy(), sec)


def test_su_edges(tmp_path):
    import ifosgpu as ig
    _note("su edges")
    # no traces: an empty file, and nothing to read (readseis.c:53)
    path = tmp_path / "empty.su"
    ig.su_write(path, None, 0, 10, np.zeros((0, 3), np.int32), 1.0, 1.0, 1.0, 1e-3)
    assert path.read_bytes() == b""
    ig.su_read(tmp_path / "absent.su", 0, 10, None)
    # one sample per trace
    sec = np.array([[0.0, 1.5], [0.0, -2.0]], np.float32)
    ig.su_write(path, ig.DBuf.from_numpy(sec), 2, 1, [[1, 2, 3], [4, 5, 6]], 1.0, 1.0, 1.0, 1e-3)
    assert path.read_bytes() == su_write(sec, [[1, 2, 3], [4, 5, 6]], 1.0, 1.0, 1.0, 1e-3)
    # more traces asked for than the file holds: the rest read as 0
    back = ig.DBuf((4, 2))
    ig.su_read(path, 4, 1, back)
    assert np.array_equal(back.numpy(), su_read(path.read_bytes(), 4, 1))
    # errors: a missing file, bad sizes
    with pytest.raises(ig.IfosError):
        ig.su_read(tmp_path / "absent.su", 2, 1, back)
    with pytest.raises(ig.IfosError):
        ig.su_write(path, ig.DBuf.from_numpy(sec), 2, 0, [[1, 2, 3], [4, 5, 6]], 1.0, 1.0, 1.0, 1e-3)
    with pytest.raises(ig.IfosError):
        ig.su_write(tmp_path / "no_such_dir" / "x.su", ig.DBuf.from_numpy(sec), 2, 1, [[1, 2, 3], [4, 5, 6]],
                    1.0, 1.0, 1.0, 1e-3)
