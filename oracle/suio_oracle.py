"""CPU restatement of the reference's SU seismogram I/O (TEST INFRASTRUCTURE).

Only the test suite imports this module, as the checker of ifos_su_write and
ifos_su_read.

* su_write follows outseis.c:30-275 for SEIS_FORMAT 1 (SU), called as
  saveseis.c:38-80 calls it (one source position, the shot's).  Per trace a
  240-byte header in the portable segy layout of segy.h:451-720 (native byte
  order) and ns native float32 samples.  The header words are computed with
  the reference's mixed FP32/FP64 arithmetic: receiver coordinates
  int * float in FP32, the offset as an FP32 sum of squares whose square
  root is taken in FP64, iround (fd.h:35) as floor(x + 0.5) in FP64, the
  sample interval as (float)NDT*DT in FP32 scaled by 1.0e6 in FP64.
  The words outseis.c never assigns (cdp, selev, swdep, unass[15]) hold
  stack contents there; they are 0 here.
* su_read follows readseis.c:53-63: skip 240 bytes per trace, read ns floats;
  samples beyond the end of the file are 0.

Sections are [ntr][ns+1] float32 arrays, sample j at index j (index 0 unused
and 0).
"""
from __future__ import annotations

import math
import struct

import numpy as np

HEADER_BYTES = 240
# byte ranges of the header words outseis.c leaves unassigned
UNASSIGNED = ((20, 24), (44, 48), (60, 64), (210, 240))

_f32 = np.float32
_f64 = np.float64


def _iround(x) -> int:
    """fd.h:35 iround: (int)floor(x + 0.5) with x + 0.5 in FP64"""
    return int(math.floor(_f64(x) + 0.5))


def su_header(t, rec_xyz, rec_id, src_xyz, dx, dy, dz, dt, ndt, ns) -> bytes:
    """outseis.c:62-101 for trace t (0-based); unset words are 0"""
    XS, YS, ZS = (_f32(v) for v in (src_xyz if src_xyz is not None else (0.0, 0.0, 0.0)))
    scalefac = _f32(math.pow(10.0, 3.0))
    xr = _f32(rec_xyz[t][0]) * _f32(dx)
    yr = _f32(rec_xyz[t][1]) * _f32(dy)
    zr = _f32(rec_xyz[t][2]) * _f32(dz)
    ssq = (XS - xr) * (XS - xr) + (YS - yr) * (YS - yr) + (ZS - zr) * (ZS - zr)  # FP32 sum
    h = bytearray(HEADER_BYTES)
    struct.pack_into("=i", h, 0, int(rec_id[t]) if rec_id is not None else t + 1)   # tracl
    struct.pack_into("=h", h, 28, 1)                                                # trid
    struct.pack_into("=i", h, 36, _iround(math.sqrt(_f64(ssq)) * _f64(scalefac)))   # offset
    struct.pack_into("=i", h, 40, _iround(yr * scalefac))                           # gelev
    struct.pack_into("=i", h, 48, _iround(YS * scalefac))                           # sdepth
    struct.pack_into("=hh", h, 68, -3, -3)                                          # scalel, scalco
    struct.pack_into("=ii", h, 72, _iround(XS * scalefac), _iround(ZS * scalefac))  # sx, sy
    struct.pack_into("=ii", h, 80, _iround(xr * scalefac), _iround(zr * scalefac))  # gx, gy
    struct.pack_into("=H", h, 114, ns)                                              # ns
    dtus = _iround(_f64(_f32(ndt) * _f32(dt)) * 1.0e6)
    struct.pack_into("=H", h, 116, dtus & 0xFFFF)                                   # dt (unsigned short cast)
    return bytes(h)


def su_write(section, rec_xyz, dx, dy, dz, dt, ndt=1, rec_id=None, src_xyz=None) -> bytes:
    """the bytes outseis.c writes for section [ntr][ns+1]"""
    sec = np.ascontiguousarray(section, np.float32)
    ntr, ns = sec.shape[0], sec.shape[1] - 1
    out = bytearray()
    for t in range(ntr):
        out += su_header(t, rec_xyz, rec_id, src_xyz, dx, dy, dz, dt, ndt, ns)
        out += sec[t, 1:].tobytes()
    return bytes(out)


def su_read(data: bytes, ntr: int, ns: int) -> np.ndarray:
    """readseis.c:53-63 over the bytes of one file"""
    out = np.zeros((ntr, ns + 1), np.float32)
    pos = 0
    for t in range(ntr):
        pos += HEADER_BYTES
        for j in range(1, ns + 1):
            if pos + 4 <= len(data):
                out[t, j] = np.frombuffer(data, np.float32, 1, pos)[0]
            pos += 4
    return out


def mask_unassigned(data: bytes, ns: int) -> bytes:
    """zero the header words outseis.c leaves unassigned, in every trace"""
    b = bytearray(data)
    step = HEADER_BYTES + 4 * ns
    for base in range(0, len(b), step):
        for lo, hi in UNASSIGNED:
            b[base + lo:base + hi] = bytes(hi - lo)
    return bytes(b)
