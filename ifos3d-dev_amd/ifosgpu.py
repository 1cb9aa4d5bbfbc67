"""Python host side of libifos3d_hip.so (ctypes), used by the tests and bench.

Mirrors the native C ABI of include/ifos3d_hip.h.  Device memory is
allocated through the library itself (DBuf), so the process runs one HIP
runtime and every launch and copy is ordered on the library's (null) stream.
The library must be built (make -C ifos3d-dev_amd); there is no fallback: a
missing library raises.

Canonical arrays (numpy, shape (ny+5, nx+5, nz+5), index [j+2, i+2, k+2])
are converted to and from the device layout (shape (ny+5, nx+5, pitch), k at
column k + KOFF) by `to_dev` / `from_dev`.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libifos3d_hip.so")
KOFF = 31

REGION_INTERIOR, REGION_CPML, REGION_ALL = 1, 2, 3
DFT_COMPAT, DFT_KAHAN = 0, 1


class IfosGeom(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("fw", C.c_int),
                ("free_surf", C.c_int), ("fdcoeff", C.c_int),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float), ("dt", C.c_float),
                ("nt", C.c_int), ("time", C.c_float), ("pitch", C.c_int)]


class IfosFields(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "syz", "sxz")]


class IfosModel(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("rho", "pi", "u", "uipjp", "ujpkp", "uipkp", "rjp", "rkp", "rip")]


class IfosPsi(C.Structure):
    _fields_ = [("p", C.c_void_p * 9)]


MAX_L = 4  # IFOS_MAX_L: relaxation mechanisms


class IfosVisco(C.Structure):
    _fields_ = [("r", C.c_void_p * 6), ("taus", C.c_void_p), ("taup", C.c_void_p), ("tausipjp", C.c_void_p),
                ("tausjpkp", C.c_void_p), ("tausipkp", C.c_void_p), ("eta", C.c_float), ("L", C.c_int),
                ("eta_more", C.c_float * (MAX_L - 1))]


def _etas(eta):
    """a scalar eta (one mechanism) or one eta per mechanism, as float32"""
    e = np.atleast_1d(np.asarray(eta, np.float32))
    if not 1 <= e.size <= MAX_L:
        raise ValueError(f"1..{MAX_L} relaxation mechanisms")
    return e


def visco_struct(r, taus, taup, tausipjp, tausjpkp, tausipkp, eta, L=None):
    """eta: scalar (one mechanism) or one per mechanism; r[n] holds the
    mechanisms' fields back to back"""
    v = IfosVisco()
    e = _etas(eta)
    for n in range(6):
        v.r[n] = r[n].data_ptr()
    v.taus, v.taup = taus.data_ptr(), taup.data_ptr()
    v.tausipjp, v.tausjpkp, v.tausipkp = tausipjp.data_ptr(), tausjpkp.data_ptr(), tausipkp.data_ptr()
    v.eta = float(e[0])
    for l in range(1, e.size):
        v.eta_more[l - 1] = float(e[l])
    v.L = e.size if L is None else L
    return v


class IfosShot(C.Structure):
    _fields_ = [("src_i", C.c_int), ("src_j", C.c_int), ("src_k", C.c_int), ("src_type", C.c_int),
                ("tshift", C.c_float), ("fc", C.c_float), ("amp", C.c_float), ("source_shape", C.c_int),
                ("ntr", C.c_int), ("rec_ijk", C.c_void_p), ("d_obs", C.c_void_p), ("want_adjoint", C.c_int)]


class IfosError(RuntimeError):
    """an IFOS_E* return code (`code`) with the library's message"""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


# return codes of include/ifos3d_hip.h
OK, EINVAL, EHIP, ENAN, ENOMEM, EUNSTABLE = 0, -1, -2, -3, -4, -5
# trace limit of ifos_residual_noalloc's static partial-sum array (ifos3d_hip.h)
RESIDUAL_NOALLOC_MAX_TRACES = 16384


_lib = None


def lib() -> C.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise IfosError(f"{LIBPATH} is missing: build it with `make -C ifos3d-dev_amd` "
                            "(or __graft_entry__.build())")
        _lib = C.CDLL(LIBPATH)
        _lib.ifos_last_error.restype = C.c_char_p
        _lib.ifos_field_elems.restype = C.c_size_t
        _lib.ifos_offset.restype = C.c_size_t
        _lib.ifos_psi_elems.restype = C.c_size_t
        _lib.ifos_batch_bytes.restype = C.c_size_t
        _lib.ifos_td_bytes.restype = C.c_size_t
        _lib.ifos_adj_bytes.restype = C.c_size_t
    return _lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().ifos_last_error().decode()
        raise IfosError(f"{what}: error {rc}: {msg}", rc)


def device_count() -> int:
    n = C.c_int(0)
    lib().ifos_device_count(C.byref(n))
    return n.value


def device_name() -> str:
    buf = C.create_string_buffer(256)
    check(lib().ifos_device_name(buf, 256), "ifos_device_name")
    return buf.value.decode()


def sync():
    check(lib().ifos_device_sync(), "ifos_device_sync")


class DBuf:
    """A device array owned by the library's HIP runtime (synchronous copies)."""

    def __init__(self, shape, dtype=np.float32, zero=True):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().ifos_dmalloc(C.byref(p), C.c_size_t(self.nbytes)), "ifos_dmalloc")
        self._p = p.value
        if zero:
            check(lib().ifos_dmemset(C.c_void_p(self._p), 0, C.c_size_t(self.nbytes)), "ifos_dmemset")

    @classmethod
    def wrap(cls, ptr, shape, dtype=np.float32):
        """a view of device memory owned elsewhere (never freed through it)"""
        b = cls.__new__(cls)
        b.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        b.dtype = np.dtype(dtype)
        b.nbytes = int(np.prod(b.shape)) * b.dtype.itemsize
        b._p = int(ptr)
        b._owned = False
        return b

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.shape, a.dtype, zero=False)
        check(lib().ifos_memcpy_h2d(C.c_void_p(b._p), a.ctypes.data_as(C.c_void_p), C.c_size_t(b.nbytes)),
              "ifos_memcpy_h2d")
        return b

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().ifos_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(self._p), C.c_size_t(self.nbytes)),
              "ifos_memcpy_d2h")
        return out

    def copy_from(self, other: "DBuf"):
        assert other.nbytes == self.nbytes
        check(lib().ifos_memcpy_d2d(C.c_void_p(self._p), C.c_void_p(other._p), C.c_size_t(self.nbytes)),
              "ifos_memcpy_d2d")
        return self

    def clone(self):
        return DBuf(self.shape, self.dtype, zero=False).copy_from(self)

    def zero_(self):
        check(lib().ifos_dmemset(C.c_void_p(self._p), 0, C.c_size_t(self.nbytes)), "ifos_dmemset")
        return self

    def data_ptr(self) -> int:
        return self._p

    def ptr(self, offset_elems=0) -> int:
        return self._p + offset_elems * self.dtype.itemsize

    def free(self):
        if not getattr(self, "_owned", True):
            self._p = None
            return
        if getattr(self, "_p", None):
            try:
                lib().ifos_dfree(C.c_void_p(self._p))
            finally:
                self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class Geom:
    """A grid (NX, NY, NZ incl. CPML) in the library's device layout."""

    def __init__(self, nx, ny, nz, fw, dx, dy, dz, dt, nt, fdcoeff=2, free_surf=0):
        self.c = IfosGeom()
        check(lib().ifos_geom_init(C.byref(self.c), nx, ny, nz, fw, C.c_float(dx), C.c_float(dy),
                                   C.c_float(dz), C.c_float(dt), nt, fdcoeff, free_surf), "ifos_geom_init")
        self.nx, self.ny, self.nz, self.fw, self.nt = nx, ny, nz, fw, nt
        self.pitch = self.c.pitch
        self.dshape = (ny + 5, nx + 5, self.pitch)
        self.cshape = (ny + 5, nx + 5, nz + 5)

    @property
    def ref(self):
        return C.byref(self.c)

    @property
    def cells(self):
        return self.nx * self.ny * self.nz

    def field_elems(self):
        return int(lib().ifos_field_elems(self.ref))

    def psi_elems(self, n):
        return int(lib().ifos_psi_elems(self.ref, n))

    def psi_shape(self, n):
        f2 = 2 * self.fw
        return [(self.ny, f2, self.nz), (f2, self.nx, self.nz), (self.ny, self.nx, f2)][n // 3]

    # ---- layout conversion (canonical numpy <-> device) ----
    def empty(self, n=None):
        return DBuf(self.dshape if n is None else (n,) + self.dshape)

    def to_host_layout(self, a):
        """canonical (..., ny+5, nx+5, nz+5) -> device layout numpy array"""
        a = np.asarray(a, dtype=np.float32)
        host = np.zeros(a.shape[:-1] + (self.pitch,), np.float32)
        host[..., KOFF - 2:KOFF - 2 + self.nz + 5] = a
        return host

    def to_dev(self, a):
        return DBuf.from_numpy(self.to_host_layout(a))

    def from_dev(self, t):
        h = t.numpy() if isinstance(t, DBuf) else np.asarray(t)
        return np.ascontiguousarray(h[..., KOFF - 2:KOFF - 2 + self.nz + 5])


def fields_struct(ts):
    return IfosFields(*[t.data_ptr() if t is not None else None for t in ts])


def model_struct(ts):
    return IfosModel(*[t.data_ptr() if t is not None else None for t in ts])


def psi_struct(ts):
    p = IfosPsi()
    for n, t in enumerate(ts):
        p.p[n] = t.data_ptr()
    return p


def cpml_profiles(g: Geom, vppml: float, fpml: float) -> np.ndarray:
    out = np.zeros(18 * (2 * g.fw + 1), np.float32)
    check(lib().ifos_cpml_profiles(g.ref, C.c_float(vppml), C.c_float(fpml),
                                   out.ctypes.data_as(C.c_void_p)), "ifos_cpml_profiles")
    return out


def wavelet(g: Geom, shape, tshift, fc, amp) -> np.ndarray:
    out = np.zeros(g.nt + 2, np.float32)
    check(lib().ifos_wavelet(g.ref, shape, C.c_float(tshift), C.c_float(fc), C.c_float(amp),
                             out.ctypes.data_as(C.c_void_p)), "ifos_wavelet")
    return out


def ntast(fmax, tast, dt) -> int:
    return int(lib().ifos_ntast(C.c_float(fmax), tast, C.c_float(dt)))


def update_v(g, region, f, m, prof, psi, stream=None):
    check(lib().ifos_update_v(g.ref, region, C.byref(fields_struct(f)), C.byref(model_struct(m)),
                              _ptr(prof), C.byref(psi_struct(psi)) if psi is not None else None, stream),
          "ifos_update_v")


def update_s(g, region, f, m, prof, psi, stream=None):
    check(lib().ifos_update_s(g.ref, region, C.byref(fields_struct(f)), C.byref(model_struct(m)),
                              _ptr(prof), C.byref(psi_struct(psi)) if psi is not None else None, stream),
          "ifos_update_s")


def update_s_visco(g, region, f, m, visco, prof, psi, stream=None):
    """visco: an IfosVisco (visco_struct)"""
    check(lib().ifos_update_s_visco(g.ref, region, C.byref(fields_struct(f)), C.byref(model_struct(m)),
                                    C.byref(visco), _ptr(prof),
                                    C.byref(psi_struct(psi)) if psi is not None else None, stream),
          "ifos_update_s_visco")


def av_tau(g, d_taus, d_out, stream=None):
    check(lib().ifos_av_tau(g.ref, _ptr(d_taus), _ptr(d_out[0]), _ptr(d_out[1]), _ptr(d_out[2]), stream),
          "ifos_av_tau")


def add_forces(g, f, m, d_src, d_sig, nsrc, nt, alpha=0.0, beta=0.0, stream=None):
    check(lib().ifos_add_forces(g.ref, C.byref(fields_struct(f)), C.byref(model_struct(m)), _ptr(d_src),
                                _ptr(d_sig), nsrc, nt, C.c_float(alpha), C.c_float(beta), stream),
          "ifos_add_forces")


def add_adjoint(g, f, d_rec, d_adj, ntr, nt, d_chain=None, stream=None):
    check(lib().ifos_add_adjoint(g.ref, C.byref(fields_struct(f)), _ptr(d_rec), _ptr(d_chain), _ptr(d_adj), ntr,
                                 nt, stream), "ifos_add_adjoint")


def rec_chain(rec_ijk) -> np.ndarray:
    rec = np.ascontiguousarray(rec_ijk, np.int32)
    out = np.zeros((len(rec), 2), np.int32)
    check(lib().ifos_rec_chain(rec.ctypes.data_as(C.c_void_p), len(rec), out.ctypes.data_as(C.c_void_p)),
          "ifos_rec_chain")
    return out


def copy(d_dst, d_src, n, stream=None):
    check(lib().ifos_copy(_ptr(d_dst), _ptr(d_src), C.c_size_t(n), stream), "ifos_copy")


def psource(g, f, d_src, d_sig, nsrc, nt, stream=None):
    check(lib().ifos_psource(g.ref, C.byref(fields_struct(f)), _ptr(d_src), _ptr(d_sig), nsrc, nt, stream),
          "ifos_psource")


def surface(g, f, m, stream=None):
    check(lib().ifos_surface(g.ref, C.byref(fields_struct(f)), C.byref(model_struct(m)), stream), "ifos_surface")


def seismo(g, f, d_rec, ntr, ins, ns, d_sec, stream=None):
    check(lib().ifos_seismo(g.ref, C.byref(fields_struct(f)), _ptr(d_rec), ntr, ins, ns, _ptr(d_sec), stream),
          "ifos_seismo")


def discfourier(g, f, F6, C6, finv, nt, ntast_, back, precision=DFT_COMPAT, stream=None):
    fv = np.asarray(finv, np.float32)
    Fp = (C.c_void_p * 6)(*[t.data_ptr() for t in F6])
    Cp = (C.c_void_p * 6)(*[t.data_ptr() for t in C6]) if C6 is not None else None
    check(lib().ifos_discfourier(g.ref, C.byref(fields_struct(f)), Fp, Cp, fv.ctypes.data_as(C.c_void_p),
                                 len(fv), nt, ntast_, back, precision, stream), "ifos_discfourier")


def gradient_F(g, Ff6, Fb6, m, finv, grads, stream=None):
    fv = np.asarray(finv, np.float32)
    Fp = (C.c_void_p * 6)(*[t.data_ptr() for t in Ff6])
    Bp = (C.c_void_p * 6)(*[t.data_ptr() for t in Fb6])
    Gp = (C.c_void_p * 3)(*[t.data_ptr() for t in grads])
    check(lib().ifos_gradient_F(g.ref, Fp, Bp, C.byref(model_struct(m)), fv.ctypes.data_as(C.c_void_p),
                                len(fv), Gp, stream), "ifos_gradient_F")


def hess_F(g, Ff6, Fb6, ntr_hess, m, finv, hess, stream=None):
    """hess_F.c: Ff6 [nf] and Fb6 [nf*(ntr_hess+1)] stacks of device fields"""
    fv = np.asarray(finv, np.float32)
    Fp = (C.c_void_p * 6)(*[t.data_ptr() for t in Ff6])
    Bp = (C.c_void_p * 6)(*[t.data_ptr() for t in Fb6])
    Hp = (C.c_void_p * 3)(*[t.data_ptr() for t in hess])
    check(lib().ifos_hess_F(g.ref, Fp, Bp, ntr_hess, C.byref(model_struct(m)), fv.ctypes.data_as(C.c_void_p),
                            len(fv), Hp, stream), "ifos_hess_F")


def su_read(path, ntr, ns, d_section):
    """readseis.c: ntr SU traces of ns samples into d_section [ntr][ns+1]"""
    check(lib().ifos_su_read(os.fsencode(path), ntr, ns, _ptr(d_section)), "ifos_su_read")


def su_write(path, d_section, ntr, ns, rec_xyz, dx, dy, dz, dt, ndt=1, rec_id=None, src_xyz=None):
    """outseis.c (SU): d_section [ntr][ns+1] to path; rec_xyz [ntr][3] grid
    indices, rec_id global receiver numbers, src_xyz the shot position (m)"""
    rx = np.ascontiguousarray(rec_xyz, np.int32).reshape(-1)
    rid = None if rec_id is None else np.ascontiguousarray(rec_id, np.int32)
    sp = None if src_xyz is None else np.ascontiguousarray(src_xyz, np.float32)
    check(lib().ifos_su_write(os.fsencode(path), _ptr(d_section), ntr, ns, rx.ctypes.data_as(C.c_void_p),
                              None if rid is None else rid.ctypes.data_as(C.c_void_p),
                              None if sp is None else sp.ctypes.data_as(C.c_void_p),
                              C.c_float(dx), C.c_float(dy), C.c_float(dz), C.c_float(dt), ndt), "ifos_su_write")


def matcopy_wrap(g, d_m, stream=None):
    check(lib().ifos_matcopy_wrap(g.ref, _ptr(d_m), stream), "ifos_matcopy_wrap")


def av_mat(g, rho, u, out6, stream=None):
    check(lib().ifos_av_mat(g.ref, _ptr(rho), _ptr(u), *[_ptr(t) for t in out6], stream), "ifos_av_mat")


def set_plain_tiles(cfg):
    """tile shape {TI, jc, group, split, pipe, kb, sb, lds} of the one-shot
    march launches (process-wide)"""
    a = (C.c_int * 8)(*[int(x) for x in cfg])
    check(lib().ifos_set_plain_tiles(a), "ifos_set_plain_tiles")


def get_plain_tiles():
    a = (C.c_int * 8)()
    check(lib().ifos_get_plain_tiles(a), "ifos_get_plain_tiles")
    return tuple(a)


def residual(g, d_obs, d_syn, d_adj, ntr, ns, stream=None) -> float:
    l2 = C.c_double(0.0)
    check(lib().ifos_residual(g.ref, _ptr(d_obs), _ptr(d_syn), _ptr(d_adj), ntr, ns, C.byref(l2), stream),
          "ifos_residual")
    return l2.value


def residual_noalloc(g, d_obs, d_syn, d_adj, ntr, ns, l2=0.0, stream=None) -> float:
    """ifos_residual_noalloc: returns l2 + the misfit, as the C call accumulates"""
    acc = C.c_double(l2)
    check(lib().ifos_residual_noalloc(g.ref, _ptr(d_obs), _ptr(d_syn), _ptr(d_adj), ntr, ns, C.byref(acc),
                                      stream), "ifos_residual_noalloc")
    return acc.value


class ShotContext:
    """The plain one-shot path: forward, residual, adjoint, gradient_F."""

    def __init__(self, g: Geom, nf_max: int, ntr_max: int, precision=DFT_COMPAT):
        self.g = g
        self.h = C.c_void_p()
        check(lib().ifos_ctx_create(C.byref(self.h), g.ref, nf_max, ntr_max, precision), "ifos_ctx_create")

    def close(self):
        if self.h:
            lib().ifos_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_model(self, d_rho, d_pi, d_u, stream=None):
        check(lib().ifos_ctx_set_model(self.h, _ptr(d_rho), _ptr(d_pi), _ptr(d_u), stream), "ifos_ctx_set_model")

    def set_cpml(self, vppml, fpml):
        check(lib().ifos_ctx_set_cpml(self.h, C.c_float(vppml), C.c_float(fpml)), "ifos_ctx_set_cpml")

    def set_freqs(self, finv, tast=100):
        fv = np.asarray(finv, np.float32)
        check(lib().ifos_ctx_set_freqs(self.h, fv.ctypes.data_as(C.c_void_p), len(fv), tast),
              "ifos_ctx_set_freqs")

    def set_visco(self, d_taus, d_taup, eta, L=None, stream=None):
        """eta: scalar (one relaxation mechanism, update_s_ssg.c) or one per
        mechanism (ifos_ctx_set_visco_n); L = 0 switches back to elastic"""
        e = _etas(eta)
        n = e.size if L is None else L
        check(lib().ifos_ctx_set_visco_n(self.h, _ptr(d_taus), _ptr(d_taup), e.ctypes.data_as(C.c_void_p), n,
                                         stream), "ifos_ctx_set_visco_n")

    def run_shot(self, src, rec_ijk, d_syn, grads, d_obs=None, shape=1, want_adjoint=True, stream=None):
        """src = (i, j, k, tshift, fc, amp, type); returns the L2 misfit (FP64 sum)."""
        rec = np.ascontiguousarray(rec_ijk, np.int32)
        s = IfosShot(src[0], src[1], src[2], src[6], src[3], src[4], src[5], shape, rec.shape[0],
                     rec.ctypes.data, d_obs.data_ptr() if d_obs is not None else None, int(want_adjoint))
        l2 = C.c_double(0.0)
        Gp = (C.c_void_p * 3)(*[t.data_ptr() for t in grads]) if grads is not None else None
        check(lib().ifos_ctx_run_shot(self.h, C.byref(s), _ptr(d_syn), Gp, C.byref(l2), stream),
              "ifos_ctx_run_shot")
        return l2.value

    def run_shot_td(self, src, rec_ijk, d_syn, grads, d_obs, every=0, shape=1, stream=None):
        """ifos_ctx_run_shot_td: the time-domain gradient with a checkpoint
        every `every` steps (0: round(sqrt(NT))); returns the L2 misfit"""
        rec = np.ascontiguousarray(rec_ijk, np.int32)
        s = IfosShot(src[0], src[1], src[2], src[6], src[3], src[4], src[5], shape, rec.shape[0],
                     rec.ctypes.data, d_obs.data_ptr(), 1)
        l2 = C.c_double(0.0)
        Gp = (C.c_void_p * 3)(*[t.data_ptr() for t in grads])
        check(lib().ifos_ctx_run_shot_td(self.h, C.byref(s), every, _ptr(d_syn), Gp, C.byref(l2), stream),
              "ifos_ctx_run_shot_td")
        return l2.value

    def run_shot_adj(self, src, rec_ijk, d_syn, grads, d_obs, every=0, shape=1, stream=None):
        """ifos_ctx_run_shot_adj: the consistent (exact discrete) adjoint;
        returns J = 1/2 sum (syn - obs)^2 and adds dJ/d(vp, vs, rho) to grads"""
        rec = np.ascontiguousarray(rec_ijk, np.int32)
        s = IfosShot(src[0], src[1], src[2], src[6], src[3], src[4], src[5], shape, rec.shape[0],
                     rec.ctypes.data, d_obs.data_ptr(), 1)
        J = C.c_double(0.0)
        Gp = (C.c_void_p * 3)(*[t.data_ptr() for t in grads])
        check(lib().ifos_ctx_run_shot_adj(self.h, C.byref(s), every, _ptr(d_syn), Gp, C.byref(J), stream),
              "ifos_ctx_run_shot_adj")
        return J.value

    @staticmethod
    def adj_bytes(g, every=0, ntr=1) -> int:
        return int(lib().ifos_adj_bytes(g.ref, every, ntr))

    @staticmethod
    def td_bytes(g, every=0, L=0) -> int:
        return int(lib().ifos_td_bytes(g.ref, every, L))

    def fields(self):
        f = IfosFields()
        check(lib().ifos_ctx_fields(self.h, C.byref(f)), "ifos_ctx_fields")
        return f

    def dft_ptrs(self, back):
        arr = (C.c_void_p * 6)()
        check(lib().ifos_ctx_dft(self.h, back, arr), "ifos_ctx_dft")
        return list(arr)


class Lbfgs:
    """L-BFGS history on the device (lbfgs.c, lbfgs_save.c) with FP64 dots.
    Vectors are compact: numpar blocks of NX*NY*NZ cells, cell
    ((j-1)*NX + i-1)*NZ + k-1."""

    def __init__(self, g, numpar=2, bfgsnum=5):
        self.g, self.numpar, self.bfgsnum = g, numpar, bfgsnum
        self.n = numpar * g.nx * g.ny * g.nz
        h = C.c_void_p()
        check(lib().ifos_lbfgs_create(g.ref, numpar, bfgsnum, C.byref(h)), "ifos_lbfgs_create")
        self.h = h

    def savegrad(self, grad1, grad2=None):
        check(lib().ifos_lbfgs_savegrad(self.h, _ptr(grad1), _ptr(grad2)), "ifos_lbfgs_savegrad")

    def update(self, grad1, grad2, iteration) -> float:
        """replaces grad1 (and grad2) by the L-BFGS direction; returns h0"""
        h0 = C.c_double()
        check(lib().ifos_lbfgs_update(self.h, _ptr(grad1), _ptr(grad2), iteration, C.byref(h0)),
              "ifos_lbfgs_update")
        return h0.value

    def _slot(self, fn, w):
        f = getattr(lib(), fn)
        f.restype = C.c_void_p
        p = f(self.h, w)
        if not p:
            raise IfosError(f"{fn}: slot {w} out of range")
        return DBuf.wrap(p, (self.n,))

    def mod(self, w) -> DBuf:
        """slot w (1..bfgsnum) of bfgsmod, the model step the caller writes"""
        return self._slot("ifos_lbfgs_mod", w)

    def grad(self, w) -> DBuf:
        return self._slot("ifos_lbfgs_grad", w)

    def scale(self):
        out = (C.c_double * self.bfgsnum)()
        check(lib().ifos_lbfgs_scale(self.h, out), "ifos_lbfgs_scale")
        return np.array(out[:], np.float64)

    def close(self):
        if self.h:
            lib().ifos_lbfgs_destroy(self.h)
            self.h = None


def conjugate(g, grads, priors, cdf):
    """conjugategrad.c on device fields: grads[3], priors[6]; returns
    (beta[3], the six FP64 sums or None for cdf 1)"""
    Gp = (C.c_void_p * 3)(*[t.data_ptr() for t in grads])
    Pp = (C.c_void_p * 6)(*[t.data_ptr() for t in priors])
    beta = (C.c_float * 3)()
    sums = (C.c_double * 6)()
    check(lib().ifos_conjugate(g.ref, Gp, Pp, beta, cdf, sums), "ifos_conjugate")
    return np.array(beta[:], np.float32), (None if cdf == 1 else np.array(sums[:], np.float64))


def precon_grad(g, grads, damptype, src_xyz=None, rec_xyz=None, nglob=None, pos=None, stream=None):
    """precongrad.c on device gradients (three device-layout fields)"""
    src = np.ascontiguousarray(src_xyz if src_xyz is not None else np.zeros((0, 3)), np.int32)
    rec = np.ascontiguousarray(rec_xyz if rec_xyz is not None else np.zeros((0, 3)), np.int32)
    Gp = (C.c_void_p * 3)(*[t.data_ptr() for t in grads])
    ng = (C.c_int * 3)(*nglob) if nglob is not None else None
    ps = (C.c_int * 3)(*pos) if pos is not None else None
    check(lib().ifos_precon_grad(g.ref, Gp, damptype, src.ctypes.data_as(C.c_void_p), len(src),
                                 rec.ctypes.data_as(C.c_void_p), len(rec), ng, ps, stream), "ifos_precon_grad")


def modelupdate(g, grads, rho, pi, u, step, weight, vp0, vs0, rho0, bfgsmod=None, numpar=2, stream=None):
    """modelupdate.c on device fields; returns the model maxima (vp, vs, rho)"""
    Gp = (C.c_void_p * 3)(*[t.data_ptr() for t in grads])
    w = (C.c_float * 3)(*weight)
    mx = (C.c_float * 3)()
    check(lib().ifos_modelupdate(g.ref, Gp, _ptr(rho), _ptr(pi), _ptr(u), _ptr(bfgsmod), numpar,
                                 int(bfgsmod is not None), C.c_float(step), w, C.c_float(vp0), C.c_float(vs0),
                                 C.c_float(rho0), mx, stream), "ifos_modelupdate")
    return np.array(mx[:], np.float32)


def steplength(L2, step, lbfgs, it_group, teststep):
    """steplength.c: returns (L2[4], step[5]) after the call"""
    l2 = (C.c_float * 4)(*[float(x) for x in L2[:4]] + [0.0] * (4 - len(L2[:4])))
    st = (C.c_float * 5)(*[float(x) for x in step[:5]] + [0.0] * (5 - len(step[:5])))
    check(lib().ifos_steplength(l2, st, int(lbfgs), int(it_group), C.c_float(teststep)), "ifos_steplength")
    return np.array(l2[:], np.float32), np.array(st[:], np.float32)


def check_stability(g: "Geom", rho, pi, u, taup=None, taus=None, L=0, stream=None):
    """checkfd_ssg.c's stability test (ifos_check_stability) on device model
    arrays: returns (code, cmax, dt_stab) -- code OK or EUNSTABLE"""
    cm, ds = C.c_float(0.0), C.c_float(0.0)
    rc = lib().ifos_check_stability(g.ref, _ptr(rho), _ptr(pi), _ptr(u), _ptr(taup), _ptr(taus), int(L),
                                    C.byref(cm), C.byref(ds), stream)
    if rc not in (OK, EUNSTABLE):
        check(rc, "ifos_check_stability")
    return rc, cm.value, ds.value


def steplength_f64(L2, step, lbfgs, it_group, teststep, stream=None):
    """the FP64 parabola fit on the device (ifos_steplength_f64): L2[0..2],
    step[0..4] in; returns (L2[4], step[5]) as float64 after the call"""
    l2 = np.zeros(4, np.float64)
    l2[:min(len(L2), 4)] = np.asarray(L2, np.float64)[:4]
    st = np.zeros(5, np.float64)
    st[:min(len(step), 5)] = np.asarray(step, np.float64)[:5]
    d_l2, d_st = DBuf.from_numpy(l2), DBuf.from_numpy(st)
    check(lib().ifos_steplength_f64(C.c_void_p(d_l2.ptr()), C.c_void_p(d_st.ptr()), int(lbfgs), int(it_group), C.c_double(teststep), stream),
          "ifos_steplength_f64")
    out = d_l2.numpy(), d_st.numpy()
    d_l2.free()
    d_st.free()
    return out


class BatchContext:
    """The batched path: several shots of one model per launch (ifos_batch_*)."""

    def __init__(self, g: Geom, nshots: int, nf_max: int, ntr_max: int, precision=DFT_COMPAT, chunk=8):
        self.g = g
        self.nshots = nshots
        self.h = C.c_void_p()
        check(lib().ifos_batch_create(C.byref(self.h), g.ref, nshots, nf_max, ntr_max, precision, chunk),
              "ifos_batch_create")

    @staticmethod
    def bytes_needed(g: Geom, nshots, nf_max, ntr_max, precision=DFT_COMPAT, chunk=8):
        return int(lib().ifos_batch_bytes(g.ref, nshots, nf_max, ntr_max, precision, chunk))

    def close(self):
        if self.h:
            lib().ifos_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_model(self, d_rho, d_pi, d_u):
        check(lib().ifos_batch_set_model(self.h, _ptr(d_rho), _ptr(d_pi), _ptr(d_u)), "ifos_batch_set_model")

    def set_cpml(self, vppml, fpml):
        check(lib().ifos_batch_set_cpml(self.h, C.c_float(vppml), C.c_float(fpml)), "ifos_batch_set_cpml")

    def set_freqs(self, finv, tast=100):
        fv = np.asarray(finv, np.float32)
        check(lib().ifos_batch_set_freqs(self.h, fv.ctypes.data_as(C.c_void_p), len(fv), tast),
              "ifos_batch_set_freqs")

    def set_visco(self, d_taus, d_taup, eta, L=None):
        """as ShotContext.set_visco, for every shot of the batch"""
        e = _etas(eta)
        n = e.size if L is None else L
        check(lib().ifos_batch_set_visco_n(self.h, _ptr(d_taus), _ptr(d_taup), e.ctypes.data_as(C.c_void_p), n),
              "ifos_batch_set_visco_n")

    def run(self, srcs, rec_ijk, d_syn=None, grads=None, d_obs=None, shape=1, want_adjoint=True, use_graph=False):
        """srcs: list of (i, j, k, tshift, fc, amp, type), one per shot; rec_ijk:
        receivers shared by the shots, or a list with one array per shot;
        d_obs: list of device buffers [3][ntr][nt+1] (one per shot);
        d_syn: device buffer [ns][3][ntr][nt+1] or None.
        Returns the per-shot L2 misfits (FP64)."""
        ns = len(srcs)
        shots, recs = self._shots(srcs, rec_ijk, d_obs, shape, want_adjoint)
        l2 = (C.c_double * ns)()
        Gp = (C.c_void_p * 3)(*[t.data_ptr() for t in grads]) if grads is not None else None
        check(lib().ifos_batch_run(self.h, shots, ns, _ptr(d_syn), Gp, l2, int(use_graph)), "ifos_batch_run")
        return list(l2)

    def _shots(self, srcs, rec_ijk, d_obs=None, shape=1, want_adjoint=True):
        ns = len(srcs)
        recs = rec_ijk if isinstance(rec_ijk, (list, tuple)) else [rec_ijk] * ns
        recs = [np.ascontiguousarray(r, np.int32) for r in recs]
        shots = (IfosShot * ns)()
        for n, src in enumerate(srcs):
            shots[n] = IfosShot(src[0], src[1], src[2], src[6], src[3], src[4], src[5], shape, recs[n].shape[0],
                                recs[n].ctypes.data, d_obs[n].data_ptr() if d_obs is not None else None,
                                int(want_adjoint))
        return shots, recs

    def dft_ptrs(self, shot, back):
        arr = (C.c_void_p * 6)()
        check(lib().ifos_batch_dft(self.h, shot, back, arr), "ifos_batch_dft")
        return list(arr)

    def set_fused(self, on=True, ti=16, jc=0):
        check(lib().ifos_batch_set_fused(self.h, int(on), ti, jc), "ifos_batch_set_fused")

    def last_fused(self):
        return bool(lib().ifos_batch_last_fused(self.h))

    TILE_DEFAULTS = (4, 0, 0, 0, 0, 1, 1, 0)

    @classmethod
    def _tile(cls, cfg):
        cfg = tuple(cfg)
        return (C.c_int * len(cls.TILE_DEFAULTS))(*(cfg + cls.TILE_DEFAULTS[len(cfg):]))

    def set_tiles(self, cfg_v=None, cfg_s=None):
        """tile shapes {TI, jc, group, split, pipe, kb, sb, lds} of the velocity / stress
        marches (missing trailing fields take TILE_DEFAULTS)"""
        cv = self._tile(cfg_v) if cfg_v is not None else None
        cs = self._tile(cfg_s) if cfg_s is not None else None
        check(lib().ifos_batch_set_tiles(self.h, cv, cs), "ifos_batch_set_tiles")

    @classmethod
    def tune_candidates(cls):
        nf = len(cls.TILE_DEFAULTS)
        buf = (C.c_int * (64 * nf))()
        n = lib().ifos_batch_tune_candidates(buf, 64)
        return [tuple(buf[nf * c:nf * c + nf]) for c in range(n)]

    def autotune(self, ns=None, reps=3):
        """times every candidate tile shape; returns (best_v, best_s, table)"""
        cands = self.tune_candidates()
        mv = (C.c_float * len(cands))()
        ms = (C.c_float * len(cands))()
        bv, bs = C.c_int(), C.c_int()
        check(lib().ifos_batch_autotune(self.h, ns or self.nshots, reps, C.byref(bv), C.byref(bs), mv, ms),
              "ifos_batch_autotune")
        table = [dict(cfg=c, v_ms=round(mv[n], 4), s_ms=round(ms[n], 4)) for n, c in enumerate(cands)]
        buf = (C.c_float * 48)()
        nfz = lib().ifos_batch_tune_fused(self.h, buf, 16)
        self.fused_table = [dict(ti=int(buf[3 * c]), jc=int(buf[3 * c + 1]), ms=round(buf[3 * c + 2], 4))
                            for c in range(nfz)]
        return cands[bv.value], cands[bs.value], table


def levels_run(levels, srcs, rec_ijk, d_syn=None, shape=1, use_graph=True):
    """ifos_levels_run: forward solves of several grid levels together.
    levels: BatchContexts (one grid each); srcs[l], rec_ijk[l]: the shots and
    receivers of level l (as BatchContext.run takes them); d_syn[l]: device
    buffer [ns][3][ntr][nt+1] or None."""
    nlev = len(levels)
    keep = []
    shot_arrs = (C.POINTER(IfosShot) * nlev)()
    ns = (C.c_int * nlev)()
    for l, b in enumerate(levels):
        shots, recs = b._shots(srcs[l], rec_ijk[l], None, shape, False)
        keep.append((shots, recs))
        shot_arrs[l] = C.cast(shots, C.POINTER(IfosShot))
        ns[l] = len(srcs[l])
    lv = (C.c_void_p * nlev)(*[b.h.value for b in levels])
    sy = None
    if d_syn is not None:
        sy = (C.c_void_p * nlev)(*[t.data_ptr() if t is not None else None for t in d_syn])
    check(lib().ifos_levels_run(lv, nlev, shot_arrs, ns, sy, int(use_graph)), "ifos_levels_run")
