"""Python host side of libifos3d_hip.so (ctypes), used by the tests and bench.

Mirrors the native C ABI of include/ifos3d_hip.h.  Device memory comes from
torch (HIP allocations); the library is called with raw device pointers and
the current torch stream.  The library must be built (make -C ifos3d-dev_amd);
there is no fallback: a missing library raises.

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


class IfosShot(C.Structure):
    _fields_ = [("src_i", C.c_int), ("src_j", C.c_int), ("src_k", C.c_int), ("src_type", C.c_int),
                ("tshift", C.c_float), ("fc", C.c_float), ("amp", C.c_float), ("source_shape", C.c_int),
                ("ntr", C.c_int), ("rec_ijk", C.c_void_p), ("d_obs", C.c_void_p), ("want_adjoint", C.c_int)]


class IfosError(RuntimeError):
    pass


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
    return _lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().ifos_last_error().decode()
        raise IfosError(f"{what}: error {rc}: {msg}")


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def stream_ptr():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


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

    # ---- layout conversion (canonical numpy <-> device torch) ----
    def empty(self, n=None, dtype=None):
        import torch
        shp = self.dshape if n is None else (n,) + self.dshape
        return torch.zeros(shp, dtype=dtype or torch.float32, device="cuda")

    def to_dev(self, a, out=None):
        import torch
        a = np.asarray(a, dtype=np.float32)
        host = np.zeros(self.dshape, np.float32)
        host[:, :, KOFF - 2:KOFF - 2 + self.nz + 5] = a
        t = torch.from_numpy(host).cuda()
        if out is not None:
            out.copy_(t)
            return out
        return t

    def from_dev(self, t):
        h = t.detach().cpu().numpy()
        return np.ascontiguousarray(h[..., KOFF - 2:KOFF - 2 + self.nz + 5])


def fields_struct(ts):
    return IfosFields(*[t.data_ptr() for t in ts])


def model_struct(ts):
    return IfosModel(*[t.data_ptr() for t in ts])


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
                              _ptr(prof), C.byref(psi_struct(psi)) if psi is not None else None,
                              stream or stream_ptr()), "ifos_update_v")


def update_s(g, region, f, m, prof, psi, stream=None):
    check(lib().ifos_update_s(g.ref, region, C.byref(fields_struct(f)), C.byref(model_struct(m)),
                              _ptr(prof), C.byref(psi_struct(psi)) if psi is not None else None,
                              stream or stream_ptr()), "ifos_update_s")


def add_forces(g, f, m, d_src, d_sig, nsrc, nt, alpha=0.0, beta=0.0, stream=None):
    check(lib().ifos_add_forces(g.ref, C.byref(fields_struct(f)), C.byref(model_struct(m)), _ptr(d_src),
                                _ptr(d_sig), nsrc, nt, C.c_float(alpha), C.c_float(beta),
                                stream or stream_ptr()), "ifos_add_forces")


def add_adjoint(g, f, d_rec, d_adj, ntr, nt, d_chain=None, stream=None):
    check(lib().ifos_add_adjoint(g.ref, C.byref(fields_struct(f)), _ptr(d_rec), _ptr(d_chain), _ptr(d_adj), ntr,
                                 nt, stream or stream_ptr()), "ifos_add_adjoint")


def rec_chain(rec_ijk) -> np.ndarray:
    rec = np.ascontiguousarray(rec_ijk, np.int32)
    out = np.zeros((len(rec), 2), np.int32)
    check(lib().ifos_rec_chain(rec.ctypes.data_as(C.c_void_p), len(rec), out.ctypes.data_as(C.c_void_p)),
          "ifos_rec_chain")
    return out


def copy(d_dst, d_src, n, stream=None):
    check(lib().ifos_copy(_ptr(d_dst), _ptr(d_src), C.c_size_t(n), stream or stream_ptr()), "ifos_copy")


def psource(g, f, d_src, d_sig, nsrc, nt, stream=None):
    check(lib().ifos_psource(g.ref, C.byref(fields_struct(f)), _ptr(d_src), _ptr(d_sig), nsrc, nt,
                             stream or stream_ptr()), "ifos_psource")


def surface(g, f, m, stream=None):
    check(lib().ifos_surface(g.ref, C.byref(fields_struct(f)), C.byref(model_struct(m)),
                             stream or stream_ptr()), "ifos_surface")


def seismo(g, f, d_rec, ntr, ins, ns, d_sec, stream=None):
    check(lib().ifos_seismo(g.ref, C.byref(fields_struct(f)), _ptr(d_rec), ntr, ins, ns, _ptr(d_sec),
                            stream or stream_ptr()), "ifos_seismo")


def discfourier(g, f, F6, C6, finv, nt, ntast_, back, precision=DFT_COMPAT, stream=None):
    fv = np.asarray(finv, np.float32)
    Fp = (C.c_void_p * 6)(*[t.data_ptr() for t in F6])
    Cp = (C.c_void_p * 6)(*[t.data_ptr() for t in C6]) if C6 is not None else None
    check(lib().ifos_discfourier(g.ref, C.byref(fields_struct(f)), Fp, Cp, fv.ctypes.data_as(C.c_void_p),
                                 len(fv), nt, ntast_, back, precision, stream or stream_ptr()),
          "ifos_discfourier")


def gradient_F(g, Ff6, Fb6, m, finv, grads, stream=None):
    fv = np.asarray(finv, np.float32)
    Fp = (C.c_void_p * 6)(*[t.data_ptr() for t in Ff6])
    Bp = (C.c_void_p * 6)(*[t.data_ptr() for t in Fb6])
    Gp = (C.c_void_p * 3)(*[t.data_ptr() for t in grads])
    check(lib().ifos_gradient_F(g.ref, Fp, Bp, C.byref(model_struct(m)), fv.ctypes.data_as(C.c_void_p),
                                len(fv), Gp, stream or stream_ptr()), "ifos_gradient_F")


def matcopy_wrap(g, d_m, stream=None):
    check(lib().ifos_matcopy_wrap(g.ref, _ptr(d_m), stream or stream_ptr()), "ifos_matcopy_wrap")


def av_mat(g, rho, u, out6, stream=None):
    check(lib().ifos_av_mat(g.ref, _ptr(rho), _ptr(u), *[_ptr(t) for t in out6], stream or stream_ptr()),
          "ifos_av_mat")


def residual(g, d_obs, d_syn, d_adj, ntr, ns, stream=None) -> float:
    l2 = C.c_double(0.0)
    check(lib().ifos_residual(g.ref, _ptr(d_obs), _ptr(d_syn), _ptr(d_adj), ntr, ns, C.byref(l2),
                              stream or stream_ptr()), "ifos_residual")
    return l2.value


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
        check(lib().ifos_ctx_set_model(self.h, _ptr(d_rho), _ptr(d_pi), _ptr(d_u), stream or stream_ptr()),
              "ifos_ctx_set_model")

    def set_cpml(self, vppml, fpml):
        check(lib().ifos_ctx_set_cpml(self.h, C.c_float(vppml), C.c_float(fpml)), "ifos_ctx_set_cpml")

    def set_freqs(self, finv, tast=100):
        fv = np.asarray(finv, np.float32)
        check(lib().ifos_ctx_set_freqs(self.h, fv.ctypes.data_as(C.c_void_p), len(fv), tast),
              "ifos_ctx_set_freqs")

    def run_shot(self, src, rec_ijk, d_syn, grads, d_obs=None, shape=1, want_adjoint=True, stream=None):
        """src = (i, j, k, tshift, fc, amp, type); returns the L2 misfit (FP64 sum)."""
        rec = np.ascontiguousarray(rec_ijk, np.int32)
        s = IfosShot(src[0], src[1], src[2], src[6], src[3], src[4], src[5], shape, rec.shape[0],
                     rec.ctypes.data, d_obs.data_ptr() if d_obs is not None else None, int(want_adjoint))
        l2 = C.c_double(0.0)
        Gp = (C.c_void_p * 3)(*[t.data_ptr() for t in grads]) if grads is not None else None
        check(lib().ifos_ctx_run_shot(self.h, C.byref(s), _ptr(d_syn), Gp, C.byref(l2),
                                      stream or stream_ptr()), "ifos_ctx_run_shot")
        return l2.value

    def fields(self):
        f = IfosFields()
        check(lib().ifos_ctx_fields(self.h, C.byref(f)), "ifos_ctx_fields")
        return f

    def dft_ptrs(self, back):
        arr = (C.c_void_p * 6)()
        check(lib().ifos_ctx_dft(self.h, back, arr), "ifos_ctx_dft")
        return list(arr)
