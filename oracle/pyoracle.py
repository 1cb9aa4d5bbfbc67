"""ctypes access to the oracle (TEST INFRASTRUCTURE).

Loads oracle/lib/liboracle_{f32,f64}.so (the CPU restatement) and, where it
has been built, oracle/_ref/libifos_ref.so (the reference's own routines
compiled from its sources).  Only the test suite, the smoke check and the
benchmark's CPU-baseline leg import this module, and only as the checker.

Arrays are numpy arrays in the canonical layout of ifos_oracle.h:
shape (ny+5, nx+5, nz+5), index [j+2, i+2, k+2].
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
REFLIB = os.path.join(HERE, "_ref", "libifos_ref.so")


class OgGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("fw", C.c_int),
                ("free_surf", C.c_int), ("fdcoeff", C.c_int),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float), ("dt", C.c_float),
                ("nt", C.c_int), ("time", C.c_float)]


MAX_L = 4  # OG_MAX_L: relaxation mechanisms


def _visco_struct(real):
    class OgVisco(C.Structure):
        _fields_ = [("r", C.c_void_p * 6), ("taus", C.c_void_p), ("taup", C.c_void_p),
                    ("tausipjp", C.c_void_p), ("tausjpkp", C.c_void_p), ("tausipkp", C.c_void_p),
                    ("eta", real * MAX_L), ("L", C.c_int)]
    return OgVisco


def _etas(eta):
    """a scalar eta (one mechanism) or a sequence of one eta per mechanism"""
    e = [float(x) for x in np.atleast_1d(np.asarray(eta, np.float64))]
    if not 1 <= len(e) <= MAX_L:
        raise ValueError(f"1..{MAX_L} relaxation mechanisms")
    return e


OgVisco32 = _visco_struct(C.c_float)
OgVisco64 = _visco_struct(C.c_double)


class OgSource(C.Structure):
    _fields_ = [("i", C.c_int), ("j", C.c_int), ("k", C.c_int),
                ("tshift", C.c_float), ("fc", C.c_float), ("amp", C.c_float), ("type", C.c_int)]


@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    fw: int
    dx: float
    dy: float
    dz: float
    dt: float
    nt: int
    fdcoeff: int = 2
    free_surf: int = 0

    @property
    def shape(self):
        return (self.ny + 5, self.nx + 5, self.nz + 5)

    @property
    def time(self):
        # TIME = (NT-1)*DT in float, ifos3d.c:177
        return float(np.float32(self.nt - 1) * np.float32(self.dt))

    def og(self) -> OgGrid:
        return OgGrid(self.nx, self.ny, self.nz, self.fw, self.free_surf, self.fdcoeff,
                      self.dx, self.dy, self.dz, self.dt, self.nt, self.time)

    def psi_shapes(self):
        f2 = 2 * self.fw
        return [(self.ny, f2, self.nz)] * 3 + [(f2, self.nx, self.nz)] * 3 + [(self.ny, self.nx, f2)] * 3

    def interior(self, a):
        """view of [1..N]^3 of a canonical array"""
        return a[3:3 + self.ny, 3:3 + self.nx, 3:3 + self.nz]


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


class Oracle:
    """The CPU restatement; precision 'f32' (compat) or 'f64' (FP64 oracle)."""

    def __init__(self, precision: str = "f32"):
        path = os.path.join(LIBDIR, f"liboracle_{precision}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        self.lib = C.CDLL(path)
        self.dtype = np.float32 if precision == "f32" else np.float64
        self.precision = precision

    def threads(self) -> int:
        """threads of the oracle's OpenMP cell loops"""
        self.lib.og_threads.restype = C.c_int
        return int(self.lib.og_threads())

    def zeros(self, g: Grid, n=None):
        return np.zeros(g.shape if n is None else (n,) + g.shape, dtype=self.dtype)

    def box(self, g: Grid):
        b = (C.c_int * 6)()
        og = g.og()
        self.lib.og_interior_box(C.byref(og), b)
        return list(b)

    def cpml_coeff(self, g: Grid, vppml: float, fpml: float):
        n = 2 * g.fw + 1
        prof = [np.zeros(n, self.dtype) for _ in range(18)]
        alpha = [np.zeros(n, self.dtype) for _ in range(6)]
        og = g.og()
        self.lib.og_cpml_coeff(C.byref(og), C.c_float(vppml), C.c_float(fpml), _ptrs(prof), _ptrs(alpha))
        return prof, alpha

    def matcopy_wrap(self, g: Grid, m):
        og = g.og()
        self.lib.og_matcopy_wrap(C.byref(og), _p(m))

    def av_mat(self, g: Grid, rho, u):
        out = [self.zeros(g) for _ in range(6)]  # uipjp ujpkp uipkp rjp rkp rip
        og = g.og()
        self.lib.og_av_mat(C.byref(og), _p(rho), _p(u), None, _p(out[0]), _p(out[1]), _p(out[2]),
                           None, None, None, _p(out[3]), _p(out[4]), _p(out[5]))
        return out

    def av_tau(self, g: Grid, rho, u, taus):
        """the arithmetic averages of taus av_mat forms when L > 0 (av_mat.c:56-60)"""
        tmp = [self.zeros(g) for _ in range(9)]
        og = g.og()
        self.lib.og_av_mat(C.byref(og), _p(rho), _p(u), _p(taus), *[_p(a) for a in tmp])
        return tmp[3:6]

    def visco(self, r, taus, taup, tausipjp, tausjpkp, tausipkp, eta, L=None):
        """an og_visco for update_s_visco / update_s_cpml_visco (keeps the
        arrays alive).  eta: scalar (one mechanism) or one per mechanism;
        r[n]: the L mechanisms' fields back to back (shape (L,) + grid or
        the grid's shape for L = 1)"""
        v = (OgVisco32 if self.precision == "f32" else OgVisco64)()
        e = _etas(eta)
        for n in range(6):
            assert r[n].size == len(e) * taus.size and r[n].flags.c_contiguous
            v.r[n] = r[n].ctypes.data
        v.taus, v.taup = taus.ctypes.data, taup.ctypes.data
        v.tausipjp, v.tausjpkp, v.tausipkp = tausipjp.ctypes.data, tausjpkp.ctypes.data, tausipkp.ctypes.data
        for l, x in enumerate(e):
            v.eta[l] = x
        v.L = len(e) if L is None else L
        v._keep = (r, taus, taup, tausipjp, tausjpkp, tausipkp)
        return v

    def update_s_visco(self, g, box, f, pi, u, uipjp, ujpkp, uipkp, V):
        og = g.og()
        b = (C.c_int * 6)(*box)
        self.lib.og_update_s_visco(C.byref(og), b, *[_p(a) for a in f], _p(pi), _p(u), _p(uipjp), _p(ujpkp),
                                   _p(uipkp), C.byref(V))

    def update_s_cpml_visco(self, g, box, f, pi, u, uipjp, ujpkp, uipkp, prof, psi, V):
        og = g.og()
        b = (C.c_int * 6)(*box)
        self.lib.og_update_s_cpml_visco(C.byref(og), b, *[_p(a) for a in f], _p(pi), _p(u), _p(uipjp),
                                        _p(ujpkp), _p(uipkp), _ptrs(prof), _ptrs(psi), C.byref(V))

    @staticmethod
    def _sources(srcs):
        arr = (OgSource * max(1, len(srcs)))()
        for n, s in enumerate(srcs):
            arr[n] = OgSource(*s)
        return arr

    def wavelet(self, g: Grid, srcs, shape: int):
        sig = np.zeros((len(srcs), g.nt + 2), np.float32)
        og = g.og()
        self.lib.og_wavelet(C.byref(og), self._sources(srcs), len(srcs), shape, _p(sig))
        return sig

    def update_v(self, g, box, nt, f, rjp, rkp, rip, srcs, sx, sy, sz, back):
        og = g.og()
        b = (C.c_int * 6)(*box)
        self.lib.og_update_v(C.byref(og), b, nt, *[_p(a) for a in f], _p(rjp), _p(rkp), _p(rip),
                             self._sources(srcs), _p(sx), _p(sy), _p(sz), len(srcs), back)

    def update_v_cpml(self, g, box, nt, f, rjp, rkp, rip, srcs, sx, sy, sz, back, prof, psi):
        og = g.og()
        b = (C.c_int * 6)(*box)
        self.lib.og_update_v_cpml(C.byref(og), b, nt, *[_p(a) for a in f], _p(rjp), _p(rkp), _p(rip),
                                  self._sources(srcs), _p(sx), _p(sy), _p(sz), len(srcs), back,
                                  _ptrs(prof), _ptrs(psi))

    def update_s(self, g, box, f, pi, u, uipjp, ujpkp, uipkp):
        og = g.og()
        b = (C.c_int * 6)(*box)
        self.lib.og_update_s_elastic(C.byref(og), b, *[_p(a) for a in f], _p(pi), _p(u),
                                     _p(uipjp), _p(ujpkp), _p(uipkp))

    def update_s_cpml(self, g, box, f, pi, u, uipjp, ujpkp, uipkp, prof, psi):
        og = g.og()
        b = (C.c_int * 6)(*box)
        self.lib.og_update_s_cpml_elastic(C.byref(og), b, *[_p(a) for a in f], _p(pi), _p(u),
                                          _p(uipjp), _p(ujpkp), _p(uipkp), _ptrs(prof), _ptrs(psi))

    def psource(self, g, nt, sxx, syy, szz, srcs, sig):
        og = g.og()
        self.lib.og_psource(C.byref(og), nt, _p(sxx), _p(syy), _p(szz), self._sources(srcs), _p(sig), len(srcs))

    def surface(self, g, u, pi, f):
        og = g.og()
        vx, vy, vz, sxx, syy, szz, sxy, syz, sxz = f
        self.lib.og_surface_elastic(C.byref(og), _p(u), _p(pi), _p(sxx), _p(syy), _p(szz), _p(sxy),
                                    _p(syz), _p(vx), _p(vy), _p(vz))

    def seismo(self, g, ins, rec_ijk, sec, f):
        """seismo_ssg.c: sample ins of sec[3] (ntr, ns+1) from vx, vy, vz"""
        og = g.og()
        rec = np.ascontiguousarray(rec_ijk, np.int32)
        self.lib.og_seismo(C.byref(og), ins, rec.shape[0], rec.ctypes.data_as(C.c_void_p), _p(sec[0]), _p(sec[1]),
                           _p(sec[2]), sec[0].shape[1] - 1, _p(f[0]), _p(f[1]), _p(f[2]))

    def residual(self, g, obs, syn):
        ntr, ns1 = obs.shape
        diff = np.zeros_like(obs)
        l2 = C.c_float(0.0)
        og = g.og()
        self.lib.og_residual(C.byref(og), _p(obs), _p(syn), _p(diff), ntr, ns1 - 1, C.byref(l2))
        return diff, l2.value

    def discfourier(self, g, nt, vx, vy, vz, Fr, Fi, finv, ntast, back):
        """Fr, Fi: lists of 3 arrays of shape (nf,)+g.shape"""
        og = g.og()
        fv = np.asarray(finv, np.float32)
        self.lib.og_discfourier(C.byref(og), nt, _p(vx), _p(vy), _p(vz), _p(Fr[0]), _p(Fr[1]), _p(Fr[2]),
                                _p(Fi[0]), _p(Fi[1]), _p(Fi[2]), _p(fv), len(fv), ntast, back)

    def hess_F(self, g, Ff, Fb, hess, rho, pi, u, finv, ntr_hess):
        """hess_F.c; Ff[6] [nf][...], Fb[6] [nf*(ntr_hess+1)][...] (receiver n,
        frequency l at n*nf + l - 1), hess[3] accumulated in place"""
        og = g.og()
        fv = np.asarray(finv, np.float32)
        self.lib.og_hess_F(C.byref(og), *[_p(a) for a in Ff], *[_p(a) for a in Fb], *[_p(a) for a in hess],
                           _p(rho), _p(pi), _p(u), fv.ctypes.data_as(C.c_void_p), len(fv), ntr_hess)

    def gradient_F(self, g, Ff, Fb, grads, rho, pi, u, finv):
        """Ff, Fb: 6 arrays each (Re x y z, Im x y z) of shape (nf,)+g.shape"""
        og = g.og()
        fv = np.asarray(finv, np.float32)
        self.lib.og_gradient_F(C.byref(og), *[_p(a) for a in Ff], *[_p(a) for a in Fb],
                               _p(grads[0]), _p(grads[1]), _p(grads[2]), _p(rho), _p(pi), _p(u),
                               _p(fv), len(fv))

    def ntast(self, fmax, tast, dt):
        self.lib.og_ntast.restype = C.c_int
        return self.lib.og_ntast(C.c_float(fmax), tast, C.c_float(dt))

    def run_shot(self, g: Grid, rho, pi, u, src, rec_ijk, finv, tast=100, vppml=4000.0, fpml=50.0,
                 shape=1, obs=None, want_adjoint=True, want_fields=False, visco=None, want_history=False,
                 grads=None):
        """visco = (taus, taup, eta): the viscoelastic update instead of the
        elastic one, eta a scalar (one relaxation mechanism) or one per
        mechanism.  grads: three
        canonical arrays the shot's gradient is added into in place (the
        reference accumulates over shots, ifos3d.c:974); fresh zeros if None"""
        class Io(C.Structure):
            _fields_ = [("rho", C.c_void_p), ("pi", C.c_void_p), ("u", C.c_void_p), ("src", OgSource),
                        ("source_shape", C.c_int), ("ntr", C.c_int), ("rec_ijk", C.c_void_p),
                        ("obs_x", C.c_void_p), ("obs_y", C.c_void_p), ("obs_z", C.c_void_p),
                        ("finv", C.c_void_p), ("nf", C.c_int), ("tast", C.c_int),
                        ("vppml", C.c_float), ("fpml", C.c_float), ("want_adjoint", C.c_int),
                        ("syn_x", C.c_void_p), ("syn_y", C.c_void_p), ("syn_z", C.c_void_p),
                        ("l2", C.c_void_p), ("grad1", C.c_void_p), ("grad2", C.c_void_p),
                        ("grad3", C.c_void_p), ("ffwd", C.c_void_p), ("fadj", C.c_void_p),
                        ("taus", C.c_void_p), ("taup", C.c_void_p),
                        ("eta", (C.c_float if self.precision == "f32" else C.c_double) * MAX_L), ("L", C.c_int),
                        ("vfwd", C.c_void_p), ("vadj", C.c_void_p)]
        rec = np.ascontiguousarray(rec_ijk, np.int32)
        ntr = rec.shape[0]
        fv = np.asarray(finv, np.float32)
        syn = [np.zeros((ntr, g.nt + 1), np.float32) for _ in range(3)]
        if grads is None:
            grads = [self.zeros(g) for _ in range(3)]
        for a in grads:
            assert a.dtype == self.dtype and a.shape == g.shape and a.flags.c_contiguous
        l2 = np.zeros(1, np.float32)
        ff = np.zeros((6, len(fv)) + g.shape, self.dtype) if want_fields else None
        fa = np.zeros((6, len(fv)) + g.shape, self.dtype) if want_fields else None
        hf = np.zeros((g.nt, 3) + g.shape, self.dtype) if want_history else None
        ha = np.zeros((g.nt, 3) + g.shape, self.dtype) if want_history else None
        o = [np.ascontiguousarray(a, np.float32) for a in obs] if obs is not None else [None] * 3
        io = Io(_p(rho).value, _p(pi).value, _p(u).value, OgSource(*src), shape, ntr, rec.ctypes.data,
                o[0].ctypes.data if o[0] is not None else None,
                o[1].ctypes.data if o[1] is not None else None,
                o[2].ctypes.data if o[2] is not None else None,
                fv.ctypes.data, len(fv), tast, vppml, fpml, int(want_adjoint),
                syn[0].ctypes.data, syn[1].ctypes.data, syn[2].ctypes.data, l2.ctypes.data,
                grads[0].ctypes.data, grads[1].ctypes.data, grads[2].ctypes.data,
                ff.ctypes.data if ff is not None else None, fa.ctypes.data if fa is not None else None)
        if want_history:
            io.vfwd, io.vadj = hf.ctypes.data, ha.ctypes.data
        if visco is not None:
            taus = np.ascontiguousarray(visco[0], self.dtype)
            taup = np.ascontiguousarray(visco[1], self.dtype)
            e = _etas(visco[2])
            io.taus, io.taup, io.L = taus.ctypes.data, taup.ctypes.data, len(e)
            for l, x in enumerate(e):
                io.eta[l] = x
        og = g.og()
        rc = self.lib.og_run_shot(C.byref(og), C.byref(io))
        if rc != 0:
            raise FloatingPointError("oracle: simulation unstable (NaN probe)")
        out = dict(syn=syn, l2=float(l2[0]), grads=grads)
        if want_fields:
            out["ffwd"] = ff
            out["fadj"] = fa
        if want_history:
            out["vfwd"] = hf
            out["vadj"] = ha
        return out

    def run_shot_adj(self, g: Grid, rho, pi, u, src, rec_ijk, obs, vppml=4000.0, fpml=50.0, shape=1,
                     grads=None):
        """og_run_shot_adj: the consistent (exact discrete) adjoint of the
        elastic forward scheme.  obs: 3 arrays (ntr, nt+1).  Returns dict(syn,
        J, grads) with J = 1/2 sum (syn - obs)^2 and grads = dJ/d(vp, vs, rho)
        on the canonical grid (added into `grads` when given)."""
        class AdjIo(C.Structure):
            _fields_ = [("rho", C.c_void_p), ("pi", C.c_void_p), ("u", C.c_void_p), ("src", OgSource),
                        ("source_shape", C.c_int), ("ntr", C.c_int), ("rec_ijk", C.c_void_p),
                        ("obs_x", C.c_void_p), ("obs_y", C.c_void_p), ("obs_z", C.c_void_p),
                        ("vppml", C.c_float), ("fpml", C.c_float),
                        ("syn_x", C.c_void_p), ("syn_y", C.c_void_p), ("syn_z", C.c_void_p),
                        ("J", C.c_void_p), ("grad1", C.c_void_p), ("grad2", C.c_void_p), ("grad3", C.c_void_p)]
        rec = np.ascontiguousarray(rec_ijk, np.int32)
        ntr = rec.shape[0]
        o = [np.ascontiguousarray(a, np.float32) for a in obs]
        syn = [np.zeros((ntr, g.nt + 1), np.float32) for _ in range(3)]
        if grads is None:
            grads = [self.zeros(g) for _ in range(3)]
        m = [np.ascontiguousarray(a, self.dtype) for a in (rho, pi, u)]
        J = np.zeros(1, np.float64)
        io = AdjIo(_p(m[0]).value, _p(m[1]).value, _p(m[2]).value, OgSource(*src), shape, ntr, rec.ctypes.data,
                   o[0].ctypes.data, o[1].ctypes.data, o[2].ctypes.data, vppml, fpml,
                   syn[0].ctypes.data, syn[1].ctypes.data, syn[2].ctypes.data, J.ctypes.data,
                   grads[0].ctypes.data, grads[1].ctypes.data, grads[2].ctypes.data)
        og = g.og()
        rc = self.lib.og_run_shot_adj(C.byref(og), C.byref(io))
        if rc == -1:
            raise FloatingPointError("oracle: simulation unstable (NaN probe)")
        if rc:
            raise ValueError("og_run_shot_adj: elastic, no free surface and an explosive source only")
        return dict(syn=syn, J=float(J[0]), grads=grads)

    def td_gradient(self, g: Grid, vfwd, vadj, rho, pi, u, grads):
        """og_td_gradient: the time-domain gradient from velocity histories
        [nt][3] (run_shot(..., want_history=True)), added to grads"""
        arr = [np.ascontiguousarray(a, self.dtype) for a in (vfwd, vadj, rho, pi, u)]
        og = g.og()
        self.lib.og_td_gradient(C.byref(og), *[_p(a) for a in arr], *[_p(t) for t in grads])


# ----------------------------------------------------------------------
class Ref:
    """The reference's own routines, compiled from its sources (oracle/_ref)."""

    FIELDS = ["vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "syz", "sxz", "rho", "pi", "u",
              "uipjp", "ujpkp", "uipkp", "rjp", "rkp", "rip", "grad1", "grad2", "grad3",
              "rxx", "ryy", "rzz", "rxy", "ryz", "rxz", "taus", "taup", "tausipjp", "tausjpkp", "tausipkp"]

    @staticmethod
    def available() -> bool:
        return os.path.exists(REFLIB)

    @staticmethod
    def outseis(path, comp, sec, rec_xyz, rec_id, src_xyz, dx, dy, dz, dt, ndt):
        """outseis.c (SU) of section [ntr][ns+1] into path"""
        lib = C.CDLL(REFLIB)
        s = np.ascontiguousarray(sec, np.float32)
        rx = np.ascontiguousarray(rec_xyz, np.int32)
        rid = np.ascontiguousarray(rec_id, np.int32)
        sp = None if src_xyz is None else np.ascontiguousarray(src_xyz, np.float32)
        lib.rg_outseis(os.fsencode(path), comp, _p(s), s.shape[0], s.shape[1] - 1, _p(rx), _p(rid),
                       None if sp is None else _p(sp), C.c_float(dx), C.c_float(dy), C.c_float(dz),
                       C.c_float(dt), ndt)

    @staticmethod
    def readseis(prefix, ishot, comp, ntr, ns):
        """readseis.c of "<prefix>_v?.su.shot<ishot>.0" -> [ntr][ns+1]"""
        lib = C.CDLL(REFLIB)
        out = np.zeros((ntr, ns + 1), np.float32)
        lib.rg_readseis(os.fsencode(prefix), ishot, comp, ntr, ns, _p(out))
        return out

    def __init__(self, g: Grid, vppml=4000.0, fpml=50.0, nfmax=1, nsrc=1, ntr=1):
        self.lib = C.CDLL(REFLIB)
        self.lib.rg_residual.restype = C.c_float
        self.lib.rg_run_shot.restype = C.c_float
        self.g = g
        self.nfmax = nfmax
        self.nsrc = nsrc
        self.ntr = ntr
        self.lib.rg_setup(g.nx, g.ny, g.nz, C.c_float(g.dx), C.c_float(g.dy), C.c_float(g.dz),
                          C.c_float(g.dt), g.nt, g.fw, g.fdcoeff, g.free_surf, C.c_float(vppml),
                          C.c_float(fpml), nfmax, nsrc, ntr)

    def fid(self, name):
        return self.FIELDS.index(name)

    def set(self, name, a):
        a = np.ascontiguousarray(a, np.float32)
        self.lib.rg_set(self.fid(name), _p(a))

    def get(self, name):
        a = np.zeros(self.g.shape, np.float32)
        self.lib.rg_get(self.fid(name), _p(a))
        return a

    def psi_set(self, n, p):
        p = np.ascontiguousarray(p, np.float32)
        self.lib.rg_psi_set(n, _p(p))

    def psi_get(self, n):
        shp = self.g.psi_shapes()[n % 9]
        p = np.zeros(shp, np.float32)
        self.lib.rg_psi_get(n, _p(p))
        return p

    def prof(self):
        """24 profiles: per axis K a b K_half a_half b_half alpha alpha_half"""
        out = []
        for n in range(24):
            a = np.zeros(2 * self.g.fw + 1, np.float32)
            self.lib.rg_prof_get(n, _p(a))
            out.append(a)
        return out

    def box(self):
        b = (C.c_int * 6)()
        self.lib.rg_box(b)
        return list(b)

    def F_set(self, n, m, a):
        a = np.ascontiguousarray(a, np.float32)
        self.lib.rg_F_set(n, m, _p(a))

    def F_get(self, n, m):
        a = np.zeros(self.g.shape, np.float32)
        self.lib.rg_F_get(n, m, _p(a))
        return a

    def set_freqs(self, finv):
        fv = np.asarray(finv, np.float32)
        self.lib.rg_set_freqs(_p(fv), len(fv))

    def set_source(self, s, i, j, k, tshift, fc, amp, typ):
        self.lib.rg_set_source(s, i, j, k, C.c_float(tshift), C.c_float(fc), C.c_float(amp), typ)

    def wavelet(self, shape):
        self.lib.rg_make_wavelet(shape)
        out = np.zeros((self.nsrc, self.g.nt + 2), np.float32)
        self.lib.rg_get_signals(_p(out))
        return out

    def set_signals(self, sig):
        sig = np.ascontiguousarray(sig, np.float32)
        self.lib.rg_set_signals(_p(sig))

    def set_receivers(self, rec_ijk):
        r = np.ascontiguousarray(rec_ijk, np.int32)
        self.lib.rg_set_receivers(_p(r))

    def set_adjoint(self, ax, ay, az):
        a = [np.ascontiguousarray(x, np.float32) for x in (ax, ay, az)]
        self.lib.rg_set_adjoint(*[_p(x) for x in a])

    def sections(self):
        s = [np.zeros((self.ntr, self.g.nt + 1), np.float32) for _ in range(3)]
        self.lib.rg_get_sections(*[_p(x) for x in s])
        return s

    def set_sections(self, syn):
        s = [np.ascontiguousarray(x, np.float32) for x in syn]
        self.lib.rg_set_sections(*[_p(x) for x in s])

    def get_adjoint(self):
        a = [np.zeros((self.ntr, self.g.nt + 2), np.float32) for _ in range(3)]
        self.lib.rg_get_adjoint(*[_p(x) for x in a])
        return a

    def residual(self, obs):
        o = [np.ascontiguousarray(x, np.float32) for x in obs]
        return float(self.lib.rg_residual(*[_p(x) for x in o], C.c_float(0.0)))

    def call(self, fn, *args):
        getattr(self.lib, "rg_" + fn)(*args)

    def set_visco(self, L, eta):
        self.lib.rg_set_visco(L, C.c_float(eta))

    def run_shot(self, shape=1, tast=100, obs=None, want_adjoint=True):
        o = [np.ascontiguousarray(x, np.float32) for x in obs] if obs is not None else None
        ptrs = [_p(x) for x in o] if o is not None else [None, None, None]
        return float(self.lib.rg_run_shot(shape, tast, *ptrs, int(want_adjoint)))


class RefOptim:
    """The reference's optimiser routines (lbfgs.c, lbfgs_save.c,
    conjugategrad.c, modelupdate.c, precongrad.c, steplength.c) compiled
    from its sources (oracle/_ref), on compact interior arrays [ny][nx][nz]
    (float32, the reference's l order)."""

    def __init__(self, ny, nx, nz):
        self.lib = C.CDLL(REFLIB)
        self.shape = (ny, nx, nz)

    def _c(self, a):
        a = np.ascontiguousarray(a, np.float32)
        assert a.shape == self.shape, a.shape
        return a.copy()

    def savegrad(self, g1, g2, numpar):
        ny, nx, nz = self.shape
        out = np.zeros(numpar * ny * nx * nz, np.float32)
        self.lib.rg_lbfgs_savegrad(nx, ny, nz, numpar, _p(self._c(g1)), _p(self._c(g2)), _p(out))
        return out

    def lbfgs(self, g1, g2, scale, mod, hist, iteration, numpar):
        """returns (direction (g1, g2), scale, hist) after lbfgs.c"""
        ny, nx, nz = self.shape
        a, b = self._c(g1), self._c(g2)
        sc = np.ascontiguousarray(scale, np.float32).copy()
        md = np.ascontiguousarray(mod, np.float32)
        hs = np.ascontiguousarray(hist, np.float32).copy()
        self.lib.rg_lbfgs(nx, ny, nz, numpar, len(sc), _p(a), _p(b), _p(sc), _p(md), _p(hs), iteration)
        return (a, b), sc, hs

    def conjugate(self, grads, priors, iteration, cdf):
        ny, nx, nz = self.shape
        g = [self._c(x) for x in grads]
        p = [self._c(x) for x in priors]
        pp = (C.c_void_p * 6)(*[x.ctypes.data for x in p])
        beta = np.zeros(3, np.float32)
        self.lib.rg_conjugate(nx, ny, nz, _p(g[0]), _p(g[1]), _p(g[2]), pp, _p(beta), iteration, cdf)
        return g, p, beta

    def modelupdate(self, grads, rho, pi, u, mod, step, it_group, lbfgs, numpar, weight, vp0, vs0, rho0):
        ny, nx, nz = self.shape
        g = [self._c(x) for x in grads]
        m = [self._c(x) for x in (rho, pi, u)]
        bfgsnum = mod.shape[0] if mod is not None else 1
        md = np.ascontiguousarray(mod if mod is not None else np.zeros((1, 1)), np.float32).copy()
        w = np.asarray(weight, np.float32)
        self.lib.rg_modelupdate(nx, ny, nz, _p(g[0]), _p(g[1]), _p(g[2]), _p(m[0]), _p(m[1]), _p(m[2]), _p(md),
                                C.c_float(step), it_group, int(lbfgs), numpar, bfgsnum, _p(w), C.c_float(vp0),
                                C.c_float(vs0), C.c_float(rho0))
        return g, m, (md if mod is not None else None)

    def precon_grad(self, grads, src_m, rec_xyz, dh, fw, damptype, free_surf):
        """src_m: [nsrc][3] source positions in metres; rec_xyz: [ntr][3] grid indices"""
        ny, nx, nz = self.shape
        g = [self._c(x) for x in grads]
        s = np.ascontiguousarray(np.asarray(src_m, np.float32).reshape(-1, 3))
        r = np.ascontiguousarray(np.asarray(rec_xyz, np.int32).reshape(-1, 3))
        self.lib.rg_precon_grad(nx, ny, nz, _p(g[0]), _p(g[1]), _p(g[2]), len(s), _p(s), len(r), _p(r),
                                C.c_float(dh[0]), C.c_float(dh[1]), C.c_float(dh[2]), fw, damptype, free_surf)
        return g

    def steplength(self, L2, step, iteration, it_group, lbfgs, teststep):
        l2 = np.zeros(4, np.float32)
        l2[:len(L2)] = L2
        st = np.zeros(5, np.float32)
        st[:len(step)] = step
        self.lib.rg_steplength(_p(l2), _p(st), iteration, it_group, int(lbfgs), C.c_float(teststep))
        return l2, st
