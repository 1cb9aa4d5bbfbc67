"""A torch.autograd.Function over the device path, checked by finite
differences of the misfit along model perturbations.

IfosMisfit.forward runs every shot through ifos_ctx_run_shot and returns the
time-domain misfit J = sum over traces and samples of (DT * running sum of
(syn - obs))^2, the quantity residual.c:43-53 accumulates; backward returns
minus the gradient gradient_F.c accumulates (grad1..3 = vp, vs, rho), since
the reference's gradient is the descent direction (lbfgs.c:46).

That gradient is a frequency-domain estimate (gradient_F.c:122-139: the
products of the forward and adjoint DFT fields at the chosen frequencies,
each weighted 1 / (4 pi^2 f^2)), not the exact derivative of the discrete J:
with the frequencies on the DFT grid 1 / (NT DT) that covers the source
band it is proportional to dJ/dm with a normalisation of its own.  The check
is therefore on proportionality: along perturbations of vp (two places), vs
and rho, the finite-difference derivative divided by <backward, delta> must
be positive and the same within 25% (measured spread about 17%, ratio about
0.01).  The misfit itself is compared with the device's FP64-reduced L2.

The same check runs with the checkpointed time-domain gradient
(ifos_ctx_run_shot_td), the full-band counterpart of those products: the
deviation from the exact derivative comes from the adjoint itself (the
reference's adjoint source and CPML), not from the frequency sampling, so
its spread is of the same size.

The consistent mode (ifos_ctx_run_shot_adj, the exact discrete adjoint of
the same forward scheme) is the exact derivative of its misfit
J = 1/2 sum (syn - obs)^2: wrapped the same way, <backward, delta> agrees
with central finite differences of J -- computed by the FP64 oracle
(og_run_shot_adj) so that FP32 rounding does not enter the difference --
to better than 1e-3 relative along a seeded smooth perturbation (seed 6,
1 % rms of vp, vs and rho) and along the single-parameter directions.
"""
import numpy as np
import pytest

from helpers import small_grid

pytestmark = pytest.mark.gpu


class _Problem:
    def __init__(self):
        import ifosgpu as ig
        self.ig = ig
        g = small_grid(nx=40, ny=40, nz=40, fw=8, nt=500, fdcoeff=2, dh=2.0, dt=2e-4)
        self.g = g
        self.G = ig.Geom(g.nx, g.ny, g.nz, g.fw, g.dx, g.dy, g.dz, g.dt, g.nt, g.fdcoeff, g.free_surf)
        shp = g.shape
        self.J, self.I, self.K = np.meshgrid(*(np.arange(s) - 2.0 for s in shp), indexing="ij")
        one = np.ones(shp)
        self.base = (4000.0 * one, 2300.0 * one, 2000.0 * one)
        blob = self.bump(24, 20, 20, 4)
        true = (4000.0 * (1 + 0.05 * blob), 2300.0 * (1 + 0.05 * blob), 2000.0 * one)
        self.srcs = [(14, 12, 14, 0.0, 50.0, 1.0, 1), (27, 12, 26, 0.0, 50.0, 1.0, 1)]
        self.rec = np.array([[10 + 2 * n, 11, 10 + 2 * m] for n in range(10) for m in range(10)], np.int32)
        self.finv = [n / (g.nt * g.dt) for n in range(1, 17)]  # 10 .. 160 Hz on the DFT grid
        self.ctx = ig.ShotContext(self.G, len(self.finv), len(self.rec), ig.DFT_COMPAT)
        self.ctx.set_cpml(4000.0, 50.0)
        self.ctx.set_freqs(self.finv, 100)
        self.d_syn = ig.DBuf((3, len(self.rec), g.nt + 1))
        self.ctx.set_model(*self.upload(true))
        self.obs = []
        for s in self.srcs:
            self.ctx.run_shot(s, self.rec, self.d_syn, None, want_adjoint=False)
            self.obs.append(self.d_syn.numpy().astype(np.float64))
        self.obs_f32 = [o.astype(np.float32) for o in self.obs]
        self.d_obs = [ig.DBuf.from_numpy(o) for o in self.obs_f32]

    def bump(self, cj, ci, ck, r):
        return np.exp(-((self.J - cj) ** 2 + (self.I - ci) ** 2 + (self.K - ck) ** 2) / (2.0 * r * r))

    def full(self, interior):
        """an interior (NY, NX, NZ) array on the canonical grid, halos by edge copy"""
        return np.pad(np.asarray(interior, np.float64), ((3, 2), (3, 2), (3, 2)), mode="edge")

    def upload(self, m):
        vp, vs, rho = m
        return [self.G.to_dev(np.asarray(a, np.float32)) for a in (rho, rho * vp * vp, rho * vs * vs)]

    def misfit(self, m):
        self.ctx.set_model(*self.upload(m))
        J = 0.0
        for n, s in enumerate(self.srcs):
            self.ctx.run_shot(s, self.rec, self.d_syn, None, want_adjoint=False)
            r = self.d_syn.numpy().astype(np.float64)[:, :, 1:] - self.obs[n][:, :, 1:]
            R = self.g.dt * np.cumsum(r, axis=2)
            J += float((R * R).sum())
        return J

    def misfit_and_gradient(self, m, mode="dft"):
        ig = self.ig
        self.ctx.set_model(*self.upload(m))
        grads = [self.G.empty() for _ in range(3)]
        l2 = 0.0
        for n, s in enumerate(self.srcs):
            if mode == "td":
                l2 += self.ctx.run_shot_td(s, self.rec, self.d_syn, grads, self.d_obs[n])
            elif mode == "adj":
                l2 += self.ctx.run_shot_adj(s, self.rec, self.d_syn, grads, self.d_obs[n])
            else:
                l2 += self.ctx.run_shot(s, self.rec, self.d_syn, grads, d_obs=self.d_obs[n])
        ig.sync()
        return l2, [self.g.interior(self.G.from_dev(t)).astype(np.float64) for t in grads]

    def misfit_fp64(self, m):
        """J = 1/2 sum (syn - obs)^2 of every shot by the FP64 oracle"""
        from pyoracle import Oracle
        o = Oracle("f64")
        vp, vs, rho = m
        model = [rho, rho * vp * vp, rho * vs * vs]
        return sum(o.run_shot_adj(self.g, *model, s, self.rec, list(self.obs_f32[n]))["J"]
                   for n, s in enumerate(self.srcs))


def _autograd_fn(problem, mode):
    """mode dft / td: backward is minus the reference's gradient (a descent
    direction, lbfgs.c:46); adj: backward is dJ/dm itself"""
    import torch
    sign = 1.0 if mode == "adj" else -1.0

    class IfosMisfit(torch.autograd.Function):
        @staticmethod
        def forward(ctx, vp, vs, rho):
            m = tuple(problem.full(t.detach().cpu().numpy()) for t in (vp, vs, rho))
            l2, grads = problem.misfit_and_gradient(m, mode)
            ctx.save_for_backward(*[torch.from_numpy(sign * gq) for gq in grads])
            return torch.tensor(l2, dtype=torch.float64)

        @staticmethod
        def backward(ctx, g_out):
            return tuple(g_out * t for t in ctx.saved_tensors)

    return IfosMisfit


@pytest.mark.parametrize("mode", ["dft", "td"])
def test_autograd_gradient_is_proportional_to_finite_differences(mode):
    """mode dft: gradient_F over 16 DFT frequencies (ifos_ctx_run_shot);
    td: the checkpointed time-domain gradient (ifos_ctx_run_shot_td)"""
    import torch
    p = _Problem()
    g = p.g
    fn = _autograd_fn(p, mode)
    interior = lambda a: torch.tensor(g.interior(a).copy(), dtype=torch.float64, requires_grad=True)
    vp, vs, rho = (interior(a) for a in p.base)
    J = fn.apply(vp, vs, rho)
    J.backward()
    J_host = p.misfit(p.base)
    print(f"{mode}: misfit: device {J.item():.6e}, host FP64 {J_host:.6e}", flush=True)
    assert J.item() == pytest.approx(J_host, rel=1e-5)
    dirs = {"vp at the anomaly": (p.bump(24, 20, 20, 4), 0, 0), "vp elsewhere": (p.bump(28, 26, 14, 3), 0, 0),
            "vs at the anomaly": (0, p.bump(24, 20, 20, 4), 0), "rho at the anomaly": (0, 0, p.bump(24, 20, 20, 4))}
    scale = (4000.0, 2300.0, 2000.0)
    ratios = []
    for name, d in dirs.items():
        dm = [1e-3 * scale[q] * np.asarray(d[q], np.float64) * np.ones(g.shape) for q in range(3)]
        fd = (p.misfit(tuple(p.base[q] + dm[q] for q in range(3))) -
              p.misfit(tuple(p.base[q] - dm[q] for q in range(3)))) / 2.0
        ad = sum(float((t.grad.numpy() * g.interior(dm[q])).sum()) for q, t in enumerate((vp, vs, rho)))
        ratios.append(fd / ad)
        print(f"{name:20s} FD {fd:+.4e}  <backward, delta> {ad:+.4e}  ratio {fd / ad:.4e}", flush=True)
    ratios = np.array(ratios)
    assert (ratios > 0).all()
    assert ratios.max() / ratios.min() < 1.25
    p.ctx.close()


def test_autograd_consistent_gradient_matches_finite_differences():
    """the consistent adjoint through the autograd wrapper: |FD - <backward,
    delta>| / |FD| < 1e-3, FD of J by the FP64 oracle"""
    import torch
    from helpers import det_uniform
    import scipy.ndimage as nd
    p = _Problem()
    g = p.g
    fn = _autograd_fn(p, "adj")
    interior = lambda a: torch.tensor(g.interior(a).copy(), dtype=torch.float64, requires_grad=True)
    vp, vs, rho = (interior(a) for a in p.base)
    J = fn.apply(vp, vs, rho)
    J.backward()
    J64 = p.misfit_fp64(p.base)
    print(f"adj: misfit: device {J.item():.8e}, FP64 oracle {J64:.8e}", flush=True)
    assert J.item() == pytest.approx(J64, rel=1e-4)

    def smooth(seed):
        d = nd.gaussian_filter(det_uniform(g.shape, seed, -1.0, 1.0).astype(np.float64), 3.0)
        return d / np.sqrt((d ** 2).mean())

    dirs = {"seed 6, 1 % rms of vp, vs, rho": [0.01 * smooth(6 + 1000 * q) for q in range(3)],
            "vp at the anomaly": [0.01 * p.bump(24, 20, 20, 4), 0 * p.base[0], 0 * p.base[0]],
            "vs at the anomaly": [0 * p.base[0], 0.01 * p.bump(24, 20, 20, 4), 0 * p.base[0]],
            "rho at the anomaly": [0 * p.base[0], 0 * p.base[0], 0.01 * p.bump(24, 20, 20, 4)]}
    eps = 1e-3
    for name, d in dirs.items():
        dm = [p.base[q] * d[q] for q in range(3)]
        fd = (p.misfit_fp64(tuple(p.base[q] + eps * dm[q] for q in range(3))) -
              p.misfit_fp64(tuple(p.base[q] - eps * dm[q] for q in range(3)))) / (2 * eps)
        ad = sum(float((t.grad.numpy() * g.interior(dm[q])).sum()) for q, t in enumerate((vp, vs, rho)))
        err = abs(fd - ad) / abs(fd)
        print(f"{name:32s} FD {fd:+.6e}  <backward, delta> {ad:+.6e}  rel {err:.2e}", flush=True)
        assert err < 1e-3, name
    p.ctx.close()
