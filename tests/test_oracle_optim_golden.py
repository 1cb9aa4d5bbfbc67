"""The optimiser oracle (oracle/optim_oracle.py) against the reference's own
optimiser routines: tests/golden/optim_golden.npz holds seeded inputs and the
outputs of lbfgs.c, lbfgs_save.c, conjugategrad.c, modelupdate.c,
precongrad.c and steplength.c compiled from source (written by
tests/golden/make_optim_golden.py).

With its sums formed as the reference forms them (dot="f32": FP32, index
order) the oracle reproduces every output bit for bit; precon_grad's tapers
come from a double `exp` that numpy and glibc may round differently in the
last double bit, so there the bar is one float ulp.  The FP64 reductions the
device uses (dot="f64") stay within the float-sum error of the reference."""
import os

import numpy as np
import pytest

import optim_oracle as oo
from helpers import max_ulp

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "optim_golden.npz")
SHAPE = (6, 5, 7)
BFGSNUM = 3
WEIGHT = (1.0, 1.0, 0.5)
SCALE0 = (4000.0, 2300.0, 2000.0)


@pytest.fixture(scope="module")
def F():
    return np.load(FIX, allow_pickle=False)


def _lbfgs_run(F, numpar, dot):
    n = int(np.prod(SHAPE))
    lb = oo.LbfgsOracle(numpar * n, BFGSNUM, dot=dot)
    rho, pi, u = [a.copy() for a in F[f"lbfgs_np{numpar}_model_in"]]
    out = []
    for it in range(1, 6):
        g = [a.copy() for a in F[f"lbfgs_np{numpar}_g_{it}"]]
        flat = np.concatenate([g[q].ravel() for q in range(numpar)])
        if it == 1:
            lb.savegrad(flat)
            d = g
        else:
            q, _ = lb.update(flat, it)
            d = [q[:n].reshape(SHAPE), (q[n:].reshape(SHAPE) if numpar > 1 else g[1]), g[2]]
        d = [np.ascontiguousarray(a, np.float32) for a in d]
        rec = {"dir": np.stack(d), "scale": np.float32(lb.scale), "hist": lb.grad.copy()}
        _, mod = oo.modelupdate(d, rho, pi, u, F[f"lbfgs_np{numpar}_step_{it}"], WEIGHT, *SCALE0, lbfgs=True,
                                numpar=numpar)
        lb.mod[lb.slot(it)] = mod
        rec["mod"] = lb.mod.copy()
        rec["model"] = np.stack([rho, pi, u])
        out.append(rec)
    return out


@pytest.mark.parametrize("numpar", [1, 2])
def test_lbfgs_sequence_bit_exact(F, numpar):
    """five iterations of lbfgs_savegrad / lbfgs / L-BFGS modelupdate"""
    for it, r in enumerate(_lbfgs_run(F, numpar, "f32"), start=1):
        key = f"lbfgs_np{numpar}"
        assert np.array_equal(r["dir"], F[f"{key}_dir_{it}"]), f"direction, it_group {it}"
        assert np.array_equal(r["scale"], F[f"{key}_scale_{it}"]), f"bfgsscale, it_group {it}"
        assert np.array_equal(r["hist"], F[f"{key}_hist_{it}"]), f"bfgsgrad, it_group {it}"
        assert np.array_equal(r["mod"], F[f"{key}_mod_{it}"]), f"bfgsmod, it_group {it}"
        assert np.array_equal(r["model"], F[f"{key}_model_{it}"]), f"model, it_group {it}"


@pytest.mark.parametrize("numpar", [1, 2])
def test_lbfgs_fp64_sums_near_reference(F, numpar):
    """the device's choice (FP64 dot products) moves the directions by no
    more than the float-sum error of the reference's dots"""
    for it, r in enumerate(_lbfgs_run(F, numpar, "f64"), start=1):
        want = F[f"lbfgs_np{numpar}_dir_{it}"].astype(np.float64)
        err = np.linalg.norm(r["dir"] - want) / np.linalg.norm(want)
        assert err < 1e-5, f"it_group {it}: rel-L2 {err:.2e}"


def test_conjugate_bit_exact(F):
    priors = [np.zeros(SHAPE, np.float32) for _ in range(6)]
    for it in range(1, 5):
        g = [a.copy() for a in F[f"cg_g_{it}"]]
        assert np.array_equal(np.stack(priors), F[f"cg_prior_in_{it}"])
        beta, _ = oo.conjugate(g, priors, 1 if it == 1 else 0, dot="f32")
        assert np.array_equal(beta, F[f"cg_beta_{it}"]), f"beta, iteration {it}: {beta} vs {F[f'cg_beta_{it}']}"
        assert np.array_equal(np.stack(g), F[f"cg_dir_{it}"]), f"direction, iteration {it}"
        assert np.array_equal(np.stack(priors), F[f"cg_prior_{it}"]), f"priors, iteration {it}"
    # the cases cover the clamp of beta[0] from below and from above
    assert F["cg_beta_3"][0] == 0.0 and F["cg_beta_4"][0] == np.float32(0.5)


@pytest.mark.parametrize("case", [0, 1])
def test_modelupdate_normalised_bit_exact(F, case):
    a = [x.copy() for x in F[f"mu_{case}_in"]]
    g, (rho, pi, u) = a[:3], a[3:]
    oo.modelupdate(g, rho, pi, u, 0.03, WEIGHT, *SCALE0, lbfgs=False)
    assert np.array_equal(np.stack(g + [rho, pi, u]), F[f"mu_{case}_out"])


@pytest.mark.parametrize("damptype", [0, 1, 2, 3])
@pytest.mark.parametrize("free_surf", [0, 1])
def test_precon_grad_matches_reference(F, damptype, free_surf):
    g = [a.copy() for a in F["pre_in"]]
    src = [(4, 3, 5), (2, 6, 1)]                     # iround(srcpos / DX) of the fixture's metres
    rec = [(2, 1, 3), (4, 2, 5), (5, 6, 7)]
    oo.precon_grad(g, damptype, src, rec, (SHAPE[1], SHAPE[0], SHAPE[2]), 2, free_surf)
    want = F[f"pre_{damptype}_{free_surf}"]
    assert max_ulp(np.stack(g), want) <= 1, f"DAMPTYPE {damptype}, FREE_SURF {free_surf}"
    if damptype == 3:
        assert not np.any(want)                       # the quirk: DAMPTYPE 3 zeroes the gradient


@pytest.mark.parametrize("case", range(5))
def test_steplength_bit_exact(F, case):
    x = F[f"step_{case}_in"]
    L2, step, lbfgs, it_group, teststep = x[:3], x[3:8], int(x[8]), int(x[9]), x[10]
    got_l2, got_step = oo.steplength(L2, step, lbfgs, it_group, teststep)
    assert np.array_equal(got_step, F[f"step_{case}_step"]), (got_step, F[f"step_{case}_step"])
    assert np.array_equal(got_l2[3:], F[f"step_{case}_L2"][3:])
