"""The adjoint of a model step on the GPU through the C-ABI
(wam_gpu_step_adjoint_dev, wam_gpu_propags_adjoint_dev, the wind map and its
adjoint), on deep-water grids of configuration 5's kind (IPHYS = 0, 24 x 25):

  - the step adjoint against the oracle's taped adjoint of the same step at
    the same state with the REAL*4 restatement's values (liborc_ad32: the
    FP32 build's values bit for bit, hence its branches, which the GPU's
    forward reproduces; derivatives in FP64; its FP64 twin is pinned by
    finite differences and the dot-product identity in
    test_oracle_adjoint.py), as relative L2 errors of each input's bar.  A
    comparison with the FP64 tape would differ wherever FP32 and FP64 take
    different sides of a kink (the limiter's floor at nearly empty bins);
  - PROPAGS's adjoint against its forward by the dot-product identity
    <P x, y> = <x, P^T y>, with and without currents and with COUNTER > 1;
  - ensemble members equal single runs bit for bit, and a checkpointed
    window equals the store-every-state window bit for bit;
  - the window's gradient of a cost of Hs misfits (through
    torch.autograd) against FP64 central finite differences of the oracle's
    forward window on 50 random directions of the winds."""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, shared_math
from wam_amd import WamGpu, workloads as W
from wam_amd.adjoint import WindowModel, energy_weights, hs_from_spectra, window
from wam_amd.model import checkpoint_recomputed_steps

pytestmark = pytest.mark.gpu


def dev():
    import torch
    return torch.device("cuda:0")


def T(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev())


def setup(ny=8, nx=10, seed=7, nsteps=4, **kw):
    cfg, mask, f, u, v, obs, tu, tv = W.c5_case(ny=ny, nx=nx, seed=seed, nsteps=nsteps, nobs=6, **kw)
    g = WamGpu(cfg, mask)
    counter, _ = g.prepare_propagation(f["depth"])
    return cfg, mask, f, u, v, obs, g, counter


def spun_up_state(torch, g, cfg, f, u, v, nspin=2):
    """A state after nspin GPU steps from the JONSWAP cold start."""
    n = g.nsea
    fa = T(W.jonswap_spectrum(cfg, f["udir"]).ravel())
    fb = torch.zeros_like(fa)
    tauw, ustar, z0 = (torch.zeros(n, device=dev()) for _ in range(3))
    ro, ws = torch.full((n,), 1.225, device=dev()), torch.zeros(n, device=dev())
    u10, udir = torch.zeros(n, device=dev()), torch.zeros(n, device=dev())
    for i in range(nspin):
        g.wind_uv_dev(T(u[i]), T(v[i]), u10, udir)
        g.propags_dev(fa, fb)
        g.implsch_dev(fb, u10, udir, tauw, ustar, z0, ro, ws)
        fa, fb = fb, fa
    torch.cuda.synchronize()
    return fa, tauw, ro, ws


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_wind_map_and_its_adjoint():
    import torch
    cfg, mask, f, u, v, obs, g, _ = setup()
    rng = np.random.default_rng(3)
    uu = rng.normal(0, 8, 4096).astype(np.float32)
    vv = rng.normal(0, 8, 4096).astype(np.float32)
    uu[:8] = 0.0
    vv[:4] = 0.0
    n = uu.size
    u10, udir = torch.zeros(n, device=dev()), torch.zeros(n, device=dev())
    g.wind_uv_dev(T(uu), T(vv), u10, udir)
    spd = np.sqrt(uu * uu + vv * vv)                       # IEEE single: the reference's SQRT(UU**2 + VV**2)
    d = np.where(spd != 0, shared_math(11, uu, vv), 0.0).astype(np.float32)
    d = np.where(d < 0, (d + np.float32(2 * math.pi)).astype(np.float32), d)     # ZPI = 2.*PI in REAL*4
    assert np.array_equal(u10.cpu().numpy().view(np.int32), spd.view(np.int32))
    assert np.array_equal(udir.cpu().numpy().view(np.int32), d.view(np.int32))
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    ub, vb = torch.zeros(n, device=dev()), torch.zeros(n, device=dev())
    g.wind_uv_adjoint_dev(T(uu), T(vv), T(a), T(b), ub, vb)
    u64, v64 = uu.astype(np.float64), vv.astype(np.float64)
    r2 = u64 * u64 + v64 * v64
    s = np.sqrt(r2)
    with np.errstate(divide="ignore", invalid="ignore"):
        eu = np.where(s > 0, a * u64 / s + b * v64 / r2, 0.0)
        ev = np.where(s > 0, a * v64 / s - b * u64 / r2, 0.0)
    assert np.allclose(ub.cpu().numpy(), eu, rtol=1e-5, atol=1e-6)
    assert np.allclose(vb.cpu().numpy(), ev, rtol=1e-5, atol=1e-6)
    assert np.all(ub.cpu().numpy()[:4] == 0) and np.all(vb.cpu().numpy()[:4] == 0)
    g.close()


@pytest.mark.parametrize("case", ["deep", "currents", "counter"])
def test_propags_adjoint_is_the_transpose(case):
    import torch
    if case == "deep":
        cfg, mask, f, *_ = W.c5_case(ny=12, nx=14, seed=5, nsteps=1, nobs=4)
        g = WamGpu(cfg, mask)
        counter, _ = g.prepare_propagation(f["depth"])
    elif case == "currents":
        cfg, mask, f = W.c2_case(ny=12, nx=14, seed=5, kl=24, ml=25)
        g = WamGpu(cfg, mask)
        counter, _ = g.prepare_propagation(f["depth"], f["u"], f["v"])
    else:
        cfg, mask, f, *_ = W.c5_case(ny=20, nx=24, seed=5, nsteps=1, nobs=4, dlat=0.05, dlon=0.05)
        g = WamGpu(cfg, mask)
        counter, _ = g.prepare_propagation(f["depth"])
        assert counter > 1
    print(f"[case] {case} NSEA={g.nsea} COUNTER={counter}", flush=True)
    nb = g.kl * g.ml * g.nsea
    rng = np.random.default_rng(11)
    for _ in range(3):
        x = rng.random(nb).astype(np.float32)
        y = rng.standard_normal(nb).astype(np.float32)
        px, pty = torch.zeros(nb, device=dev()), torch.zeros(nb, device=dev())
        g.propags_dev(T(x), px)
        g.propags_adjoint_dev(T(y), pty)
        torch.cuda.synchronize()
        lhs = float(np.dot(px.cpu().numpy().astype(np.float64), y.astype(np.float64)))
        rhs = float(np.dot(x.astype(np.float64), pty.cpu().numpy().astype(np.float64)))
        assert abs(lhs - rhs) <= 2e-6 * (abs(lhs) + np.abs(y).sum()), (lhs, rhs)
    g.close()


def oracle_step_adjoint(cfg, mask, f, fl3, tauw, u, v, ybar):
    o = Oracle(cfg, mask, ad32=True)
    o.prepare_propagation(f["depth"])
    z0 = np.full(tauw.size, 1e-3)
    _, xbar = o.step_ad(fl3, tauw, z0, u, v, f["depth"], ybar=ybar)
    return xbar


@pytest.mark.parametrize("which,kl", [("fl3", 24), ("tauw", 24), ("z0", 24), ("all", 24), ("all", 30)])
def test_step_adjoint_matches_the_oracle(which, kl):
    """Bars of a step's inputs for random bars of its outputs (FL3', TAUW',
    Z0' all, or one at a time), against the oracle's REAL*4-valued tape; 24
    directions run the split kernels, 30 (not a multiple of 4) the
    one-workgroup-per-point kernel."""
    import torch
    cfg, mask, f, u, v, obs, g, counter = setup(ny=8, nx=10, seed=7, nsteps=4, kl=kl)
    n, nb = g.nsea, g.kl * g.ml * g.nsea
    fl3, tauw, ro, ws = spun_up_state(torch, g, cfg, f, u, v)
    rng = np.random.default_rng(17)
    ybar = rng.standard_normal(nb + 2 * n)
    if which == "fl3":
        ybar[nb:] = 0
    elif which == "tauw":
        ybar[:nb] = 0
        ybar[nb + n:] = 0
    elif which == "z0":
        ybar[:nb + n] = 0
    # a bar of FL3' proportional to the spectrum's scale keeps the three parts comparable
    ybar[:nb] *= 1e3
    ui, vi = u[2], v[2]
    u10, udir = torch.zeros(n, device=dev()), torch.zeros(n, device=dev())
    g.wind_uv_dev(T(ui), T(vi), u10, udir)
    fbar = T(ybar[:nb])
    tb, zb = T(ybar[nb:nb + n]), T(ybar[nb + n:])
    fbo, u10b, udirb, tbo = (torch.zeros_like(x) for x in (fbar, u10, udir, tb))
    g.step_adjoint_dev(fl3, u10, udir, tauw, ro, ws, fbar, fbo, u10b, udirb, tbo, tauw_bar=tb, z0_bar=zb)
    ub, vb = torch.zeros(n, device=dev()), torch.zeros(n, device=dev())
    g.wind_uv_adjoint_dev(T(ui), T(vi), u10b, udirb, ub, vb)
    torch.cuda.synchronize()
    xbar = oracle_step_adjoint(cfg, mask, f, fl3.cpu().numpy(), tauw.cpu().numpy(), ui, vi, ybar)
    parts = dict(fl3=(fbo, xbar[:nb]), tauw=(tbo, xbar[nb:nb + n]), u=(ub, xbar[nb + 2 * n:nb + 3 * n]),
                 v=(vb, xbar[nb + 3 * n:]))
    errs = {k: rel(a.cpu().numpy(), b) for k, (a, b) in parts.items()}
    print(f"[case] {which} KL={kl} NSEA={n} COUNTER={counter} rel errors {errs}", flush=True)
    # FL3's bar to FP32 round-off; TAUW's and the winds' bars pass through
    # AIRSEA's iteration derivatives, accumulated in FP32
    bars = dict(fl3=1e-6, tauw=2e-4, u=1e-4, v=1e-4)
    for k, e in errs.items():
        assert e < bars[k], (k, e)
    g.close()


def test_ensemble_members_equal_single_runs():
    import torch
    cfg, mask, f, u, v, obs, g, counter = setup(ny=8, nx=10, seed=9, nsteps=4)
    n, nb = g.nsea, g.kl * g.ml * g.nsea
    f1, t1, ro, ws = spun_up_state(torch, g, cfg, f, u, v, 1)
    f2, t2, _, _ = spun_up_state(torch, g, cfg, f, u, v, 3)
    rng = np.random.default_rng(5)
    fb = T(rng.standard_normal(2 * nb))
    tb = T(rng.standard_normal(2 * n))
    u10, udir = torch.zeros(2 * n, device=dev()), torch.zeros(2 * n, device=dev())
    g.wind_uv_dev(T(np.concatenate([u[1], u[3]])), T(np.concatenate([v[1], v[3]])), u10, udir)
    fl3 = torch.cat([f1, f2])
    tauw = torch.cat([t1, t2])
    ro2, ws2 = torch.cat([ro, ro]), torch.cat([ws, ws])
    out = [torch.zeros(2 * nb, device=dev())] + [torch.zeros(2 * n, device=dev()) for _ in range(3)]
    g.step_adjoint_dev(fl3, u10, udir, tauw, ro2, ws2, fb, out[0], out[1], out[2], out[3], tauw_bar=tb, nmem=2)
    for m in range(2):
        o = [torch.zeros(nb, device=dev())] + [torch.zeros(n, device=dev()) for _ in range(3)]
        sl, sb = slice(m * n, (m + 1) * n), slice(m * nb, (m + 1) * nb)
        g.step_adjoint_dev(fl3[sb].contiguous(), u10[sl].contiguous(), udir[sl].contiguous(),
                           tauw[sl].contiguous(), ro, ws, fb[sb].contiguous(), o[0], o[1], o[2], o[3],
                           tauw_bar=tb[sl].contiguous())
        torch.cuda.synchronize()
        assert torch.equal(o[0], out[0][sb]), m
        for i in range(1, 4):
            assert torch.equal(o[i], out[i][sl]), (m, i)
    g.close()


def window_grad(torch, g, cfg, f, u, v, obs, nsteps, obs_every, nsnap, target):
    n = g.nsea
    fl3 = T(W.jonswap_spectrum(cfg, f["udir"]).ravel())
    tauw = torch.zeros(n, device=dev())
    uu = T(u[:nsteps]).requires_grad_(True)
    vv = T(v[:nsteps]).requires_grad_(True)
    model = WindowModel(g, nsteps, obs, obs_every=obs_every, nsnap=nsnap)
    spec = window(uu, vv, fl3, tauw, model)
    hs = hs_from_spectra(spec, energy_weights(g))
    j = ((hs - target) ** 2).sum()
    j.backward()
    torch.cuda.synchronize()
    return float(j), uu.grad.clone(), vv.grad.clone(), hs.detach(), model


def test_checkpointed_window_equals_store_all():
    import torch
    cfg, mask, f, u, v, obs, g, counter = setup(ny=8, nx=10, seed=13, nsteps=10)
    target = torch.full((5, obs.size), 0.1, device=dev())
    j0, gu0, gv0, _, m0 = window_grad(torch, g, cfg, f, u, v, obs, 10, 2, None, target)
    j1, gu1, gv1, _, m1 = window_grad(torch, g, cfg, f, u, v, obs, 10, 2, 2, target)
    assert j0 == j1
    assert torch.equal(gu0, gu1) and torch.equal(gv0, gv1)
    assert float(gu0.abs().sum()) > 0
    # forward steps: the window once plus what the schedule recomputes after
    # its first adjoint step (the stored-every-state schedule recomputes none)
    assert m0.forward_steps == 10
    assert m1.forward_steps <= 10 + checkpoint_recomputed_steps(10, 2)
    g.close()


def oracle_window_cost(o, cfg, f, u, v, obs, nsteps, obs_every, weights, target, fl0):
    """J of the FP64 oracle's window: the wind map, PROPAGS, IMPLSCH per step."""
    n = u.shape[1]
    fl3 = fl0.astype(np.float64)
    tauw = np.zeros(n)
    z0 = np.zeros(n)
    j = 0.0
    mijs = []
    for i in range(nsteps):
        u10 = np.sqrt(u[i] * u[i] + v[i] * v[i])
        udir = np.where(u10 != 0, np.arctan2(u[i], v[i]), 0.0)
        udir = np.where(udir < 0, udir + 2 * math.pi, udir)
        r = o.implsch(o.propags(fl3), u10, udir, tauw, z0, f["depth"])
        fl3, tauw, z0 = r["fl3"], r["tauw"], r["z0"]
        mijs.append(r["mij"])
        if (i + 1) % obs_every == 0:
            k = (i + 1) // obs_every - 1
            e = weights @ fl3.reshape(-1, n)[:, obs]
            hs = 4.0 * np.sqrt(np.maximum(e, 1e-12))
            j += float(((hs - target[k]) ** 2).sum())
    return j, np.stack(mijs)


def test_window_gradient_matches_fp64_finite_differences():
    """dJ/d(u, v) of J = sum of squared Hs misfits at 6 points every 3 steps
    over 12 steps, along 50 random directions: the GPU adjoint (FP32 state,
    checkpointed with 3 slots) against central differences of the FP64
    oracle's window, |g.d - FD| <= 1e-3 |FD| where the perturbations flip no
    cut-off index MIJ.  The step h = 1e-5 m/s per unit of the direction: the
    model is piecewise smooth (MAX/MIN in the limiter, the floor FLMINFR
    chosen by the rounded wind speed), and differences over 1e-3 m/s already
    straddle kinks on a grid of this size."""
    import torch
    nsteps, every = 12, 3
    cfg, mask, f, u, v, obs, g, counter = setup(ny=10, nx=12, seed=21, nsteps=nsteps)
    w = energy_weights(g)
    fl0 = W.jonswap_spectrum(cfg, f["udir"]).ravel()
    o = Oracle(cfg, mask, fp64=True)
    o.prepare_propagation(f["depth"])
    rng = np.random.default_rng(23)
    target = np.full((nsteps // every, obs.size), 1.0) + 0.1 * rng.standard_normal((nsteps // every, obs.size))
    j, gu, gv, hs, _ = window_grad(torch, g, cfg, f, u, v, obs, nsteps, every, 3, T(target))
    gu, gv = gu.cpu().numpy().astype(np.float64), gv.cpu().numpy().astype(np.float64)
    h = 1e-5
    ok, flips = 0, 0
    worst = 0.0
    for d in range(50):
        du = rng.standard_normal(u.shape)
        dv = rng.standard_normal(v.shape)
        jp, mp = oracle_window_cost(o, cfg, f, u + h * du, v + h * dv, obs, nsteps, every, w, target, fl0)
        jm, mm = oracle_window_cost(o, cfg, f, u - h * du, v - h * dv, obs, nsteps, every, w, target, fl0)
        if not np.array_equal(mp, mm):
            flips += 1
            continue
        fd = (jp - jm) / (2 * h)
        ad = float((gu * du).sum() + (gv * dv).sum())
        err = abs(ad - fd) / max(abs(fd), 1e-30)
        worst = max(worst, err)
        ok += 1
    print(f"[case] NSEA={g.nsea} J={j:.6g} directions compared {ok}, skipped for MIJ flips {flips}, "
          f"worst relative error {worst:.3g}", flush=True)
    assert ok >= 40
    assert worst <= 1e-3
    g.close()
