"""Ensembles on the GPU through the C-ABI: members advanced together
(wam_gpu_propags_ensemble_dev, wam_gpu_implsch_ensemble_dev, the captured
time loop wam_gpu_ensemble_run_dev) must equal plain runs of each member
(wam_gpu_propags_dev, wam_gpu_implsch_dev) bit for bit, and those are
checked against the oracle; every launch shape the autotuner may choose must
give the same results; the captured loop must give the eager loop's results on
every replay, with the program's own eager work between replays."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from wam_amd import WamGpu, workloads as W

pytestmark = pytest.mark.gpu

NMEM = 4


def bits(a):
    return np.ascontiguousarray(a.cpu().numpy() if hasattr(a, "cpu") else a, np.float32).view(np.int32)


def setup(iphys=0, kl=36, ml=36, seed=51):
    import torch
    cfg, mask, f = W.c2_case(ny=60, nx=90, seed=seed, iphys=iphys, kl=kl, ml=ml)
    g = WamGpu(cfg, mask)
    counter, _ = g.prepare_propagation(f["depth"], f["u"], f["v"])
    n = g.nsea
    rng = np.random.default_rng(seed)
    u10 = np.stack([f["u10"] * (1.0 + 0.1 * rng.standard_normal(n)) for _ in range(NMEM)]).astype(np.float32)
    u10 = np.clip(u10, 0.5, None)
    udir = np.stack([f["udir"] + 0.2 * m for m in range(NMEM)]).astype(np.float32)
    fl3 = np.stack([W.jonswap_spectrum(cfg, udir[m]).ravel() * (1.0 + 0.1 * m) for m in range(NMEM)]).astype(
        np.float32)
    dev = torch.device("cuda:0")
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    st = dict(fa=t(fl3.ravel()), fb=torch.zeros(fl3.size, device=dev), u10=t(u10.ravel()), udir=t(udir.ravel()),
              tauw=torch.zeros(NMEM * n, device=dev), ustar=torch.zeros(NMEM * n, device=dev),
              z0=torch.zeros(NMEM * n, device=dev), roairn=torch.full((NMEM * n,), 1.225, device=dev),
              wstar=torch.zeros(NMEM * n, device=dev))
    return cfg, mask, f, g, n, counter, st, fl3, u10, udir


def ensemble_steps(torch, g, n, st, nsteps, lcflx=False):
    fa, fb = st["fa"].clone(), st["fb"].clone()
    tauw, ustar, z0 = st["tauw"].clone(), st["ustar"].clone(), st["z0"].clone()
    phiaw = torch.zeros(NMEM * n, device=fa.device)
    flux = torch.zeros(6 * NMEM * n, device=fa.device)
    mij = torch.zeros(NMEM * n, dtype=torch.int32, device=fa.device)
    for _ in range(nsteps):
        g.propags_ensemble_dev(NMEM, fa, fb)
        g.implsch_ensemble_dev(NMEM, fb, st["u10"], st["udir"], tauw, ustar, z0, st["roairn"], st["wstar"],
                               lcflx=lcflx, phiaw=phiaw, fluxes=flux if lcflx else None, mij=mij)
        fa, fb = fb, fa
    torch.cuda.synchronize()
    return dict(fl3=fa, tauw=tauw, ustar=ustar, z0=z0, phiaw=phiaw, flux=flux.view(6, NMEM * n), mij=mij)


def member_plain(torch, g, n, st, m, nsteps, lcflx=False):
    nb = st["fa"].numel() // NMEM
    fa = st["fa"][m * nb:(m + 1) * nb].clone()
    fb = torch.zeros_like(fa)
    sl = slice(m * n, (m + 1) * n)
    tauw, ustar, z0 = st["tauw"][sl].clone(), st["ustar"][sl].clone(), st["z0"][sl].clone()
    phiaw = torch.zeros(n, device=fa.device)
    flux = torch.zeros(6 * n, device=fa.device)
    mij = torch.zeros(n, dtype=torch.int32, device=fa.device)
    u10, udir = st["u10"][sl].contiguous(), st["udir"][sl].contiguous()
    ro, ws = st["roairn"][sl].contiguous(), st["wstar"][sl].contiguous()
    for _ in range(nsteps):
        g.propags_dev(fa, fb)
        g.implsch_dev(fb, u10, udir, tauw, ustar, z0, ro, ws, lcflx=lcflx, phiaw=phiaw,
                      fluxes=flux if lcflx else None, mij=mij)
        fa, fb = fb, fa
    torch.cuda.synchronize()
    return dict(fl3=fa, tauw=tauw, ustar=ustar, z0=z0, phiaw=phiaw, flux=flux.view(6, n), mij=mij)


def same_as_members(torch, g, n, st, e, nsteps, lcflx=False):
    nb = st["fa"].numel() // NMEM
    for m in range(NMEM):
        p = member_plain(torch, g, n, st, m, nsteps, lcflx)
        sl = slice(m * n, (m + 1) * n)
        assert np.array_equal(bits(e["fl3"][m * nb:(m + 1) * nb]), bits(p["fl3"])), m
        for k in ("tauw", "ustar", "z0", "phiaw"):
            assert np.array_equal(bits(e[k][sl]), bits(p[k])), (m, k)
        assert np.array_equal(e["mij"][sl].cpu().numpy(), p["mij"].cpu().numpy()), m
        if lcflx:
            assert np.array_equal(bits(e["flux"][:, sl]), bits(p["flux"])), m


@pytest.mark.parametrize("iphys,kl,ml", [(0, 36, 36), (2, 24, 25)])
def test_members_equal_plain_runs_and_the_oracle(iphys, kl, ml):
    """Each member of a 4-member ensemble equals a plain run of that member
    after two steps (the second with LCFLX), and the plain run of members 0
    and 3 equals the oracle."""
    import torch
    cfg, mask, f, g, n, counter, st, fl3, u10, udir = setup(iphys, kl, ml)
    print(f"[case] iphys={iphys} {kl}x{ml} NSEA={n} COUNTER={counter}", flush=True)
    e = ensemble_steps(torch, g, n, st, 1)
    same_as_members(torch, g, n, st, e, 1)
    e2 = ensemble_steps(torch, g, n, st, 2, lcflx=True)
    same_as_members(torch, g, n, st, e2, 2, lcflx=True)
    o = Oracle(cfg, mask)
    o.prepare_propagation(f["depth"], f["u"], f["v"])
    nb = fl3.shape[1]
    for m in (0, NMEM - 1):
        z = np.zeros(n, np.float32)
        r = o.implsch(o.propags(fl3[m]), u10[m], udir[m], z, z, f["depth"])
        assert np.array_equal(bits(e["fl3"][m * nb:(m + 1) * nb]), bits(r["fl3"])), m
        assert np.array_equal(bits(e["ustar"][m * n:(m + 1) * n]), bits(r["ustar"])), m
    g.close()


def test_every_launch_shape_gives_the_same_results():
    """Launch shapes the autotuner chooses among, including IMPLSCH chunks that
    split members (1000 member-points), give the default shapes' results."""
    import torch
    cfg, mask, f, g, n, counter, st, *_ = setup()
    ref = ensemble_steps(torch, g, n, st, 2)
    for shape in ((64, 128, 1, 1000), (128, 256, 2, 4096), (256, 64, 0, 3 * n + 17)):
        g.set_tuning(*shape)
        assert g.tuning() == shape
        e = ensemble_steps(torch, g, n, st, 2)
        for k in ("fl3", "tauw", "ustar", "z0"):
            assert np.array_equal(bits(e[k]), bits(ref[k])), (shape, k)
    g.close()


def test_autotuner_picks_shapes_with_the_same_results():
    import torch
    cfg, mask, f, g, n, counter, st, *_ = setup()
    ref = ensemble_steps(torch, g, n, st, 1)
    fb = st["fb"].clone()
    shapes = g.autotune(NMEM, st["fa"], fb, st["u10"], st["udir"], st["tauw"], st["z0"], st["roairn"], st["wstar"],
                        reps=2)
    log = g.tuning_log()
    print("[autotune]", shapes, log, flush=True)
    assert shapes[0] in (64, 128, 256) and shapes[1] in (64, 128, 256) and shapes[2] in (0, 1, 2)
    assert {w for w, _, _ in log} >= {"prop_wg", "chunk", "src_wg", "snl_per_cu"}
    e = ensemble_steps(torch, g, n, st, 1)
    for k in ("fl3", "tauw", "ustar", "z0"):
        assert np.array_equal(bits(e[k]), bits(ref[k])), k
    g.close()


def test_captured_loop_equals_eager_on_every_replay():
    """The HIP-graph time loop (3 steps) against the same loop issued launch
    by launch, on four replays with eager work of the program in between."""
    import torch
    cfg, mask, f, g, n, counter, st, *_ = setup()
    nsteps = 3
    eager = ensemble_steps(torch, g, n, st, nsteps)
    fa, fb = st["fa"].clone(), st["fb"].clone()
    tauw, ustar, z0 = st["tauw"].clone(), st["ustar"].clone(), st["z0"].clone()
    for rep in range(4):
        fa.copy_(st["fa"])
        fb.zero_()
        tauw.copy_(st["tauw"])
        z0.copy_(st["z0"])
        ustar.zero_()
        torch.cuda.synchronize()
        out = g.ensemble_run_dev(NMEM, nsteps, fa, fb, st["u10"], st["udir"], tauw, ustar, z0, st["roairn"],
                                 st["wstar"], graph=True)
        torch.cuda.synchronize()
        for k, v in (("fl3", out), ("tauw", tauw), ("ustar", ustar), ("z0", z0)):
            assert np.array_equal(bits(v), bits(eager[k])), (rep, k)
        # eager work between replays: a plain step of member 0 and a reduction
        _ = member_plain(torch, g, n, st, 0, 1)
        _ = float(fa.sum())
    # the same loop without the graph
    fa.copy_(st["fa"]); tauw.copy_(st["tauw"]); z0.copy_(st["z0"])
    out = g.ensemble_run_dev(NMEM, nsteps, fa, fb, st["u10"], st["udir"], tauw, ustar, z0, st["roairn"],
                             st["wstar"], graph=False)
    torch.cuda.synchronize()
    assert np.array_equal(bits(out), bits(eager["fl3"]))
    g.close()
