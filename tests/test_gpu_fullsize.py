"""One model step (PROPAGS, then IMPLSCH) of the production workload at full
size on the GPU: the 0.1 deg 901 x 1600 grid, 1 009 120 sea points, 36 x 36,
COUNTER = 2, spectra resident in HBM.  It exercises what only exists at that
size: the 32-bit buffer offsets over three frequency planes of
propags_fused_sc_kernel, implsch_snl_kernel's per-XCD point tickets over a
million points and the 15.7 GB of IMPLSCH work arrays.

The oracle checks it on windows of the grid.  A window is the grid cut to a
rectangle (cells outside it become land); at points three or more cells
inside its edge the neighbours, the depth and current gradients and so the
propagation weights are the full grid's, and IMPLSCH is pointwise, so the
oracle's step on the window equals the full grid's step there, bit for bit.
test_windows_reproduce_the_full_grid checks that claim on the CPU against
the oracle's own full-grid step; the GPU test then compares the full-size
GPU step with the oracle on four windows (over 20 000 interior sea points,
at the grid's southern and northern edges, in mid-basin and near the eastern
edge), bit for bit."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from wam_amd import workloads as W

MARGIN = 3


def window(cfg, mask, f, r0, c0, h, w):
    """The window rows r0..r0+h-1, columns c0..c0+w-1: (cfg, mask, fields,
    global sea-point indices of its sea points, which of them are interior)."""
    wcfg = dict(cfg)
    wcfg.update(nx=w, ny=h, amosop=cfg["amosop"] + r0 * cfg["xdella"], amowep=cfg["amowep"] + c0 * cfg["xdello"])
    wcfg["amonop"] = wcfg["amosop"] + (h - 1) * wcfg["xdella"]
    wcfg["amoeap"] = wcfg["amowep"] + (w - 1) * wcfg["xdello"]
    idx = np.full(mask.shape, -1, np.int64)
    idx[mask.astype(bool)] = np.arange(int(mask.sum()))
    wm = mask[r0:r0 + h, c0:c0 + w]
    gidx = idx[r0:r0 + h, c0:c0 + w][wm.astype(bool)]
    rr, cc = np.nonzero(wm)
    inner = (rr >= MARGIN) & (rr < h - MARGIN) & (cc >= MARGIN) & (cc < w - MARGIN)
    # rows at the grid's own edge have no neighbour beyond it in the full grid either
    inner |= ((rr < MARGIN) & (r0 == 0) | (rr >= h - MARGIN) & (r0 + h == mask.shape[0])) & \
        (cc >= MARGIN) & (cc < w - MARGIN)
    wf = {k: (None if v is None else np.asarray(v)[gidx]) for k, v in f.items()}
    return wcfg, np.ascontiguousarray(wm), wf, gidx, inner


def oracle_step(cfg, mask, f, fl3, counter):
    """PROPAGS (COUNTER forced to the full grid's) then IMPLSCH, from zero TAUW and Z0."""
    o = Oracle(cfg, mask)
    o.force_counter(counter)
    c, _ = o.prepare_propagation(f["depth"], f["u"], f["v"])
    assert c == counter
    n = o.nsea
    p = o.propags(fl3)
    z = np.zeros(n, np.float32)
    r = o.implsch(p, f["u10"], f["udir"], z, z, f["depth"], nthreads=16)
    return p, r


def test_windows_reproduce_the_full_grid():
    """The window method on the CPU: the oracle on windows of a 200 x 300 grid
    equals the oracle on the whole grid at the windows' interior points."""
    cfg, mask, f = W.c2_case(ny=200, nx=300, seed=41)
    o = Oracle(cfg, mask)
    counter, _ = o.prepare_propagation(f["depth"], f["u"], f["v"])
    n = o.nsea
    fl3 = W.jonswap_spectrum(cfg, f["udir"]).reshape(-1, n)
    fl3 = (fl3 * np.linspace(0.5, 1.5, n, dtype=np.float32)[None, :]).astype(np.float32)
    pf = o.propags(fl3.ravel()).reshape(-1, n)
    z = np.zeros(n, np.float32)
    rf = o.implsch(pf.ravel(), f["u10"], f["udir"], z, z, f["depth"], nthreads=8)
    checked = 0
    for r0, c0 in ((0, 0), (70, 120), (200 - 60, 300 - 80)):
        wcfg, wm, wf, gidx, inner = window(cfg, mask, f, r0, c0, 60, 80)
        p, r = oracle_step(wcfg, wm, wf, fl3[:, gidx].ravel(), counter)
        ni = gidx.size
        assert np.array_equal(p.reshape(-1, ni)[:, inner].view(np.int32), pf[:, gidx[inner]].view(np.int32))
        assert np.array_equal(r["fl3"].reshape(-1, ni)[:, inner].view(np.int32),
                              rf["fl3"].reshape(-1, n)[:, gidx[inner]].view(np.int32))
        for k in ("ustar", "tauw", "z0"):
            assert np.array_equal(r[k][inner].view(np.int32), rf[k][gidx[inner]].view(np.int32)), k
        checked += int(inner.sum())
    assert checked > 5000


@pytest.mark.gpu
@pytest.mark.parametrize("iphys", [0, 2])          # Janssen; ST6 (both in IMPLSCH's three-kernel form here)
def test_full_size_step_on_windows(iphys):
    import torch
    from wam_amd import WamGpu
    cfg, mask, f = W.c2_case(iphys=iphys)
    g = WamGpu(cfg, mask)
    n = g.nsea
    assert n == 1009120
    counter, _ = g.prepare_propagation(f["depth"], f["u"], f["v"])
    print(f"[full] IPHYS={iphys} NSEA={n} COUNTER={counter}", flush=True)
    assert counter == 2
    dev = torch.device("cuda:0")
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    u10, udir = t(f["u10"]), t(f["udir"])
    # the initial spectrum, built on the device: JONSWAP x cos^2 spreading
    # about the local wind, scaled by a smooth factor per point
    fr = torch.tensor(W.frequencies(cfg), dtype=torch.float64, device=dev)
    th = (torch.arange(cfg["kl"], dtype=torch.float64, device=dev) + 0.5) * (2 * np.pi / cfg["kl"])
    et = 0.018 * 9.806 ** 2 / (2 * np.pi) ** 4 * fr ** -5 * torch.exp(-torch.clamp(1.25 * (0.3 / fr) ** 4, max=50.0))
    sig = torch.where(fr > 0.3, 0.09, 0.07)
    et = et * 3.3 ** torch.exp(-((fr / 0.3 - 1.0) ** 2) / (2 * sig * sig))
    st = (2.0 / np.pi) * torch.clamp(torch.cos(th[:, None] - udir.double()[None, :]), min=0.0) ** 2
    scale = torch.linspace(0.5, 1.5, n, dtype=torch.float64, device=dev)
    fa = (et[:, None, None].float() * (st * scale[None, :]).float()[None, :, :]).reshape(-1).contiguous()
    del st
    fb = torch.empty_like(fa)
    z = torch.zeros(n, device=dev)
    tauw, ustar, z0 = z.clone(), z.clone(), z.clone()
    roairn, wstar = torch.full((n,), 1.225, device=dev), z.clone()
    wins = [(0, 0), (380, 700), (901 - 120, 300), (450, 1600 - 120)]
    inputs, after_prop = {}, {}
    for r0, c0 in wins:
        _, _, _, gidx, _ = window(cfg, mask, f, r0, c0, 120, 120)
        gi = torch.from_numpy(gidx).to(dev)
        inputs[(r0, c0)] = fa.view(-1, n).index_select(1, gi).cpu().numpy()
    g.propags_dev(fa, fb)
    for r0, c0 in wins:
        _, _, _, gidx, _ = window(cfg, mask, f, r0, c0, 120, 120)
        after_prop[(r0, c0)] = fb.view(-1, n).index_select(1, torch.from_numpy(gidx).to(dev)).cpu().numpy()
    g.implsch_dev(fb, u10, udir, tauw, ustar, z0, roairn, wstar)
    torch.cuda.synchronize()
    print("[full] one step done on the GPU", flush=True)
    checked = 0
    for r0, c0 in wins:
        wcfg, wm, wf, gidx, inner = window(cfg, mask, f, r0, c0, 120, 120)
        gi = torch.from_numpy(gidx).to(dev)
        got = fb.view(-1, n).index_select(1, gi).cpu().numpy()
        p, r = oracle_step(wcfg, wm, wf, inputs[(r0, c0)].ravel(), counter)
        ni = gidx.size
        assert np.array_equal(after_prop[(r0, c0)][:, inner].view(np.int32),
                              p.reshape(-1, ni)[:, inner].view(np.int32)), (r0, c0, "PROPAGS")
        assert np.array_equal(got[:, inner].view(np.int32), r["fl3"].reshape(-1, ni)[:, inner].view(np.int32)), \
            (r0, c0, "IMPLSCH")
        for k, d in (("ustar", ustar), ("tauw", tauw), ("z0", z0)):
            assert np.array_equal(d[gi].cpu().numpy()[inner].view(np.int32), r[k][inner].view(np.int32)), (r0, c0, k)
        checked += int(inner.sum())
        print(f"[full] window {r0},{c0}: {int(inner.sum())} interior sea points bit-identical", flush=True)
    assert checked >= 10000
    assert bool(torch.isfinite(fb).all())
    g.close()
