"""Error paths and robustness of the C-ABI on the GPU, each condition forced
from the host: fields that are not finite, a CFL number no sub-step count can
cover, a NaN wind (which must stay at its own sea point), a failed device
allocation (reported, and the context usable afterwards), and device tensors
written on a torch side stream (the library's stream is ordered after it)."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from wam_amd import WamGpu, WamError, workloads as W

pytestmark = pytest.mark.gpu

WAM_E_ARG, WAM_E_HIP, WAM_E_CFL = 1, 2, 3


def bits(a):
    a = a.cpu().numpy() if hasattr(a, "cpu") else a
    return np.ascontiguousarray(a, np.float32).view(np.int32)


def case(kl=24, ml=25, ny=30, nx=40, seed=31):
    return W.c2_case(ny=ny, nx=nx, seed=seed, kl=kl, ml=ml)


def tdev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).to("cuda:0")


@pytest.mark.parametrize("path", ["host", "device"])
def test_nonfinite_fields_refused(path):
    cfg, mask, f = case()
    g = WamGpu(cfg, mask)
    n = g.nsea
    prep = (lambda d, u, v: g.prepare_propagation(d, u, v)) if path == "host" else \
        (lambda d, u, v: g.prepare_propagation_dev(tdev(d), tdev(u), tdev(v)))
    bad = []
    d = f["depth"].copy(); d[n // 3] = np.inf
    bad.append((d, f["u"], f["v"]))
    u = f["u"].copy(); u[n // 2] = np.nan
    bad.append((f["depth"], u, f["v"]))
    v = f["v"].copy(); v[7] = -np.inf
    bad.append((f["depth"], f["u"], v))
    for d, u, v in bad:
        with pytest.raises(WamError) as e:
            prep(d, u, v)
        assert e.value.code == WAM_E_ARG, e.value
    # a CFL number beyond any sub-step count (a current of 1e30 m/s at one point)
    u = f["u"].copy(); u[n // 2] = 1e30
    with pytest.raises(WamError) as e:
        prep(f["depth"], u, f["v"])
    assert e.value.code == WAM_E_CFL, e.value
    # the context is usable afterwards, bit for bit
    o = Oracle(cfg, mask)
    o.prepare_propagation(f["depth"], f["u"], f["v"])
    prep(f["depth"], f["u"], f["v"])
    fl3 = W.jonswap_spectrum(cfg, f["udir"]).ravel()
    assert np.array_equal(bits(g.propags(fl3)), bits(o.propags(fl3)))
    z = np.zeros(n, np.float32)
    rg = g.implsch(fl3, f["u10"], f["udir"], z, z)
    ro = o.implsch(fl3, f["u10"], f["udir"], z, z, f["depth"])
    assert np.array_equal(bits(rg["fl3"]), bits(ro["fl3"]))


@pytest.mark.parametrize("stored", [False, True])
def test_nan_weight_refused(stored):
    """Finite but extreme fields whose propagation weights are NaN at one sea
    point: a depth step from 1 m to 3e38 m on both of its axes and a current
    of 1e5 m/s, so that TEMPD = V*DDPHI + U*DDLAM is +Inf - Inf there.  The
    oracle's weights there are NaN (no Inf); CHECK_CFL refuses them, in both
    weight modes, before its CFL bounds."""
    cfg, mask, f = case()
    o = Oracle(cfg, mask)
    n = o.nsea
    klat, klon = o.get("KLAT"), o.get("KLON")
    s_, n_, w_, e_ = klat[:n] - 1, klat[n:2 * n] - 1, klon[:n] - 1, klon[n:2 * n] - 1
    inner = np.flatnonzero((s_ >= 0) & (n_ >= 0) & (w_ >= 0) & (e_ >= 0))
    p = int(inner[inner.size // 2])
    d, u, v = f["depth"].copy(), f["u"].copy(), f["v"].copy()
    d[[p, n_[p], e_[p]]] = 3e38
    d[[s_[p], w_[p]]] = 1.0
    u[p], v[p] = -1e5, 1e5
    o.prepare_propagation(d, u, v)
    w = o.get("COEF").reshape(-1, n)[:, p]
    assert np.isnan(w).any() and not np.isinf(w).any()
    g = WamGpu(cfg, mask)
    g.set_propagation_weights(stored)
    with pytest.raises(WamError, match="not a number") as e:
        g.prepare_propagation(d, u, v)
    assert e.value.code == WAM_E_CFL, e.value


@pytest.mark.parametrize("kl", [24, 30])          # the three-kernel IMPLSCH and the single kernel
def test_nan_wind_stays_at_its_point(kl):
    cfg, mask, f = case(kl=kl, seed=32)
    g = WamGpu(cfg, mask)
    g.prepare_propagation(f["depth"], f["u"], f["v"])
    n = g.nsea
    fl3 = W.jonswap_spectrum(cfg, f["udir"]).ravel()
    z = np.zeros(n, np.float32)
    clean = g.implsch(fl3, f["u10"], f["udir"], z, z)
    u10 = f["u10"].copy()
    i = n // 2
    u10[i] = np.nan
    r = g.implsch(fl3, u10, f["udir"], z, z)
    keep = np.ones(n, bool)
    keep[i] = False
    nb = cfg["kl"] * cfg["ml"]
    assert np.array_equal(bits(r["fl3"].reshape(nb, n)[:, keep]), bits(clean["fl3"].reshape(nb, n)[:, keep]))
    for k in ("ustar", "tauw", "z0"):
        assert np.array_equal(bits(r[k][keep]), bits(clean[k][keep])), k
    # at the point itself the device follows the oracle's REAL*4 evaluation,
    # NaNs included (MAX and MIN return the other operand of a NaN)
    o = Oracle(cfg, mask)
    o.prepare_propagation(f["depth"], f["u"], f["v"])
    ro = o.implsch(fl3, u10, f["udir"], z, z, f["depth"])
    for k in ("fl3", "ustar", "tauw", "z0"):
        assert np.array_equal(r[k], ro[k], equal_nan=True), k


def test_allocation_failure_reported():
    """With the device memory taken, IMPLSCH's first call cannot allocate its
    work arrays: WAM_E_HIP with a message, no partial set left behind, and
    the same call succeeds once memory is free again."""
    import torch
    cfg, mask, f = W.c2_case(ny=240, nx=400, seed=33)     # 36 x 36: about 1 GB of work arrays
    g = WamGpu(cfg, mask)
    g.prepare_propagation(f["depth"], f["u"], f["v"])
    n = g.nsea
    fl3 = W.jonswap_spectrum(cfg, f["udir"]).ravel()
    args = [tdev(f["u10"]), tdev(f["udir"]), tdev(np.zeros(n)), tdev(np.zeros(n)), tdev(np.zeros(n)),
            tdev(np.full(n, 1.225)), tdev(np.zeros(n))]
    a = tdev(fl3)
    torch.cuda.synchronize()
    hold = []
    try:
        free, _ = torch.cuda.mem_get_info()
        chunk = 8 << 30
        while free > (96 << 20) and chunk >= (16 << 20):
            try:
                hold.append(torch.empty(min(free - (64 << 20), chunk), dtype=torch.uint8, device="cuda:0"))
            except torch.cuda.OutOfMemoryError:
                chunk //= 2
            free, _ = torch.cuda.mem_get_info()
        assert free < (256 << 20), free
        print(f"device memory left: {free >> 20} MiB", flush=True)
        with pytest.raises(WamError) as e:
            g.implsch_dev(a, *args)
        assert e.value.code == WAM_E_HIP and "out of memory" in str(e.value), e.value
    finally:
        hold.clear()
        torch.cuda.empty_cache()
    g.implsch_dev(a, *args)
    torch.cuda.synchronize()
    fresh = WamGpu(cfg, mask)
    fresh.prepare_propagation(f["depth"], f["u"], f["v"])
    z = np.zeros(n, np.float32)
    r = fresh.implsch(fl3, f["u10"], f["udir"], z, z)
    assert np.array_equal(bits(a), bits(r["fl3"]))


def test_inputs_written_on_a_side_stream():
    """The spectrum is written on a torch side stream behind a long matrix
    product, and IMPLSCH is called from that stream without a host
    synchronisation: the context's stream waits for it, and the side stream
    then reads the result after IMPLSCH."""
    import torch
    cfg, mask, f = case(kl=36, ml=36, ny=40, nx=60, seed=34)
    g = WamGpu(cfg, mask)
    g.prepare_propagation(f["depth"], f["u"], f["v"])
    n = g.nsea
    fl3 = W.jonswap_spectrum(cfg, f["udir"]).ravel()
    z = np.zeros(n, np.float32)
    want = g.implsch(fl3, f["u10"], f["udir"], z, z)["fl3"]
    src = tdev(fl3)
    args = [tdev(f["u10"]), tdev(f["udir"]), tdev(z), tdev(z), tdev(z), tdev(np.full(n, 1.225)), tdev(z)]
    dst = torch.zeros_like(src)
    m = torch.randn(2048, 2048, device="cuda:0")
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        for _ in range(20):
            m = m @ m / 2048.0                     # keeps the side stream busy before the copy
        dst.copy_(src)
        g.implsch_dev(dst, *args)
        out = dst.clone()                          # read on the side stream after IMPLSCH
    torch.cuda.synchronize()
    assert np.array_equal(bits(out), bits(want))
