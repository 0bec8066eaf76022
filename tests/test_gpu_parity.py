"""GPU parity of the HIP path against the CPU oracle (through the C-ABI).

The kernels evaluate the reference's REAL*4 arithmetic in the same order as
the oracle with the same elementary functions, so spectra, USTAR, TAUW, Z0,
PHIAW, the WNFLUXES side outputs, the cut-off index MIJ, the propagation
weights and PROPAGS are required to be bit-identical to the oracle.  The
north-star tolerance for spectra after one step (1e-5 relative) is asserted
as well, so a deliberate reordering would still be judged on that bar.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from wam_amd import WamGpu, WamError, workloads as W

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.array_equal(a.view(np.int32), b.view(np.int32))


def maxrel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30))) if a.size else 0.0


CASES = {
    "c1_st6_deep": lambda: W.c1_case(),
    "c1_janssen_deep": lambda: _with(W.c1_case(), iphys=0),
    "c2cut_janssen_shallow_curr": lambda: W.c2_case(ny=40, nx=60, seed=5),
    "c2cut_st6_shallow_curr": lambda: W.c2_case(ny=40, nx=60, seed=5, iphys=2, kl=24, ml=25),
    "c2cut_janssen_shallow_nocurr": lambda: _with(W.c2_case(ny=30, nx=50, seed=8, kl=24, ml=30), refr_c=0),
    "c2cut_janssen_shallow_norefr": lambda: _with(W.c2_case(ny=30, nx=50, seed=8, kl=12, ml=25), refr_c=0,
                                                    refr_d=0),
    "c2cut_janssen_shallow_curr_kl12": lambda: W.c2_case(ny=30, nx=40, seed=6, kl=12, ml=25),
    # 30 directions (not a multiple of 4): IMPLSCH's single-kernel form and the generic PROP_CURR kernel
    "c2cut_janssen_shallow_curr_kl30": lambda: W.c2_case(ny=30, nx=40, seed=7, kl=30, ml=25),
    # 16 directions (not a multiple of 3): PROPAGS a direction at a time instead of in blocks of three
    "c2cut_janssen_shallow_curr_kl16": lambda: W.c2_case(ny=30, nx=40, seed=9, kl=16, ml=25),
    "deep_current": lambda: _with(W.c2_case(ny=30, nx=40, seed=4, kl=24, ml=25), shallow=0, deep_depth=True),
}


def _with(case, deep_depth=False, **kw):
    cfg, mask, f = case
    cfg = dict(cfg)
    cfg.update(kw)
    f = dict(f)
    if deep_depth:
        f["depth"] = np.full_like(f["depth"], 999.0)
    if not cfg.get("refr_c"):
        f["u"], f["v"] = None, None
    return cfg, mask, f


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    cfg, mask, f = CASES[request.param]()
    print(f"case {request.param}: iphys={cfg['iphys']} shallow={cfg['shallow']} refr_c={cfg['refr_c']} "
          f"{cfg['kl']}x{cfg['ml']}", flush=True)
    o = Oracle(cfg, mask)
    g = WamGpu(cfg, mask)                   # weights computed in the PROPAGS stencil
    g.set_propagation_weights(False)
    gs = WamGpu(cfg, mask)                  # weight planes stored in HBM
    gs.set_propagation_weights(True)
    co = o.prepare_propagation(f["depth"], f["u"], f["v"])
    cg = g.prepare_propagation(f["depth"], f["u"], f["v"])
    cs = gs.prepare_propagation(f["depth"], f["u"], f["v"])
    fl3 = W.jonswap_spectrum(cfg, f["udir"]).ravel()
    return dict(name=request.param, cfg=cfg, mask=mask, f=f, o=o, g=g, gs=gs, co=co, cg=cg, cs=cs, fl3=fl3)


def test_propagation_weights_and_cfl(case):
    o, gs = case["o"], case["gs"]
    for c in (case["cg"], case["cs"]):                          # both weight modes
        assert case["co"][0] == c[0]                            # COUNTER
        assert np.float32(case["co"][1]) == np.float32(c[1])    # max CFL
    assert bits_equal(gs.table("COEF"), o.get("COEF"))
    if case["cfg"]["shallow"]:
        assert np.array_equal(case["g"].table("INDEP"), o.get("INDEP"))


def test_prepare_propagation_on_device(case):
    """PREPARE_PROPAGATION from fields in HBM: the per-point tables (INDEP,
    INDEP_G, gradients, DELLA0, DRGM), COUNTER and the CFL maximum equal the
    host-prepared context's bit for bit, and PROPAGS and IMPLSCH then match
    the oracle bit for bit."""
    torch = pytest.importorskip("torch")
    cfg, f, o = case["cfg"], case["f"], case["o"]
    gd = WamGpu(cfg, case["mask"])
    gd.set_propagation_weights(False)
    dev = torch.device("cuda:0")
    t = lambda x: None if x is None else torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    cd = gd.prepare_propagation_dev(t(f["depth"]), t(f["u"]), t(f["v"]))
    assert cd[0] == case["cg"][0]
    assert np.float32(cd[1]) == np.float32(case["cg"][1])
    for name in ("INDEP", "INDEP_G"):
        assert np.array_equal(gd.table(name), case["g"].table(name)), name
    for name in ("DDPHI", "DDLAM", "DUPHI", "DULAM", "DVPHI", "DVLAM", "DELLA0", "DRGM"):
        assert bits_equal(gd.table(name), case["g"].table(name)), name
    assert bits_equal(gd.propags(case["fl3"]), o.propags(case["fl3"]))
    n = gd.nsea
    fl3 = o.propags(case["fl3"])
    z = np.zeros(n, np.float32)
    ro = o.implsch(fl3, f["u10"], f["udir"], z, z, f["depth"])
    rg = gd.implsch(fl3, f["u10"], f["udir"], z, z)
    for k in ("fl3", "ustar", "tauw", "z0"):
        assert bits_equal(rg[k], ro[k]), (k, maxrel(rg[k], ro[k]))
    # a non-positive depth is refused, as on the host path
    bad = np.array(f["depth"], np.float32)
    bad[n // 2] = 0.0
    with pytest.raises(WamError):
        gd.prepare_propagation_dev(t(bad), t(f["u"]), t(f["v"]))


@pytest.mark.parametrize("mode", ["computed", "stored"])
def test_propags_bit_exact(case, mode):
    a = case["o"].propags(case["fl3"])
    b = (case["g"] if mode == "computed" else case["gs"]).propags(case["fl3"])
    assert bits_equal(b, a)


def test_implsch_one_step(case):
    o, g, f = case["o"], case["g"], case["f"]
    n = g.nsea
    fl3 = o.propags(case["fl3"])
    z = np.zeros(n, np.float32)
    ro = o.implsch(fl3, f["u10"], f["udir"], z, z, f["depth"], lcflx=True)
    rg = g.implsch(fl3, f["u10"], f["udir"], z, z, lcflx=True)
    assert maxrel(rg["fl3"], ro["fl3"]) <= 1e-5
    for k in ("fl3", "ustar", "tauw", "z0", "phiaw", "flux"):
        assert bits_equal(rg[k], ro[k]), (k, maxrel(rg[k], ro[k]))


@pytest.mark.parametrize("three", [False, True])
def test_implsch_forms(case, three):
    """Both IMPLSCH forms (three kernels; one kernel with one workgroup per
    sea point), forced whatever the automatic choice, over three steps with
    LCFLX on the first: each bit-identical to the oracle.  Where the three-
    kernel form does not apply (30 directions) the single kernel runs."""
    o, f, cfg = case["o"], case["f"], case["cfg"]
    g = WamGpu(cfg, case["mask"])
    g.set_implsch_form(three)
    g.prepare_propagation(f["depth"], f["u"], f["v"])
    n = g.nsea
    fo = fg = o.propags(case["fl3"])
    zo = zg = two = twg = np.zeros(n, np.float32)
    for step in range(3):
        lc = step == 0
        ro = o.implsch(fo, f["u10"], f["udir"], two, zo, f["depth"], lcflx=lc)
        rg = g.implsch(fg, f["u10"], f["udir"], twg, zg, lcflx=lc)
        for k in ("fl3", "ustar", "tauw", "z0", "phiaw") + (("flux",) if lc else ()):
            assert bits_equal(rg[k], ro[k]), (step, k, maxrel(rg[k], ro[k]))
        fo, two, zo = ro["fl3"], ro["tauw"], ro["z0"]
        fg, twg, zg = rg["fl3"], rg["tauw"], rg["z0"]
    g.close()


def test_output_parameters_bit_exact(case):
    """The integrated output parameters on the GPU (host and device-resident
    forms) against the oracle, both direction conventions."""
    o, g, f = case["o"], case["g"], case["f"]
    n = g.nsea
    rng = np.random.default_rng(3)
    fl3 = o.propags(case["fl3"])
    ust = (0.05 + rng.random(n)).astype(np.float32)
    tauw = (0.5 * ust * ust * rng.random(n)).astype(np.float32)
    z0 = (1e-4 + 1e-3 * rng.random(n)).astype(np.float32)
    for orient in (1, 0):
        ro = o.output_parameters(fl3, f["u10"], f["udir"], ust, tauw, z0, f["depth"], f["u"], f["v"], orient)
        rg = g.output_parameters(fl3, f["u10"], f["udir"], ust, tauw, z0, orient)
        for i, name in enumerate(g.OUT_NAMES):
            assert bits_equal(rg[i], ro[i]), (orient, name, maxrel(rg[i], ro[i]))
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    out = torch.empty(g.NOUT * n, device=dev)
    g.set_stream(torch.cuda.current_stream().cuda_stream)
    g.output_parameters_dev(t(fl3), t(f["u10"]), t(f["udir"]), t(ust), t(tauw), t(z0), out, orient=1)
    torch.cuda.synchronize()
    g.set_stream(0)
    ro = o.output_parameters(fl3, f["u10"], f["udir"], ust, tauw, z0, f["depth"], f["u"], f["v"], 1)
    assert bits_equal(out.cpu().numpy().reshape(g.NOUT, n), ro)


def test_multi_step_device_resident(case):
    """Six propagation + source steps on torch tensors resident in HBM,
    through wam_gpu_propags_dev / wam_gpu_implsch_dev, against the oracle."""
    torch = pytest.importorskip("torch")
    o, g, f, cfg = case["o"], case["g"], case["f"], case["cfg"]
    n = g.nsea
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    g.set_stream(stream.cuda_stream)
    fa = torch.from_numpy(case["fl3"].copy()).to(dev)
    fb = torch.empty_like(fa)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    u10, udir = t(f["u10"]), t(f["udir"])
    tauw, z0, ust = t(np.zeros(n)), t(np.zeros(n)), t(np.zeros(n))
    ro_, ws = t(np.full(n, 1.225)), t(np.zeros(n))
    mij = torch.zeros(n, dtype=torch.int32, device=dev)
    fo = case["fl3"].copy()
    tw_o = np.zeros(n, np.float32)
    z0_o = np.zeros(n, np.float32)
    for step in range(6):
        g.propags_dev(fa, fb)
        fa, fb = fb, fa
        g.implsch_dev(fa, u10, udir, tauw, ust, z0, ro_, ws, mij=mij)
        fo = o.propags(fo)
        r = o.implsch(fo, f["u10"], f["udir"], tw_o, z0_o, f["depth"])
        fo, tw_o, z0_o = r["fl3"], r["tauw"], r["z0"]
    torch.cuda.synchronize()
    g.set_stream(0)
    assert bits_equal(fa.cpu().numpy(), fo)
    assert bits_equal(ust.cpu().numpy(), r["ustar"])
    assert np.array_equal(mij.cpu().numpy(), r["mij"])


def test_host_argument_errors_are_reported(case):
    g = case["g"]
    bad = np.zeros(g.nsea * g.kl * g.ml + 1, np.float32)
    with pytest.raises(ValueError):
        g.implsch(bad, np.zeros(g.nsea), np.zeros(g.nsea), np.zeros(g.nsea), np.zeros(g.nsea))
    rc = g.lib.wam_gpu_propags(g.ctx, 1, g.nsea + 1, g.kl, g.ml, None)
    assert rc == 1


def test_call_order_is_checked():
    """IMPLSCH / PROPAGS before PREPARE_PROPAGATION report WAM_E_STATE."""
    cfg, mask, f = W.c1_case()
    g = WamGpu(cfg, mask)
    n = g.nsea
    fl3 = W.jonswap_spectrum(cfg, f["udir"]).ravel()
    with pytest.raises(WamError) as e:
        g.implsch(fl3, f["u10"], f["udir"], np.zeros(n), np.zeros(n))
    assert e.value.code == 4
    with pytest.raises(WamError) as e:
        g.propags(fl3)
    assert e.value.code == 4


def test_depth_and_indep_must_match_the_prepared_fields(case):
    """IMPLSCH's DEPTH and INDEP arguments (wamodel.f90:354 passes the module
    arrays PREPARE_PROPAGATION used): the prepared ones are accepted and give
    the same result as none; a changed depth, or a changed INDEP in shallow
    water, reports WAM_E_STATE instead of silently using the old fields."""
    g, o, cfg, f = case["g"], case["o"], case["cfg"], case["f"]
    n = g.nsea
    z = np.zeros(n, np.float32)
    depth = np.asarray(f["depth"], np.float32)
    indep = g.table("INDEP").astype(np.int32)
    a = g.implsch(case["fl3"], f["u10"], f["udir"], z, z)
    b = g.implsch(case["fl3"], f["u10"], f["udir"], z, z, depth=depth, indep=indep)
    assert bits_equal(a["fl3"], b["fl3"])
    with pytest.raises(WamError) as e:
        g.implsch(case["fl3"], f["u10"], f["udir"], z, z, depth=depth * np.float32(1.5), indep=indep)
    assert e.value.code == 4
    if cfg["shallow"]:
        with pytest.raises(WamError) as e:
            g.implsch(case["fl3"], f["u10"], f["udir"], z, z, depth=depth, indep=indep + 1)
        assert e.value.code == 4
