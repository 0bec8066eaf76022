"""The off-by-default sources on the GPU (SET_MODEL_OPTION's WAVE_BREAKING:
SDEPTHLIM and SFBRK with CMPQB; PHILLIPS: SOURCE_PHILLIPS;
wam_source_module.f90:513-515,639-646) against the oracle, bit for bit,
through the C-ABI: the three-kernel IMPLSCH (Janssen, KL % 4 == 0), the
single kernel (30 directions; ST6), ensembles and the no-allocation entry
point.  The oracle's forms of these sources are pinned by the reference-built
goldens (tests/golden/ref_implsch_*break*, *phillips*); the GPU is held to
the same goldens in test_gpu_ref_golden.py."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from wam_amd import WamGpu, WamError, workloads as W

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.array_equal(a.view(np.int32), b.view(np.int32))


def shallow_case(iphys=0, kl=24, ml=25, seed=12):
    """A coastal sub-box whose depths run from 0.6 to 24 m, so that the depth
    limit caps the energy of the shallowest points and a part of the waves
    break everywhere else."""
    cfg, mask, f = W.c2_case(ny=24, nx=32, seed=seed, iphys=iphys, kl=kl, ml=ml)
    n = f["depth"].size
    rng = np.random.default_rng(seed)
    f = dict(f, depth=np.geomspace(0.6, 24.0, n).astype(np.float32)[rng.permutation(n)])
    return cfg, mask, f


def calm_deep_case():
    """Deep water (999 m) and a calm sea: the JONSWAP spectrum scaled by 1e-4,
    where the Phillips term is a visible part of the growth."""
    cfg, mask, f = W.c2_case(ny=20, nx=24, seed=21, iphys=0, kl=24, ml=25)
    cfg = dict(cfg, shallow=0, refr_d=0, refr_c=0)
    f = dict(f, depth=np.full_like(f["depth"], 999.0), u=None, v=None)
    return cfg, mask, f


CASES = {
    "split_shallow": lambda: shallow_case(),
    "split_shallow_36x36": lambda: shallow_case(kl=36, ml=36, seed=13),
    "single_kernel_kl30": lambda: shallow_case(kl=30, ml=25, seed=14),
    "st6_shallow": lambda: shallow_case(iphys=2, seed=15),
    "st6_shallow_three_kernels": lambda: shallow_case(iphys=2, seed=17),
    "split_deep_calm": calm_deep_case,
}
OPTIONS = {"breaking": (1, 0), "phillips": (0, 1), "both": (1, 1)}


def run_pair(cfg, mask, f, brk, phil, nstep, scale=1.0, lcflx_first=True, three=None):
    cfg = dict(cfg, wave_breaking=brk, phillips=phil)
    o = Oracle(cfg, mask)
    g = WamGpu(cfg, mask)
    g.set_implsch_form(three)
    u = f["u"] if cfg["refr_c"] else None
    v = f["v"] if cfg["refr_c"] else None
    o.prepare_propagation(f["depth"], u, v)
    g.prepare_propagation(f["depth"], u, v)
    n = g.nsea
    fo = fg = (W.jonswap_spectrum(cfg, f["udir"]).ravel() * scale).astype(np.float32)
    two = tzo = tzg = twg = np.zeros(n, np.float32)
    for step in range(nstep):
        lc = lcflx_first and step == 0
        ro = o.implsch(fo, f["u10"], f["udir"], two, tzo, f["depth"], lcflx=lc)
        rg = g.implsch(fg, f["u10"], f["udir"], twg, tzg, lcflx=lc)
        for k in ("fl3", "ustar", "tauw", "z0", "phiaw"):
            assert bits_equal(rg[k], ro[k]), (step, k)
        if lc:
            assert bits_equal(rg["flux"], ro["flux"]), step
        fo, two, tzo = ro["fl3"], ro["tauw"], ro["z0"]
        fg, twg, tzg = rg["fl3"], rg["tauw"], rg["z0"]
    g.close()
    return fo, fg


@pytest.mark.parametrize("opt", sorted(OPTIONS))
@pytest.mark.parametrize("name", sorted(CASES))
def test_sources_bit_exact(name, opt):
    cfg, mask, f = CASES[name]()
    brk, phil = OPTIONS[opt]
    scale = 1e-4 if name == "split_deep_calm" else 1.0
    print(f"[case] {name} {opt} iphys={cfg['iphys']} shallow={cfg['shallow']} {cfg['kl']}x{cfg['ml']}", flush=True)
    three = True if name.endswith("three_kernels") else None
    fo, fg = run_pair(cfg, mask, f, brk, phil, nstep=3, scale=scale, three=three)
    # the sources change the result (a switch that did nothing would pass the comparison)
    plain, _ = run_pair(cfg, mask, f, 0, 0, nstep=3, scale=scale, lcflx_first=False)
    if cfg["shallow"] or phil:
        assert not bits_equal(fo, plain)


def test_breaking_stays_off_in_deep_water():
    cfg, mask, f = calm_deep_case()
    g = WamGpu(cfg, mask)
    assert g.set_source_options(True, True) == (False, True)
    cs, ms, _ = shallow_case()
    gs = WamGpu(cs, ms)
    assert gs.set_source_options(True, False) == (True, False)
    assert gs.set_source_options(False, False) == (False, False)


def test_sources_in_ensembles_and_no_alloc():
    """Members advanced together with the sources on equal plain runs of each
    member, and the no-allocation IMPLSCH equals the allocating one."""
    import torch
    cfg, mask, f = shallow_case(kl=36, ml=36, seed=16)
    g = WamGpu(dict(cfg, wave_breaking=1, phillips=1), mask)
    g.prepare_propagation(f["depth"], f["u"], f["v"])
    n = g.nsea
    nmem = 3
    dev = torch.device("cuda:0")
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    rng = np.random.default_rng(3)
    u10 = np.clip(np.stack([f["u10"] * (1.0 + 0.15 * rng.standard_normal(n)) for _ in range(nmem)]), 0.5, None)
    udir = np.stack([f["udir"] + 0.3 * m for m in range(nmem)])
    fl3 = np.stack([W.jonswap_spectrum(cfg, udir[m]).ravel() * (1.0 + 0.2 * m) for m in range(nmem)])
    z = lambda k=1: torch.zeros(k * n, device=dev)
    ro = torch.full((nmem * n,), 1.225, device=dev)
    # the ensemble
    fe, tw, us, z0 = t(fl3.ravel()), z(nmem), z(nmem), z(nmem)
    g.implsch_ensemble_dev(nmem, fe, t(u10.ravel()), t(udir.ravel()), tw, us, z0, ro, z(nmem))
    for m in range(nmem):
        for entry in ("plain", "noalloc"):
            fm, twm, usm, z0m = t(fl3[m]), z(), z(), z()
            call = g.implsch_dev if entry == "plain" else g.implsch_dev_noalloc
            call(fm, t(u10[m]), t(udir[m]), twm, usm, z0m, ro[:n], z())
            torch.cuda.synchronize()
            nb = fm.numel()
            assert bits_equal(fm.cpu().numpy(), fe[m * nb:(m + 1) * nb].cpu().numpy()), (m, entry)
            for a, b in ((twm, tw), (usm, us), (z0m, z0)):
                assert bits_equal(a.cpu().numpy(), b[m * n:(m + 1) * n].cpu().numpy()), (m, entry)
    g.close()


def test_adjoint_refuses_phillips():
    import torch
    cfg, mask, f = calm_deep_case()
    g = WamGpu(dict(cfg, phillips=1), mask)
    g.prepare_propagation(f["depth"])
    n, nb = g.nsea, g.nsea * cfg["kl"] * cfg["ml"]
    dev = torch.device("cuda:0")
    z = lambda k: torch.zeros(k, device=dev)
    with pytest.raises(WamError, match="PHILLIPS"):
        g.step_adjoint_dev(z(nb), z(n), z(n), z(n), torch.full((n,), 1.225, device=dev), z(n), z(nb), z(nb), z(n),
                           z(n), z(n), z(n))
