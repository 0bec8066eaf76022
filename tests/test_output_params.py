"""The oracle's integrated output parameters (oracle/orc_output.c, after
COMPUTE_OUTPUT_PARAMETER, wam_output_module.f90:481-601) pinned by closed
forms: a spectrum in one bin, where every period is 1/FR and the direction is
the bin's; cos^2 spreading, whose circular spread is known analytically; the
axis cases of the current direction; and the FP64 restatement."""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from wam_amd import workloads as W

OUT = ("u10", "udir", "ustar", "cd", "charnock", "depth", "current_speed", "current_dir", "hs", "tp", "tm", "tm1",
       "tm2", "mdir", "spread", "tauwn", "z0")
IX = {n: i for i, n in enumerate(OUT)}


def deep_case():
    cfg = W.base_config(kl=24, ml=25, nx=4, ny=4)
    return cfg, np.ones((4, 4), np.uint8)


def run(o, fl3, n, u10=10.0, udir=0.5, ustar=0.3, tauw=0.05, z0=1e-3, depth=50.0, u=None, v=None, orient=1):
    f = lambda x: np.full(n, x, np.float32) if np.isscalar(x) else np.asarray(x, np.float32)
    return o.output_parameters(fl3, f(u10), f(udir), f(ustar), f(tauw), f(z0), f(depth), u, v, orient)


@pytest.mark.parametrize("k0,m0", [(3, 7), (17, 12), (0, 1)])
def test_single_bin_spectrum(k0, m0):
    cfg, mask = deep_case()
    o = Oracle(cfg, mask)
    n, kl, ml = o.nsea, cfg["kl"], cfg["ml"]
    a = np.float32(0.25)
    fl3 = np.zeros((ml, kl, n), np.float32)
    fl3[m0, k0, :] = a
    out = run(o, fl3.ravel(), n)
    fr, dfim, th = o.get("FR"), o.get("DFIM"), o.get("TH")
    e = float(a) * float(dfim[m0])
    assert np.allclose(out[IX["hs"]], 4.0 * math.sqrt(e), rtol=1e-6)
    assert np.all(out[IX["tp"]] == np.float32(1.0) / fr[m0])
    for p in ("tm", "tm1", "tm2"):            # E / (E*FR), E / (E*FR), sqrt(E / (E*FR^2))
        assert np.allclose(out[IX[p]], 1.0 / float(fr[m0]), rtol=2e-6), p
    assert np.allclose(out[IX["mdir"]], math.degrees(float(th[k0])), atol=2e-4)
    assert np.all(out[IX["spread"]] < 0.05)


def test_top_frequency_tails():
    """Energy in the top frequency ML only: every moment then carries the
    f**-5 tail beyond FR(ML) (wam_fre_dir_module.f90:406-409, EX_TAIL = -5),
    M0 = a (DFIM + DELTH FR/4), M1 = a (DFIM FR + DELTH FR**2/3),
    M2 = a (DFIM FR**2 + DELTH FR**3/2), so TM1 = M0/M1 and TM2 = SQRT(M0/M2)
    (TM1_TM2_PERIODS_B, wam_interface_module.f90:1210-1243); a dropped tail or a
    wrong tail exponent moves them by several per cent."""
    cfg, mask = deep_case()
    o = Oracle(cfg, mask)
    n, kl, ml = o.nsea, cfg["kl"], cfg["ml"]
    a = 0.04
    fl3 = np.zeros((ml, kl, n), np.float32)
    fl3[ml - 1, 5, :] = a
    out = run(o, fl3.ravel(), n)
    fr, dfim = float(o.get("FR")[ml - 1]), float(o.get("DFIM")[ml - 1])
    delth = 2.0 * math.pi / kl
    m0 = a * (dfim + delth * fr / 4.0)
    m1 = a * (dfim * fr + delth * fr ** 2 / 3.0)
    m2 = a * (dfim * fr ** 2 + delth * fr ** 3 / 2.0)
    assert np.allclose(out[IX["hs"]], 4.0 * math.sqrt(m0), rtol=2e-6)
    assert np.allclose(out[IX["tm1"]], m0 / m1, rtol=2e-6)
    assert np.allclose(out[IX["tm2"]], math.sqrt(m0 / m2), rtol=2e-6)
    # without the tail TM2 would be 1/FR: the tail's share is not negligible
    assert abs(float(out[IX["tm2"]][0]) * fr - 1.0) > 0.02


def test_cos2_spreading_about_the_wind():
    """(2/pi) cos^2 spreading: mean resultant length 8/(3 pi), spread
    sqrt(2 (1 - 8/(3 pi))) rad, mean direction the wind's."""
    cfg, mask, f = W.c1_case()
    o = Oracle(cfg, mask)
    n = o.nsea
    udir = np.full(n, np.float32(1.1), np.float32)
    fl3 = W.jonswap_spectrum(cfg, udir).ravel()
    out = run(o, fl3, n, udir=udir)
    assert np.allclose(out[IX["spread"]], math.degrees(math.sqrt(2 * (1 - 8 / (3 * math.pi)))), rtol=2e-3)
    # on the discrete direction grid the mean direction is the resultant of
    # the sampled weights max(0, cos(th - u))^2
    th = o.get("TH").astype(np.float64)
    w = np.maximum(0.0, np.cos(th - 1.1)) ** 2
    mdir = math.degrees(math.atan2((w * np.sin(th)).sum(), (w * np.cos(th)).sum()))
    assert abs(mdir - math.degrees(1.1)) < 0.05
    assert np.allclose(out[IX["mdir"]], mdir, atol=2e-4)
    assert np.allclose(out[IX["udir"]], math.degrees(1.1), atol=1e-4)
    # the same field "going to": +180 degrees
    out0 = run(o, fl3, n, udir=udir, orient=0)
    assert np.allclose(out0[IX["mdir"]], mdir + 180.0, atol=2e-4)


def test_wind_and_current_fields():
    cfg, mask = deep_case()
    cfg = dict(cfg, shallow=1)
    o = Oracle(cfg, mask)
    n = o.nsea
    fl3 = np.full(cfg["kl"] * cfg["ml"] * n, 0.01, np.float32)
    u = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    u[0], v[1], v[2], u[3] = 1.0, -1.0, 1.0, -1.0
    out = run(o, fl3, n, u10=0.0, ustar=0.0, tauw=0.0, z0=2e-3, depth=37.5, u=u, v=v)
    assert np.all(out[IX["cd"]] == np.float32(0.0001) / np.float32(0.01))
    assert np.all(out[IX["charnock"]] == (np.float32(9.806) * np.float32(2e-3)) / np.float32(1e-6))
    assert np.all(out[IX["depth"]] == np.float32(37.5))
    assert np.allclose(out[IX["current_speed"]][:4], 1.0)
    assert np.allclose(out[IX["current_dir"]][:4], [90.0, 180.0, 0.0, 270.0], atol=1e-4)
    out0 = run(o, fl3, n, u10=0.0, ustar=0.0, tauw=0.0, z0=2e-3, u=u, v=v, orient=0)
    assert np.allclose(out0[IX["current_dir"]][:4], [270.0, 0.0, 180.0, 90.0], atol=1e-4)
    deep = Oracle(deep_case()[0], mask)
    assert np.all(run(deep, fl3, n)[IX["depth"]] == 999.0)


def test_fp64_restatement_agrees():
    cfg, mask, f = W.c2_case(ny=20, nx=30, seed=3)
    o32, o64 = Oracle(cfg, mask), Oracle(cfg, mask, fp64=True)
    n = o32.nsea
    rng = np.random.default_rng(1)
    fl3 = (rng.random(cfg["kl"] * cfg["ml"] * n) * 0.1).astype(np.float32)
    args = dict(u10=f["u10"], udir=f["udir"], ustar=0.4, tauw=0.1, z0=1e-3, depth=f["depth"], u=f["u"], v=f["v"])
    a = run(o32, fl3, n, **args)
    b = o64.output_parameters(fl3.astype(np.float64), f["u10"], f["udir"], np.full(n, 0.4), np.full(n, 0.1),
                              np.full(n, 1e-3), f["depth"], f["u"], f["v"], 1)
    for i, name in enumerate(OUT):
        if name in ("mdir", "current_dir"):
            # random spectra are nearly isotropic: the mean direction is the
            # angle of a short resultant, so FP32 sums move it by ~1e-3 deg
            d = np.abs(a[i] - b[i])
            assert np.max(np.minimum(d, 360.0 - d)) < 1e-2, name
        else:
            assert np.allclose(a[i], b[i], rtol=2e-5, atol=1e-7), name
