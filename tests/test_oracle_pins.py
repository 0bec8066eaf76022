"""Pins the CPU oracle (oracle/) against the reference's own golden data,
closed forms, invariants and brute force, so that the GPU parity tests
compare against something trustworthy.  CPU only."""
import json
import math
import os

import numpy as np
import pytest

from oracle.oracle import Oracle, shared_math
from wam_amd import workloads as W

GOLD = os.path.join(os.path.dirname(__file__), "golden", "dia_12x25.json")


def regression_cfg():
    """The reference's regression case grid and spectrum (const/Preproc_User:36,79-90)."""
    cfg = W.base_config(kl=12, ml=25, fr1=0.04177248, iphys=2, shallow=1, idelpro=300, idelt=600, cdfac=1.08)
    return W.grid_config(cfg, 41, 41, -9.5, 0.0, 0.5, 0.5)


@pytest.fixture(scope="module")
def gold():
    with open(GOLD) as fh:
        return json.load(fh)


@pytest.fixture(scope="module")
def reg_oracle():
    cfg = regression_cfg()
    return Oracle(cfg, np.ones((41, 41), np.uint8))


def test_dia_tables_match_reference_protocol(gold, reg_oracle):
    """NLWEIGT/INIT_SNONLIN tables vs dayfiles/wam_prot_coarse_ST6:878-937:
    index tables bit-exact, weights to the printed 8 decimals (+/- a few ulp
    of the Intel runtime's arithmetic)."""
    o = reg_oracle
    for name in ("acl1", "acl2", "cl11", "cl21", "dal1", "dal2"):
        assert abs(o.scalar(name.upper()) - gold[name]) <= 2e-7 * max(1.0, abs(gold[name])), name
    assert o.scalar("ISP") == 2 and o.scalar("ISM") == -4 and o.scalar("MFRSTLW") == -3
    assert o.scalar("MLSTHG") == 29 and o.scalar("KFRH") == 8
    rows = np.array(gold["freq_rows"])
    assert np.array_equal(rows[:, 0], np.arange(-3, 30))
    for col, name in ((1, "IKP"), (2, "IKP1"), (3, "IKM"), (4, "IKM1")):
        assert np.array_equal(o.get(name), rows[:, col].astype(int)), name
    for col, name in ((5, "FKLAP"), (6, "FKLAP1"), (7, "FKLAM"), (8, "FKLAM1")):
        assert np.max(np.abs(o.get(name) - rows[:, col])) <= 3.5e-7, name
    af = o.get("AF11")
    assert np.allclose(af, rows[:, 9], rtol=6e-3), "AF11 (printed with 3 significant digits)"
    ang = np.array(gold["angular_rows"])
    for kh in range(2):
        for j, name in enumerate(("K1W", "K2W", "K11W", "K21W")):
            assert np.array_equal(o.get(name).reshape(2, 12)[kh], ang[:, 1 + 4 * kh + j]), (name, kh)
    assert np.max(np.abs(o.get("FRH") - np.array(gold["frh"]))) <= 4e-7


def test_frequency_direction_grid_closed_forms(reg_oracle):
    o = reg_oracle
    fr = o.get("FR").astype(np.float64)
    assert np.allclose(fr, 0.04177248 * 1.1 ** np.arange(25), rtol=2e-6)
    delth = 2 * math.pi / 12
    assert np.allclose(o.get("TH"), (np.arange(12) + 0.5) * delth, rtol=1e-6)
    df = o.get("DF").astype(np.float64)
    ref = 0.05 * np.r_[fr[0], fr[1:-1] + fr[:-2], fr[-2]]
    assert np.allclose(df, ref, rtol=2e-6)
    assert np.allclose(o.get("DFIM"), df * delth, rtol=2e-6)
    assert abs(o.scalar("MO_TAIL") - delth / 4 * fr[-1]) < 1e-6 * fr[-1]
    assert abs(o.scalar("MM1_TAIL") - delth / 5) < 1e-7
    assert np.allclose(o.get("GOM"), 9.806 / (4 * math.pi * fr), rtol=2e-6)


def test_shallow_tables_satisfy_dispersion(reg_oracle):
    """AKI (wam_general_module.f90:301-364) converges to omega^2 = g k tanh(kd)
    to its stopping tolerance; TCGOND is the analytic group velocity."""
    o = reg_oracle
    nd, ml = 69, 25
    k = o.get("TFAK").reshape(ml, nd).astype(np.float64)
    cg = o.get("TCGOND").reshape(ml, nd).astype(np.float64)
    fr = o.get("FR").astype(np.float64)
    d = 1.0 * 1.1 ** np.arange(nd)
    om = 2 * math.pi * fr[:, None]
    kd = k * d[None, :]
    deep = kd > 40.0
    lhs = om ** 2
    rhs = np.where(deep, 9.806 * k, 9.806 * k * np.tanh(kd))
    assert np.max(np.abs(lhs / rhs - 1.0)) < 3e-4
    cg_ref = np.where(kd <= 10.0, 0.5 * np.sqrt(9.806 * np.tanh(kd) / k) * (1 + 2 * kd / np.sinh(np.minimum(2 * kd, 700))),
                      9.806 / (4 * math.pi * fr[:, None]))
    assert np.max(np.abs(cg / cg_ref - 1.0)) < 2e-5


def test_wavnu2_dispersion():
    cfg, mask, f = W.c2_case(ny=12, nx=16, seed=11, iphys=2, kl=24, ml=25)
    o = Oracle(cfg, mask)
    wn, cg = o.wavnu2(f["depth"])
    wn = wn.astype(np.float64)
    fr = o.get("FR").astype(np.float64)
    h = f["depth"].astype(np.float64)[None, :]
    om2 = (2 * math.pi * fr[:, None]) ** 2
    assert np.max(np.abs(9.806 * wn * np.tanh(wn * h) / om2 - 1)) < 2e-6


def test_flminfr_is_one_percent_jonswap_floor(reg_oracle):
    """MIN_ENERGY: JONSWAP from the fetch law at 1 % alpha, floor 1e-6
    (wam_tables_module.f90:315-385), recomputed in float64."""
    o = reg_oracle
    fl = o.get("FLMINFR").reshape(25, 300).astype(np.float64)
    fr = o.get("FR").astype(np.float64)
    g = 9.806
    delphi = 0.5 * 40000000.0 / 360.0
    fetch = min(0.5 * delphi, 25000.0)
    u = np.arange(1, 301) * 0.25
    ug = g / u
    fp = np.maximum(0.13, 2.84 * ((g * fetch) / u ** 2) ** -0.3)
    fp = np.minimum(fp, fr[-1] / ug)
    al = 0.01 * np.maximum(0.0081, 0.033 * fp ** (2.0 / 3.0))
    fp = fp * ug
    f = fr[:, None]
    arg = 1.25 * (fp[None, :] / f) ** 4
    et = al[None, :] * g * g / (2 * math.pi) ** 4 * f ** -5 * np.exp(-arg)
    sig = np.where(f > fp[None, :], 0.09, 0.07)
    a2 = (f / fp[None, :] - 1) ** 2 / (2 * sig ** 2)
    et = np.where(a2 < 99, et * 3.0 ** np.exp(-a2), et)
    et = np.where(arg < 50, et, 0.0)
    ref = np.maximum(et, 1e-6)
    assert np.max(np.abs(fl / ref - 1)) < 5e-5


def test_neighbour_tables_brute_force():
    """PREPARE_GRID / CLOSEST_GP_LAT on a regular grid: W/E neighbours on the
    row (periodic if IPER), N/S neighbours in the same column, 0 for land."""
    rng = np.random.default_rng(7)
    for iper in (0, 1):
        ny, nx = 9, 13
        cfg = W.grid_config(W.base_config(iper=iper), nx, ny, -10.0, 0.0, 1.0, 1.0)
        if iper:
            cfg["amoeap"] = cfg["amowep"] + (nx - 1) * cfg["xdello"]
        mask = (rng.random((ny, nx)) > 0.3).astype(np.uint8)
        o = Oracle(cfg, mask)
        num = np.zeros((ny, nx), int)
        num[mask.astype(bool)] = np.arange(1, int(mask.sum()) + 1)
        klat = o.get("KLAT").reshape(2, 2, -1)
        klon = o.get("KLON").reshape(2, -1)
        ns = int(mask.sum())
        for ip in range(1, ns + 1):
            k, i = map(int, np.argwhere(num == ip)[0])
            w = num[k, i - 1] if i > 0 else (num[k, nx - 1] if iper else 0)
            e = num[k, i + 1] if i < nx - 1 else (num[k, 0] if iper else 0)
            s = num[k - 1, i] if k > 0 else 0
            n = num[k + 1, i] if k < ny - 1 else 0
            assert (klon[0, ip - 1], klon[1, ip - 1]) == (w, e)
            assert (klat[0, 0, ip - 1], klat[0, 1, ip - 1]) == (s, n)
            assert (klat[1, 0, ip - 1], klat[1, 1, ip - 1]) == (s, n)


def _interior_spectrum(o, seed=3):
    """A smooth, positive spectrum supported on interior frequencies only."""
    rng = np.random.default_rng(seed)
    ml, kl = o.ml, o.kl
    m = np.arange(ml)[:, None]
    k = np.arange(kl)[None, :]
    f = np.exp(-0.5 * ((m - ml / 2) / 2.5) ** 2) * (1.2 + np.cos(2 * math.pi * (k - 3) / kl)) * 0.5
    f = f * (1 + 0.1 * rng.random((ml, kl)))
    f[:6] = 0.0
    f[-6:] = 0.0
    return f


@pytest.mark.parametrize("kl,ml", [(12, 25), (24, 25), (36, 36)])
def test_snonlin_conserves_and_fl_is_its_diagonal_derivative(kl, ml):
    """DIA (SNONLIN, wam_source_module.f90:2742-3047): for a spectrum supported
    inside the resolved band the transfer sums to zero over the bins, and FL is
    the exact diagonal of dSL/dF (checked by central differences in REAL*8)."""
    cfg = W.base_config(kl=kl, ml=ml, nx=4, ny=4)
    mask = np.ones((4, 4), np.uint8)
    o = Oracle(cfg, mask, fp64=True)
    f = _interior_spectrum(o)
    sl, fl = o.snonlin(f)
    assert abs(sl.sum()) < 1e-10 * np.abs(sl).sum()
    rng = np.random.default_rng(1)
    idx = [tuple(x) for x in np.argwhere(f > 0.05 * f.max())]
    for j in rng.choice(len(idx), 12, replace=False):
        m, k = idx[j]
        eps = 1e-6 * f[m, k]
        fp_ = f.copy(); fp_[m, k] += eps
        fm_ = f.copy(); fm_[m, k] -= eps
        d = (o.snonlin(fp_)[0][m, k] - o.snonlin(fm_)[0][m, k]) / (2 * eps)
        assert abs(d - fl[m, k]) <= 1e-6 * (abs(fl[m, k]) + 1e-12 * np.abs(fl).max()), (m, k, d, fl[m, k])


def test_snonlin_fp32_against_fp64():
    cfg = W.base_config(kl=24, ml=25, nx=4, ny=4)
    mask = np.ones((4, 4), np.uint8)
    o32, o64 = Oracle(cfg, mask), Oracle(cfg, mask, fp64=True)
    f = _interior_spectrum(o64)
    s32, _ = o32.snonlin(f)
    s64, _ = o64.snonlin(f)
    assert np.max(np.abs(s32 - s64)) < 1e-4 * np.abs(s64).max()


def test_prep_spher_deep_closed_form():
    """PREP_SPHER_DEEP (wam_propagation_module.f90:1248-1327) for the 3600-point
    deep-water grid, recomputed in float64."""
    cfg, mask, fields = W.c1_case()
    o = Oracle(cfg, mask)
    counter, cfl = o.prepare_propagation(fields["depth"])
    assert counter == 1
    ns, kl, ml = o.nsea, cfg["kl"], cfg["ml"]
    coef = o.get("COEF").reshape(9, ml, kl, ns).astype(np.float64)
    g, circ = 9.806, 40000000.0
    fr = 0.0418 * 1.1 ** np.arange(ml)
    gom = g / (4 * math.pi * fr)
    delth = 2 * math.pi / kl
    th = (np.arange(kl) + 0.5) * delth
    lat = np.radians(-15.0 + 0.5 * (np.arange(ns) // 60))
    dco = 1 / np.cos(lat)
    delphi = 0.5 * circ / 360.0
    dt = 600.0
    dellam = delphi
    row = np.arange(ns) // 60
    dpsn_s = np.where(row > 0, dco / (1 / np.cos(np.radians(-15.0 + 0.5 * (row - 1)))), 1.0)
    dpsn_n = np.where(row < 59, dco / (1 / np.cos(np.radians(-15.0 + 0.5 * (row + 1)))), 1.0)
    della0 = dt * dco / dellam
    drgm = np.sin(lat) * dco
    delth0 = 0.5 * dt / (delth * circ / (2 * math.pi))
    delph0 = 0.5 * dt / delphi
    for kk in range(kl):
        sd, cd = math.sin(th[kk]), math.cos(th[kk])
        sp = delth0 * (sd + math.sin(th[(kk + 1) % kl]))
        sm = delth0 * (sd + math.sin(th[(kk - 1) % kl]))
        up, dn = (dpsn_s, dpsn_n) if cd >= 0 else (dpsn_n, dpsn_s)   # upwind side is JYO(K,1)
        dpn = abs(cd * delph0) * (up + 1)
        dle = abs(sd) * della0
        dtp, dtm = drgm * sp, drgm * sm
        dtc = abs(cd * delph0) * (dn + 1) + dle + np.maximum(0, dtp) - np.minimum(0, dtm)
        for mm in (0, 12, 24):
            assert np.allclose(coef[1, mm, kk], gom[mm] * dpn, rtol=3e-5, atol=1e-9)
            assert np.allclose(coef[3, mm, kk], gom[mm] * dle, rtol=3e-5, atol=1e-9)
            assert np.allclose(coef[0, mm, kk], 1 - gom[mm] * dtc, rtol=3e-5, atol=1e-6)
            assert np.allclose(coef[5, mm, kk], -gom[mm] * np.minimum(0, dtp), rtol=3e-5, atol=1e-9)
            assert np.allclose(coef[6, mm, kk], gom[mm] * np.maximum(0, dtm), rtol=3e-5, atol=1e-9)
    assert abs(cfl - np.max(1 - coef[0])) < 1e-6


def test_propags_constant_spectrum_and_land_slot():
    """With no refraction the update is linear and local: a spectrum that is zero
    everywhere stays zero, and a point whose neighbours are all land only keeps
    its own weighted value (the land slot holds zero, PROPAGS :712)."""
    cfg, mask, fields = W.c2_case(ny=20, nx=30, seed=2, iphys=0, kl=12, ml=12)
    o = Oracle(cfg, mask)
    o.prepare_propagation(fields["depth"], fields["u"], fields["v"])
    n = o.nsea
    z = np.zeros(n * 144, np.float32)
    assert not o.propags(z).any()
    # linearity: P(a x + y) = a P(x) + P(y) (exactly for a power of two)
    rng = np.random.default_rng(0)
    x = rng.random(n * 144).astype(np.float32)
    y = rng.random(n * 144).astype(np.float32)
    lhs = o.propags(2.0 * x + y).astype(np.float64)
    rhs = 2.0 * o.propags(x).astype(np.float64) + o.propags(y)
    assert np.max(np.abs(lhs - rhs)) < 1e-5 * np.abs(rhs).max()


def test_shared_math_accuracy():
    """wam_math.h single-precision functions vs float64 numpy: within half an
    ulp (correctly rounded; the float64 reference's own error is ~1e-16)."""
    rng = np.random.default_rng(5)

    def ulp_err(got, ref):
        ref32 = ref.astype(np.float32)
        u = np.spacing(np.abs(ref32)).astype(np.float64)
        return np.max(np.abs(got.astype(np.float64) - ref) / u)

    half = 0.5 + 1e-6
    x = rng.uniform(-85, 85, 200000).astype(np.float32)
    assert ulp_err(shared_math(0, x), np.exp(x.astype(np.float64))) <= half
    xp = np.exp(rng.uniform(-80, 80, 200000)).astype(np.float32)
    assert ulp_err(shared_math(1, xp), np.log(xp.astype(np.float64))) <= half
    assert ulp_err(shared_math(2, xp), np.log10(xp.astype(np.float64))) <= half
    xt = rng.uniform(-12, 12, 200000).astype(np.float32)
    assert ulp_err(shared_math(3, xt), np.tanh(xt.astype(np.float64))) <= half
    xc = rng.uniform(-50, 50, 200000).astype(np.float32)
    assert ulp_err(shared_math(4, xc), np.cosh(xc.astype(np.float64))) <= half
    assert ulp_err(shared_math(5, xc), np.sinh(xc.astype(np.float64))) <= half
    xs = rng.uniform(-15, 15, 200000).astype(np.float32)
    assert ulp_err(shared_math(6, xs), np.cos(xs.astype(np.float64))) <= half
    assert ulp_err(shared_math(7, xs), np.sin(xs.astype(np.float64))) <= half
    xw = np.exp(rng.uniform(-9, 13, 200000)).astype(np.float32)
    yw = rng.uniform(-4, 4, 200000).astype(np.float32)
    assert ulp_err(shared_math(8, xw, yw), np.power(xw.astype(np.float64), yw.astype(np.float64))) <= half
    ya = (rng.normal(size=200000) * 10.0 ** rng.uniform(-3, 3, 200000)).astype(np.float32)
    xa = (rng.normal(size=200000) * 10.0 ** rng.uniform(-3, 3, 200000)).astype(np.float32)
    a = np.arctan2(ya.astype(np.float64), xa.astype(np.float64))
    assert ulp_err(shared_math(11, ya, xa), a) <= half
    axes = shared_math(11, np.array([0, 1, -1, 0, 0], np.float32), np.array([1, 0, 0, -1, 0], np.float32))
    assert list(axes) == [0.0, np.float32(np.pi / 2), np.float32(-np.pi / 2), np.float32(np.pi), 0.0]
    # IEEE signed zeros: atan2(-0, -1) = -pi, sin(-0) = -0, tanh(-0) = -0
    neg = shared_math(11, np.array([-0.0], np.float32), np.array([-1.0], np.float32))
    assert neg[0] == np.float32(-np.pi)
    for fn in (3, 5, 7):
        assert np.signbit(shared_math(fn, np.array([-0.0], np.float32))[0])
    # exact cases: log10 of powers of ten, exp(0), cos(0)
    p10 = np.array([1, 10, 100, 1000, 1e4, 1e5, 1e6, 1e7, 1e8, 1e9, 1e10], np.float32)
    assert list(shared_math(2, p10)) == list(range(11))
    assert shared_math(0, np.zeros(1, np.float32))[0] == 1.0 and shared_math(6, np.zeros(1, np.float32))[0] == 1.0
    # Fortran NINT: halves away from zero
    v = np.array([2.5, -2.5, 0.49999997, -0.5, 1.5, 3.4999998], np.float32)
    assert list(shared_math(9, v)) == [3, -3, 0, -1, 2, 3]
