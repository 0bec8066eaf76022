"""The oracle's nest-exchange arithmetic (oracle/orc_nest.c) against golden
vectors of the reference's own TOTAL_ENERGY, FEMEAN, MEAN_DIRECTION, INTSPEC,
ROTSPEC and STRSPEC (wam_interface_module.f90), built from the reference
sources with flang (tests/golden/make_ref_nest.py).  The reference's run used
flang's runtime for ATAN2, SIN, COS, LOG10 and REAL**INTEGER, the oracle
wam_math.h: single values may differ in the last bit, hence the tolerances."""
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from wam_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["kl24_ml25", "kl36_ml36"]


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-6 * np.max(np.abs(b)) + 1e-30)))


@pytest.fixture(scope="module", params=CASES)
def case(request):
    g = np.load(os.path.join(HERE, "golden", f"ref_nest_{request.param}.npz"))
    cfg = W.base_config(kl=int(g["kl"]), ml=int(g["ml"]))
    assert cfg["fr1"] == pytest.approx(float(g["fr1"])) and cfg["iref"] == int(g["iref"])
    o = Oracle(cfg, np.ones((cfg["ny"], cfg["nx"]), np.uint8))
    return o, g


def oracle_means(o, f):
    n = f.shape[0]
    fl3 = np.zeros((o.ml, o.kl, o.nsea), np.float32)
    fl3[:, :, :n] = f.transpose(1, 2, 0)
    return o.boundary_output(fl3, np.arange(1, n + 1))[1]


def test_mean_parameters(case):
    o, g = case
    m = oracle_means(o, g["f"])
    # EMEAN and FMEAN are sums and quotients in the reference's order: exact;
    # THQ goes through ATAN2
    assert np.array_equal(m[0], g["mean"][0]) and np.array_equal(m[2], g["mean"][2])
    assert rel(m[1], g["mean"][1]) < 1e-6


def test_rotspec_strspec(case):
    o, g = case
    f = g["f"]
    for (i, x), want in zip(g["rots"], g["rotspec"]):
        assert np.array_equal(o.rotspec(f[int(i) - 1], np.float32(x)), want), x
    for (i, x), want in zip(g["strs"], g["strspec"]):
        got = o.strspec(f[int(i) - 1], np.float32(x))
        assert rel(got, want) < 2e-6, x


def test_intspec(case):
    o, g = case
    f, mean = g["f"], g["mean"]
    worst = 0.0
    for (i1, i2, d12, d1l), want, wm in zip(g["pairs"], g["intspec"], g["intspec_mean"]):
        a, b = int(i1) - 1, int(i2) - 1
        got, mo = o.intspec(np.float32(d12), np.float32(d1l), f[a], mean[:, a], f[b], mean[:, b])
        # an ulp of ATAN2/SIN/COS moves ROTSPEC's and STRSPEC's weights by about
        # 1e-7, visible where a bin is empty and its neighbour is not: an
        # absolute floor of 2e-7 of the spectrum's maximum
        err = np.abs(got.astype(np.float64) - want) - 2e-7 * np.abs(want).max()
        worst = max(worst, float(np.max(err / np.maximum(np.abs(want), 1e-30))))
        assert np.all(err <= 5e-5 * np.abs(want)), (i1, i2, d12, d1l)
        # (EMEAN, THQ, FMEAN) against the reference's FMEAN, EMEAN, THETM
        np.testing.assert_allclose(mo, [wm[1], wm[2], wm[0]], rtol=1e-6, atol=1e-7)
    print(f"INTSPEC worst relative difference {worst:.2e}")
