"""GPU parity of the nest boundary exchange (BOUNDARY_OUTPUT's gather and
mean parameters, BOUNDARY_INPUT's time and space interpolation by INTSPEC,
wam_boundary_module.f90:168-397) against the oracle, through the C-ABI: the
spectra, the mean values and FL3 after the insertion are bit-identical."""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from wam_amd import WamError, WamGpu, workloads as W

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.array_equal(a.view(np.int32), b.view(np.int32))


def field(cfg, n, seed):
    """FL3 [ML][KL][NSEA]: JONSWAP spectra of three peak frequencies, random
    directions and scales, and a few empty points."""
    rng = np.random.default_rng(seed)
    f = np.empty((cfg["ml"], cfg["kl"], n), np.float32)
    for j, fp in enumerate((0.12, 0.2, 0.33)):
        sl = slice(j, n, 3)
        k = len(range(*sl.indices(n)))
        f[:, :, sl] = W.jonswap_spectrum(cfg, rng.uniform(0, 2 * math.pi, k), fp=fp) * \
            rng.uniform(0.2, 3.0, k).astype(np.float32)
    f[:, :, rng.integers(0, n, 3)] = 0.0
    return f


@pytest.fixture(scope="module", params=[(36, 36), (24, 25)], ids=["kl36_ml36", "kl24_ml25"])
def pair(request):
    kl, ml = request.param
    cfg, mask, _ = W.c2_case(ny=30, nx=40, seed=9, kl=kl, ml=ml)
    g = WamGpu(cfg, mask)
    o = Oracle(cfg, mask)
    assert g.nsea == o.nsea
    yield cfg, g, o
    g.close()


def test_boundary_output(pair):
    cfg, g, o = pair
    n = g.nsea
    fl3 = field(cfg, n, 1)
    rng = np.random.default_rng(2)
    ij = np.concatenate([rng.integers(1, n + 1, 150), [1, n, 7, 7]]).astype(np.int32)
    print(f"boundary_output kl={cfg['kl']} ml={cfg['ml']} nb={ij.size}", flush=True)
    fbc, mean = g.boundary_output(fl3, ij)
    rf, rm = o.boundary_output(fl3, ij)
    assert bits_equal(fbc, rf)
    assert bits_equal(mean, rm)
    import torch
    d = torch.from_numpy(fl3).cuda()
    fbc2, mean2 = g.boundary_output(d, ij)
    assert bits_equal(fbc2, rf) and bits_equal(mean2, rm)
    with pytest.raises(WamError) as e:
        g.boundary_output(fl3, np.array([0, 1], np.int32))
    assert e.value.code == 1


@pytest.mark.parametrize("which,del1l", [(1, 0.0), (2, 0.0), (0, 900.0), (0, 1800.0), (0, 3599.0)])
def test_boundary_input(pair, which, del1l):
    cfg, g, o = pair
    n = g.nsea
    fl3 = field(cfg, n, 3)
    src = field(cfg, n, 4)
    rng = np.random.default_rng(5)
    nbinp = 48
    p1, p2 = rng.integers(0, n, nbinp), rng.integers(0, n, nbinp)
    f1 = np.ascontiguousarray(src[:, :, p1].transpose(2, 0, 1))
    f2 = np.ascontiguousarray(src[:, :, p2].transpose(2, 0, 1))
    _, m1 = o.boundary_output(src, p1 + 1)
    _, m2 = o.boundary_output(src, p2 + 1)
    m1, m2 = m1.copy(), m2.copy()
    # the special branches: no energy on one side, a mean frequency that
    # stretches out of the range, directions either side of zero
    f1[0] = 0.0
    m1[0, 0] = 0.0
    m2[0, 1] = 10.0 * m2[0, 1]
    m2[2, 2] = 50.0 * m2[2, 2]
    m1[2, 3] = 0.02 * m1[2, 3]
    m1[1, 4], m2[1, 4] = 0.01, 2 * math.pi - 0.01
    nbounf = 300
    ijarf = rng.permutation(n)[:nbounf] + 1
    ijarf[-1] = ijarf[10]                          # a point listed twice: the last entry stays
    ibfl = rng.integers(1, nbinp + 1, nbounf)
    ibfr = rng.integers(1, nbinp + 1, nbounf)
    bfw = rng.uniform(0, 1, nbounf).astype(np.float32)
    bfw[:4] = [0.0, 1.0, 0.0, 1.0]
    print(f"boundary_input kl={cfg['kl']} ml={cfg['ml']} which={which} del1l={del1l}", flush=True)
    want, rfi, rmi = o.boundary_input(fl3, f1, m1, f2, m2, which, 3600.0, del1l, ijarf, ibfl, ibfr, bfw)
    got = fl3.copy()
    fi, mi = g.boundary_input(got, f1 if which != 2 else None, m1 if which != 2 else None,
                              f2 if which != 1 else None, m2 if which != 1 else None, which, 3600.0, del1l,
                              ijarf, ibfl, ibfr, bfw)
    assert bits_equal(fi, rfi)
    assert bits_equal(mi, rmi)
    assert bits_equal(got, want)
    # the device form on FL3 in HBM
    import torch
    d = torch.from_numpy(fl3).cuda()
    g.boundary_input(d, f1, m1, f2, m2, which, 3600.0, del1l, ijarf, ibfl, ibfr, bfw)
    assert bits_equal(d.cpu().numpy(), want)
    with pytest.raises(WamError):
        g.boundary_input(fl3.copy(), f1, m1, f2, m2, which, 3600.0, del1l, ijarf, ibfl, ibfr + nbinp, bfw)


@pytest.mark.parametrize("name", ["kl24_ml25", "kl36_ml36"])
def test_against_reference_golden(name):
    """The HIP path against the reference's own BOUNDARY_OUTPUT means and
    INTSPEC (tests/golden/ref_nest_*.npz, the flang-built reference): the
    bars of the oracle's golden test."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", f"ref_nest_{name}.npz"))
    cfg = W.base_config(kl=int(g["kl"]), ml=int(g["ml"]))
    mask = np.ones((cfg["ny"], cfg["nx"]), np.uint8)
    gpu = WamGpu(cfg, mask)
    f, mean = g["f"], g["mean"]
    nb = f.shape[0]
    fl3 = np.zeros((gpu.ml, gpu.kl, gpu.nsea), np.float32)
    fl3[:, :, :nb] = f.transpose(1, 2, 0)
    print(f"reference golden {name}", flush=True)
    _, m = gpu.boundary_output(fl3, np.arange(1, nb + 1, dtype=np.int32))
    assert np.array_equal(m[0], mean[0]) and np.array_equal(m[2], mean[2])
    np.testing.assert_allclose(m[1], mean[1], rtol=1e-6)
    for (i1, i2, d12, d1l), want in zip(g["pairs"], g["intspec"]):
        a, b = int(i1) - 1, int(i2) - 1
        out = np.zeros_like(fl3)
        if d12 == 1.0:            # the space step: INTSPEC(1, BFW, FI(IBFL), FI(IBFR))
            gpu.boundary_input(out, f, mean, None, None, 1, 3600.0, 0.0, np.array([1], np.int32),
                               np.array([a + 1], np.int32), np.array([b + 1], np.int32), np.array([d1l], np.float32))
            got = out[:, :, 0]
        else:                     # the time step: INTSPEC(DEL12, DEL1L, F1, F2)
            fi, _ = gpu.boundary_input(out, f[a:a + 1], mean[:, a:a + 1], f[b:b + 1], mean[:, b:b + 1], 0, d12, d1l,
                                       np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32),
                                       np.zeros(0, np.float32))
            got = fi[0]
        err = np.abs(got.astype(np.float64) - want) - 2e-7 * np.abs(want).max()
        assert np.all(err <= 5e-5 * np.abs(want)), (i1, i2, d12, d1l)
    gpu.close()
