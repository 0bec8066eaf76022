"""The device build of wam_math.h against the host build (the oracle's
orc_api_math), bit for bit: every function on strided sweeps of the argument
ranges of tests/test_math_cr.py plus special values, through the C-ABI
wam_gpu_eval_math.  Host and device evaluate the same double-precision
operations, so any difference is a compiler or hardware difference."""
import ctypes

import numpy as np
import pytest

from oracle.oracle import shared_math

pytestmark = pytest.mark.gpu

SPECIAL = np.array([0.0, -0.0, 1.0, -1.0, 0.5, 2.0, 1e-45, 1.17549435e-38, 3.4028235e38, np.inf, -np.inf, np.nan,
                    88.72283935546875, 88.7228394, -103.97208, -87.33655, 9.01, 10.0, 89.41598, 0.6931472,
                    0.34657359, 0.17328680, 1.5707964, 3.1415927, 0.7853982, 65535.0, 1e-8, -1e-8], np.float32)


def floats(lo, hi, stride, signed=True):
    b = np.arange(np.float32(lo).view(np.int32), np.float32(hi).view(np.int32), stride, dtype=np.int64)
    x = b.astype(np.int32).view(np.float32)
    return np.concatenate([x, -x, SPECIAL]) if signed else np.concatenate([x, SPECIAL])


def device(fn, x, y=None):
    from wam_amd import _lib
    lib = _lib.load()
    x = np.ascontiguousarray(x, np.float32)
    y = None if y is None else np.ascontiguousarray(y, np.float32)
    out = np.empty_like(x)
    p = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    assert lib.wam_gpu_eval_math(fn, p(x), p(y), p(out), x.size) == 0
    return out


def same(a, b, x=None, y=None):
    ok = (a.view(np.int32) == b.view(np.int32)) | (np.isnan(a) & np.isnan(b))
    if not ok.all() and x is not None:
        bad = np.flatnonzero(~ok)
        print(f"{bad.size} differ; first:", flush=True)
        for i in bad[:12]:
            extra = "" if y is None else f" y={y[i]!r}"
            print(f"  x={x[i]!r} ({x[i].view(np.int32):#010x}){extra} device={a[i]!r} host={b[i]!r}", flush=True)
    return bool(ok.all())


@pytest.mark.parametrize("fn,lo,hi", [(0, 0.0, 104.0), (1, 1e-45, 3.4e38), (2, 1e-45, 3.4e38), (3, 0.0, 10.0),
                                      (4, 0.0, 90.0), (5, 0.0, 90.0), (6, 0.0, 65536.0), (7, 0.0, 65536.0),
                                      (9, 0.0, 1e6), (12, 1e-45, 3.4e38)])
def test_device_math_equals_host(fn, lo, hi):
    x = floats(lo, hi, 257, signed=fn not in (1, 2))
    print(f"[case] fn={fn} n={x.size}", flush=True)
    assert same(device(fn, x), shared_math(fn, x), x), fn


@pytest.mark.parametrize("fn", [8, 10, 11])
def test_device_math_two_arguments(fn):
    rng = np.random.default_rng(11 + fn)
    n = 2_000_000
    if fn == 8:
        x = np.exp(rng.uniform(-12, 14, n)).astype(np.float32)
        y = rng.uniform(-4, 4, n).astype(np.float32)
    elif fn == 10:
        x = rng.uniform(-30, 30, n).astype(np.float32)
        y = rng.integers(0, 12, n).astype(np.float32)
    else:
        x = (rng.normal(size=n) * 10.0 ** rng.uniform(-6, 6, n)).astype(np.float32)
        y = (rng.normal(size=n) * 10.0 ** rng.uniform(-6, 6, n)).astype(np.float32)
        x[:4], y[:4] = [0.0, -0.0, 0.0, -0.0], [-1.0, -1.0, 0.0, 0.0]
    print(f"[case] fn={fn} n={n}", flush=True)
    assert same(device(fn, x, y), shared_math(fn, x, y), x, y), fn
