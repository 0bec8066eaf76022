"""The oracle's linearised model step (liborc_ad: the FP64 restatement of
PROPAGS then IMPLSCH recorded on a tape, oracle/orc_ad.cpp), pinned three
ways on a 33-point 24 x 25 grid:
  - its value equals the plain FP64 restatement's step (to FP64 round-off),
  - its tangent-linear model equals central finite differences of the step,
  - its adjoint satisfies the dot-product identity <L dx, ybar> = <dx, L^T ybar>.
A dropped term or a wrong sign in either sweep fails the second or third
check; the reference's own tests hold no tangent-linear or adjoint vectors."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from wam_amd import workloads as W


@pytest.fixture(scope="module")
def case():
    cfg, mask, f = W.c2_case(ny=6, nx=8, seed=3, kl=24, ml=25)
    o = Oracle(cfg, mask, ad=True)
    o.prepare_propagation(f["depth"], f["u"], f["v"])
    n = o.nsea
    fl3 = W.jonswap_spectrum(cfg, f["udir"]).ravel().astype(np.float64)
    u = f["u10"] * np.sin(f["udir"])
    v = f["u10"] * np.cos(f["udir"])
    tw = np.full(n, 0.05)
    z0 = np.full(n, 1e-3)
    return cfg, mask, f, o, n, fl3, u, v, tw, z0


def test_value_is_the_fp64_step(case):
    cfg, mask, f, o, n, fl3, u, v, tw, z0 = case
    y = o.step_ad(fl3, tw, z0, u, v, f["depth"])
    o64 = Oracle(cfg, mask, fp64=True)
    o64.prepare_propagation(f["depth"], f["u"], f["v"])
    u10 = np.sqrt(u * u + v * v)
    ud = np.arctan2(u, v)
    ud = np.where(ud < 0, ud + 2 * np.pi, ud)
    r = o64.implsch(o64.propags(fl3), u10, ud, tw.copy(), z0.copy(), f["depth"])
    nb = fl3.size
    assert np.max(np.abs(y[:nb] - r["fl3"]) / np.maximum(np.abs(r["fl3"]), 1e-30)) < 1e-12
    assert np.max(np.abs(y[nb:nb + n] - r["tauw"])) < 1e-14
    assert np.max(np.abs(y[nb + n:] - r["z0"]) / r["z0"]) < 1e-12


def test_tangent_linear_matches_finite_differences_and_adjoint_is_its_transpose(case):
    cfg, mask, f, o, n, fl3, u, v, tw, z0 = case
    nb = fl3.size
    rng = np.random.default_rng(1)
    x = np.concatenate([fl3, tw, z0, u, v])
    scale = np.concatenate([fl3 * 1e-2 + 1e-12, np.full(n, 1e-3), z0 * 1e-2, np.full(2 * n, 0.1)])
    dx = rng.standard_normal(x.size) * scale
    y0, dy = o.step_ad(fl3, tw, z0, u, v, f["depth"], dx=dx)

    def step(x):
        return o.step_ad(x[:nb], x[nb:nb + n], x[nb + n:nb + 2 * n], x[nb + 2 * n:nb + 3 * n], x[nb + 3 * n:],
                         f["depth"])

    h = 1e-4
    fd = (step(x + h * dx) - step(x - h * dx)) / (2 * h)
    assert np.linalg.norm(fd - dy) / np.linalg.norm(dy) < 1e-7
    # the TLM is not the identity: a check that passes for any map it is not
    assert np.linalg.norm(dy[:nb] - dx[:nb]) > 1e-3 * np.linalg.norm(dx[:nb])
    ybar = rng.standard_normal(y0.size)
    _, xbar = o.step_ad(fl3, tw, z0, u, v, f["depth"], ybar=ybar)
    lhs, rhs = np.dot(dy, ybar), np.dot(dx, xbar)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # and column by column on a few inputs of each kind
    for k in (0, nb // 2, nb + 3, nb + n + 5, nb + 2 * n + 7, nb + 3 * n + 11):
        e = np.zeros(x.size)
        e[k] = 1.0
        _, col = o.step_ad(fl3, tw, z0, u, v, f["depth"], dx=e)
        assert abs(np.dot(col, ybar) - xbar[k]) <= 1e-12 * max(1.0, np.abs(col).sum() * np.abs(ybar).max())


def test_fp32_tape_is_the_fp32_step_and_its_adjoint_the_transpose(case):
    """The REAL*4-valued tape (liborc_ad32): its value is the FP32
    restatement's step bit for bit (so its branches are that build's, which
    the GPU kernels reproduce), and its adjoint is the transpose of its
    tangent-linear model."""
    cfg, mask, f, o, n, fl3, u, v, tw, z0 = case
    o32 = Oracle(cfg, mask, ad32=True)
    o32.prepare_propagation(f["depth"])
    f32 = fl3.astype(np.float32)
    u32, v32 = u.astype(np.float32), v.astype(np.float32)
    y = o32.step_ad(f32, tw.astype(np.float32), z0.astype(np.float32), u32, v32, f["depth"])
    p = Oracle(cfg, mask)
    p.prepare_propagation(f["depth"])
    u10 = np.sqrt(u32 * u32 + v32 * v32)
    from oracle.oracle import shared_math
    ud = np.where(u10 != 0, shared_math(11, u32, v32), np.float32(0)).astype(np.float32)
    ud = np.where(ud < 0, (ud + np.float32(2 * np.pi)).astype(np.float32), ud)
    r = p.implsch(p.propags(f32), u10, ud, tw.astype(np.float32), z0.astype(np.float32), f["depth"])
    nb = fl3.size
    assert np.array_equal(y[:nb].astype(np.float32).view(np.int32), r["fl3"].view(np.int32))
    assert np.array_equal(y[nb:nb + n].astype(np.float32).view(np.int32), r["tauw"].view(np.int32))
    assert np.array_equal(y[nb + n:].astype(np.float32).view(np.int32), r["z0"].view(np.int32))
    rng = np.random.default_rng(2)
    dx = rng.standard_normal(nb + 4 * n)
    _, dy = o32.step_ad(f32, tw, z0, u32, v32, f["depth"], dx=dx)
    ybar = rng.standard_normal(nb + 2 * n)
    _, xbar = o32.step_ad(f32, tw, z0, u32, v32, f["depth"], ybar=ybar)
    lhs, rhs = np.dot(dy, ybar), np.dot(dx, xbar)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
