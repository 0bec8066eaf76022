"""CPU checks of the product library: every entry point declared in
include/wam_gpu.h is exported, and the host-built tables (the module data
IMPLSCH and PROPAGS read) are bit-identical to the oracle's."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle.oracle import Oracle
from wam_amd import WamGpu, WamError, workloads as W
from wam_amd import _lib


def header_functions(root):
    text = open(os.path.join(root, "include", "wam_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wam_gpu_\w+)\s*\(", text)))


def test_library_exports_every_declared_entry_point(root):
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = header_functions(root)
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, n


INT_TABLES = ["IKP", "IKP1", "IKM", "IKM1", "K1W", "K2W", "K11W", "K21W", "INLCOEF", "IXLG", "KXLT", "KLAT", "KLON"]
FLOAT_TABLES = ["FR", "DF", "DFIM", "DFIMOFR", "DFIM_FR", "FR5", "FRM5", "RHOWG_DFIM", "GOM", "TH", "COSTH", "SINTH",
                "FLMINFR", "TFAK", "TCGOND", "TSIHKD", "AF11", "FKLAP", "FKLAP1", "FKLAM", "FKLAM1", "FRH", "FTRF",
                "RNLCOEF", "SINPH", "COSPH", "DELLAM", "DPSN"]


@pytest.mark.parametrize("case", ["c1", "c2cut", "c2cut_st6", "regression_12x25"])
def test_host_tables_bit_exact_vs_oracle(case):
    if case == "c1":
        cfg, mask, _ = W.c1_case()
    elif case == "c2cut":
        cfg, mask, _ = W.c2_case(ny=50, nx=70, seed=9)
    elif case == "c2cut_st6":
        cfg, mask, _ = W.c2_case(ny=50, nx=70, seed=9, iphys=2, kl=24, ml=25)
    else:
        cfg = W.grid_config(W.base_config(kl=12, ml=25, fr1=0.04177248, iphys=2, shallow=1, idelpro=300,
                                          idelt=600), 41, 41, -9.5, 0.0, 0.5, 0.5)
        mask = np.ones((41, 41), np.uint8)
    o = Oracle(cfg, mask)
    p = WamGpu(cfg, mask, host_only=True)
    assert p.nsea == o.nsea
    for n in INT_TABLES:
        assert np.array_equal(p.table(n), o.get(n)), n
    for n in FLOAT_TABLES:
        assert np.array_equal(p.table(n).view(np.int32), o.get(n).view(np.int32)), n
    for s in ["ACL1", "ACL2", "CL11", "CL21", "DAL1", "DAL2", "ISP", "ISM", "MFRSTLW", "MLSTHG", "KFRH", "X0TAUHF",
              "DELPHI", "DELU", "INV_LOG_CO", "MO_TAIL", "MM1_TAIL", "MP1_TAIL", "ALPHA"]:
        assert np.float32(p.scalar(s)) == np.float32(o.scalar(s)), s


def test_snonlin_gather_lists_cover_every_scatter():
    """The ordered gather lists hold one entry per update the reference's
    SNONLIN scatter performs: 18 per interior quadruplet."""
    cfg = W.base_config(kl=36, ml=36, nx=4, ny=4)
    p = WamGpu(cfg, np.ones((4, 4), np.uint8), host_only=True)
    n = p.table("NCONTRIB")
    kl, mlsthg = 36, int(p.scalar("MLSTHG"))
    assert n.sum() <= 2 * kl * mlsthg * 9
    interior = n.reshape(36, 36)[8:-8]
    assert (interior == 18).all()


@pytest.mark.parametrize("kl,ml", [(36, 36), (24, 25), (12, 25)])
def test_chunked_gather_lists_replay_the_scatter_order(kl, ml):
    """The banded per-chunk lists the kernel reads, decoded and concatenated
    chunk by chunk, give back every bin's ordered contribution list."""
    cfg = W.base_config(kl=kl, ml=ml, nx=4, ny=4)
    p = WamGpu(cfg, np.ones((4, 4), np.uint8), host_only=True)
    nb = kl * ml
    n = np.asarray(p.table("NCONTRIB"))
    ncmax = int(p.scalar("NCMAX"))
    lists = np.asarray(p.table("CONTRIB")).view(np.int32).reshape(nb, ncmax)
    ch = np.asarray(p.table("SN_CH")).view(np.int32).reshape(-1, 5)
    e = np.asarray(p.table("SN_E")).view(np.int32)
    ww = np.asarray(p.table("SN_WW")).view(np.int32)
    chunk = int(p.scalar("CHUNK"))
    assert 2 * kl * chunk <= 511          # quadruplet slot 511 stays free for the padding entry
    got = [[] for _ in range(nb)]
    for c, (blo, bhi, w, off, wwoff) in enumerate(ch):
        nbc = bhi - blo
        assert w % 6 == 0
        for b in range(blo, bhi):
            ended = False
            for j in range(w):
                v = int(e[off + j * nbc + (b - blo)])
                if v == 0x01FF:
                    ended = True
                    continue
                assert not ended, "entry after the padding"
                assert j < ww[wwoff + (b - blo) // 64], "wavefront list length too short"
                ql, wi = v & 511, v >> 9
                mcl, ty = divmod(wi, 9)
                assert ql // (2 * kl) == mcl and ql < 511
                q = c * chunk * 2 * kl + ql
                got[b].append(q | (ty << 16))
    for b in range(nb):
        assert got[b] == list(lists[b, :n[b]]), b


@pytest.mark.parametrize("kl,ml", [(36, 36), (24, 25), (12, 25)])
def test_products_gather_lists_replay_the_scatter_order(kl, ml):
    """The per-run lists of the products form (byte offsets of update slots,
    two entries per 32-bit word), decoded chunk by chunk, give back every
    bin's ordered contribution list; padding is the zero slot only."""
    cfg = W.base_config(kl=kl, ml=ml, nx=4, ny=4)
    p = WamGpu(cfg, np.ones((4, 4), np.uint8), host_only=True)
    nb = kl * ml
    n = np.asarray(p.table("NCONTRIB"))
    ncmax = int(p.scalar("NCMAX"))
    lists = np.asarray(p.table("CONTRIB")).view(np.int32).reshape(nb, ncmax)
    rw = np.asarray(p.table("SP_RW")).view(np.int32).reshape(-1, 2)
    e = np.asarray(p.table("SP_E")).view(np.int32)
    chunk, nchunk, nrun = int(p.scalar("SP_CHUNK")), int(p.scalar("SP_NCHUNK")), int(p.scalar("SP_NRUN"))
    slots, stride = 512, 516                      # quadruplet slots of a chunk, update slots per type
    assert 2 * kl * chunk < slots                 # the last slot of type 0 stays free for the padding
    pad = (slots - 1) * 8
    assert nrun == (nb + 63) // 64 and len(rw) == nchunk * nrun
    got = [[] for _ in range(nb)]
    for c in range(nchunk):
        for r in range(nrun):
            off, w = rw[c * nrun + r]
            assert w % 2 == 0
            for l in range(64):
                b = 64 * r + l
                ended = False
                for i in range(w):
                    v = int(e[off + ((i // 2) * 64 + l) * 2 + (i % 2)])
                    if v == pad:
                        ended = True
                        continue
                    assert not ended, "entry after the padding"
                    assert b < nb and v % 8 == 0
                    ty, ql = divmod(v // 8, stride)
                    assert ty < 9 and ql < 2 * kl * chunk
                    got[b].append((c * chunk * 2 * kl + ql) | (ty << 16))
    for b in range(nb):
        assert got[b] == list(lists[b, :n[b]]), b


def test_bad_options_raise():
    cfg = W.base_config(iphys=1)
    with pytest.raises(WamError) as e:
        WamGpu(cfg, np.ones((60, 60), np.uint8), host_only=True)
    assert e.value.code == 1
    cfg = W.base_config()
    with pytest.raises(WamError) as e:
        WamGpu(cfg, np.zeros((60, 60), np.uint8), host_only=True)
    assert e.value.code == 5


def test_binding_arity_matches_header(root):
    """Each ctypes signature has as many parameters as the C prototype."""
    text = open(os.path.join(root, "include", "wam_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    for name, params in re.findall(r"\b(wam_gpu_\w+)\s*\(([^)]*)\)\s*;", text):
        n = 0 if params.strip() in ("", "void") else params.count(",") + 1
        assert len(_lib.SIGNATURES[name][1]) == n, name
