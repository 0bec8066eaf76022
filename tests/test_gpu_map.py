"""GPU gridding of the integrated parameters for the MAP file
(WRITE_INT_PAR_OUTPUT, wam_output_module.f90:786-911): the grids written by
wam_gpu_write_map equal the oracle's parameters placed at GRID(IXLG, KXLT),
ZMISS on land, bit for bit."""
import struct

import numpy as np
import pytest

from oracle.oracle import Oracle
from wam_amd import WamError, WamGpu, workloads as W

pytestmark = pytest.mark.gpu
PARAMS = list(range(1, 17)) + [32]


def records(path):
    recs = []
    with open(path, "rb") as fh:
        while True:
            head = fh.read(4)
            if not head:
                return recs
            n = struct.unpack("<i", head)[0]
            recs.append(fh.read(n))
            assert struct.unpack("<i", fh.read(4))[0] == n


def test_write_map_matches_oracle(tmp_path):
    import torch
    cfg, mask, f = W.c2_case(ny=30, nx=40, seed=12, kl=24, ml=25)
    g, o = WamGpu(cfg, mask), Oracle(cfg, mask)
    n = g.nsea
    g.prepare_propagation(f["depth"], f["u"], f["v"])
    o.prepare_propagation(f["depth"], f["u"], f["v"])
    rng = np.random.default_rng(3)
    fl3 = W.jonswap_spectrum(cfg, f["udir"])
    ustar, tauw, z0 = (rng.uniform(0.1, 1.0, n).astype(np.float32) * s for s in (1.0, 0.5, 1e-3))
    want = o.output_parameters(fl3, f["u10"], f["udir"], ustar, tauw, z0, f["depth"], f["u"], f["v"], orient=1)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda()
    out = torch.empty(17 * n, device="cuda")
    g.output_parameters_dev(t(fl3), t(f["u10"]), t(f["udir"]), t(ustar), t(tauw), t(z0), out, orient=1)
    p = str(tmp_path / "MAP")
    print("write_map on a 30 x 40 grid", flush=True)
    g.write_map(p, "20240101060000", "20240101000000", PARAMS, out)
    g.write_map(p, "20240101120000", "20240101000000", [9, 32], out, append=True)
    recs = records(p)
    assert len(recs) == 3 + 17 + 3 + 2
    ny, nx = mask.shape
    rows, cols = np.nonzero(mask)
    for r0, params in ((0, PARAMS), (20, [9, 32])):
        cd, fnx, fny, w, s, e, nn, ca = struct.unpack("<14sff4i14s", recs[r0])
        assert (fnx, fny, [w, s, e, nn]) == (nx, ny, [cfg["amowep"], cfg["amosop"], cfg["amoeap"], cfg["amonop"]])
        assert list(np.flatnonzero(np.frombuffer(recs[r0 + 2], np.int32)) + 1) == params
        for j, ip in enumerate(params):
            grid = np.full((ny, nx), -9999999.0, np.float32)
            grid[rows, cols] = want[ip - 1 if ip <= 16 else 16]
            got = np.frombuffer(recs[r0 + 3 + j], np.float32)
            assert np.array_equal(got.view(np.int32), grid.ravel().view(np.int32)), ip
    with pytest.raises(WamError):
        g.write_map(p, "20240101180000", "20240101000000", [17], out, append=True)
    g.close()
