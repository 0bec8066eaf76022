"""The device-resident restart entry points (wam_gpu_save_restart,
wam_gpu_connect_restart): fields in HBM staged through pinned host memory,
against the host reader and writer bit for bit."""
import numpy as np
import pytest

import wam_amd
from wam_amd import WamError, WamGpu, workloads as W

pytestmark = pytest.mark.gpu

DATES = ["20240101000000", "20240101003000", "20240101000000", "20231231000000", "20231231120000"]


def test_save_and_connect_restart(tmp_path):
    torch = pytest.importorskip("torch")
    cfg, mask, f = W.c2_case(ny=40, nx=60, seed=5)
    g = WamGpu(cfg, mask)
    n, kl, ml = g.nsea, cfg["kl"], cfg["ml"]
    print(f"restart case nsea={n} {kl}x{ml}", flush=True)
    dev = torch.device("cuda:0")
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    fl3 = W.jonswap_spectrum(cfg, f["udir"]).reshape(ml, kl, n)
    rng = np.random.default_rng(2)
    tauw = rng.random(n, dtype=np.float32)
    p = str(tmp_path / "rst")
    with pytest.raises(WamError) as e:                 # depth and currents come from PREPARE_PROPAGATION
        g.save_restart(p, DATES, t(fl3), t(f["u10"]), t(f["udir"]), t(tauw))
    assert e.value.code == 4
    g.prepare_propagation(f["depth"], f["u"], f["v"])
    g.save_restart(p, DATES, t(fl3), t(f["u10"]), t(f["udir"]), t(tauw))
    dates, r = wam_amd.restart_read(p, n, kl, ml)
    assert dates == DATES
    for k, w in (("fl3", fl3), ("u10", f["u10"]), ("udir", f["udir"]), ("tauw", tauw), ("depth", f["depth"]),
                 ("u", f["u"]), ("v", f["v"])):
        assert np.array_equal(r[k], np.asarray(w, np.float32)), k
    # back into HBM, from a file of the host writer
    q = str(tmp_path / "rst2")
    fl3b = (fl3 * 2).astype(np.float32)
    wam_amd.restart_write(q, DATES[:3] + ["20231231000000", ""], fl3b, f["u10"], f["udir"], tauw, f["depth"])
    o = {k: torch.empty(sz, device=dev) for k, sz in (("fl3", n * kl * ml), ("u10", n), ("udir", n), ("tauw", n),
                                                      ("depth", n), ("u", n), ("v", n))}
    dates2, pres = g.connect_restart(q, o["fl3"], o["u10"], o["udir"], o["tauw"], o["depth"], o["u"], o["v"])
    assert dates2 == DATES[:3] + ["20231231000000", ""]
    assert pres == dict(depth=True, current=False)
    assert np.array_equal(o["fl3"].cpu().numpy(), fl3b.ravel())
    for k, w in (("u10", f["u10"]), ("udir", f["udir"]), ("tauw", tauw), ("depth", f["depth"])):
        assert np.array_equal(o[k].cpu().numpy(), np.asarray(w, np.float32)), k
    mask2 = np.array(mask, copy=True)                  # one more sea point: another grid
    iy, ix = np.argwhere(mask2 == 0)[0]
    mask2[iy, ix] = 1
    g2 = WamGpu(cfg, mask2)
    assert g2.nsea == n + 1
    o2 = {k: torch.empty(sz, device=dev) for k, sz in (("fl3", (n + 1) * kl * ml), ("u10", n + 1), ("udir", n + 1),
                                                       ("tauw", n + 1))}
    with pytest.raises(WamError) as e:
        g2.connect_restart(q, o2["fl3"], o2["u10"], o2["udir"], o2["tauw"])
    assert e.value.code == 6
