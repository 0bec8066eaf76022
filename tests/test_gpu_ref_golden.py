"""The HIP path against the reference's own IMPLSCH: the same golden vectors
(tests/golden/ref_implsch_*.npz, written by running
wam_source_module.f90:374-700 built from the reference sources) and the same
tolerances as the oracle's check in test_ref_golden.py, through the C-ABI
drop-in entry point wam_gpu_implsch."""
import os

import numpy as np
import pytest

from test_ref_golden import FIXTURES, OUT_FLUX, PROP_FIXTURES, check, check_propags, load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["auto", "three_kernels"])
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[12:-4] for p in FIXTURES])
def test_gpu_matches_reference_implsch(path, form):
    from wam_amd import WamGpu
    z, cfg = load(path)
    print(f"[case] {os.path.basename(path)} {form}", flush=True)
    g = WamGpu(cfg, z["mask"])
    g.set_implsch_form(True if form == "three_kernels" else None)
    n = z["in_u10"].size
    cur = np.zeros(n, np.float32) if cfg["refr_c"] else None      # IMPLSCH does not read currents
    g.prepare_propagation(z["in_depth"], cur, cur)
    fl3, tw, z0 = z["in_fl3"].copy(), z["in_tauw"].copy(), z["in_z0"].copy()
    for _ in range(int(z["nstep"])):
        q = g.implsch(fl3, z["in_u10"], z["in_udir"], tw, z0, roairn=z["in_roairn"], wstar=z["in_wstar"],
                      lcflx=bool(int(z["lcflx"])))
        fl3, tw, z0 = q["fl3"], q["tauw"], q["z0"]
    flux = dict(zip(OUT_FLUX, q["flux"]))
    check(z, cfg, lambda k: flux[k] if k in flux else q[k], g.table("DFIM"), g.table("FR"))
    g.close()


@pytest.mark.parametrize("path", PROP_FIXTURES, ids=[os.path.basename(p)[12:-4] for p in PROP_FIXTURES])
def test_gpu_matches_reference_propags(path):
    from wam_amd import WamGpu
    z, cfg = load(path)
    print(f"[case] {os.path.basename(path)}", flush=True)
    g = WamGpu(cfg, z["mask"])
    cur = (z["in_u"], z["in_v"]) if cfg["refr_c"] else (None, None)
    counter, cfl = g.prepare_propagation(z["in_depth"], *cur)
    f = z["in_fl3"]
    for _ in range(int(z["nprop"])):
        f = g.propags(f)
    check_propags(z, counter, cfl, g.table("INDEP") if cfg["shallow"] else None, f)
    g.close()
