"""Golden vectors of the reference's own PREPARE_PROPAGATION and PROPAGS
(myWAveModel/WAM src/mod/wam_propagation_module.f90:279-598, 635-829, with
PREPARE_PROPAGATION_COEF :837-1913 and CHECK_CFL :1917-2138), built from the
reference sources by oracle/Makefile.ref into oracle/_ref/ and driven by
oracle/ref_propags_driver.f90.  Run in the container that holds the
reference checkout:

    python tests/golden/make_ref_propags.py

Each case writes ref_propags_<case>.npz: the model options and sea mask
(the neighbour tables and latitude arrays are rebuilt from them and are what
the driver was given), depth and currents, the input spectrum, and the
reference's outputs after NPROP PROPAGS calls: the spectrum, INDEP, the
sub-step count and the max CFL number that PRINT_PROPAGATION_STATUS prints.
"""
import json
import os
import re
import struct
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
ORACLE = os.path.join(ROOT, "oracle")
DRIVER = os.path.join(ORACLE, "_ref", "ref_propags")


def _with(case, deep_depth=False, **kw):
    cfg, mask, f = case
    cfg = dict(cfg)
    cfg.update(kw)
    f = dict(f)
    if deep_depth:
        f["depth"] = np.full_like(f["depth"], 999.0)
    if not cfg.get("refr_c"):
        f["u"], f["v"] = None, None
    return cfg, mask, f


def cases():
    from wam_amd import workloads as W
    return {
        "shallow_curr_36x36": W.c2_case(ny=8, nx=12, seed=5),
        "shallow_curr_24x25": W.c2_case(ny=10, nx=14, seed=6, kl=24, ml=25),
        "shallow_depthrefr_24x30": _with(W.c2_case(ny=10, nx=14, seed=8, kl=24, ml=30), refr_c=0),
        "shallow_norefr_12x25": _with(W.c2_case(ny=10, nx=14, seed=8, kl=12, ml=25), refr_c=0, refr_d=0),
        "deep_curr_24x25": _with(W.c2_case(ny=10, nx=14, seed=4, kl=24, ml=25), shallow=0, deep_depth=True),
        "deep_c1grid_24x25": c1_corner(W),
    }


def c1_corner(W):
    """The 0.5 deg deep-water case's options on a 16 x 12 all-sea corner of its grid."""
    cfg, _, _ = W.c1_case()
    cfg = W.grid_config(cfg, 16, 12, -15.0, 0.0, 0.5, 0.5)
    mask = np.ones((12, 16), np.uint8)
    n = int(mask.sum())
    return cfg, mask, dict(depth=np.full(n, 999.0, np.float32), u=None, v=None,
                           udir=np.full(n, np.pi / 2, np.float32))


def inputs(cfg, mask, f):
    from oracle.oracle import Oracle
    from wam_amd import workloads as W
    o = Oracle(cfg, mask)
    n = o.nsea
    depth = np.asarray(f["depth"], np.float32)
    u = np.zeros(n, np.float32) if f.get("u") is None else np.asarray(f["u"], np.float32)
    v = np.zeros(n, np.float32) if f.get("v") is None else np.asarray(f["v"], np.float32)
    udir = f["udir"] if "udir" in f else np.ones(n, np.float32)
    rng = np.random.default_rng(1)
    f3 = (W.jonswap_spectrum(cfg, udir).reshape(-1, n) * rng.uniform(0.5, 1.5, n)[None, :]).astype(np.float32)
    return o, depth, u, v, f3.ravel()


def run_reference(cfg, o, depth, u, v, f3, nprop, tag):
    n = o.nsea
    fin, fout, flog = (f"/tmp/ref_propags_{tag}.{x}" for x in ("in", "out", "log"))
    with open(fin, "wb") as fh:
        fh.write(struct.pack("12i", cfg["kl"], cfg["ml"], cfg["iref"], cfg["ndepth"], cfg["idelpro"], cfg["idelt"],
                             n, cfg["ny"], cfg["shallow"], cfg["refr_d"], cfg["refr_c"], nprop))
        fh.write(struct.pack("4f", cfg["fr1"], cfg["deptha"], cfg["depthd"], o.scalar("DELPHI")))
        for name in ("SINPH", "COSPH", "DELLAM"):
            fh.write(np.asarray(o.get(name), np.float32).tobytes())
        for name in ("KXLT", "KLAT", "KLON"):                     # KXLT is the reference's KFROMIJ
            fh.write(np.asarray(o.get(name), np.int32).tobytes())
        fh.write(np.ones(2 * n, np.float32).tobytes())             # WLAT = 1 on a regular grid
        for a in (depth, u, v, f3):
            fh.write(np.asarray(a, np.float32).tobytes())
    subprocess.check_call(f"ulimit -s unlimited && {DRIVER} {fin} {fout} {flog}", shell=True)
    d = np.fromfile(fout, np.int32)
    log = open(flog).read()
    cfl = float(re.search(r"total_cfl\s*=\s*([0-9.Ee+-]+)", log).group(1))
    m = re.search(r"DIVIDED INTO\s+(\d+)", log)
    for p in (fin, fout, flog):
        os.remove(p)
    return dict(indep=d[:n].copy(), fl3=d[n:].view(np.float32).copy(), counter=int(m.group(1)) if m else 1,
                cfl=cfl)


def main():
    subprocess.check_call(["make", "-s", "-C", ORACLE, "-f", "Makefile.ref"])
    for name, (cfg, mask, f) in cases().items():
        o, depth, u, v, f3 = inputs(cfg, mask, f)
        nprop = 2
        out = run_reference(cfg, o, depth, u, v, f3, nprop, name)
        np.savez_compressed(os.path.join(HERE, f"ref_propags_{name}.npz"), cfg=json.dumps(cfg), mask=mask,
                            nprop=nprop, in_depth=depth, in_u=u, in_v=v, in_fl3=f3, out_fl3=out["fl3"],
                            out_indep=out["indep"], out_counter=out["counter"], out_cfl=out["cfl"])
        print(name, o.nsea, "points, COUNTER", out["counter"], "CFL", out["cfl"])


if __name__ == "__main__":
    main()
