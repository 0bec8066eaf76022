"""Golden vectors of the reference's own coupled time step, PROPAGS then
IMPLSCH (myWAveModel/WAM src/chief/wamodel.f90:332,354), built from the
reference sources by oracle/Makefile.ref into oracle/_ref/ and driven by
oracle/ref_step_driver.f90.  Run in the container that holds the reference
checkout:

    python tests/golden/make_ref_step.py

Each case writes ref_step_<case>.npz: the model options and sea mask (the
neighbour tables and latitude arrays are rebuilt from them and are what the
driver was given), depth, currents, winds, the initial spectrum, and the
reference's FL3, TAUW, USTAR, Z0 after the first step and after NSTEP steps.
"""
import json
import os
import struct
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
ORACLE = os.path.join(ROOT, "oracle")
DRIVER = os.path.join(ORACLE, "_ref", "ref_step")


def cases():
    from wam_amd import workloads as W
    c1cfg, _, _ = W.c1_case()
    c1corner = W.grid_config(dict(c1cfg), 16, 12, -15.0, 0.0, 0.5, 0.5)
    return {
        # 0.1 deg shallow water with depth and current refraction, 36 x 36, 300 s, 24 h
        "c2sub_janssen_24h": dict(gen="c2", ny=20, nx=28, seed=31, iphys=0, nstep=288),
        # the same physics with ST6, 24 x 25, 12 h
        "c2sub_st6_12h": dict(gen="c2", ny=16, nx=22, seed=32, iphys=2, kl=24, ml=25, nstep=144),
        # the 0.5 deg deep-water ST6 case (U10 20 m/s from the west), 600 s, 6 h
        "c1corner_st6_6h": dict(gen="c1", cfg=c1corner, nstep=36),
    }


def inputs(case):
    from oracle.oracle import Oracle
    from wam_amd import workloads as W
    if case["gen"] == "c2":
        kw = {k: case[k] for k in ("kl", "ml") if k in case}
        cfg, mask, f = W.c2_case(ny=case["ny"], nx=case["nx"], seed=case["seed"], iphys=case["iphys"], **kw)
        depth, u, v = f["depth"], f["u"], f["v"]
        u10, udir = f["u10"], f["udir"]
    else:
        cfg = case["cfg"]
        mask = np.ones((cfg["ny"], cfg["nx"]), np.uint8)
        n = int(mask.sum())
        rng = np.random.default_rng(33)
        depth = np.full(n, 999.0, np.float32)
        u = v = np.zeros(n, np.float32)
        u10 = (20.0 + rng.uniform(-4, 4, n)).astype(np.float32)
        udir = (np.pi / 2 + rng.uniform(-0.4, 0.4, n)).astype(np.float32)
    o = Oracle(cfg, mask)
    n = o.nsea
    fl3 = W.jonswap_spectrum(cfg, udir).ravel().astype(np.float32)
    z = np.zeros(n, np.float32)
    inp = dict(depth=np.asarray(depth, np.float32), u=np.asarray(u, np.float32), v=np.asarray(v, np.float32),
               fl3=fl3, u10=np.asarray(u10, np.float32), udir=np.asarray(udir, np.float32), tauw=z, z0=z,
               roairn=np.full(n, 1.225, np.float32), wstar=z)
    return cfg, mask, o, inp


def run_reference(cfg, o, inp, nstep, tag):
    n = o.nsea
    fin, fout = f"/tmp/ref_step_{tag}.in", f"/tmp/ref_step_{tag}.out"
    with open(fin, "wb") as fh:
        fh.write(struct.pack("13i", cfg["kl"], cfg["ml"], cfg["iref"], cfg["ndepth"], cfg["idelpro"], cfg["idelt"],
                             n, cfg["ny"], cfg["shallow"], cfg["refr_d"], cfg["refr_c"], cfg["iphys"], nstep))
        fh.write(struct.pack("8f", cfg["fr1"], cfg["deptha"], cfg["depthd"], o.scalar("DELPHI"), cfg["alpha"],
                             cfg["betamax"], cfg["zalp"], cfg["cdfac"]))
        for name in ("SINPH", "COSPH", "DELLAM"):
            fh.write(np.asarray(o.get(name), np.float32).tobytes())
        for name in ("KXLT", "KLAT", "KLON"):                     # KXLT is the reference's KFROMIJ
            fh.write(np.asarray(o.get(name), np.int32).tobytes())
        fh.write(np.ones(2 * n, np.float32).tobytes())             # WLAT = 1 on a regular grid
        for k in ("depth", "u", "v", "fl3", "u10", "udir", "tauw", "z0", "roairn", "wstar"):
            fh.write(np.ascontiguousarray(inp[k], np.float32).tobytes())
    subprocess.check_call(f"ulimit -s unlimited && {DRIVER} {fin} {fout}", shell=True)
    d = np.fromfile(fout, np.float32)
    nb = inp["fl3"].size
    out = {}
    off = 0
    for when in ("1", "n"):
        out[f"fl3_{when}"] = d[off:off + nb]
        off += nb
        for k in ("tauw", "ustar", "z0"):
            out[f"{k}_{when}"] = d[off:off + n]
            off += n
    os.remove(fin)
    os.remove(fout)
    return out


def main():
    subprocess.check_call(["make", "-s", "-C", ORACLE, "-f", "Makefile.ref"])
    for name, case in cases().items():
        cfg, mask, o, inp = inputs(case)
        out = run_reference(cfg, o, inp, case["nstep"], name)
        np.savez_compressed(os.path.join(HERE, f"ref_step_{name}.npz"), cfg=json.dumps(cfg), mask=mask,
                            nstep=case["nstep"], **{"in_" + k: v for k, v in inp.items()},
                            **{"out_" + k: v for k, v in out.items()})
        print(name, o.nsea, "points,", case["nstep"], "steps")


if __name__ == "__main__":
    main()
