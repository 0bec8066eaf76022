"""Golden vectors of the reference's own IMPLSCH (myWAveModel/WAM
src/mod/wam_source_module.f90:374-700), built from the reference sources by
oracle/Makefile.ref into oracle/_ref/ and driven by oracle/ref_implsch_driver.f90.
Run in the container that holds the reference checkout:

    python tests/golden/make_ref_implsch.py

Each case writes ref_implsch_<case>.npz: the model options, the grid (so the
tables, including MIN_ENERGY's fetch-dependent floor, can be rebuilt), the
inputs of IMPLSCH for a few seeded sea points, and the reference's outputs
after NSTEP consecutive IMPLSCH calls (no propagation in between).  Case
names given on the command line write only those cases.

The *_break_* and *_phillips_* cases switch on the off-by-default sources
(SET_MODEL_OPTION's WAVE_BREAKING: SDEPTHLIM and SFBRK; PHILLIPS:
SOURCE_PHILLIPS).  Their shallow points get depths of 0.6 to 24 m, so that
the depth limit (0.25 GAMD DEPTH)**2 caps the energy of the shallowest ones,
and the Phillips cases start from a calm sea (the JONSWAP spectrum scaled by
1e-4), where the linear growth term is a visible part of the first steps.
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
DRIVER = os.path.join(ORACLE, "_ref", "ref_implsch")

OUTPUTS = ["tauw", "ustar", "z0", "phiaw", "phioc", "tauoc_x", "tauoc_y", "phibot", "taubot_x", "taubot_y"]


def cases():
    from wam_amd import workloads as W
    c1cfg, _, _ = W.c1_case()
    # the 60 x 60 deep-water case's options and grid spacing on a 4 x 4 corner
    # (IMPLSCH is pointwise; the tables depend on the spacing only)
    c1small = W.grid_config(dict(c1cfg), 4, 4, -15.0, 0.0, 0.5, 0.5)
    return {
        "c2sub_janssen_1step": dict(gen="c2", iphys=0, nstep=1, lcflx=1),
        "c2sub_st6_1step": dict(gen="c2", iphys=2, nstep=1, lcflx=1),
        "c2sub_janssen_24h": dict(gen="c2", iphys=0, nstep=288, lcflx=0),
        "c1_st6_1step": dict(gen="c1", cfg=c1small, nstep=1, lcflx=0),
        "c1_st6_6h": dict(gen="c1", cfg=c1small, nstep=36, lcflx=0),
        "c1_janssen_1step": dict(gen="c1", cfg=dict(c1small, iphys=0), nstep=1, lcflx=0),
        "c2sub_break_phillips_1step": dict(gen="c2", iphys=0, nstep=1, lcflx=1, breaking=1, phillips=1,
                                           shallow_depths=True),
        "c2sub_break_phillips_6h": dict(gen="c2", iphys=0, nstep=72, lcflx=0, breaking=1, phillips=1,
                                        shallow_depths=True),
        "c2sub_st6_break_phillips_1step": dict(gen="c2", iphys=2, nstep=1, lcflx=1, breaking=1, phillips=1,
                                               shallow_depths=True),
        # deep water: the reference switches breaking off there (wam_timopt_module.f90:213-227)
        "c1_janssen_phillips_calm_1step": dict(gen="c1", cfg=dict(c1small, iphys=0), nstep=1, lcflx=0,
                                               breaking=1, phillips=1, scale=1e-4),
        "c1_janssen_phillips_calm_6h": dict(gen="c1", cfg=dict(c1small, iphys=0), nstep=36, lcflx=0,
                                            breaking=1, phillips=1, scale=1e-4),
    }


def inputs(case):
    from oracle.oracle import Oracle
    from wam_amd import workloads as W
    if case["gen"] == "c2":
        cfg, mask, f = W.c2_case(ny=6, nx=8, seed=3, iphys=case["iphys"])
        depth = f["depth"]
        u10, udir = f["u10"], f["udir"]
        if case.get("shallow_depths"):
            depth = np.geomspace(0.6, 24.0, depth.size).astype(np.float32)[np.random.default_rng(5).permutation(depth.size)]
    else:
        cfg = case["cfg"]
        mask = np.ones((cfg["ny"], cfg["nx"]), np.uint8)
        n = int(mask.sum())
        rng = np.random.default_rng(17)
        depth = np.full(n, 999.0, np.float32)
        u10 = (20.0 + rng.uniform(-6, 6, n)).astype(np.float32)
        udir = (np.pi / 2 + rng.uniform(-0.5, 0.5, n)).astype(np.float32)
    if case.get("breaking") or case.get("phillips"):
        cfg = dict(cfg, wave_breaking=int(case.get("breaking", 0)), phillips=int(case.get("phillips", 0)))
    o = Oracle(cfg, mask)
    o.prepare_propagation(depth, None, None)
    n = o.nsea
    fl3 = W.jonswap_spectrum(cfg, udir).ravel() * case.get("scale", 1.0)
    indep = o.get("INDEP").astype(np.int32) if cfg["shallow"] else np.zeros(n, np.int32)
    return cfg, mask, dict(fl3=fl3.astype(np.float32), u10=np.asarray(u10, np.float32),
                           udir=np.asarray(udir, np.float32), tauw=np.zeros(n, np.float32),
                           z0=np.zeros(n, np.float32), roairn=np.full(n, 1.225, np.float32),
                           wstar=np.zeros(n, np.float32), depth=np.asarray(depth, np.float32),
                           indep=indep), float(o.scalar("DELPHI"))


def run_reference(cfg, inp, delphi, nstep, lcflx, tag):
    n = inp["u10"].size
    fin, fout = f"/tmp/ref_implsch_{tag}.in", f"/tmp/ref_implsch_{tag}.out"
    with open(fin, "wb") as fh:
        fh.write(struct.pack("13i", cfg["kl"], cfg["ml"], cfg["iref"], cfg["iphys"], cfg["shallow"], cfg["ndepth"],
                             cfg["idelpro"], cfg["idelt"], n, lcflx, nstep, cfg.get("wave_breaking", 0),
                             cfg.get("phillips", 0)))
        fh.write(struct.pack("8f", cfg["fr1"], cfg["deptha"], cfg["depthd"], cfg["alpha"], cfg["betamax"],
                             cfg["zalp"], cfg["cdfac"], delphi))
        for k in ("fl3", "u10", "udir", "tauw", "z0", "roairn", "wstar", "depth"):
            fh.write(np.ascontiguousarray(inp[k], np.float32).tobytes())
        fh.write(np.ascontiguousarray(inp["indep"], np.int32).tobytes())
    # IMPLSCH's automatic arrays live on the stack
    subprocess.check_call(f"ulimit -s unlimited && {DRIVER} {fin} {fout}", shell=True)
    d = np.fromfile(fout, np.float32)
    nb = inp["fl3"].size
    out = dict(fl3=d[:nb])
    for i, k in enumerate(OUTPUTS):
        out[k] = d[nb + i * n: nb + (i + 1) * n]
    os.remove(fin)
    os.remove(fout)
    return out


def main():
    subprocess.check_call(["make", "-s", "-C", ORACLE, "-f", "Makefile.ref"])
    only = set(sys.argv[1:])
    for name, case in cases().items():
        if only and name not in only:
            continue
        cfg, mask, inp, delphi = inputs(case)
        out = run_reference(cfg, inp, delphi, case["nstep"], case["lcflx"], name)
        np.savez_compressed(os.path.join(HERE, f"ref_implsch_{name}.npz"),
                            cfg=json.dumps(cfg), mask=mask, nstep=case["nstep"], lcflx=case["lcflx"], delphi=delphi,
                            **{"in_" + k: v for k, v in inp.items()}, **{"out_" + k: v for k, v in out.items()})
        print(name, inp["u10"].size, "points,", case["nstep"], "steps")


if __name__ == "__main__":
    main()
