"""ORACLE -- test infrastructure only.  Nothing in the product uses it.

Calibration of the CPU baseline against the reference itself: the time of
one model step (PROPAGS then IMPLSCH, wamodel.f90:332,354) of the reference
built from its own sources (oracle/_ref/ref_step, oracle/Makefile.ref; one
process, flang -O2) against the oracle's REAL*4 restatement (liborc32) on the
same inputs and the same host, one core each, and the oracle on every core
(OpenMP over sea points, as bench.py's cpu_baseline runs it).  The ratio
R = t_reference / t_oracle turns an oracle time into a reference-equivalent
time.  Run where the reference checkout was built:

    python oracle/calibrate_ref.py [case ...]      (cases: c1, c2cut)

The reference's per-step time is the difference of two runs with different
step counts (set-up and file I/O cancel); it prints one JSON line per case.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
# run as a script, this folder would shadow the oracle package
sys.path = [p for p in sys.path if os.path.abspath(p or ".") != HERE]
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def cases():
    from wam_amd import workloads as W
    return {
        # configuration 1: 0.5 deg 60 x 60 deep water, ST6, 24 x 25
        "c1": dict(make=lambda: W.c1_case(), steps=(1, 11), osteps=10),
        # a 100 800-point cut of configuration 2: 0.1 deg 300 x 480, shallow water with currents, 36 x 36, IPHYS = 0
        "c2cut": dict(make=lambda: W.c2_case(ny=300, nx=480), steps=(1, 3), osteps=2),
    }


def inputs(cfg, mask, f):
    from oracle.oracle import Oracle
    from wam_amd import workloads as W
    o = Oracle(cfg, mask)
    n = o.nsea
    z = np.zeros(n, np.float32)
    u = f["u"] if f.get("u") is not None else z
    v = f["v"] if f.get("v") is not None else z
    inp = dict(depth=np.asarray(f["depth"], np.float32), u=np.asarray(u, np.float32), v=np.asarray(v, np.float32),
               fl3=W.jonswap_spectrum(cfg, f["udir"]).ravel().astype(np.float32),
               u10=np.asarray(f["u10"], np.float32), udir=np.asarray(f["udir"], np.float32), tauw=z, z0=z,
               roairn=np.full(n, 1.225, np.float32), wstar=z)
    return o, inp


def time_reference(cfg, o, inp, steps, tag):
    from make_ref_step import run_reference
    t = []
    for k in steps:
        t0 = time.perf_counter()
        run_reference(cfg, o, inp, k, f"cal_{tag}_{k}")
        t.append(time.perf_counter() - t0)
    return (t[1] - t[0]) / (steps[1] - steps[0])


def time_oracle(cfg, mask, inp, nsteps, nthreads):
    from oracle.oracle import Oracle
    o = Oracle(cfg, mask)
    cur = cfg["refr_c"] != 0
    o.prepare_propagation(inp["depth"], inp["u"] if cur else None, inp["v"] if cur else None)
    f, tw, z0 = inp["fl3"], inp["tauw"], inp["z0"]
    t0 = time.perf_counter()
    for _ in range(nsteps):
        f = o.propags(f)
        r = o.implsch(f, inp["u10"], inp["udir"], tw, z0, inp["depth"], nthreads=nthreads)
        f, tw, z0 = r["fl3"], r["tauw"], r["z0"]
    return (time.perf_counter() - t0) / nsteps


def main(names):
    ncpu = os.cpu_count()
    model = next((l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")), "?")
    for name in names:
        c = cases()[name]
        cfg, mask, f = c["make"]()
        o, inp = inputs(cfg, mask, f)
        t_ref = time_reference(cfg, o, inp, c["steps"], name)
        t_orc = time_oracle(cfg, mask, inp, c["osteps"], 1)
        t_all = time_oracle(cfg, mask, inp, c["osteps"], ncpu)
        print(json.dumps(dict(case=name, nsea=o.nsea, kl=cfg["kl"], ml=cfg["ml"], iphys=cfg["iphys"],
                              shallow=cfg["shallow"], cpu=model, cores=ncpu,
                              ref_s_per_step_1core=round(t_ref, 4), oracle_s_per_step_1core=round(t_orc, 4),
                              oracle_s_per_step_all_cores=round(t_all, 4), R=round(t_ref / t_orc, 3),
                              ref_updates_per_s_1core=round(o.nsea / t_ref, 1))), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(cases()))
