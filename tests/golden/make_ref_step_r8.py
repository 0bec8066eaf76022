"""Golden vectors of the reference's coupled time step in REAL*8: the same
PROPAGS then IMPLSCH loop (myWAveModel/WAM src/chief/wamodel.f90:332,354),
the same cases and inputs as make_ref_step.py, with the reference's modules
and oracle/ref_step_driver.f90 built by `make -f oracle/Makefile.ref r8`
(flang -fdefault-real-8 -fdefault-double-8, into oracle/_ref/r8/).  They
give the error bar of REAL*4 arithmetic, which the product reproduces.  Run
in the container that holds the reference checkout:

    python tests/golden/make_ref_step_r8.py

The REAL*8 run is given the inputs stored in the REAL*4 goldens
(ref_step_<case>.npz) exactly, every field a float32 value widened, the
grid's latitude arrays from the FP64 oracle and the configuration's decimal
constants in double.  Each case writes
ref_step_r8_<case>.npz: FL3 after the first and the last step rounded to
float32 (a relative 3e-8 below the error bars it serves), and TAUW, USTAR, Z0
in float64.
"""
import json
import os
import struct
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_ref_step import ORACLE, cases  # noqa: E402

DRIVER8 = os.path.join(ORACLE, "_ref", "r8", "ref_step")


def run_reference_r8(cfg, mask, inp, nstep, tag):
    from oracle.oracle import Oracle
    o = Oracle(cfg, mask, fp64=True)
    n = o.nsea
    fin, fout = f"/tmp/ref_step_r8_{tag}.in", f"/tmp/ref_step_r8_{tag}.out"
    with open(fin, "wb") as fh:
        fh.write(struct.pack("13i", cfg["kl"], cfg["ml"], cfg["iref"], cfg["ndepth"], cfg["idelpro"], cfg["idelt"],
                             n, cfg["ny"], cfg["shallow"], cfg["refr_d"], cfg["refr_c"], cfg["iphys"], nstep))
        fh.write(struct.pack("8d", cfg["fr1"], cfg["deptha"], cfg["depthd"], o.scalar("DELPHI"), cfg["alpha"],
                             cfg["betamax"], cfg["zalp"], cfg["cdfac"]))
        for name in ("SINPH", "COSPH", "DELLAM"):
            fh.write(np.asarray(o.get(name), np.float64).tobytes())
        for name in ("KXLT", "KLAT", "KLON"):                     # KXLT is the reference's KFROMIJ
            fh.write(np.asarray(o.get(name), np.int32).tobytes())
        fh.write(np.ones(2 * n, np.float64).tobytes())             # WLAT = 1 on a regular grid
        for k in ("depth", "u", "v", "fl3", "u10", "udir", "tauw", "z0", "roairn", "wstar"):
            fh.write(np.ascontiguousarray(inp[k], np.float32).astype(np.float64).tobytes())
    subprocess.check_call(f"ulimit -s unlimited && {DRIVER8} {fin} {fout}", shell=True)
    d = np.fromfile(fout, np.float64)
    nb = inp["fl3"].size
    out = {}
    off = 0
    for when in ("1", "n"):
        out[f"fl3_{when}"] = d[off:off + nb].astype(np.float32)
        off += nb
        for k in ("tauw", "ustar", "z0"):
            out[f"{k}_{when}"] = d[off:off + n].copy()
            off += n
    assert off == d.size
    os.remove(fin)
    os.remove(fout)
    return out


def main():
    subprocess.check_call(["make", "-s", "-C", ORACLE, "-f", "Makefile.ref", "r8"])
    for name in cases():
        z = np.load(os.path.join(HERE, f"ref_step_{name}.npz"))
        cfg, mask, nstep = json.loads(str(z["cfg"])), z["mask"], int(z["nstep"])
        inp = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
        out = run_reference_r8(cfg, mask, inp, nstep, name)
        np.savez_compressed(os.path.join(HERE, f"ref_step_r8_{name}.npz"), nstep=nstep,
                            **{"out_" + k: v for k, v in out.items()})
        print(name, inp["u10"].size, "points,", nstep, "steps")


if __name__ == "__main__":
    main()
