"""Golden vectors of the reference's own nest-exchange arithmetic
(myWAveModel/WAM src/mod/wam_interface_module.f90: TOTAL_ENERGY, FEMEAN,
MEAN_DIRECTION as BOUNDARY_OUTPUT calls them, INTSPEC :290-408 with ROTSPEC
:891-941 and STRSPEC :945-1006), built from the reference sources by
oracle/Makefile.ref into oracle/_ref/ and driven by
oracle/ref_nest_driver.f90.  Run in the container that holds the reference
checkout:

    python tests/golden/make_ref_nest.py

Each case writes ref_nest_<case>.npz: the grid (KL, ML, FR1, IREF), seeded
spectra F[NB][ML][KL] (JONSWAP of random peak frequency and direction, an
empty one, a scaled copy), the INTSPEC pairs (indices, DEL12, DEL1L: end
points, a side without energy, interior weights), the ROTSPEC angles and
STRSPEC factors, and the reference's outputs.
"""
import math
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_nest")

CASES = {"kl24_ml25": (24, 25), "kl36_ml36": (36, 36)}


def make_case(kl, ml, seed):
    from wam_amd import workloads as W
    cfg = W.base_config(kl=kl, ml=ml)
    rng = np.random.default_rng(seed)
    nb = 12
    f = np.empty((nb, ml, kl), np.float32)
    for i in range(nb):
        s = W.jonswap_spectrum(cfg, np.array([rng.uniform(0, 2 * math.pi)]), fp=rng.uniform(0.1, 0.4))
        f[i] = s[:, :, 0] * np.float32(rng.uniform(0.3, 3.0))
    f[10] = 0.0                                   # no energy
    f[11] = np.float32(2.5) * f[0]                # a scaled copy
    pairs = [(1, 2, 3600.0, 0.0), (1, 2, 3600.0, 3600.0), (11, 3, 3600.0, 1800.0), (3, 11, 1.0, 0.25),
             (1, 12, 1.0, 0.5)]
    for _ in range(20):
        i1, i2 = rng.integers(1, 11, 2)
        pairs.append((int(i1), int(i2), 1.0, float(np.float32(rng.uniform(0.0, 1.0)))))
    for _ in range(6):
        i1, i2 = rng.integers(1, 11, 2)
        pairs.append((int(i1), int(i2), 3600.0, float(rng.integers(1, 3600))))
    dth = 2 * math.pi / kl
    rots = [(1, 0.0), (2, dth), (3, -2 * dth), (4, 0.5 * dth)] + \
        [(int(rng.integers(1, 11)), float(np.float32(rng.uniform(-math.pi, math.pi)))) for _ in range(8)]
    co = 1.1
    strs = [(1, 1.0), (2, co ** 2), (3, co ** -3), (4, co ** (ml + 1)), (5, co ** -(ml + 1))] + \
        [(int(rng.integers(1, 11)), float(np.float32(math.exp(rng.uniform(-0.6, 0.6))))) for _ in range(8)]
    return cfg, f, pairs, rots, strs


def run(cfg, f, pairs, rots, strs):
    kl, ml = cfg["kl"], cfg["ml"]
    nb = f.shape[0]
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in"), os.path.join(d, "out")
        with open(fin, "wb") as fh:
            fh.write(struct.pack("<7i", kl, ml, cfg["iref"], nb, len(pairs), len(rots), len(strs)))
            fh.write(struct.pack("<f", cfg["fr1"]))
            fh.write(np.ascontiguousarray(f, np.float32).tobytes())      # F(KL,ML,NB)
            for i1, i2, d12, d1l in pairs:
                fh.write(struct.pack("<2i2f", i1, i2, d12, d1l))
            for i, x in rots:
                fh.write(struct.pack("<if", i, x))
            for i, x in strs:
                fh.write(struct.pack("<if", i, x))
        subprocess.check_call([DRIVER, fin, fout])
        out = np.fromfile(fout, np.float32)
    nbs = kl * ml
    o = 0
    mean = out[o:o + 3 * nb].reshape(3, nb)
    o += 3 * nb
    fl, mo = [], []
    for _ in pairs:
        fl.append(out[o:o + nbs].reshape(ml, kl))
        mo.append(out[o + nbs:o + nbs + 3])          # FMEAN, EMEAN, THETM
        o += nbs + 3
    rot = [out[o + i * nbs:o + (i + 1) * nbs].reshape(ml, kl) for i in range(len(rots))]
    o += len(rots) * nbs
    st = [out[o + i * nbs:o + (i + 1) * nbs].reshape(ml, kl) for i in range(len(strs))]
    o += len(strs) * nbs
    assert o == out.size
    return mean, np.array(fl), np.array(mo), np.array(rot), np.array(st)


def main():
    for name, (kl, ml) in CASES.items():
        cfg, f, pairs, rots, strs = make_case(kl, ml, seed=kl * 100 + ml)
        mean, fl, mo, rot, st = run(cfg, f, pairs, rots, strs)
        np.savez_compressed(os.path.join(HERE, f"ref_nest_{name}.npz"), kl=kl, ml=ml, fr1=cfg["fr1"],
                            iref=cfg["iref"], f=f, pairs=np.array(pairs, np.float64), rots=np.array(rots, np.float64),
                            strs=np.array(strs, np.float64), mean=mean, intspec=fl, intspec_mean=mo, rotspec=rot,
                            strspec=st)
        print(name, "written")


if __name__ == "__main__":
    main()
