"""The Fortran host (fortran/): ISO_C_BINDING shims with the reference's
IMPLSCH/PROPAGS argument lists driving the HIP library, as WAMODEL does
(wamodel.f90:286-360).  The CPU test checks that flang builds and links it;
the GPU test runs a few steps (the last with LCFLX, returning PHIAW and the
WNFLUXES fields), then the integrated output parameters, and compares all of
them with the oracle bit-for-bit."""
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import Oracle
from wam_amd import workloads as W

FDIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fortran")


def build_driver():
    subprocess.check_call(["make", "-s", "-C", FDIR])
    exe = os.path.join(FDIR, "wam_gpu_driver")
    assert os.path.exists(exe)
    return exe


def write_case(d, cfg, mask, f, fl3, nstep):
    os.makedirs(d, exist_ok=True)
    cur = 1 if cfg["refr_c"] else 0
    n = int(mask.sum())
    keys = ["kl", "ml", "fr1", "iref", "iphys", "shallow", "spherical", "refr_d", "refr_c", "iper", "isnonlin",
            "idelpro", "idelt", "ndepth", "deptha", "depthd", "betamax", "zalp", "alpha", "cdfac", "nx", "ny",
            "amosop", "amonop", "amowep", "amoeap", "xdella", "xdello"]
    with open(os.path.join(d, "case.nml"), "w") as fh:
        fh.write("&WAMCASE\n")
        for k in keys:
            v = cfg[k]
            fh.write(f"  {k.upper()} = {repr(float(v)) + 'D0' if isinstance(v, float) else int(v)},\n")
        fh.write(f"  NSTEP = {nstep}, CURRENT = {cur}, NSEA = {n},\n")
        fh.write(f"  WAVE_BREAKING = {int(cfg.get('wave_breaking', 0))}, PHILLIPS = {int(cfg.get('phillips', 0))}\n/\n")
    # MASK(NX,NY) in Fortran order is the row-major [NY][NX] array
    np.ascontiguousarray(mask, np.int8).tofile(os.path.join(d, "mask.i8"))
    for k in ("depth", "u10", "udir"):
        np.ascontiguousarray(f[k], np.float32).tofile(os.path.join(d, f"{k}.f32"))
    if cur:
        np.ascontiguousarray(f["u"], np.float32).tofile(os.path.join(d, "u.f32"))
        np.ascontiguousarray(f["v"], np.float32).tofile(os.path.join(d, "v.f32"))
    np.ascontiguousarray(fl3, np.float32).tofile(os.path.join(d, "fl3.f32"))


def test_fortran_host_builds():
    build_driver()


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["c1", "c2cut", "c2cut_sources"])
def test_fortran_host_matches_oracle(tmp_path, which):
    exe = build_driver()
    if which == "c1":
        cfg, mask, f = W.c1_case()
    else:
        cfg, mask, f = W.c2_case(ny=30, nx=40, seed=12, kl=24, ml=25)
    if which == "c2cut_sources":
        # WAVE_BREAKING and PHILLIPS through WAM_GPU_SET_SOURCE_OPTIONS, on depths of 0.6 to 24 m
        cfg = dict(cfg, wave_breaking=1, phillips=1)
        n = f["depth"].size
        f = dict(f, depth=np.geomspace(0.6, 24.0, n).astype(np.float32)[np.random.default_rng(4).permutation(n)])
    nstep = 3
    fl3 = W.jonswap_spectrum(cfg, f["udir"]).ravel()
    d = str(tmp_path / which)
    write_case(d, cfg, mask, f, fl3, nstep)
    print(f"running {exe} on {which}", flush=True)
    subprocess.check_call([exe, d], timeout=120)
    out = np.fromfile(os.path.join(d, "fl3_out.f32"), np.float32)
    ust = np.fromfile(os.path.join(d, "ustar_out.f32"), np.float32)
    o = Oracle(cfg, mask)
    o.prepare_propagation(f["depth"], f["u"], f["v"])
    n = o.nsea
    fo, tw, z0 = fl3, np.zeros(n, np.float32), np.zeros(n, np.float32)
    for istep in range(nstep):
        fo = o.propags(fo)
        # the last step is at an output time: LCFLX (wamodel.f90:342)
        r = o.implsch(fo, f["u10"], f["udir"], tw, z0, f["depth"], lcflx=istep == nstep - 1)
        fo, tw, z0 = r["fl3"], r["tauw"], r["z0"]
    assert np.array_equal(out.view(np.int32), fo.view(np.int32))
    assert np.array_equal(ust.view(np.int32), r["ustar"].view(np.int32))
    # PHIAW and the six WNFLUXES fields returned through IMPLSCH_GPU's optional arguments
    phiaw = np.fromfile(os.path.join(d, "phiaw_out.f32"), np.float32)
    flx = np.fromfile(os.path.join(d, "flux_out.f32"), np.float32).reshape(6, n)
    assert np.array_equal(phiaw.view(np.int32), np.asarray(r["phiaw"], np.float32).view(np.int32))
    assert np.array_equal(flx.view(np.int32), np.asarray(r["flux"], np.float32).reshape(6, n).view(np.int32))
    assert np.abs(flx).max() > 0
    # BOUT(NSEA,17) of OUTPUT_PARAMETERS_GPU (directions "coming from")
    bout = np.fromfile(os.path.join(d, "bout.f32"), np.float32).reshape(17, n)
    cur = cfg["refr_c"] != 0
    ro = o.output_parameters(fo, f["u10"], f["udir"], r["ustar"], tw, z0, f["depth"],
                             f["u"] if cur else None, f["v"] if cur else None, orient=1)
    assert np.array_equal(bout.view(np.int32), ro.view(np.int32))
    # the nest exchange of BOUNDARY_OUTPUT_GPU and BOUNDARY_INPUT_GPU on the result
    kl, ml = cfg["kl"], cfg["ml"]
    nb, nbf = min(n, 50), min(n, 80)
    ijarc = 1 + (np.arange(nb) * n) // nb
    fbc = np.fromfile(os.path.join(d, "fbc.f32"), np.float32).reshape(nb, ml, kl)
    bmean = np.fromfile(os.path.join(d, "bmean.f32"), np.float32).reshape(3, nb)
    rf, rm = o.boundary_output(fo, ijarc)
    assert np.array_equal(fbc.view(np.int32), rf.view(np.int32))
    assert np.array_equal(bmean.view(np.int32), rm.view(np.int32))
    i = np.arange(1, nbf + 1)
    ijarf = n - ((i - 1) * n) // nbf
    ibfl = 1 + (i - 1) % (nb - 1)
    bfw = (np.mod(i, 7).astype(np.float32) / np.float32(7.0)).astype(np.float32)
    want, _, _ = o.boundary_input(fo.reshape(ml, kl, n), rf, rm, rf[::-1], rm[:, ::-1], 0, 3600.0, 1200.0, ijarf,
                                  ibfl, ibfl + 1, bfw)
    got = np.fromfile(os.path.join(d, "fl3_nest.f32"), np.float32)
    assert np.array_equal(got.view(np.int32), want.ravel().view(np.int32))
