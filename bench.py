"""Benchmark of WAM's per-step hot path (PROPAGS + IMPLSCH) on one MI355X.

Workload: the production-shaped case -- 0.1 deg spherical grid 901 x 1600
with a seeded 30 % land mask (about 1.0 M sea points), 36 directions x 36
frequencies, shallow water with depth and current refraction, Janssen
physics (IPHYS=0), 300 s steps.  A step is one PROPAGS call (with the COUNTER
sub-steps CHECK_CFL prescribes) followed by one IMPLSCH call, as in
wamodel.f90:332,354, on spectra resident in HBM.  The metric is sea-point
spectrum updates per second: NSEA / t_step.

Prints one JSON line.  `--gpus N` with N > 1 is not supported: this build
runs on a single GPU.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# floating-point work of IMPLSCH per sea point, counted by the oracle's
# op-counting build on a seeded sub-box of this workload
OPCOUNT_FILE = os.path.join(ROOT, "tests", "golden", "implsch_opcount.json")
# HBM bytes per dispatch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (scaled by a plain copy kernel of known size), per kernel and workload
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_c2.json")
# kernels the library times (wam_gpu_get_timing)
KERNELS = ("propags", "implsch_pre", "implsch_snl", "implsch_post", "implsch")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def device_spectrum(torch, cfg, udir_d, fp=0.3, alpha=0.018, gamma=3.3):
    """JONSWAP x (2/pi) cos^2 spreading about the local wind, built on the device."""
    from wam_amd import workloads as W
    g = 9.806
    fr = W.frequencies(cfg)
    arg = 1.25 * (fp / fr) ** 4
    et = alpha * g * g / (2 * math.pi) ** 4 * fr ** -5 * np.exp(-np.minimum(arg, 50.0))
    sig = np.where(fr > fp, 0.09, 0.07)
    et = et * gamma ** np.exp(-((fr / fp - 1.0) ** 2) / (2 * sig * sig))
    kl = cfg["kl"]
    th = torch.tensor((np.arange(kl) + 0.5) * (2 * math.pi / kl), dtype=torch.float32, device=udir_d.device)
    st = (2.0 / math.pi) * torch.clamp(torch.cos(th[:, None] - udir_d[None, :]), min=0.0) ** 2
    e = torch.tensor(et, dtype=torch.float32, device=udir_d.device)
    return (e[:, None, None] * st[None, :, :]).contiguous()


def host_cpu():
    """nproc and the CPU model of the host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def cpu_baseline(args):
    """The oracle (C restatement, REAL*4) on the host cores, on a bounded
    sub-box of the same workload (same statistics, 36 x 36, shallow, currents).
    Threads: OMP_NUM_THREADS when set (the host's CPU share per GPU), else
    every core."""
    from oracle.oracle import Oracle
    from wam_amd import workloads as W
    nproc, model = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc
    os.environ["OMP_NUM_THREADS"] = str(threads)
    cfg, mask, f = W.c2_case(ny=args.cpu_ny, nx=args.cpu_nx)
    o = Oracle(cfg, mask)
    counter, _ = o.prepare_propagation(f["depth"], f["u"], f["v"])
    n = o.nsea
    fl3 = W.jonswap_spectrum(cfg, f["udir"]).ravel()
    tauw = np.zeros(n, np.float32)
    z0 = np.zeros(n, np.float32)
    r = o.implsch(o.propags(fl3), f["u10"], f["udir"], tauw, z0, f["depth"], nthreads=threads)   # warm-up
    fl3, tauw, z0 = r["fl3"], r["tauw"], r["z0"]
    steps = args.cpu_steps
    t0 = time.perf_counter()
    for _ in range(steps):
        fl3 = o.propags(fl3)
        r = o.implsch(fl3, f["u10"], f["udir"], tauw, z0, f["depth"], nthreads=threads)
        fl3, tauw, z0 = r["fl3"], r["tauw"], r["z0"]
    dt = (time.perf_counter() - t0) / steps
    # the reference's own step is R = 1.111 times the oracle's on one core for
    # this physics and size (a 100 800-point cut, profiles/cpu_reference_calibration.md)
    r_ref = 1.111
    return dict(value=n / dt, unit="sea-point updates/s", cores=threads, nproc=nproc, cpu_model=model, kind="port",
                sample=f"oracle REAL*4 on a {args.cpu_ny}x{args.cpu_nx} sub-box of the same workload "
                       f"({n} sea points, 36x36, COUNTER={counter}), {steps} steps, OpenMP over points",
                seconds_per_step=dt,
                reference_equivalent=dict(value=n / (dt * r_ref), R=r_ref,
                                          note="oracle time x R, R = t(reference step) / t(oracle step) on one core "
                                               "of the build container (flang -O2 build of the reference's sources)"))


def nest_rings(mask):
    """The nest boundary point lists on one grid, in the sea-point order:
    BOUNDARY_OUTPUT's points on the edge of an inner rectangle (rows ny/8 and
    7ny/8, columns nx/8 to 7nx/8) and BOUNDARY_INPUT's points on the grid's
    outer edge, each interpolated between two of NBINP = NBOUNF/5 input
    spectra (a 5:1 nest) with weights 0, 0.2, ..., 0.8."""
    ny, nx = mask.shape
    sea = mask.astype(bool)
    idx = np.cumsum(sea.ravel()).reshape(sea.shape)          # 1-based sea-point number of each cell

    def ring(r0, r1, c0, c1):
        sel = np.zeros(sea.shape, bool)
        sel[[r0, r1], c0:c1 + 1] = True
        sel[r0:r1 + 1, [c0, c1]] = True
        return idx[sel & sea].astype(np.int32)

    ijarc = ring(ny // 8, 7 * ny // 8, nx // 8, 7 * nx // 8)
    ijarf = ring(0, ny - 1, 0, nx - 1)
    nbinp = max(2, ijarf.size // 5)
    j = np.arange(ijarf.size)
    ibfl = (np.minimum(j // 5, nbinp - 2) + 1).astype(np.int32)
    bfw = ((j % 5) / 5.0).astype(np.float32)
    return ijarc, nbinp, ijarf, ibfl, ibfl + 1, bfw


def bench_nest(args, torch, g, cfg, mask, fl3):
    """The nest boundary exchange (BOUNDARY_OUTPUT, then BOUNDARY_INPUT's time
    and space interpolation) on FL3 resident in HBM, the grid acting as the
    coarse grid for the output and as the fine grid for the input (point
    lists of nest_rings).  One exchange = wam_gpu_boundary_output_dev +
    wam_gpu_boundary_input_dev (host arrays for the boundary spectra, as the
    file I/O between them)."""
    n, kl, ml = g.nsea, cfg["kl"], cfg["ml"]
    nbs = kl * ml
    ijarc, NB_INP, ijarf, ibfl, ibfr, bfw = nest_rings(mask)
    NB_OUT, NB_FINE = ijarc.size, ijarf.size
    f1, m1 = g.boundary_output(fl3, ijarc[:NB_INP])
    f2, m2 = g.boundary_output(fl3, ijarc[::-1][:NB_INP].copy())

    def exchange():
        g.boundary_output(fl3, ijarc)
        g.boundary_input(fl3, f1, m1, f2, m2, 0, 3600.0, 1800.0, ijarf, ibfl, ibfr, bfw, return_set=False)

    for _ in range(args.warmup):
        exchange()
    g.set_timing(True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for _ in range(args.steps):
        exchange()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - w0) / args.steps
    kt = {}
    for name in ("boundary_gather", "boundary_time", "boundary_space"):
        ms, calls = g.kernel_time(name)
        kt[name] = ms / calls
    g.set_timing(False)
    # algorithmic bytes: every input value read once and every output value
    # written once (spectra of KL*ML floats, three mean values and the indices)
    nbytes = {"boundary_gather": NB_OUT * (nbs * 8 + 4 + 12),
              "boundary_time": NB_INP * (nbs * 12 + 36),
              "boundary_space": NB_INP * (nbs * 4 + 12) + NB_FINE * (nbs * 4 + 16)}
    roofs = {}
    for name, ms in kt.items():
        gbs = nbytes[name] / (ms * 1e-3) / 1e9
        roofs[name] = dict(kernel=f"{name}_kernel", bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS,
                           unit="GB/s", frac=round(gbs / HBM_PEAK_GBS, 4), traffic=None,
                           algorithmic_bytes_per_launch=nbytes[name], avg_launch_ms=round(ms, 4))
    dominant = max(kt, key=kt.get)
    out = {
        "metric": "nest boundary exchanges/s (BOUNDARY_OUTPUT + BOUNDARY_INPUT)",
        "value": 1.0 / wall, "unit": "exchanges/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (spectra of the resident c2 state; seeded boundary point lists)",
        "config": {"workload": f"nest exchange on c2 FL3: {NB_OUT} output points (inner rectangle), {NB_INP} "
                               f"input spectra, {NB_FINE} fine boundary points (outer edge, 5:1), {kl}x{ml}",
                   "nsea": n},
        "kernel_ms": {k: round(v, 4) for k, v in kt.items()},
        "gpu_ms_per_exchange": round(sum(kt.values()), 4),
        "roofline": roofs[dominant], "roofline_all": roofs,
    }
    if not args.no_cpu_baseline:
        from oracle.oracle import Oracle
        from wam_amd import workloads as W
        log("[bench] CPU baseline (oracle, one exchange) ...")
        cfg_, cmask, _ = W.c2_case(ny=args.cpu_ny, nx=args.cpu_nx, iphys=args.iphys)
        o = Oracle(cfg_, cmask)
        ns = o.nsea
        h = np.broadcast_to(f1[:1].reshape(ml, kl, 1), (ml, kl, ns)).copy()
        cij, cnb, cif, cl, cr, cw = nest_rings(cmask)
        c0 = time.perf_counter()
        o.boundary_output(h, cij)
        o.boundary_input(h, f1[:cnb], m1[:, :cnb], f2[:cnb], m2[:, :cnb], 0, 3600.0, 1800.0, cif, cl, cr, cw,
                         inplace=True)
        ct = time.perf_counter() - c0
        out["cpu_baseline"] = {"value": 1.0 / ct, "unit": "exchanges/s", "cores": 1, "kind": "port",
                               "sample": f"one exchange by the oracle on the {args.cpu_ny}x{args.cpu_nx} "
                                         f"sub-box ({ns} sea points; the same rings there: {cij.size} output, "
                                         f"{cnb} input, {cif.size} fine points; single thread, FL3 updated in "
                                         "place); rate scaled to the GPU's point counts",
                               "raw_exchanges_per_s": 1.0 / ct}
        # the oracle's time is linear in the point counts: scale to this workload's
        scale = (cij.size + 2 * cnb + cif.size) / (NB_OUT + 2 * NB_INP + NB_FINE)
        out["cpu_baseline"]["value"] = scale / ct
    print(json.dumps(out), flush=True)


def bench_st6(args, torch):
    """The ST6 physics (IPHYS = 2): configuration 2's production grid with ST6
    (its secondary physics) and configuration 1 (0.5 deg 60 x 60 deep water,
    24 x 25, ST6), each a plain PROPAGS + IMPLSCH step on spectra in HBM.
    IMPLSCH takes its automatic form: three kernels at c2's size, the single
    kernel (one workgroup per sea point) at c1's 3600 points."""
    from wam_amd import WamGpu, workloads as W
    dev = torch.device("cuda:0")
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    res = {}
    for name, make, steps in (("c2", lambda: W.c2_case(ny=args.ny, nx=args.nx, iphys=2), args.st6_steps),
                              ("c1", lambda: W.c1_case(), 20)):
        cfg, mask, f = make()
        g = WamGpu(cfg, mask)
        n = g.nsea
        counter, cfl = g.prepare_propagation(f["depth"], f["u"] if cfg["refr_c"] else None,
                                             f["v"] if cfg["refr_c"] else None)
        stream = torch.cuda.current_stream()
        g.set_stream(stream.cuda_stream)
        u10, udir = t(f["u10"]), t(f["udir"])
        tauw, ustar, z0 = (torch.zeros(n, device=dev) for _ in range(3))
        roairn, wstar = torch.full((n,), 1.225, device=dev), torch.zeros(n, device=dev)
        fa = device_spectrum(torch, cfg, udir).reshape(-1)
        fb = torch.empty_like(fa)

        def step():
            nonlocal fa, fb
            g.propags_dev(fa, fb)
            fa, fb = fb, fa
            g.implsch_dev(fa, u10, udir, tauw, ustar, z0, roairn, wstar)

        step()
        torch.cuda.synchronize()
        g.set_timing(True)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - w0) / steps * 1e3
        kt = {}
        for k in KERNELS:
            kms, calls = g.kernel_time(k)
            if calls:
                kt[k] = round(kms / calls, 3)
        g.set_timing(False)
        res[name] = {"nsea": n, "kl": cfg["kl"], "ml": cfg["ml"], "shallow": cfg["shallow"], "counter": counter,
                     "steps": steps, "ms_per_step": round(ms, 3), "value": n / (ms * 1e-3),
                     "unit": "sea-point updates/s", "kernel_ms": kt,
                     "state_finite": bool(torch.isfinite(fa).all().item())}
        log(f"[bench] ST6 {name}: NSEA={n}, {ms:.2f} ms/step, kernels {kt}")
        del fa, fb, u10, udir, tauw, ustar, z0, roairn, wstar
        g.close()
        torch.cuda.empty_cache()
    res["workload"] = ("ST6 (IPHYS=2): c2's 0.1deg 901x1600 grid, 36x36, shallow + currents, 300 s; c1: 0.5deg "
                       "60x60 deep water, 24x25, 600 s; PROPAGS + IMPLSCH per step (three-kernel IMPLSCH at c2, single kernel "
                       "at c1)")
    return res


def bench_ensemble(args, torch):
    """The batched ensemble (configuration 3): NMEM members on the 0.1 deg
    450 x 800 sub-box (252 000 sea points, 36 x 36, shallow with currents),
    members differing by their winds.  The launch shapes are chosen by the
    library's autotuner on the state, then K steps (PROPAGS + IMPLSCH of every
    member) run as one captured HIP graph and, for comparison, launch by
    launch; a plain run of one member on the same grid gives the per-member
    reference cost."""
    from wam_amd import WamGpu, workloads as W
    t0 = time.time()
    nmem, K = args.ens_members, args.ens_steps
    cfg, mask, f, u10m, udirm = W.c3_case(nmem=nmem)
    g = WamGpu(cfg, mask)
    n = g.nsea
    counter, cfl = g.prepare_propagation(f["depth"], f["u"], f["v"])
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    g.set_stream(stream.cuda_stream)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    u10, udir = t(u10m.ravel()), t(udirm.ravel())
    nmp = nmem * n
    tauw, ustar, z0 = (torch.zeros(nmp, device=dev) for _ in range(3))
    roairn, wstar = torch.full((nmp,), 1.225, device=dev), torch.zeros(nmp, device=dev)
    base = device_spectrum(torch, cfg, udir[:n]).reshape(-1)
    nb = base.numel()
    fa = base.repeat(nmem)
    fb = torch.empty_like(fa)
    del base
    torch.cuda.synchronize()
    log(f"[ensemble] {nmem} members x {n} sea points, COUNTER={counter}, state {2 * fa.numel() * 4 / 1e9:.1f} GB "
        f"({time.time() - t0:.1f}s)")
    out = g.ensemble_run_dev(nmem, 1, fa, fb, u10, udir, tauw, ustar, z0, roairn, wstar, graph=False)
    if out is fb:
        fa, fb = fb, fa
    torch.cuda.synchronize()
    a0 = time.time()
    shapes = g.autotune(nmem, fa, fb, u10, udir, tauw, z0, roairn, wstar, reps=1)
    log(f"[ensemble] autotuned in {time.time() - a0:.1f}s: {shapes}")
    # eager: K steps launch by launch, per-kernel HIP-event times
    g.set_timing(True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    out = g.ensemble_run_dev(nmem, K, fa, fb, u10, udir, tauw, ustar, z0, roairn, wstar, graph=False)
    torch.cuda.synchronize()
    t_eager = (time.perf_counter() - w0) / K
    kt = {}
    for name in KERNELS:
        ms, calls = g.kernel_time(name)
        if calls:
            kt[name] = ms / K
    g.set_timing(False)
    if out is fb:
        fa, fb = fb, fa
    # the captured loop: built and run once, then timed on a replay
    g.ensemble_run_dev(nmem, K, fa, fb, u10, udir, tauw, ustar, z0, roairn, wstar, graph=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    out = g.ensemble_run_dev(nmem, K, fa, fb, u10, udir, tauw, ustar, z0, roairn, wstar, graph=True)
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - w0) / K
    finite = bool(torch.isfinite(out).all().item())
    log(f"[ensemble] {K} steps: graph {t_graph * 1e3:.1f} ms/step, eager {t_eager * 1e3:.1f} ms/step")
    # the plain run of member 0 on the same grid, K steps after one warm-up step
    p_a, p_b = fa[:nb].clone(), torch.empty(nb, device=dev)
    sl = slice(0, n)
    pt_tw, pt_us, pt_z0 = tauw[sl].clone(), ustar[sl].clone(), z0[sl].clone()
    pu, pd, pr, pw = u10[sl].contiguous(), udir[sl].contiguous(), roairn[sl].contiguous(), wstar[sl].contiguous()

    def plain_step():
        nonlocal p_a, p_b
        g.propags_dev(p_a, p_b)
        g.implsch_dev(p_b, pu, pd, pt_tw, pt_us, pt_z0, pr, pw)
        p_a, p_b = p_b, p_a

    plain_step()
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for _ in range(K):
        plain_step()
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - w0) / K
    opc = json.load(open(OPCOUNT_FILE))["c2_sub_iphys0"]
    snl = kt.get("implsch_snl")
    res = {
        "workload": f"c3: {nmem} members x {n} sea points (0.1deg 450x800 sub-box, 36x36, shallow + currents, "
                    f"IPHYS=0, 300 s), members differ by 10% wind perturbations (200 km scale)",
        "members": nmem, "nsea": n, "member_points": nmp, "counter": counter, "steps": K,
        "value": nmp / t_graph, "unit": "member-point updates/s",
        "ms_per_step_graph": round(t_graph * 1e3, 2), "ms_per_step_eager": round(t_eager * 1e3, 2),
        "plain_member_ms_per_step": round(t_plain * 1e3, 3),
        "cost_per_member_vs_plain": round(t_graph / nmem / t_plain, 4),
        "kernel_ms_per_step": {k: round(v, 3) for k, v in kt.items()},
        "launch_shapes": dict(zip(("prop_wg", "src_wg", "snl_per_cu", "chunk"), shapes)),
        "state_finite": finite,
    }
    if snl:
        tf = opc["stages"]["snl"]["flops_per_point"] * nmp / (snl * 1e-3) / 1e12
        res["roofline_snl"] = dict(kernel="implsch_snl_kernel<1>", bound="valu", achieved=round(tf, 3),
                                   peak=FP32_PEAK_TFLOPS, unit="TFLOP/s", frac=round(tf / FP32_PEAK_TFLOPS, 4))
    g.close()
    return res


def bench_nested(args, torch):
    """Configuration 4: three grids (0.5 deg 24 x 25, 0.2 deg and 0.05 deg
    36 x 36, about 50 k, 200 k and 800 k sea points), the last two a nest with
    the boundary exchange after every step, advanced K steps grid after grid
    on one stream (sequential launches) and as one graph whose grids are
    concurrent branches (fused)."""
    from wam_amd import WamGpu, WamMultiGrid, workloads as W
    t0 = time.time()
    K = args.nest_steps
    grids_in, (c, f, link) = W.c4_case()
    dev = torch.device("cuda:0")
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    grids = []
    for cfg, mask, fl in grids_in:
        g = WamGpu(cfg, mask)
        g.prepare_propagation(fl["depth"], fl["u"], fl["v"])
        n = g.nsea
        st = dict(fa=device_spectrum(torch, cfg, t(fl["udir"])).reshape(-1), u10=t(fl["u10"]), udir=t(fl["udir"]),
                  roairn=torch.full((n,), 1.225, device=dev), wstar=torch.zeros(n, device=dev),
                  tauw=torch.zeros(n, device=dev), ustar=torch.zeros(n, device=dev), z0=torch.zeros(n, device=dev))
        st["fb"] = torch.empty_like(st["fa"])
        grids.append((g, st))
    nsea = [g.nsea for g, _ in grids]
    torch.cuda.synchronize()
    log(f"[nested] grids {nsea}, nest {link[0].size} -> {link[1].size} points ({time.time() - t0:.1f}s)")
    mg = WamMultiGrid(grids)
    mg.add_nest(c, f, *link)
    res = {"workload": "c4: 0.5deg 24x25, 0.2deg 36x36 and 0.05deg 36x36 grids (shallow + currents, IPHYS=0, "
                       "300 s); 0.2deg -> 0.05deg nest exchanged every step",
           "nsea": nsea, "nest_points": [int(link[0].size), int(link[1].size)], "steps": K,
           "counter": [int(g.scalar("COUNTER")) for g, _ in grids]}
    for name, fused in (("sequential", False), ("fused", True)):
        mg.run(K, fused=fused)                 # warm-up (and the capture of the graph)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        mg.run(K, fused=fused)
        torch.cuda.synchronize()
        res[f"ms_per_step_{name}"] = round((time.perf_counter() - w0) / K * 1e3, 3)
    res["value"] = sum(nsea) / (res["ms_per_step_fused"] * 1e-3)
    res["unit"] = "sea-point updates/s (all grids)"
    res["fused_vs_sequential"] = round(res["ms_per_step_sequential"] / res["ms_per_step_fused"], 4)
    res["state_finite"] = all(bool(torch.isfinite(st[mg.state(i)]).all()) for i, (_, st) in enumerate(grids))
    log(f"[nested] {K} steps: sequential {res['ms_per_step_sequential']} ms/step, fused {res['ms_per_step_fused']}")
    mg.close()
    for g, _ in grids:
        g.close()
    return res


def bench_adjoint(args, torch):
    """Configuration 5: the gradient of J = sum of squared Hs misfits at 500
    seeded points every hour of a 12 h window (144 steps of 300 s) with
    respect to the wind components u, v of every step, on a deep-water 0.25
    deg grid of 100 800 sea points (24 x 25, IPHYS = 0).  Observations are
    the Hs of a truth run with perturbed winds plus N(0, 0.1 m).  Timed: one
    forward step (wind map, PROPAGS, IMPLSCH) against one adjoint step (wind
    map, PROPAGS recomputed, IMPLSCH's adjoint, PROPAGS transposed, the wind
    map's adjoint) at the same state; the whole window's forward and gradient
    with every state stored and with the binomial schedule at args.adj_snaps
    slots (the two gradients must be equal bit for bit); and the gradient of
    a batch of args.adj_members members over args.adj_member_steps steps."""
    from wam_amd import WamGpu, workloads as W
    from wam_amd.adjoint import WindowModel, energy_weights, hs_from_spectra, window
    from wam_amd.model import checkpoint_recomputed_steps
    t0 = time.time()
    nsteps = args.adj_steps
    cfg, mask, f, u, v, obs, tu, tv = W.c5_case(nsteps=nsteps)
    g = WamGpu(cfg, mask)
    n = g.nsea
    counter, cfl = g.prepare_propagation(f["depth"])
    dev = torch.device("cuda:0")
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    U, V = t(u), t(v)
    fl0 = t(W.jonswap_spectrum(cfg, f["udir"]).ravel())
    tw0 = torch.zeros(n, device=dev)
    w = energy_weights(g)
    every = f["obs_every"]
    log(f"[adjoint] {n} sea points, COUNTER={counter}, {nsteps} steps, {obs.size} observed points "
        f"({time.time() - t0:.1f}s)")
    # observations: the truth run (perturbed winds) plus N(0, 0.1 m)
    truth = WindowModel(g, nsteps, obs, obs_every=every, nsnap=0)
    with torch.no_grad():
        hs_t = hs_from_spectra(truth.forward(t(tu), t(tv), fl0, tw0), w)
    rng = np.random.default_rng(20261207)
    obs_v = hs_t + t(rng.normal(0.0, f["obs_sd"], tuple(hs_t.shape)))
    del truth
    # one forward step and one adjoint step at the same state, K times each
    ro, ws = torch.full((n,), 1.225, device=dev), torch.zeros(n, device=dev)
    fa, fb = fl0.clone(), torch.empty_like(fl0)
    tauw, ustar, z0, u10, udir = (torch.zeros(n, device=dev) for _ in range(5))
    for i in range(12):                                 # an hour of spin-up
        g.wind_uv_dev(U[i], V[i], u10, udir)
        g.propags_dev(fa, fb)
        g.implsch_dev(fb, u10, udir, tauw, ustar, z0, ro, ws)
        fa, fb = fb, fa
    K = args.adj_reps
    fs, ts = fa.clone(), tauw.clone()
    fbar = torch.randn_like(fa)
    fbo, u10b, udirb, twb, ub, vb = torch.empty_like(fa), *(torch.empty(n, device=dev) for _ in range(5))

    def fwd_step():
        nonlocal fa, fb
        g.wind_uv_dev(U[12], V[12], u10, udir)
        g.propags_dev(fa, fb)
        g.implsch_dev(fb, u10, udir, tauw, ustar, z0, ro, ws)
        fa, fb = fb, fa

    def adj_step():
        g.wind_uv_dev(U[12], V[12], u10, udir)
        g.step_adjoint_dev(fs, u10, udir, ts, ro, ws, fbar, fbo, u10b, udirb, twb)
        g.wind_uv_adjoint_dev(U[12], V[12], u10b, udirb, ub, vb)

    res = {"workload": "c5: 0.25deg 400x360 deep water (999 m), 30% land, 24x25, IPHYS=0, 300 s; "
                       f"J = sum (Hs - obs)^2 at {obs.size} points every {every} steps over {nsteps} steps, "
                       "gradient w.r.t. u, v of every step",
           "nsea": n, "counter": counter, "steps": nsteps, "obs_points": int(obs.size)}
    timing = {}
    for name, fn in (("forward", fwd_step), ("adjoint", adj_step)):
        fn()
        torch.cuda.synchronize()
        g.set_timing(True)
        w0 = time.perf_counter()
        for _ in range(K):
            fn()
        torch.cuda.synchronize()
        timing[name] = (time.perf_counter() - w0) / K
        for kname in ("propags", "implsch_pre", "implsch_snl", "implsch_post", "implsch_adjoint", "propags_adjoint"):
            ms, calls = g.kernel_time(kname)
            if calls:
                res.setdefault("kernel_ms_" + name, {})[kname] = round(ms / K, 3)
        g.set_timing(False)
    res["ms_forward_step"] = round(timing["forward"] * 1e3, 3)
    res["ms_adjoint_step"] = round(timing["adjoint"] * 1e3, 3)
    res["adjoint_over_forward_step"] = round(timing["adjoint"] / timing["forward"], 3)
    ka = res.get("kernel_ms_adjoint", {}).get("implsch_adjoint")
    if ka:
        # compulsory HBM traffic of the kernel: F and the bar of FL3 out read,
        # the input spectrum's bar written (4 bytes each per bin), per-point
        # fields negligible
        gbs = 12.0 * g.kl * g.ml * n / (ka * 1e-3) / 1e9
        res["roofline_implsch_adjoint"] = dict(kernel="implsch_adj_kernel", bound="valu (per-point scalar chains, "
                                               "LDS)", hbm_achieved=round(gbs, 1), hbm_peak=HBM_PEAK_GBS,
                                               unit="GB/s", hbm_frac=round(gbs / HBM_PEAK_GBS, 4))
    del fa, fb, fs, ts, fbar, fbo
    # the whole window: forward, then the gradient, every state stored and binomial
    grads = {}
    for label, nsnap in (("store_all", None), ("binomial", args.adj_snaps)):
        uu, vv = U.clone().requires_grad_(True), V.clone().requires_grad_(True)
        model = WindowModel(g, nsteps, obs, obs_every=every, nsnap=nsnap)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        spec = window(uu, vv, fl0, tw0, model)
        hs = hs_from_spectra(spec, w)
        j = ((hs - obs_v) ** 2).sum()
        torch.cuda.synchronize()
        t_f = time.perf_counter() - w0
        w0 = time.perf_counter()
        j.backward()
        torch.cuda.synchronize()
        t_b = time.perf_counter() - w0
        grads[label] = (uu.grad, vv.grad)
        res[label] = {"nsnap": model.nsnap, "forward_s": round(t_f, 3), "reverse_s": round(t_b, 3),
                      "reverse_over_forward": round(t_b / t_f, 3),
                      "gradient_over_forward": round((t_f + t_b) / t_f, 3),
                      "forward_steps_run": model.forward_steps,
                      "recomputed_steps_schedule": checkpoint_recomputed_steps(nsteps, model.nsnap),
                      "J": float(j.detach())}
        log(f"[adjoint] {label}: forward {t_f:.2f}s, reverse {t_b:.2f}s, J = {float(j):.6g}")
        del model, spec
        torch.cuda.empty_cache()
    gu0, gv0 = grads["store_all"]
    gu1, gv1 = grads["binomial"]
    res["binomial_equals_store_all"] = bool(torch.equal(gu0, gu1) and torch.equal(gv0, gv1))
    res["gradient_norm"] = float(torch.sqrt((gu0.double() ** 2).sum() + (gv0.double() ** 2).sum()))
    res["gradient_finite"] = bool(torch.isfinite(gu0).all() and torch.isfinite(gv0).all())
    del grads, gu0, gv0, gu1, gv1
    torch.cuda.empty_cache()
    # a batch of members, binomial checkpointing
    nm, ms = args.adj_members, args.adj_member_steps
    if nm > 1:
        rngm = np.random.default_rng(20261208)
        fac = 1.0 + 0.05 * rngm.standard_normal((1, nm, 1))
        um = t((u[:ms, None, :] * fac).reshape(ms, -1))
        vm = t((v[:ms, None, :] * fac).reshape(ms, -1))
        uu, vv = um.requires_grad_(True), vm.requires_grad_(True)
        model = WindowModel(g, ms, obs, obs_every=every, nsnap=args.adj_snaps, nmem=nm)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        spec = window(uu, vv, fl0.repeat(nm), torch.zeros(nm * n, device=dev), model)
        hs = hs_from_spectra(spec, w)
        j = ((hs - obs_v[:ms // every, None, :]) ** 2).sum()
        torch.cuda.synchronize()
        t_f = time.perf_counter() - w0
        w0 = time.perf_counter()
        j.backward()
        torch.cuda.synchronize()
        t_b = time.perf_counter() - w0
        res["batched"] = {"members": nm, "steps": ms, "nsnap": model.nsnap, "forward_s": round(t_f, 3),
                          "reverse_s": round(t_b, 3), "reverse_over_forward": round(t_b / t_f, 3),
                          "gradient_over_forward": round((t_f + t_b) / t_f, 3),
                          "member_point_steps_per_s_gradient": round(nm * n * ms / (t_f + t_b), 1),
                          "gradient_finite": bool(torch.isfinite(uu.grad).all() and torch.isfinite(vv.grad).all())}
        log(f"[adjoint] {nm} members x {ms} steps: forward {t_f:.2f}s, reverse {t_b:.2f}s")
        del model, spec, uu, vv
    g.close()
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ny", type=int, default=901)
    ap.add_argument("--nx", type=int, default=1600)
    ap.add_argument("--iphys", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-ny", type=int, default=360)
    ap.add_argument("--cpu-nx", type=int, default=480)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--nest", action="store_true",
                    help="measure the nest boundary exchange on the resident state instead (one JSON line)")
    ap.add_argument("--no-ensemble", action="store_true", help="leave out the batched ensemble measurement")
    ap.add_argument("--ens-members", type=int, default=32)
    ap.add_argument("--ens-steps", type=int, default=4)
    ap.add_argument("--no-st6", action="store_true", help="leave out the ST6 measurement")
    ap.add_argument("--st6-steps", type=int, default=3)
    ap.add_argument("--no-nested", action="store_true", help="leave out the nested-grid measurement")
    ap.add_argument("--nest-steps", type=int, default=6)
    ap.add_argument("--no-adjoint", action="store_true", help="leave out the adjoint (gradient) measurement")
    ap.add_argument("--adj-steps", type=int, default=144, help="steps of the adjoint window (300 s each)")
    ap.add_argument("--adj-reps", type=int, default=10)
    ap.add_argument("--adj-snaps", type=int, default=12, help="checkpoint slots of the binomial schedule")
    ap.add_argument("--adj-members", type=int, default=16)
    ap.add_argument("--adj-member-steps", type=int, default=24)
    ap.add_argument("--only", choices=("ensemble", "nested", "adjoint", "st6"),
                    help="run one of the configurations alone and print its object as the JSON line (profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one plain 1 GiB device copy before the timed steps (reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import torch
    if args.only:
        fn = {"ensemble": bench_ensemble, "nested": bench_nested, "adjoint": bench_adjoint, "st6": bench_st6}[args.only]
        print(json.dumps(fn(args, torch)), flush=True)
        return
    from wam_amd import WamGpu, workloads as W

    t0 = time.time()
    cfg, mask, f = W.c2_case(ny=args.ny, nx=args.nx, iphys=args.iphys)
    g = WamGpu(cfg, mask)
    n = g.nsea
    log(f"[bench] grid {args.ny}x{args.nx}, NSEA={n}, {cfg['kl']}x{cfg['ml']}, inputs {time.time() - t0:.1f}s")
    counter, cfl = g.prepare_propagation(f["depth"], f["u"], f["v"])
    log(f"[bench] PREPARE_PROPAGATION done: COUNTER={counter} max CFL={cfl:.4f} ({time.time() - t0:.1f}s)")
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    g.set_stream(stream.cuda_stream)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    u10, udir = t(f["u10"]), t(f["udir"])
    tauw, ustar, z0 = (torch.zeros(n, device=dev) for _ in range(3))
    roairn, wstar = torch.full((n,), 1.225, device=dev), torch.zeros(n, device=dev)
    fa = device_spectrum(torch, cfg, udir).reshape(-1)
    fb = torch.empty_like(fa)
    torch.cuda.synchronize()
    log(f"[bench] state resident in HBM ({fa.numel() * 4 / 1e9:.2f} GB spectrum), {time.time() - t0:.1f}s")

    if args.nest:
        bench_nest(args, torch, g, cfg, mask, fa)
        return

    def step(ev=None):
        nonlocal fa, fb
        if ev:
            ev[0].record(stream)
        g.propags_dev(fa, fb)
        fa, fb = fb, fa
        if ev:
            ev[1].record(stream)
        g.implsch_dev(fa, u10, udir, tauw, ustar, z0, roairn, wstar)
        if ev:
            ev[2].record(stream)

    if args.copy_calibration:
        src = torch.ones(1 << 28, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)          # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        log("[bench] copy calibration: 1 GiB read + 1 GiB written by one elementwise kernel")
        del src, dst

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"[bench] warm-up step {i + 1} done")
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    g.set_timing(True)            # HIP events around every kernel of the timed steps, on the library's stream
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for i in range(args.steps):
        step(evs[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ktime = {}
    for name in KERNELS:
        ms, calls = g.kernel_time(name)
        if calls:
            ktime[name] = ms / calls
    g.set_timing(False)
    t_prop = np.array([a.elapsed_time(b) for a, b, c in evs]) * 1e-3
    t_src = np.array([b.elapsed_time(c) for a, b, c in evs]) * 1e-3
    ms_step = wall / args.steps * 1e3
    log(f"[bench] {args.steps} steps: {ms_step:.2f} ms/step; PROPAGS {t_prop.mean() * 1e3:.2f} ms "
        f"({counter} sub-steps), IMPLSCH {t_src.mean() * 1e3:.2f} ms")
    finite = bool(torch.isfinite(fa).all().item())
    # significant wave height 4*sqrt(E), E = sum over m of DFIM(m) * sum over k of F (no tail)
    dfim = torch.tensor(g.table("DFIM"), device=dev)
    energy = (fa.view(cfg["ml"], cfg["kl"], n).sum(1) * dfim[:, None]).sum(0)
    hs = 4.0 * torch.sqrt(torch.clamp(energy, min=0))

    kl, ml = cfg["kl"], cfg["ml"]
    traffic = {}
    if os.path.exists(TRAFFIC_FILE):
        traffic = json.load(open(TRAFFIC_FILE))
    # PROPAGS algorithmic bytes per sub-step.  The weights are computed inside
    # the stencil (the library's default for shallow water with currents), so
    # the traffic is the spectrum read and written once plus the per-point
    # fields: 4 neighbour indices, U, V, INDEP, INDEP_G, the six depth and
    # current gradients, DRGM, DELLA0, DPSN_S, DPSN_N (19 x 4 B).
    prop_bytes = n * kl * ml * 4 * 2 + n * 4 * 19
    prop_ms = ktime["propags"]
    prop_gbs = prop_bytes / (prop_ms * 1e-3) / 1e9
    # KL a multiple of 3: the register-block kernel, else a direction at a time
    prop_kernel = "propags_blk_kernel<1, 3>" if kl % 3 == 0 else "propags_fused_sc_kernel<1>"
    roofs = {"propags": dict(kernel=prop_kernel, bound="hbm", achieved=round(prop_gbs, 1),
                             peak=HBM_PEAK_GBS, unit="GB/s", frac=round(prop_gbs / HBM_PEAK_GBS, 4),
                             traffic=traffic.get(prop_kernel.split("<")[0], {}).get("bytes_per_dispatch"),
                             algorithmic_bytes_per_launch=prop_bytes, avg_launch_ms=round(prop_ms, 3),
                             note="one sub-step; weights computed in the stencil (no 9-plane weight stream)")}
    # IMPLSCH: compute bound (VALU FP32).  Algorithmic flops per sea point
    # from the oracle's op count, per stage for the three kernels of IPHYS=0
    # (pre: up to SDISSIP; snl: SNONLIN, SBOTTOM and the limited update;
    # post: new mean parameters and tail).
    opc = json.load(open(OPCOUNT_FILE))[f"c2_sub_iphys{args.iphys}"]
    stage_of = {"implsch_pre": "pre", "implsch_snl": "snl", "implsch_post": "post"}
    for name, ms in ktime.items():
        if name == "propags":
            continue
        if name == "implsch":
            fpp, tpp, kname = opc["flops_per_point"], opc["trans_per_point"], f"implsch_kernel<{args.iphys},1>"
        else:
            st = opc["stages"][stage_of[name]]
            fpp, tpp, kname = st["flops_per_point"], st["trans_per_point"], f"{name}_kernel<1>"
        tflops = fpp * n / (ms * 1e-3) / 1e12
        roofs[name] = dict(kernel=kname, bound="valu", achieved=round(tflops, 3), peak=FP32_PEAK_TFLOPS,
                           unit="TFLOP/s", frac=round(tflops / FP32_PEAK_TFLOPS, 4),
                           traffic=traffic.get(kname, {}).get("bytes_per_dispatch"),
                           algorithmic_flops_per_launch=round(fpp * n), flops_per_point=round(fpp, 1),
                           transcendentals_per_point=round(tpp, 1), avg_launch_ms=round(ms, 3),
                           note="FP32 vector (VALU) work, no GEMM-shaped arithmetic, so no MFMA; peak is "
                                "MI355X's FP32 vector 157.3 TFLOP/s; unfused add/mul/div/sqrt as counted by the "
                                "oracle's op counter")
    # the dominant kernel: the largest share of the step's GPU time
    share = {k: v * (counter if k == "propags" else 1) for k, v in ktime.items()}
    dominant = max(share, key=share.get)
    others = {k: v for k, v in roofs.items() if k != dominant}
    out = {
        "metric": "sea-point spectrum updates/s per step (src+prop); batched & adjoint vs plain",
        "value": n / (ms_step * 1e-3),
        "unit": "sea-point updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded land mask, bathymetry, winds, currents; JONSWAP initial spectra)",
        "config": {"workload": "c2: 0.1deg 901x1600 spherical, 30% land, 36x36, shallow + depth/current "
                               "refraction, IPHYS=0, 300 s step (PROPAGS + IMPLSCH)",
                   "nsea": n, "kl": kl, "ml": ml, "iphys": args.iphys, "counter": counter,
                   "max_cfl": round(cfl, 4), "parallelism": "single GPU"},
        "t_propags_ms": round(t_prop.mean() * 1e3, 3), "t_implsch_ms": round(t_src.mean() * 1e3, 3),
        "kernel_ms": {k: round(v, 3) for k, v in ktime.items()},
        "roofline": roofs[dominant],
        "roofline_other": others.get("propags", next(iter(others.values()), None)),
        "roofline_all": others,
        "state_finite": finite, "mean_hs_m": round(float(hs.mean()), 4),
    }
    # the single-grid state and its context are released before the other configurations
    del fa, fb, u10, udir, tauw, ustar, z0, roairn, wstar, dfim, energy, hs
    g.close()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    if not args.no_ensemble:
        out["batched"] = bench_ensemble(args, torch)
    if not args.no_st6:
        torch.cuda.empty_cache()
        out["st6"] = bench_st6(args, torch)
    if not args.no_nested:
        torch.cuda.empty_cache()
        out["nested"] = bench_nested(args, torch)
    if not args.no_adjoint:
        torch.cuda.empty_cache()
        out["adjoint"] = bench_adjoint(args, torch)
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
