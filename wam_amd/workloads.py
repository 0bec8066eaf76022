"""Model configurations and seeded synthetic inputs.

The two single-GPU workloads of the hot path:
  * ``c1``: spherical 60x60 all-sea grid at 0.5 deg (3600 points), deep
    water, 24 directions x 25 frequencies (f1 = 0.0418 Hz), ST6 physics
    (IPHYS=2), uniform 20 m/s wind from 270 deg, JONSWAP initial spectra,
    600 s steps.
  * ``c2``: 0.1 deg spherical grid 901 x 1600 with a seeded fractal land mask
    (30 % land, about 1.0 M sea points), 36 x 36 spectra, log-normal fractal
    bathymetry 5-5000 m with depth and current refraction, Gaussian random
    wind field (mean 10, sd 5 m/s, k**-3 spectrum, 500 km decorrelation),
    divergence-free currents up to 1 m/s,
    300 s steps, Janssen physics (IPHYS=0).
Any grid can be cut to a smaller sub-box for parity runs.  Coordinates are in
the reference's M_SEC units (1 degree = 360000, wam_coordinate_module.f90:41-43).
"""
import math

import numpy as np

M_DEGREE = 360000


def base_config(**kw):
    cfg = dict(kl=24, ml=25, fr1=0.0418, iref=1, iphys=0, shallow=0, spherical=1, refr_d=0, refr_c=0,
               iper=0, isnonlin=0, idelpro=600, idelt=600, ndepth=69, deptha=1.0, depthd=1.1,
               betamax=1.20, zalp=0.008, alpha=0.006, cdfac=1.13,
               nx=60, ny=60, amosop=-15 * M_DEGREE, amonop=int(14.5 * M_DEGREE), amowep=0,
               amoeap=int(29.5 * M_DEGREE), xdella=M_DEGREE // 2, xdello=M_DEGREE // 2)
    cfg.update(kw)
    return cfg


def grid_config(cfg, nx, ny, south_deg, west_deg, dlat_deg, dlon_deg):
    c = dict(cfg)
    c.update(nx=nx, ny=ny, amosop=int(round(south_deg * M_DEGREE)), amowep=int(round(west_deg * M_DEGREE)),
             xdella=int(round(dlat_deg * M_DEGREE)), xdello=int(round(dlon_deg * M_DEGREE)))
    c["amonop"] = c["amosop"] + (ny - 1) * c["xdella"]
    c["amoeap"] = c["amowep"] + (nx - 1) * c["xdello"]
    return c


def _fractal(ny, nx, beta, rng):
    """Gaussian random field with an isotropic k**-beta power spectrum, unit variance."""
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    amp = k ** (-beta / 2.0)
    amp[0, 0] = 0.0
    noise = rng.standard_normal((ny, nx // 2 + 1)) + 1j * rng.standard_normal((ny, nx // 2 + 1))
    f = np.fft.irfft2(noise * amp, s=(ny, nx))
    f = (f - f.mean()) / (f.std() + 1e-30)
    return f


def _smooth_blur(x, cells):
    """Gaussian blur of a 2-D field (periodic FFT convolution)."""
    ky = np.fft.fftfreq(x.shape[0])[:, None]
    kx = np.fft.rfftfreq(x.shape[1])[None, :]
    g = np.exp(-2.0 * (math.pi * cells) ** 2 * (kx * kx + ky * ky))
    return np.fft.irfft2(np.fft.rfft2(x) * g, s=x.shape)


def _smooth_field(ny, nx, scale_cells, rng):
    """Gaussian random field with a Gaussian correlation of the given length (cells)."""
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    k2 = kx * kx + ky * ky
    amp = np.exp(-0.5 * k2 * (2 * math.pi * scale_cells) ** 2 / 2.0)
    noise = rng.standard_normal((ny, nx // 2 + 1)) + 1j * rng.standard_normal((ny, nx // 2 + 1))
    f = np.fft.irfft2(noise * amp, s=(ny, nx))
    return (f - f.mean()) / (f.std() + 1e-30)


def _kminus3_field(ny, nx, scale_cells, rng):
    """Gaussian random field, unit variance, with the isotropic power spectrum
    P(kappa) ~ (kappa**2 + 1/L**2)**(-3/2): k**-3 above 1/L and flat below.  In
    two dimensions its covariance is exp(-r/L), so L (cells) is the e-folding
    decorrelation length."""
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    kap2 = (2 * math.pi) ** 2 * (kx * kx + ky * ky)
    amp = (kap2 + 1.0 / (scale_cells * scale_cells)) ** -0.75
    amp[0, 0] = 0.0
    noise = rng.standard_normal((ny, nx // 2 + 1)) + 1j * rng.standard_normal((ny, nx // 2 + 1))
    f = np.fft.irfft2(noise * amp, s=(ny, nx))
    return (f - f.mean()) / (f.std() + 1e-30)


def c1_case():
    """3600-point deep-water ST6 case: (cfg, mask, fields)."""
    cfg = base_config(iphys=2, kl=24, ml=25, idelpro=600, idelt=600)
    mask = np.ones((cfg["ny"], cfg["nx"]), np.uint8)
    n = int(mask.sum())
    fields = dict(depth=np.full(n, 999.0, np.float32), u=None, v=None,
                  u10=np.full(n, 20.0, np.float32), udir=np.full(n, math.pi / 2, np.float32))
    return cfg, mask, fields


def c2_case(ny=901, nx=1600, seed=20261203, iphys=0, kl=36, ml=36, dlat=0.1, dlon=0.1, south=-45.0, west=0.0,
            land_frac=0.30):
    """Production-shaped shallow/current case; sub-boxes keep the same statistics."""
    cfg = base_config(iphys=iphys, kl=kl, ml=ml, shallow=1, refr_d=1, refr_c=1, idelpro=300, idelt=300)
    cfg = grid_config(cfg, nx, ny, south, west, dlat, dlon)
    rng = np.random.default_rng(seed)
    land = _fractal(ny, nx, 3.0, rng)
    thr = np.quantile(land, 1.0 - land_frac)
    mask = (land < thr).astype(np.uint8)
    sea = mask.astype(bool)
    n = int(sea.sum())
    # log-normal bathymetry: a smooth basin-scale field with weak fractal
    # detail, shoaling log-linearly over ~200 km shelves towards the coast
    from scipy.ndimage import distance_transform_edt
    dist = distance_transform_edt(sea)
    shelf = np.clip(dist / 18.0, 0.0, 1.0)
    logd = np.log(2500.0) + 0.5 * _smooth_field(ny, nx, 25.0, rng) + 0.05 * _fractal(ny, nx, 4.0, rng)
    depth2 = np.exp(np.log(5.0) + (logd - np.log(5.0)) * shelf)
    depth2 = np.clip(depth2, 5.0, 5000.0)
    # wind speed: Gaussian random field, mean 10 and sd 5 m/s, k**-3 spectrum
    # with a 500 km decorrelation length (0.1 deg ~ 11 km)
    cells = 500e3 / (dlat * 111e3)
    wspd = np.clip(10.0 + 5.0 * _kminus3_field(ny, nx, cells, rng), 0.5, None)
    wdir = math.pi / 2 + 0.8 * _smooth_field(ny, nx, cells / 2.0, rng)
    # divergence-free current from a stream function, max 1 m/s
    psi = _smooth_field(ny, nx, cells / 8.0, rng)
    cu = -(np.roll(psi, -1, 0) - np.roll(psi, 1, 0))
    cv = np.roll(psi, -1, 1) - np.roll(psi, 1, 1)
    sc = 1.0 / max(np.abs(cu).max(), np.abs(cv).max(), 1e-30)
    fields = dict(depth=depth2[sea].astype(np.float32), u=(sc * cu)[sea].astype(np.float32),
                  v=(sc * cv)[sea].astype(np.float32), u10=wspd[sea].astype(np.float32),
                  udir=np.mod(wdir[sea], 2 * math.pi).astype(np.float32))
    assert fields["depth"].shape[0] == n
    return cfg, mask, fields


def c3_case(nmem=32, ny=450, nx=800, seed=20261204, iphys=0, kl=36, ml=36, amp=0.10, scale_km=200.0):
    """Ensemble of NMEM members on a 0.1 deg sub-box of the production-shaped
    grid (450 x 800 cells, about 250 k sea points): the members share the
    grid, depth and currents and differ by their winds, the base wind times
    (1 + amp * GRF_m), GRF_m a unit Gaussian random field with a Gaussian
    correlation of scale_km, seeded per member.  Returns (cfg, mask, fields,
    u10[nmem][nsea], udir[nmem][nsea])."""
    cfg, mask, f = c2_case(ny=ny, nx=nx, seed=seed, iphys=iphys, kl=kl, ml=ml)
    sea = mask.astype(bool)
    cells = scale_km * 1e3 / (0.1 * 111e3)
    u10 = np.empty((nmem, int(sea.sum())), np.float32)
    for m in range(nmem):
        g = _smooth_field(ny, nx, cells, np.random.default_rng(seed + 1 + m))
        u10[m] = np.clip(f["u10"] * (1.0 + amp * g[sea]), 0.5, None).astype(np.float32)
    udir = np.broadcast_to(f["udir"], u10.shape).astype(np.float32)
    return cfg, mask, f, u10, udir


def grid_lat_lon(cfg):
    """Latitudes and longitudes [deg] of the rows and columns."""
    lat = (cfg["amosop"] + np.arange(cfg["ny"]) * cfg["xdella"]) / M_DEGREE
    lon = (cfg["amowep"] + np.arange(cfg["nx"]) * cfg["xdello"]) / M_DEGREE
    return lat, lon


def nest_points(ccfg, cmask, fcfg, fmask):
    """The tables of a nest, as a pre-processor sets them up: the fine grid's
    sea points on its outer edge (IJARF, 1-based) take the spectrum between two
    coarse-grid sea points (IBFL, IBFR, 1-based into IJARC, the coarse output
    points, themselves 1-based sea-point numbers of the coarse grid), at the
    fraction BFW of the way from the first to the second: along the fine
    grid's southern and northern edges the coarse columns either side on the
    nearest coarse row, along its western and eastern edges the coarse rows
    either side on the nearest column.  A coarse cell on land gives way to
    the other one; where both are land the point is left out."""
    clat, clon = grid_lat_lon(ccfg)
    flat, flon = grid_lat_lon(fcfg)
    cidx = np.full(cmask.shape, 0, np.int64)
    cidx[cmask.astype(bool)] = np.arange(1, int(cmask.sum()) + 1)
    fidx = np.full(fmask.shape, 0, np.int64)
    fidx[fmask.astype(bool)] = np.arange(1, int(fmask.sum()) + 1)
    dla, dlo = ccfg["xdella"] / M_DEGREE, ccfg["xdello"] / M_DEGREE
    ny, nx = fmask.shape
    out_pts, ijarf, left, right, w = {}, [], [], [], []

    def coarse(r, c):
        if 0 <= r < cmask.shape[0] and 0 <= c < cmask.shape[1] and cidx[r, c] > 0:
            return int(cidx[r, c])
        return 0

    for r in range(ny):
        for c in range(nx):
            if fidx[r, c] == 0 or not (r in (0, ny - 1) or c in (0, nx - 1)):
                continue
            if r in (0, ny - 1):            # along a row: bracket in longitude
                cr = int(round((flat[r] - clat[0]) / dla))
                x = (flon[c] - clon[0]) / dlo
                c0 = int(np.floor(x))
                a, b, f = coarse(cr, c0), coarse(cr, c0 + 1), x - c0
            else:                           # along a column: bracket in latitude
                cc = int(round((flon[c] - clon[0]) / dlo))
                y = (flat[r] - clat[0]) / dla
                r0 = int(np.floor(y))
                a, b, f = coarse(r0, cc), coarse(r0 + 1, cc), y - r0
            if a == 0 and b == 0:
                continue
            if a == 0:
                a, f = b, 0.0
            if b == 0:
                b, f = a, 0.0
            for q in (a, b):
                out_pts.setdefault(q, len(out_pts) + 1)
            ijarf.append(int(fidx[r, c]))
            left.append(out_pts[a])
            right.append(out_pts[b])
            w.append(f)
    ijarc = np.array(sorted(out_pts, key=out_pts.get), np.int32)
    return (ijarc, np.array(ijarf, np.int32), np.array(left, np.int32), np.array(right, np.int32),
            np.array(w, np.float32))


def c4_case(scale=1.0, seed=20261205):
    """Three grids of configuration 4: 0.5 deg 24 x 25 (about 50 k sea points),
    0.2 deg 36 x 36 (about 200 k) and 0.05 deg 36 x 36 (about 800 k) inside
    the 0.2 deg one, generated as c2_case (bathymetry, winds, currents), 300 s
    steps; the last two form a nest (coarse 0.2 deg -> fine 0.05 deg).  The
    0.5 deg grid's spectra differ, which the reference's boundary input does
    not accept (read_boundary_input.f90:136-156), so it runs beside them.
    scale < 1 shrinks every grid's rows and columns by that factor.  Returns
    ([(cfg, mask, fields)] * 3, (coarse, fine, nest_points(...)))."""
    s = lambda v: max(8, int(round(v * scale)))
    g1 = c2_case(ny=s(220), nx=s(325), seed=seed, kl=24, ml=25, dlat=0.5, dlon=0.5, south=-55.0, west=0.0)
    g2 = c2_case(ny=s(400), nx=s(715), seed=seed + 1, dlat=0.2, dlon=0.2, south=-40.0, west=0.0)
    # the fine grid: rows and columns of 0.05 deg inside the 0.2 deg grid's box
    ny3, nx3 = s(900), s(1270)
    south3 = -40.0 + 0.2 * s(400) / 2 - 0.05 * ny3 / 2
    west3 = 0.2 * s(715) / 2 - 0.05 * nx3 / 2
    g3 = c2_case(ny=ny3, nx=nx3, seed=seed + 2, dlat=0.05, dlon=0.05, south=round(south3, 2), west=round(west3, 2))
    link = nest_points(g2[0], g2[1], g3[0], g3[1])
    return [g1, g2, g3], (1, 2, link)


def c5_case(ny=400, nx=360, seed=20261206, nsteps=144, dlat=0.25, dlon=0.25, south=-50.0, west=0.0,
            land_frac=0.30, nobs=500, obs_every=12, truth_amp=0.10, obs_sd=0.1, kl=24, ml=25):
    """Adjoint workload (configuration 5): a deep-water grid (DEPTH = 999 m as
    in a run with SHALLOW_RUN = F, no currents), IPHYS = 0, KL x ML = 24 x 25 spectra,
    300 s steps; at the default size 400 x 360 cells of 0.25 deg with a
    seeded fractal land mask, about 100 k sea points.  The winds are u, v
    components for each of NSTEPS steps, blending two seeded k**-3 random
    fields (mean speed 10 m/s, sd 5 m/s, 500 km decorrelation) linearly over
    the window.  NOBS seeded sea points are observed every OBS_EVERY steps.
    Returns (cfg, mask, fields, u[nsteps][nsea], v[nsteps][nsea], obs_points,
    truth_u, truth_v): the truth winds are the winds times (1 + truth_amp *
    GRF) with a 200 km Gaussian random field; the caller makes observations
    from the truth run (Hs plus N(0, obs_sd) noise)."""
    cfg = base_config(iphys=0, kl=kl, ml=ml, shallow=0, refr_d=0, refr_c=0, idelpro=300, idelt=300)
    cfg = grid_config(cfg, nx, ny, south, west, dlat, dlon)
    rng = np.random.default_rng(seed)
    land = _fractal(ny, nx, 3.0, rng)
    mask = (land < np.quantile(land, 1.0 - land_frac)).astype(np.uint8)
    sea = mask.astype(bool)
    n = int(sea.sum())
    cells = 500e3 / (dlat * 111e3)

    def wind():
        spd = np.clip(10.0 + 5.0 * _kminus3_field(ny, nx, cells, rng), 0.5, None)
        wdir = math.pi / 2 + 0.8 * _smooth_field(ny, nx, cells / 2.0, rng)
        return (spd * np.sin(wdir))[sea], (spd * np.cos(wdir))[sea]

    (ua, va), (ub, vb) = wind(), wind()
    w = (np.arange(nsteps, dtype=np.float64) / max(nsteps - 1, 1))[:, None]
    u = ((1.0 - w) * ua[None, :] + w * ub[None, :]).astype(np.float32)
    v = ((1.0 - w) * va[None, :] + w * vb[None, :]).astype(np.float32)
    g = _smooth_field(ny, nx, 200e3 / (dlat * 111e3), rng)[sea]
    tu = (u * (1.0 + truth_amp * g)[None, :]).astype(np.float32)
    tv = (v * (1.0 + truth_amp * g)[None, :]).astype(np.float32)
    obs = np.sort(rng.choice(n, size=min(nobs, n), replace=False)).astype(np.int64)
    spd0 = np.hypot(u[0], v[0])
    udir0 = np.mod(np.arctan2(u[0], v[0]), 2 * math.pi)
    fields = dict(depth=np.full(n, 999.0, np.float32), u=None, v=None, u10=spd0.astype(np.float32),
                  udir=udir0.astype(np.float32), obs_every=obs_every, obs_sd=obs_sd)
    return cfg, mask, fields, u, v, obs, tu, tv


def frequencies(cfg):
    fr = np.empty(cfg["ml"], np.float64)
    fr[0] = cfg["fr1"] * 1.1 ** (1 - cfg["iref"])
    for m in range(1, cfg["ml"]):
        fr[m] = fr[m - 1] * 1.1
    return fr


def jonswap_spectrum(cfg, udir, alpha=0.018, fp=0.3, gamma=3.3, sa=0.07, sb=0.09):
    """Initial spectra F(IJ,K,M) as [ML][KL][NSEA] float32: JONSWAP E(f)
    (wam_jonswap_module.f90:136-213) times (2/pi) cos^2 spreading about the wind
    direction (wam_coldstart_module.f90:381-412)."""
    g = 9.806
    fr = frequencies(cfg)
    kl = cfg["kl"]
    th = (np.arange(kl) + 0.5) * (2 * math.pi / kl)
    arg = 1.25 * (fp / fr) ** 4
    et = alpha * g * g / (2 * math.pi) ** 4 * fr ** -5 * np.exp(-np.minimum(arg, 50.0))
    sig = np.where(fr > fp, sb, sa)
    et = et * gamma ** np.exp(-((fr / fp - 1.0) ** 2) / (2 * sig * sig))
    st = (2.0 / math.pi) * np.maximum(0.0, np.cos(th[:, None] - np.asarray(udir, np.float64)[None, :])) ** 2
    st[st < 1e-9] = 0.0
    f = et[:, None, None] * st[None, :, :]
    return f.astype(np.float32)
