"""Host-side mirror of WAM's interface for the hot path.

``WamGpu`` mirrors the reference's call sequence (wamodel.f90:286-451):
``prepare_propagation`` stands for PREPARE_PROPAGATION (initmdl.f90:204,
wamodel.f90:326), ``propags`` for PROPAGS(FL3) (wamodel.f90:332) and
``implsch`` for IMPLSCH(FL3, U10, UDIR, TAUW, USTAR, Z0, ROAIRN, WSTAR, DEPTH,
INDEP) (wamodel.f90:354).  Host numpy arrays go through the drop-in entry
points; torch tensors on the GPU go through the device-resident ones.
Errors raise ``WamError`` carrying the library's code and message, where the
reference calls ABORT1.
"""
import ctypes
import functools
import os
import sys

import numpy as np

from . import _lib


class WamError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"WAM error {code}: {msg}")
        self.code = code


def _stream_ordered(fn):
    """Device-memory entry points run on the context's stream.  When torch is
    in use and its current stream is another one, the context's stream waits
    for it before the call (inputs written there) and it waits for the
    context's stream after the call (outputs read there); both waits are on
    the device, the host does not block."""
    @functools.wraps(fn)
    def wrapper(self, *args, **kwargs):
        streams = self._torch_streams()
        if streams is None:
            return fn(self, *args, **kwargs)
        cur, own = streams
        own.wait_stream(cur)
        try:
            return fn(self, *args, **kwargs)
        finally:
            cur.wait_stream(own)
    return wrapper


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(ctypes.c_void_p)
    return ctypes.c_void_p(a.data_ptr())      # torch tensor


def _dates(dates):
    """The five restart dates (CDTPRO, CDTSOU, CDA, CDTA, CDCA) as the file's
    70 blank-padded characters; a blank CDTA or CDCA leaves out the depth or
    current record."""
    if len(dates) != 5:
        raise ValueError("five dates: CDTPRO, CDTSOU, CDA, CDTA, CDCA")
    out = b""
    for d in dates:
        d = (d or "").encode("ascii")
        if len(d) > 14:
            raise ValueError("a date has at most 14 characters")
        out += d.ljust(14, b" ")
    return out


def _undates(raw):
    return [raw[14 * i:14 * i + 14].decode("ascii").rstrip(" ") for i in range(5)]


def restart_write(path, dates, fl3, u10, udir, tauw, depth=None, u=None, v=None):
    """SAVE_RESTART_FILE's binary form from host arrays (wam_restart_write);
    fl3 is (NSEA, KL, ML) in Fortran order, i.e. a numpy array [ML][KL][NSEA]."""
    from . import _lib
    lib = _lib.load()
    f = np.ascontiguousarray(fl3, np.float32)
    if f.ndim != 3:
        raise ValueError("fl3 must be [ML][KL][NSEA]")
    ml, kl, n = f.shape
    a = [np.ascontiguousarray(x, np.float32) if x is not None else None for x in (u10, udir, tauw, depth, u, v)]
    rc = lib.wam_restart_write(os.fsencode(path), n, kl, ml, _dates(dates), _ptr(f), *[_ptr(x) for x in a])
    if rc:
        raise WamError(rc, f"wam_restart_write({path})")


def restart_read(path, nsea, kl, ml, depth=True, current=True):
    """CONNECT_RESTART's binary form into host arrays (wam_restart_read):
    returns the dates and a dict of fl3 [ML][KL][NSEA], u10, udir, tauw and,
    when the file holds them, depth, u, v."""
    from . import _lib
    lib = _lib.load()
    out = dict(fl3=np.empty((ml, kl, nsea), np.float32), u10=np.empty(nsea, np.float32),
               udir=np.empty(nsea, np.float32), tauw=np.empty(nsea, np.float32),
               depth=np.empty(nsea, np.float32), u=np.empty(nsea, np.float32), v=np.empty(nsea, np.float32))
    raw = ctypes.create_string_buffer(70)
    present = ctypes.c_int32(0)
    rc = lib.wam_restart_read(os.fsencode(path), nsea, kl, ml, raw, _ptr(out["fl3"]), _ptr(out["u10"]),
                              _ptr(out["udir"]), _ptr(out["tauw"]), _ptr(out["depth"]) if depth else None,
                              _ptr(out["u"]) if current else None, _ptr(out["v"]) if current else None,
                              ctypes.byref(present))
    if rc:
        raise WamError(rc, f"wam_restart_read({path})")
    if not (present.value & 1) or not depth:
        out.pop("depth")
    if not (present.value & 2) or not current:
        out.pop("u"), out.pop("v")
    return _undates(raw.raw), out


def boundary_write_header(path, kl, ml, th1, fr1, co, nbounc, idel_b_out, idel_bo_file):
    """SAVE_BOUNDARY_FILE's header record (wam_boundary_write_header): a new
    file with the eight REALs KL, ML, TH(1), FR(1), CO, NBOUNC, IDEL_B_OUT,
    IDEL_BO_FILE."""
    lib = _lib.load()
    hdr = np.array([kl, ml, th1, fr1, co, nbounc, idel_b_out, idel_bo_file], np.float32)
    rc = lib.wam_boundary_write_header(os.fsencode(path), _ptr(hdr))
    if rc:
        raise WamError(rc, f"wam_boundary_write_header({path})")


def boundary_append(path, lon, lat, cdtpro, mean, depth, fbc):
    """BOUNDARY_OUTPUT's records for one output time (wam_boundary_append):
    lon, lat [nb] (INTEGER, M_SEC), the date CDTPRO, mean [3][nb] (EMEAN, THQ,
    FMEAN), depth [nb] and the spectra fbc [nb][ML][KL]."""
    lib = _lib.load()
    f = np.ascontiguousarray(fbc, np.float32)
    nb, ml, kl = f.shape
    lo, la = (np.ascontiguousarray(x, np.int32) for x in (lon, lat))
    m = np.ascontiguousarray(mean, np.float32)
    d = np.ascontiguousarray(depth, np.float32)
    if lo.size != nb or la.size != nb or m.shape != (3, nb) or d.size != nb:
        raise ValueError("lon, lat, depth [nb] and mean [3][nb] for fbc [nb][ML][KL]")
    c = (cdtpro or "").encode("ascii").ljust(14, b" ")[:14]
    rc = lib.wam_boundary_append(os.fsencode(path), nb, kl, ml, _ptr(lo), _ptr(la), c, _ptr(m), _ptr(d), _ptr(f))
    if rc:
        raise WamError(rc, f"wam_boundary_append({path})")


def boundary_read_header(path):
    """READ_BOUNDARY_INPUT's header (wam_boundary_read_header): the eight
    header values and the byte offset of the first point record."""
    lib = _lib.load()
    hdr = np.zeros(8, np.float32)
    pos = ctypes.c_int64(0)
    rc = lib.wam_boundary_read_header(os.fsencode(path), _ptr(hdr), ctypes.byref(pos))
    if rc:
        raise WamError(rc, f"wam_boundary_read_header({path})")
    return hdr, pos.value


def boundary_read(path, pos, nb, kl, ml):
    """One set of boundary spectra from byte offset pos
    (wam_boundary_read): returns a dict of lon, lat, dates, mean [3][nb],
    depth, fbc [nb][ML][KL] and the offset of the next set."""
    lib = _lib.load()
    out = dict(lon=np.empty(nb, np.int32), lat=np.empty(nb, np.int32), mean=np.empty((3, nb), np.float32),
               depth=np.empty(nb, np.float32), fbc=np.empty((nb, ml, kl), np.float32))
    raw = ctypes.create_string_buffer(14 * nb)
    p = ctypes.c_int64(pos)
    rc = lib.wam_boundary_read(os.fsencode(path), ctypes.byref(p), nb, kl, ml, _ptr(out["lon"]), _ptr(out["lat"]),
                               raw, _ptr(out["mean"]), _ptr(out["depth"]), _ptr(out["fbc"]))
    if rc:
        raise WamError(rc, f"wam_boundary_read({path})")
    out["dates"] = [raw.raw[14 * i:14 * i + 14].decode("ascii") for i in range(nb)]
    return out, p.value


MAP_NOUT = 70


def _fflag(params):
    """FFLAG_P(1:70) as int32 from the 1-based parameter numbers to write."""
    f = np.zeros(MAP_NOUT, np.int32)
    for ip in params:
        f[int(ip) - 1] = 1
    return f


def _date14(d):
    return (d or "").encode("ascii").ljust(14, b" ")[:14]


def map_write(path, cdtpro, cdatea, nx, ny, corners, nlon_rg, zdello, params, grids, append=False):
    """WRITE_INT_PAR_OUTPUT's records for one output time (wam_map_write):
    corners = AMOWEP, AMOSOP, AMOEAP, AMONOP [M_SEC]; grids [len(params)][NY][NX]
    for the parameter numbers params in increasing order."""
    lib = _lib.load()
    c = np.ascontiguousarray(corners, np.int32)
    nl, zd = np.ascontiguousarray(nlon_rg, np.int32), np.ascontiguousarray(zdello, np.int32)
    g = np.ascontiguousarray(grids, np.float32)
    if list(params) != sorted(params) or g.shape != (len(params), ny, nx) or nl.size != ny or zd.size != ny:
        raise ValueError("params increasing, grids [len(params)][NY][NX], nlon_rg and zdello [NY]")
    rc = lib.wam_map_write(os.fsencode(path), int(bool(append)), _date14(cdtpro), _date14(cdatea), nx, ny, _ptr(c),
                           _ptr(nl), _ptr(zd), _ptr(_fflag(params)), _ptr(g))
    if rc:
        raise WamError(rc, f"wam_map_write({path})")


class WamGpu:
    def __init__(self, cfg, mask, host_only=False):
        self.lib = _lib.load()
        self.cfg = dict(cfg)
        c = _lib.WamConfig()
        for name, _ in _lib.WamConfig._fields_:
            setattr(c, name, cfg[name])
        self._c = c
        m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
        if m.size != cfg["nx"] * cfg["ny"]:
            raise ValueError("mask size must be nx*ny")
        h = ctypes.c_void_p()
        init = self.lib.wam_gpu_init_host if host_only else self.lib.wam_gpu_init
        rc = init(ctypes.byref(h), ctypes.byref(c), _ptr(m))
        self.ctx = h
        self._check(rc)
        self.kl, self.ml = cfg["kl"], cfg["ml"]
        self.nsea = int(self.scalar("NSEA"))
        if not host_only and (cfg.get("wave_breaking") or cfg.get("phillips")):
            self.set_source_options(bool(cfg.get("wave_breaking")), bool(cfg.get("phillips")))

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.wam_gpu_last_error(self.ctx)
            raise WamError(rc, msg.decode() if msg else "")

    def _torch_streams(self):
        """(torch's current stream, the context's stream as a torch stream),
        or None when there is nothing to order (torch not in use on the
        device, or the context already runs on torch's current stream)."""
        torch = sys.modules.get("torch")
        if torch is None or not torch.cuda.is_initialized():
            return None
        h = ctypes.c_void_p()
        self._check(self.lib.wam_gpu_get_stream(self.ctx, ctypes.byref(h)))
        cur = torch.cuda.current_stream()
        handle = h.value or 0
        if handle == cur.cuda_stream:
            return None
        cache = self.__dict__.setdefault("_own_streams", {})
        if handle not in cache:
            cache[handle] = torch.cuda.ExternalStream(handle)
        return cur, cache[handle]

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.wam_gpu_finalize(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- set-up -------------------------------------------------------------
    def set_propagation_weights(self, stored):
        """stored=True: weight planes in HBM (the reference's COEF_* arrays);
        False: computed inside the PROPAGS stencil; None (default): computed
        for shallow water with currents, stored otherwise.  Applies from the
        next prepare_propagation."""
        mode = -1 if stored is None else int(bool(stored))
        self._check(self.lib.wam_gpu_set_propagation_weights(self.ctx, mode))

    def set_implsch_form(self, three_kernels):
        """IMPLSCH in three kernels (True), one kernel (False) or the automatic
        choice (None); the results are bit-identical."""
        form = -1 if three_kernels is None else int(bool(three_kernels))
        self._check(self.lib.wam_gpu_set_implsch_form(self.ctx, form))

    def set_source_options(self, wave_breaking=False, phillips=False):
        """SET_MODEL_OPTION's WAVE_BREAKING (SDEPTHLIM and SFBRK; stays off in
        deep water) and PHILLIPS (SOURCE_PHILLIPS) switches, off by default;
        they apply from the next IMPLSCH call.  Returns the switches as set."""
        self._check(self.lib.wam_gpu_set_source_options(self.ctx, int(bool(wave_breaking)), int(bool(phillips))))
        b, p = ctypes.c_int32(), ctypes.c_int32()
        self._check(self.lib.wam_gpu_get_source_options(self.ctx, ctypes.byref(b), ctypes.byref(p)))
        return bool(b.value), bool(p.value)

    def prepare_propagation(self, depth, u=None, v=None):
        d = np.ascontiguousarray(depth, dtype=np.float32)
        uu = None if u is None else np.ascontiguousarray(u, dtype=np.float32)
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float32)
        for a in (d, uu, vv):
            if a is not None and a.size != self.nsea:
                raise ValueError("per-point arrays must have NSEA elements")
        self._check(self.lib.wam_gpu_prepare_propagation(self.ctx, _ptr(d), _ptr(uu), _ptr(vv)))
        return int(self.scalar("COUNTER")), float(self.scalar("CFL_MAX"))

    @_stream_ordered
    def prepare_propagation_dev(self, depth, u=None, v=None):
        """PREPARE_PROPAGATION from fields already in HBM (torch tensors of
        NSEA float32 on the context's device, or device pointers as ints)."""
        def dptr(x):
            if x is None:
                return None
            if isinstance(x, int):
                return x
            if x.dtype.itemsize != 4 or x.numel() != self.nsea or not x.is_contiguous():
                raise ValueError("per-point fields must be contiguous float32 tensors of NSEA elements")
            return x.data_ptr()
        self._check(self.lib.wam_gpu_prepare_propagation_dev(self.ctx, dptr(depth), dptr(u), dptr(v)))
        return int(self.scalar("COUNTER")), float(self.scalar("CFL_MAX"))

    def table(self, name):
        n = self.lib.wam_gpu_get_table(self.ctx, name.encode(), None, 0)
        if n < 0:
            raise KeyError(name)
        ints = name in ("IKP", "IKP1", "IKM", "IKM1", "K1W", "K2W", "K11W", "K21W", "INLCOEF", "IXLG", "KXLT",
                        "INDEP", "INDEP_G", "KLAT", "KLON", "NCONTRIB")
        out = np.zeros(n, np.int32 if ints else np.float32)
        self.lib.wam_gpu_get_table(self.ctx, name.encode(), _ptr(out), n)
        return out

    def scalar(self, name):
        return self.lib.wam_gpu_get_scalar(self.ctx, name.encode())

    NOUT = 17
    OUT_NAMES = ("u10", "udir", "ustar", "cd", "charnock", "depth", "current_speed", "current_dir", "hs", "tp", "tm",
                 "tm1", "tm2", "mdir", "spread", "tauwn", "z0")

    def output_parameters(self, fl3, u10, udir, ustar, tauw, z0, orient=1):
        """Integrated output parameters (host arrays): [NOUT][NSEA] in the
        order of OUT_NAMES (WAM_OUT_* of include/wam_gpu.h)."""
        n = self.nsea
        a = [np.ascontiguousarray(x, dtype=np.float32) for x in (fl3, u10, udir, ustar, tauw, z0)]
        out = np.zeros((self.NOUT, n), np.float32)
        self._check(self.lib.wam_gpu_output_parameters(self.ctx, n, self.cfg["kl"], self.cfg["ml"],
                                                       *[_ptr(x) for x in a], int(orient), _ptr(out)))
        return out

    @_stream_ordered
    def output_parameters_dev(self, fl3, u10, udir, ustar, tauw, z0, out, orient=1):
        """Device-resident form on torch CUDA tensors; out holds NOUT * NSEA floats."""
        p = lambda x: ctypes.c_void_p(x.data_ptr())
        self._check(self.lib.wam_gpu_output_parameters_dev(self.ctx, p(fl3), p(u10), p(udir), p(ustar), p(tauw), p(z0),
                                                           int(orient), p(out)))

    @_stream_ordered
    def save_restart(self, path, dates, fl3, u10, udir, tauw):
        """SAVE_RESTART_FILE from device tensors (wam_gpu_save_restart); the
        depth and current records are the context's prepared fields."""
        self._check(self.lib.wam_gpu_save_restart(self.ctx, os.fsencode(path), _dates(dates), _ptr(fl3), _ptr(u10),
                                                  _ptr(udir), _ptr(tauw)))

    @_stream_ordered
    def connect_restart(self, path, fl3, u10, udir, tauw, depth=None, u=None, v=None):
        """CONNECT_RESTART into device tensors (wam_gpu_connect_restart);
        returns the dates and which optional records the file held."""
        raw = ctypes.create_string_buffer(70)
        present = ctypes.c_int32(0)
        self._check(self.lib.wam_gpu_connect_restart(self.ctx, os.fsencode(path), raw, _ptr(fl3), _ptr(u10),
                                                     _ptr(udir), _ptr(tauw), _ptr(depth), _ptr(u), _ptr(v),
                                                     ctypes.byref(present)))
        return _undates(raw.raw), dict(depth=bool(present.value & 1), current=bool(present.value & 2))

    @_stream_ordered
    def boundary_output(self, fl3, ijarc):
        """BOUNDARY_OUTPUT's gather and mean parameters at the 1-based sea
        points ijarc: FL3 as a host array [ML][KL][NSEA] or a device tensor;
        returns fbc [nb][ML][KL] and mean [3][nb] (EMEAN, THQ, FMEAN)."""
        ij = np.ascontiguousarray(ijarc, np.int32)
        nb = ij.size
        fbc = np.empty((nb, self.ml, self.kl), np.float32)
        mean = np.empty((3, nb), np.float32)
        if isinstance(fl3, np.ndarray):
            f = np.ascontiguousarray(fl3, np.float32)
            self._check(self.lib.wam_gpu_boundary_output(self.ctx, self.nsea, self.kl, self.ml, _ptr(f), nb, _ptr(ij),
                                                         _ptr(fbc), _ptr(mean)))
        else:
            self._check(self.lib.wam_gpu_boundary_output_dev(self.ctx, nb, _ptr(ij), _ptr(fl3), _ptr(fbc),
                                                             _ptr(mean)))
        return fbc, mean

    @_stream_ordered
    def boundary_input(self, fl3, f1, m1, f2, m2, which, del12, del1l, ijarf, ibfl, ibfr, bfw, return_set=True):
        """BOUNDARY_INPUT's time and space interpolation into FL3 (a host
        array [ML][KL][NSEA], updated in place, or a device tensor): which = 1
        takes (f1, m1), 2 takes (f2, m2), 0 interpolates at del1l of del12;
        f* are [nbinp][ML][KL], m* [3][nbinp] (EMEAN, THQ, FMEAN).  Returns
        the set at the model time (FI, MI), or nothing with return_set=False
        (no copy back)."""
        a = lambda x: None if x is None else np.ascontiguousarray(x, np.float32)
        i = lambda x: np.ascontiguousarray(x, np.int32)
        f1, m1, f2, m2, w = a(f1), a(m1), a(f2), a(m2), a(bfw)
        ij, il, ir = i(ijarf), i(ibfl), i(ibfr)
        nbinp = (m1 if m1 is not None else m2).shape[-1]
        fi = np.empty((nbinp, self.ml, self.kl), np.float32) if return_set else None
        mi = np.empty((3, nbinp), np.float32) if return_set else None
        args = [_ptr(f1), _ptr(m1), _ptr(f2), _ptr(m2), int(which), float(del12), float(del1l), ij.size, _ptr(ij),
                _ptr(il), _ptr(ir), _ptr(w)]
        if isinstance(fl3, np.ndarray):
            if fl3.dtype != np.float32 or not fl3.flags.c_contiguous:
                raise ValueError("fl3 must be a C-contiguous float32 array [ML][KL][NSEA]")
            self._check(self.lib.wam_gpu_boundary_input(self.ctx, self.nsea, self.kl, self.ml, _ptr(fl3), nbinp,
                                                        *args, _ptr(fi), _ptr(mi)))
        else:
            self._check(self.lib.wam_gpu_boundary_input_dev(self.ctx, nbinp, *args, _ptr(fl3), _ptr(fi), _ptr(mi)))
        return (fi, mi) if return_set else None

    @_stream_ordered
    def output_grids_dev(self, params, out, grids):
        """The gridded fields of the parameter numbers params (1-16, 32) from
        out (device tensor, [NOUT][NSEA]) into grids ([len(params)][NY][NX])."""
        self._check(self.lib.wam_gpu_output_grids_dev(self.ctx, _ptr(_fflag(params)), _ptr(out), _ptr(grids)))

    @_stream_ordered
    def write_map(self, path, cdtpro, cdatea, params, out, append=False):
        """One output time of the MAP file from the device parameters out
        ([NOUT][NSEA], output_parameters_dev) for the parameter numbers params."""
        self._check(self.lib.wam_gpu_write_map(self.ctx, os.fsencode(path), int(bool(append)), _date14(cdtpro),
                                               _date14(cdatea), _ptr(_fflag(params)), _ptr(out)))

    def set_timing(self, on):
        """Records a HIP event pair around every PROPAGS / IMPLSCH kernel
        launch from now on (on=True; earlier records are dropped) or stops."""
        self._check(self.lib.wam_gpu_set_timing(self.ctx, int(bool(on))))

    def kernel_time(self, name):
        """(total GPU ms, launches) of one kernel since set_timing(True):
        "propags", "implsch_pre", "implsch_snl", "implsch_post", "implsch",
        "boundary_gather", "boundary_time" or "boundary_space"."""
        ms = ctypes.c_double(0.0)
        n = ctypes.c_int64(0)
        self._check(self.lib.wam_gpu_get_timing(self.ctx, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def set_stream(self, stream_handle):
        self._check(self.lib.wam_gpu_set_stream(self.ctx, ctypes.c_void_p(stream_handle)))

    def sync(self):
        self._check(self.lib.wam_gpu_sync(self.ctx))

    # -- drop-in (host memory) ---------------------------------------------
    def propags(self, fl3):
        f = np.ascontiguousarray(fl3, dtype=np.float32).copy()
        self._check(self.lib.wam_gpu_propags(self.ctx, 1, self.nsea, self.kl, self.ml, _ptr(f)))
        return f

    def implsch(self, fl3, u10, udir, tauw, z0, depth=None, indep=None, roairn=None, wstar=None, lcflx=False):
        n = self.nsea
        f = np.ascontiguousarray(fl3, dtype=np.float32).copy()
        tw = np.ascontiguousarray(tauw, dtype=np.float32).copy()
        z = np.ascontiguousarray(z0, dtype=np.float32).copy()
        us = np.zeros(n, np.float32)
        ph = np.zeros(n, np.float32)
        fx = np.zeros((6, n), np.float32)
        ro = np.full(n, 1.225, np.float32) if roairn is None else np.ascontiguousarray(roairn, np.float32)
        ws = np.zeros(n, np.float32) if wstar is None else np.ascontiguousarray(wstar, np.float32)
        u = np.ascontiguousarray(u10, np.float32)
        d = np.ascontiguousarray(udir, np.float32)
        if f.size != n * self.kl * self.ml:
            raise ValueError("FL3 must have NSEA*KL*ML elements")
        # DEPTH and INDEP, when given, must be the fields of the last
        # prepare_propagation (the library checks them, WAM_E_STATE otherwise)
        dp = None if depth is None else np.ascontiguousarray(depth, np.float32)
        ip = None if indep is None else np.ascontiguousarray(indep, np.int32)
        self._check(self.lib.wam_gpu_implsch(self.ctx, n, self.kl, self.ml, _ptr(f), _ptr(u), _ptr(d), _ptr(tw),
                                             _ptr(us), _ptr(z), _ptr(ro), _ptr(ws), _ptr(dp), _ptr(ip),
                                             int(bool(lcflx)), _ptr(ph), _ptr(fx)))
        return dict(fl3=f, tauw=tw, ustar=us, z0=z, phiaw=ph, flux=fx)

    # -- device resident (torch tensors in HBM) ----------------------------
    @_stream_ordered
    def propags_dev(self, d_in, d_out):
        self._check(self.lib.wam_gpu_propags_dev(self.ctx, _ptr(d_in), _ptr(d_out)))

    @_stream_ordered
    def implsch_dev(self, fl3, u10, udir, tauw, ustar, z0, roairn, wstar, lcflx=False, phiaw=None, fluxes=None,
                    mij=None):
        self._check(self.lib.wam_gpu_implsch_dev(self.ctx, _ptr(fl3), _ptr(u10), _ptr(udir), _ptr(tauw),
                                                 _ptr(ustar), _ptr(z0), _ptr(roairn), _ptr(wstar),
                                                 int(bool(lcflx)), _ptr(phiaw), _ptr(fluxes), _ptr(mij)))

    # -- ensembles: NMEM members on this grid, member arrays concatenated --
    @_stream_ordered
    def propags_ensemble_dev(self, nmem, d_in, d_out):
        self._check(self.lib.wam_gpu_propags_ensemble_dev(self.ctx, int(nmem), _ptr(d_in), _ptr(d_out)))

    @_stream_ordered
    def implsch_ensemble_dev(self, nmem, fl3, u10, udir, tauw, ustar, z0, roairn, wstar, lcflx=False, phiaw=None,
                             fluxes=None, mij=None):
        self._check(self.lib.wam_gpu_implsch_ensemble_dev(self.ctx, int(nmem), _ptr(fl3), _ptr(u10), _ptr(udir),
                                                          _ptr(tauw), _ptr(ustar), _ptr(z0), _ptr(roairn),
                                                          _ptr(wstar), int(bool(lcflx)), _ptr(phiaw), _ptr(fluxes),
                                                          _ptr(mij)))

    @_stream_ordered
    def ensemble_run_dev(self, nmem, nsteps, fa, fb, u10, udir, tauw, ustar, z0, roairn, wstar, graph=True):
        """nsteps PROPAGS + IMPLSCH steps of the ensemble; returns the buffer
        holding the result (fa after an even number of steps, else fb)."""
        self._check(self.lib.wam_gpu_ensemble_run_dev(self.ctx, int(nmem), int(nsteps), _ptr(fa), _ptr(fb),
                                                      _ptr(u10), _ptr(udir), _ptr(tauw), _ptr(ustar), _ptr(z0),
                                                      _ptr(roairn), _ptr(wstar), int(bool(graph))))
        return fa if nsteps % 2 == 0 else fb

    @_stream_ordered
    def autotune(self, nmem, fa, fb, u10, udir, tauw, z0, roairn, wstar, reps=2):
        self._check(self.lib.wam_gpu_autotune(self.ctx, int(nmem), _ptr(fa), _ptr(fb), _ptr(u10), _ptr(udir),
                                              _ptr(tauw), _ptr(z0), _ptr(roairn), _ptr(wstar), int(reps)))
        return self.tuning()

    def tuning(self):
        """(PROPAGS workgroup width, per-lane IMPLSCH width, SNONLIN workgroups per CU, member-points per chunk)."""
        out = np.zeros(4, np.int64)
        self._check(self.lib.wam_gpu_get_tuning(self.ctx, _ptr(out)))
        return tuple(int(v) for v in out)

    def set_tuning(self, prop_wg=256, src_wg=64, snl_per_cu=0, chunk=0):
        self._check(self.lib.wam_gpu_set_tuning(self.ctx, _ptr(np.array([prop_wg, src_wg, snl_per_cu, chunk],
                                                                         np.int64))))

    def tuning_log(self):
        n = self.lib.wam_gpu_get_tuning_log(self.ctx, None, 0)
        buf = ctypes.create_string_buffer(int(n) + 1)
        self.lib.wam_gpu_get_tuning_log(self.ctx, buf, n + 1)
        return [(w, int(v), float(ms)) for w, v, ms in (l.split() for l in buf.value.decode().splitlines())]

    @_stream_ordered
    def implsch_dev_noalloc(self, fl3, u10, udir, tauw, ustar, z0, roairn, wstar, lcflx=False, phiaw=None,
                            fluxes=None, mij=None):
        """implsch_dev on the library's static device scratch array: allocates nothing."""
        self._check(self.lib.wam_gpu_implsch_dev_noalloc(self.ctx, _ptr(fl3), _ptr(u10), _ptr(udir), _ptr(tauw),
                                                         _ptr(ustar), _ptr(z0), _ptr(roairn), _ptr(wstar),
                                                         int(bool(lcflx)), _ptr(phiaw), _ptr(fluxes), _ptr(mij)))

    @_stream_ordered
    def step_adjoint_dev(self, fl3, u10, udir, tauw, roairn, wstar, fl3_bar, fl3_bar_out, u10_bar, udir_bar,
                         tauw_bar_out, tauw_bar=None, ustar_bar=None, z0_bar=None, nmem=1):
        """Adjoint of one PROPAGS + IMPLSCH step at the input state (fl3, tauw)
        with its forcing: from the bars of the step's outputs (fl3_bar and the
        optional tauw_bar, ustar_bar, z0_bar) into the bars of its inputs
        (fl3_bar_out, which may be fl3_bar, u10_bar, udir_bar, tauw_bar_out)."""
        self._check(self.lib.wam_gpu_step_adjoint_dev(
            self.ctx, int(nmem), _ptr(fl3), _ptr(u10), _ptr(udir), _ptr(tauw), _ptr(roairn), _ptr(wstar),
            _ptr(fl3_bar), _ptr(tauw_bar), _ptr(ustar_bar), _ptr(z0_bar), _ptr(fl3_bar_out), _ptr(u10_bar),
            _ptr(udir_bar), _ptr(tauw_bar_out)))

    @_stream_ordered
    def propags_adjoint_dev(self, d_in_bar, d_out_bar, nmem=1):
        """Adjoint of PROPAGS: the transposed stencil, COUNTER sub-steps."""
        self._check(self.lib.wam_gpu_propags_adjoint_dev(self.ctx, int(nmem), _ptr(d_in_bar), _ptr(d_out_bar)))

    @_stream_ordered
    def wind_uv_dev(self, u, v, u10, udir):
        """U10 = SQRT(U**2 + V**2), UDIR = ATAN2(U, V) (+ ZPI where negative)."""
        self._check(self.lib.wam_gpu_wind_uv_dev(self.ctx, _ptr(u), _ptr(v), _ptr(u10), _ptr(udir), int(u.numel())))

    @_stream_ordered
    def wind_uv_adjoint_dev(self, u, v, u10_bar, udir_bar, u_bar, v_bar):
        self._check(self.lib.wam_gpu_wind_uv_adjoint_dev(self.ctx, _ptr(u), _ptr(v), _ptr(u10_bar), _ptr(udir_bar),
                                                         _ptr(u_bar), _ptr(v_bar), int(u.numel())))


def checkpoint_schedule(nsteps, nsnap):
    """The binomial checkpointing schedule of wam_checkpoint_schedule: a list
    of (op, a, b) with op one of CK_ADVANCE, CK_STORE, CK_RESTORE, CK_ADJOINT."""
    lib = _lib.load()
    n = lib.wam_checkpoint_schedule(int(nsteps), int(nsnap), None, 0)
    if n < 0:
        raise WamError(-n, "bad checkpointing request")
    buf = np.zeros(3 * n, np.int32)
    lib.wam_checkpoint_schedule(int(nsteps), int(nsnap), _ptr(buf), n)
    return [tuple(int(x) for x in buf[3 * i:3 * i + 3]) for i in range(n)]


def checkpoint_recomputed_steps(nsteps, nsnap):
    n = _lib.load().wam_checkpoint_recomputed_steps(int(nsteps), int(nsnap))
    if n < 0:
        raise WamError(-n, "bad checkpointing request")
    return int(n)


CK_ADVANCE, CK_STORE, CK_RESTORE, CK_ADJOINT = 0, 1, 2, 3


class _GridState(ctypes.Structure):
    """struct wam_grid_state of include/wam_gpu.h."""
    _fields_ = [("ctx", ctypes.c_void_p), ("d_fa", ctypes.c_void_p), ("d_fb", ctypes.c_void_p),
                ("d_u10", ctypes.c_void_p), ("d_udir", ctypes.c_void_p), ("d_roairn", ctypes.c_void_p),
                ("d_wstar", ctypes.c_void_p), ("d_tauw", ctypes.c_void_p), ("d_ustar", ctypes.c_void_p),
                ("d_z0", ctypes.c_void_p)]


class WamMultiGrid:
    """Several grids advanced together with nests (wam_gpu_multigrid_*).
    grids: list of (WamGpu, dict of device tensors fa, fb, u10, udir, roairn,
    wstar, tauw, ustar, z0)."""

    def __init__(self, grids):
        self.lib = _lib.load()
        self._keep = grids
        arr = (_GridState * len(grids))()
        for i, (g, st) in enumerate(grids):
            arr[i] = _GridState(g.ctx, *[st[k].data_ptr() for k in ("fa", "fb", "u10", "udir", "roairn", "wstar",
                                                                      "tauw", "ustar", "z0")])
        h = ctypes.c_void_p()
        rc = self.lib.wam_gpu_multigrid_create(ctypes.byref(h), len(grids), arr)
        self.mg = h.value
        if rc:
            msg = self.lib.wam_gpu_multigrid_last_error(self.mg) if self.mg else b""
            self.close()
            raise WamError(rc, (msg or b"").decode())

    def _check(self, rc):
        if rc:
            raise WamError(rc, self.lib.wam_gpu_multigrid_last_error(self.mg).decode())

    def add_nest(self, coarse, fine, ijarc, ijarf, ibfl, ibfr, bfw):
        a = [np.ascontiguousarray(x, np.int32) for x in (ijarc, ijarf, ibfl, ibfr)]
        w = np.ascontiguousarray(bfw, np.float32)
        self._check(self.lib.wam_gpu_multigrid_add_nest(self.mg, int(coarse), int(fine), a[0].size, _ptr(a[0]),
                                                        a[1].size, _ptr(a[1]), _ptr(a[2]), _ptr(a[3]), _ptr(w)))

    def run(self, nsteps, fused=True):
        self._check(self.lib.wam_gpu_multigrid_run(self.mg, int(nsteps), int(bool(fused))))

    def state(self, grid):
        """'fa' or 'fb': the tensor of grid i that holds its state."""
        p = ctypes.c_void_p()
        self._check(self.lib.wam_gpu_multigrid_state(self.mg, int(grid), ctypes.byref(p)))
        st = self._keep[grid][1]
        return "fa" if p.value == st["fa"].data_ptr() else "fb"

    def close(self):
        if getattr(self, "mg", None):
            self.lib.wam_gpu_multigrid_free(self.mg)
            self.mg = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
