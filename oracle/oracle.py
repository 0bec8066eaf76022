"""ORACLE -- test infrastructure only.

ctypes wrapper of the CPU restatement in this directory (liborc32.so for
REAL*4, liborc64.so for REAL*8).  Only tests/, the smoke check and bench.py's
cpu_baseline leg may import this module; the product package never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("kl", ctypes.c_int), ("ml", ctypes.c_int), ("fr1", ctypes.c_double), ("iref", ctypes.c_int),
        ("iphys", ctypes.c_int), ("shallow", ctypes.c_int), ("spherical", ctypes.c_int),
        ("refr_d", ctypes.c_int), ("refr_c", ctypes.c_int), ("iper", ctypes.c_int),
        ("isnonlin", ctypes.c_int), ("idelpro", ctypes.c_int), ("idelt", ctypes.c_int),
        ("ndepth", ctypes.c_int), ("deptha", ctypes.c_double), ("depthd", ctypes.c_double),
        ("betamax", ctypes.c_double), ("zalp", ctypes.c_double), ("alpha", ctypes.c_double),
        ("cdfac", ctypes.c_double),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("amosop", ctypes.c_int64), ("amonop", ctypes.c_int64), ("amowep", ctypes.c_int64),
        ("amoeap", ctypes.c_int64), ("xdella", ctypes.c_int64), ("xdello", ctypes.c_int64),
    ]


def make_config(cfg):
    """cfg: dict with the model options (see wam_amd.config.ModelConfig.as_dict)."""
    c = OrcConfig()
    for name, _ in OrcConfig._fields_:
        setattr(c, name, cfg[name])
    return c


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_LIBS = {}


def _lib(fp64, opcount=False, ad=False):
    if ad:
        key = "liborc_ad.so" if fp64 else "liborc_ad32.so"
    else:
        key = "liborc_opcount.so" if opcount else ("liborc64.so" if fp64 else "liborc32.so")
    if key not in _LIBS:
        path = os.path.join(_HERE, key)
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        lib.orc_api_create.restype = ctypes.c_void_p
        lib.orc_api_create.argtypes = [ctypes.POINTER(OrcConfig), ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        lib.orc_api_free.argtypes = [ctypes.c_void_p]
        lib.orc_api_get.restype = ctypes.c_void_p
        lib.orc_api_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_long),
                                    ctypes.POINTER(ctypes.c_int)]
        lib.orc_api_scalar.restype = ctypes.c_double
        lib.orc_api_scalar.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        lib.orc_api_prepare_propagation.restype = ctypes.c_int
        lib.orc_api_prepare_propagation.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 3
        lib.orc_api_propags.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        lib.orc_api_force_counter.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.orc_api_source_options.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        lib.orc_api_implsch.argtypes = ([ctypes.c_void_p] + [ctypes.c_void_p] * 9 +
                                        [ctypes.c_int] + [ctypes.c_void_p] * 3 + [ctypes.c_int])
        lib.orc_api_wavnu2.argtypes = [ctypes.c_void_p] * 4
        lib.orc_api_output_parameters.argtypes = [ctypes.c_void_p] * 10 + [ctypes.c_int, ctypes.c_void_p]
        lib.orc_api_snonlin.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_double, ctypes.c_double]
        lib.orc_api_rotspec.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_double]
        lib.orc_api_strspec.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_double]
        lib.orc_api_intspec.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double] + [ctypes.c_void_p] * 6
        lib.orc_api_boundary_output.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 3
        lib.orc_api_boundary_input.argtypes = ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 4 +
                                               [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int] +
                                               [ctypes.c_void_p] * 6)
        lib.orc_api_math.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
        lib.orc_api_aki.restype = ctypes.c_double
        lib.orc_api_aki.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        if ad:
            lib.orc_ad_step.restype = ctypes.c_long
            lib.orc_ad_step.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 11
        if opcount:
            lib.orc_opcount_implsch.argtypes = ([ctypes.c_void_p] + [ctypes.c_void_p] * 9 +
                                                [ctypes.c_int] + [ctypes.c_void_p] * 5)
        _LIBS[key] = lib
    return _LIBS[key]


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """One model configuration on one regular grid (sea mask nx*ny, row-major from the south-west)."""

    def __init__(self, cfg, mask, fp64=False, opcount=False, ad=False, ad32=False):
        """ad=True: the FP64 restatement together with one model step's
        tangent-linear model and adjoint (step_ad); ad32=True: the same with
        the REAL*4 restatement's values and branches (derivatives in FP64)."""
        self.cfg = dict(cfg)
        ad = ad or ad32
        self.fp64 = (fp64 or ad) and not opcount and not ad32
        self.dtype = np.float64 if self.fp64 else np.float32
        self.opcount = opcount
        self.ad = ad
        self.lib = _lib(self.fp64, opcount, ad)
        self._c = make_config(cfg)
        m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
        st = ctypes.c_int(0)
        self.ctx = self.lib.orc_api_create(ctypes.byref(self._c), _p(m), ctypes.byref(st))
        if st.value != 0:
            raise ValueError(f"oracle create failed with status {st.value}")
        self.kl, self.ml = cfg["kl"], cfg["ml"]
        self.nsea = int(self.scalar("NSEA"))
        # the off-by-default sources: WAVE_BREAKING_RUN (SDEPTHLIM, SFBRK) and PHILLIPS_RUN
        if cfg.get("wave_breaking") or cfg.get("phillips"):
            self.lib.orc_api_source_options(self.ctx, int(bool(cfg.get("wave_breaking"))),
                                            int(bool(cfg.get("phillips"))))

    def __del__(self):
        if getattr(self, "ctx", None):
            self.lib.orc_api_free(self.ctx)
            self.ctx = None

    def get(self, name):
        n = ctypes.c_long(0)
        t = ctypes.c_int(0)
        p = self.lib.orc_api_get(self.ctx, name.encode(), ctypes.byref(n), ctypes.byref(t))
        if not p or n.value == 0:
            raise KeyError(name)
        ct = ctypes.c_int if t.value == 1 else (ctypes.c_double if self.fp64 else ctypes.c_float)
        arr = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=(n.value,))
        return arr.copy()

    def scalar(self, name):
        return self.lib.orc_api_scalar(self.ctx, name.encode())

    def prepare_propagation(self, depth, u=None, v=None):
        d = np.ascontiguousarray(depth, dtype=self.dtype)
        uu = None if u is None else np.ascontiguousarray(u, dtype=self.dtype)
        vv = None if v is None else np.ascontiguousarray(v, dtype=self.dtype)
        st = self.lib.orc_api_prepare_propagation(self.ctx, _p(d), _p(uu), _p(vv))
        if st != 0:
            raise ValueError(f"oracle prepare_propagation status {st}")
        return int(self.scalar("COUNTER")), float(self.scalar("CFL_MAX"))

    def step_ad(self, fl3, tauw, z0, u, v, depth, roairn=None, wstar=None, dx=None, ybar=None):
        """One model step, PROPAGS then IMPLSCH, in FP64 (or REAL*4) from the wind
        components u, v (U10 = sqrt(u^2 + v^2), UDIR = atan2(u, v)):
        returns y = (FL3', TAUW', Z0') packed, and with dx (packed as
        x = (FL3, TAUW, Z0, U, V)) also the tangent-linear dy, or with ybar
        (packed as y) also the adjoint xbar."""
        assert self.ad
        n = self.nsea
        d = lambda a: np.ascontiguousarray(a, np.float64)
        ro = np.full(n, 1.225) if roairn is None else d(roairn)
        ws = np.zeros(n) if wstar is None else d(wstar)
        nb = self.kl * self.ml * n
        y = np.zeros(nb + 2 * n)
        mode, arg, out = 0, None, None
        if dx is not None:
            mode, arg, out = 1, d(dx), np.zeros(nb + 2 * n)
        elif ybar is not None:
            mode, arg, out = 2, d(ybar), np.zeros(nb + 4 * n)
        self.lib.orc_ad_step(self.ctx, mode, _p(d(fl3)), _p(d(tauw)), _p(d(z0)), _p(d(u)), _p(d(v)), _p(ro), _p(ws),
                             _p(d(depth)), _p(arg), _p(y), _p(out))
        return (y, out) if mode else y

    def force_counter(self, counter):
        """For a window of a larger grid: CHECK_CFL's sub-step count is that
        grid's (the weights rescaled by it), from the next prepare_propagation."""
        self.lib.orc_api_force_counter(self.ctx, int(counter))

    def propags(self, fl3):
        f = np.array(fl3, dtype=self.dtype, copy=True).ravel()
        self.lib.orc_api_propags(self.ctx, _p(f))
        return f

    def implsch(self, fl3, u10, udir, tauw, z0, depth, roairn=None, wstar=None, lcflx=False, nthreads=0):
        n = self.nsea
        dt = self.dtype
        f = np.ascontiguousarray(fl3, dtype=dt).copy()
        u10 = np.ascontiguousarray(u10, dtype=dt)
        udir = np.ascontiguousarray(udir, dtype=dt)
        tauw = np.ascontiguousarray(tauw, dtype=dt).copy()
        z0 = np.ascontiguousarray(z0, dtype=dt).copy()
        depth = np.ascontiguousarray(depth, dtype=dt)
        roairn = np.full(n, 1.225, dt) if roairn is None else np.ascontiguousarray(roairn, dtype=dt)
        wstar = np.zeros(n, dt) if wstar is None else np.ascontiguousarray(wstar, dtype=dt)
        ustar = np.zeros(n, dt)
        phiaw = np.zeros(n, dt)
        flux = np.zeros((6, n), dt)
        mij = np.zeros(n, np.int32)
        self.lib.orc_api_implsch(self.ctx, _p(f), _p(u10), _p(udir), _p(tauw), _p(ustar), _p(z0),
                                 _p(roairn), _p(wstar), _p(depth), int(bool(lcflx)), _p(phiaw),
                                 _p(flux), _p(mij), int(nthreads))
        return dict(fl3=f, tauw=tauw, ustar=ustar, z0=z0, phiaw=phiaw, flux=flux, mij=mij)

    def implsch_opcount(self, fl3, u10, udir, tauw, z0, depth, roairn=None, wstar=None, lcflx=False):
        """IMPLSCH of the REAL*4 restatement with every floating-point
        operation counted (orc_opcount.h), on one thread.  Returns the same
        fields as implsch() and the totals over all sea points:
        add, mul, div, sqrt, trans (elementary-function calls), with the same
        per stage in "stages" (0: up to SDISSIP; 1: SNONLIN, SBOTTOM,
        WNFLUXES and the limited update; 2: new mean parameters and tail)."""
        if not self.opcount:
            raise ValueError("construct the Oracle with opcount=True")
        n = self.nsea
        dt = np.float32
        f = np.ascontiguousarray(fl3, dtype=dt).copy()
        u10 = np.ascontiguousarray(u10, dtype=dt)
        udir = np.ascontiguousarray(udir, dtype=dt)
        tauw = np.ascontiguousarray(tauw, dtype=dt).copy()
        z0 = np.ascontiguousarray(z0, dtype=dt).copy()
        depth = np.ascontiguousarray(depth, dtype=dt)
        roairn = np.full(n, 1.225, dt) if roairn is None else np.ascontiguousarray(roairn, dtype=dt)
        wstar = np.zeros(n, dt) if wstar is None else np.ascontiguousarray(wstar, dtype=dt)
        ustar = np.zeros(n, dt)
        phiaw = np.zeros(n, dt)
        flux = np.zeros((6, n), dt)
        mij = np.zeros(n, np.int32)
        counts = np.zeros(5, np.uint64)
        stages = np.zeros((3, 5), np.uint64)
        self.lib.orc_opcount_implsch(self.ctx, _p(f), _p(u10), _p(udir), _p(tauw), _p(ustar), _p(z0),
                                     _p(roairn), _p(wstar), _p(depth), int(bool(lcflx)), _p(phiaw),
                                     _p(flux), _p(mij), _p(counts), _p(stages))
        res = dict(fl3=f, tauw=tauw, ustar=ustar, z0=z0, phiaw=phiaw, flux=flux, mij=mij)
        names = ["add", "mul", "div", "sqrt", "trans"]
        tot = dict(zip(names, (int(x) for x in counts)))
        tot["stages"] = [dict(zip(names, (int(x) for x in row))) for row in stages]
        return res, tot

    def output_parameters(self, fl3, u10, udir, ustar, tauw, z0, depth, u=None, v=None, orient=1):
        """COMPUTE_OUTPUT_PARAMETER's integrated parameters (orc_output.c):
        [17][NSEA] in the order of the WAM_OUT_* indices of include/wam_gpu.h."""
        n = self.nsea
        a = lambda x: None if x is None else np.ascontiguousarray(x, dtype=self.dtype)
        args = [a(x) for x in (fl3, u10, udir, ustar, tauw, z0, depth, u, v)]   # alive across the call
        out = np.zeros((17, n), self.dtype)
        self.lib.orc_api_output_parameters(self.ctx, *[_p(x) for x in args], int(orient), _p(out))
        return out

    def rotspec(self, f, rthet):
        """ROTSPEC of F[ML][KL] by rthet radians."""
        f = np.ascontiguousarray(f, dtype=self.dtype)
        out = np.zeros_like(f)
        self.lib.orc_api_rotspec(self.ctx, _p(f), _p(out), float(rthet))
        return out

    def strspec(self, f, gamma):
        """STRSPEC of F[ML][KL] by the factor gamma."""
        f = np.ascontiguousarray(f, dtype=self.dtype)
        out = np.zeros_like(f)
        self.lib.orc_api_strspec(self.ctx, _p(f), _p(out), float(gamma))
        return out

    def intspec(self, del12, del1l, f1, m1, f2, m2):
        """INTSPEC (orc_nest.c) of two spectra F[ML][KL] with their (EMEAN,
        THQ, FMEAN); returns the spectrum and its (EMEAN, THQ, FMEAN)."""
        a = lambda x: np.ascontiguousarray(x, dtype=self.dtype)
        f1, m1, f2, m2 = a(f1), a(m1), a(f2), a(m2)
        fl = np.zeros_like(f1)
        mo = np.zeros(3, self.dtype)
        self.lib.orc_api_intspec(self.ctx, float(del12), float(del1l), _p(f1), _p(m1), _p(f2), _p(m2), _p(fl), _p(mo))
        return fl, mo

    def boundary_output(self, fl3, ijarc):
        """BOUNDARY_OUTPUT's spectra and mean parameters at the 1-based sea
        points ijarc: FBC[nb][ML][KL] and [3][nb] (EMEAN, THQ, FMEAN)."""
        fl3 = np.ascontiguousarray(fl3, dtype=self.dtype)
        ij = np.ascontiguousarray(ijarc, dtype=np.int32)
        nb = ij.size
        fbc = np.zeros((nb, self.ml, self.kl), self.dtype)
        mean = np.zeros((3, nb), self.dtype)
        self.lib.orc_api_boundary_output(self.ctx, _p(fl3), nb, _p(ij), _p(fbc), _p(mean))
        return fbc, mean

    def boundary_input(self, fl3, f1, m1, f2, m2, which, del12, del1l, ijarf, ibfl, ibfr, bfw, inplace=False):
        """BOUNDARY_INPUT's time then space interpolation into a copy of FL3
        [ML][KL][NSEA] (into FL3 itself with inplace=True); returns (FL3,
        FI[nbinp][ML][KL], MI[3][nbinp])."""
        a = lambda x: np.ascontiguousarray(x, dtype=self.dtype)
        i = lambda x: np.ascontiguousarray(x, dtype=np.int32)
        if inplace:
            if fl3.dtype != self.dtype or not fl3.flags.c_contiguous:
                raise ValueError("in place: FL3 must be C-contiguous in the oracle's precision")
            out = fl3
        else:
            out = np.array(fl3, dtype=self.dtype, copy=True, order="C")
        f1, m1, f2, m2, bfw = a(f1), a(m1), a(f2), a(m2), a(bfw)
        ijarf, ibfl, ibfr = i(ijarf), i(ibfl), i(ibfr)
        nbinp = m1.shape[-1]
        fi = np.zeros((nbinp, self.ml, self.kl), self.dtype)
        mi = np.zeros((3, nbinp), self.dtype)
        self.lib.orc_api_boundary_input(self.ctx, _p(out), nbinp, _p(f1), _p(m1), _p(f2), _p(m2), int(which),
                                        float(del12), float(del1l), ijarf.size, _p(ijarf), _p(ibfl), _p(ibfr),
                                        _p(bfw), _p(fi), _p(mi))
        return out, fi, mi

    def wavnu2(self, depth):
        d = np.ascontiguousarray(depth, dtype=self.dtype)
        wn = np.zeros(self.ml * self.nsea, self.dtype)
        cg = np.zeros(self.ml * self.nsea, self.dtype)
        self.lib.orc_api_wavnu2(self.ctx, _p(d), _p(wn), _p(cg))
        return wn.reshape(self.ml, self.nsea), cg.reshape(self.ml, self.nsea)

    def snonlin(self, F, depth=999.0, akmean=0.0):
        """SNONLIN alone on one spectrum F[ML][KL]; returns (SL, FL) increments."""
        f = np.ascontiguousarray(F, dtype=self.dtype).ravel().copy()
        sl = np.zeros_like(f)
        fl = np.zeros_like(f)
        self.lib.orc_api_snonlin(self.ctx, _p(f), _p(sl), _p(fl), float(depth), float(akmean))
        return sl.reshape(self.ml, self.kl), fl.reshape(self.ml, self.kl)

    def aki(self, om, depth):
        return self.lib.orc_api_aki(self.ctx, float(om), float(depth))


def shared_math(fn, x, y=None):
    """Evaluates a wam_math.h function (see orc_api_math) on float32 arrays."""
    lib = _lib(False)
    x = np.ascontiguousarray(x, dtype=np.float32)
    yy = None if y is None else np.ascontiguousarray(y, dtype=np.float32)
    out = np.zeros_like(x)
    lib.orc_api_math(int(fn), _p(x), _p(yy), _p(out), x.size)
    return out
