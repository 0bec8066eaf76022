"""Gradients through a window of model steps with respect to the winds.

``WindowModel`` runs NSTEPS model steps as WAMODEL's time loop does them
(wamodel.f90:332,354): each step maps the step's wind components u, v to
U10 and UDIR (WAM_WIND, wam_wind_module.f90:1157-1167), then PROPAGS and
IMPLSCH, all through the library's device entry points.  Its reverse sweep
runs the step adjoints (wam_gpu_step_adjoint_dev, the wind map's adjoint) in
reverse order with the binomial checkpointing schedule of
wam_checkpoint_schedule, so that only NSNAP + 1 states are held in HBM.
``window(u, v, fl3, tauw, model)`` is the ``torch.autograd.Function`` form:
it returns the spectra at the observation points after every observed step,
so that any cost written with torch (``hs_from_spectra`` gives significant
wave heights) can be differentiated with ``backward()``.  The reference has
no adjoint model; there is no counterpart.
"""
import numpy as np

from .model import CK_ADJOINT, CK_ADVANCE, CK_RESTORE, CK_STORE, checkpoint_schedule


class WindowModel:
    """A window of nsteps steps on the prepared context ``gpu`` (WamGpu),
    observed at the sea points ``obs_points`` (0-based) after every
    ``obs_every``-th step; ``nsnap`` checkpoint slots besides the initial
    state (None: one per step, nothing recomputed).  ``nmem`` > 1 advances an
    ensemble (member-major arrays, as the ensemble entry points): winds
    [nsteps][nmem*NSEA], spectra [nmem][ML][KL][NSEA].  ROAIRN and WSTAR are
    the constants of the reference's runs without air-density or gustiness
    input (1.225 and 0, wam_initial_module.f90:193-194) unless given."""

    def __init__(self, gpu, nsteps, obs_points, obs_every=1, nsnap=None, roairn=None, wstar=None, nmem=1):
        import torch
        self.gpu, self.nsteps, self.obs_every = gpu, int(nsteps), int(obs_every)
        self.nsnap = self.nsteps - 1 if nsnap is None else int(nsnap)
        self.nmem = int(nmem)
        dev = torch.device("cuda:0")
        self.dev = dev
        n = gpu.nsea * self.nmem
        self.nbin = gpu.kl * gpu.ml
        self.n, self.nb = n, self.nbin * n
        self.pts = torch.as_tensor(np.asarray(obs_points, np.int64), device=dev)
        f = lambda: torch.zeros(self.nb, device=dev)
        p = lambda: torch.zeros(n, device=dev)
        self.fa, self.fb = f(), f()
        self.tauw, self.ustar, self.z0, self.u10, self.udir = p(), p(), p(), p(), p()
        self.roairn = torch.full((n,), 1.225, device=dev) if roairn is None else roairn
        self.wstar = torch.zeros(n, device=dev) if wstar is None else wstar
        self.actions = checkpoint_schedule(self.nsteps, self.nsnap)
        self.slots = {}
        self.forward_steps = 0

    @property
    def nobs_t(self):
        return self.nsteps // self.obs_every

    # ---- forward ----
    def _load(self, fl3, tauw):
        self.fa.copy_(fl3)
        self.tauw.copy_(tauw)

    def _store(self, slot):
        import torch
        if slot not in self.slots:
            self.slots[slot] = (torch.empty(self.nb, device=self.dev), torch.empty(self.n, device=self.dev))
        f, t = self.slots[slot]
        f.copy_(self.fa)
        t.copy_(self.tauw)

    def _restore(self, slot):
        f, t = self.slots[slot]
        self.fa.copy_(f)
        self.tauw.copy_(t)

    def _step(self, u, v, i):
        g = self.gpu
        g.wind_uv_dev(u[i], v[i], self.u10, self.udir)
        if self.nmem == 1:
            g.propags_dev(self.fa, self.fb)
            g.implsch_dev(self.fb, self.u10, self.udir, self.tauw, self.ustar, self.z0, self.roairn, self.wstar)
        else:
            g.propags_ensemble_dev(self.nmem, self.fa, self.fb)
            g.implsch_ensemble_dev(self.nmem, self.fb, self.u10, self.udir, self.tauw, self.ustar, self.z0,
                                   self.roairn, self.wstar)
        self.fa, self.fb = self.fb, self.fa
        self.forward_steps += 1

    def _sample(self):
        """[KL*ML][npts] (one member) or [nmem][KL*ML][npts]."""
        s = self.fa.view(self.nmem * self.nbin, self.gpu.nsea).index_select(1, self.pts)
        return s if self.nmem == 1 else s.view(self.nmem, self.nbin, -1)

    def forward(self, u, v, fl3, tauw):
        """Runs the window; returns [nobs_t][KL*ML][npts] spectra at the
        observation points.  Stores the checkpoints the schedule takes before
        its first adjoint step, and the state before the last step."""
        import torch
        self._load(fl3, tauw)
        lead = []
        for op, a, b in self.actions:
            if op == CK_ADJOINT:
                break
            if op == CK_STORE:
                lead.append((a, b))
        self.first = next(i for i, (op, _, _) in enumerate(self.actions) if op == CK_ADJOINT)
        shape = (self.nbin, self.pts.numel()) if self.nmem == 1 else (self.nmem, self.nbin, self.pts.numel())
        out = torch.empty((self.nobs_t,) + shape, device=self.dev)
        self.last = (torch.empty(self.nb, device=self.dev), torch.empty(self.n, device=self.dev))
        for i in range(self.nsteps):
            for a, b in lead:
                if a == i:
                    self._store(b)
            if i == self.nsteps - 1:
                self.last[0].copy_(self.fa)
                self.last[1].copy_(self.tauw)
            self._step(u, v, i)
            if (i + 1) % self.obs_every == 0:
                out[(i + 1) // self.obs_every - 1] = self._sample()
        return out

    # ---- reverse ----
    def backward(self, u, v, g_out):
        """Given the bars of the sampled spectra, returns the bars of u, v
        ([nsteps][nsea]), of the initial spectrum and of the initial TAUW."""
        import torch
        gpu = self.gpu
        ub = torch.zeros_like(u)
        vb = torch.zeros_like(v)
        fbar = torch.zeros(self.nb, device=self.dev)
        twb, twb_out = torch.zeros(self.n, device=self.dev), torch.zeros(self.n, device=self.dev)
        u10b, udirb = torch.zeros(self.n, device=self.dev), torch.zeros(self.n, device=self.dev)
        self.fa.copy_(self.last[0])
        self.tauw.copy_(self.last[1])
        ns = self.gpu.nsea
        for op, a, b in self.actions[self.first:]:
            if op == CK_ADVANCE:
                for i in range(a, b):
                    self._step(u, v, i)
            elif op == CK_STORE:
                self._store(b)
            elif op == CK_RESTORE:
                self._restore(b)
            else:
                if (a + 1) % self.obs_every == 0:
                    k = (a + 1) // self.obs_every - 1
                    fv = fbar.view(self.nmem * self.nbin, ns)
                    fv.index_add_(1, self.pts, g_out[k].reshape(self.nmem * self.nbin, -1))
                gpu.wind_uv_dev(u[a], v[a], self.u10, self.udir)
                gpu.step_adjoint_dev(self.fa, self.u10, self.udir, self.tauw, self.roairn, self.wstar, fbar, fbar,
                                     u10b, udirb, twb_out, tauw_bar=twb, nmem=self.nmem)
                twb, twb_out = twb_out, twb
                gpu.wind_uv_adjoint_dev(u[a], v[a], u10b, udirb, ub[a], vb[a])
        return ub, vb, fbar, twb


def _function():
    import torch

    class _Window(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u, v, fl3, tauw, model):
            ctx.model = model
            ctx.save_for_backward(u, v)
            return model.forward(u, v, fl3, tauw)

        @staticmethod
        def backward(ctx, g):
            u, v = ctx.saved_tensors
            ub, vb, fb, tb = ctx.model.backward(u, v, g.contiguous())
            return ub, vb, fb, tb, None

    return _Window


_WINDOW = None


def window(u, v, fl3, tauw, model):
    """torch.autograd form of WindowModel: differentiable in u, v, fl3, tauw."""
    global _WINDOW
    if _WINDOW is None:
        _WINDOW = _function()
    return _WINDOW.apply(u, v, fl3, tauw, model)


def energy_weights(gpu):
    """TOTAL_ENERGY's weights per bin (wam_interface_module.f90:1342-1410):
    DFIM(M) plus the tail factor MO_TAIL on the last frequency, [KL*ML]
    (bin index m*KL + k), from the context's tables."""
    dfim = gpu.table("DFIM").astype(np.float64)
    w = np.repeat(dfim, gpu.kl)
    w[-gpu.kl:] += float(gpu.scalar("MO_TAIL"))
    return w


def hs_from_spectra(spec, weights, emin=1.0e-12):
    """Significant wave height 4 * SQRT(MAX(E, EMIN)) of [..., KL*ML, npts]
    spectra with the per-bin energy weights (torch, differentiable)."""
    import torch
    w = torch.as_tensor(weights, dtype=spec.dtype, device=spec.device)
    e = (spec * w[:, None]).sum(-2)
    return 4.0 * torch.sqrt(torch.clamp(e, min=emin))
