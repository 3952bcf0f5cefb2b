"""``torch.autograd.Function`` around the solver's forward and exact adjoint.

``iso_acoustic_traces(vp, rho, alpha, wavelets, geometry)`` returns the modelled
traces ``(n_rec, nt)`` as a torch tensor; its backward runs the exact discrete
adjoint (the transpose of the forward time-stepping map) and returns dJ/dvp,
dJ/drho, dJ/dalpha and dJ/dwavelets for whichever inputs require grad, so
``torch.autograd.gradcheck``-style finite-difference checks hold to FP32
rounding. Medium tensors live on the extended grid, like
``ScalarField.extended_data``.

Two paths, picked by where vp lives:

- host tensors: the medium and traces go through host arrays (``sa_set_medium``,
  ``sa_forward``, ``sa_adjoint_exact``) and the chain rule runs in numpy;
- GPU tensors: nothing of the size of the grid leaves HBM. The medium is set from device
  pointers (``sa_set_medium_dev``), the traces are written to a device tensor
  (``sa_forward_dev``), the adjoint writes its accumulators to device tensors
  (``sa_adjoint_exact_dev``) and the chain rule to vp / rho / alpha and the wavelets runs on
  the device (``sa_exact_gradients_dev``). The shot's plan (taps, boundary tables, scheme,
  source samples) is built once and reused while the wavelets, the kind of medium (which of
  rho / alpha are given, constant or variable rho) and the velocity bound of the scheme choice
  stay the same; per call only a few scalars (max vp, min / max rho) reach the host.

The forward keeps either every time level (``checkpoint_every=None``) or only
checkpoints every K steps, recomputing each K-step segment during the adjoint
(``checkpoint_every=K``), which bounds memory at (2 nt/K + K) wavefields.
"""

import numpy as np
import torch

from ._native import Solver
from .physics.boundaries import SPONGE, PML
from .physics.gradients import exact_gradients
from .physics.plan import build_plan, adjoint_source_samples, RHO_VARIABLE, _db_to_neper
from .utils import tukey

__all__ = ['Geometry', 'iso_acoustic_traces']


class Geometry:
    """Everything of a shot that is not differentiated: grid, time axis, source and
    receiver coordinates and the scheme options of IsoAcousticDevito."""

    def __init__(self, space, time, source_coordinates, receiver_coordinates, device=0, checkpoint_every=None,
                 **options):
        self.space = space
        self.time = time
        self.source_coordinates = np.asarray(source_coordinates)
        self.receiver_coordinates = np.asarray(receiver_coordinates)
        self.device = device
        self.checkpoint_every = checkpoint_every
        self.options = options
        self._solver = None
        # identifies the forward run whose medium and saved wavefield the solver holds
        self.state_token = None
        # device path: the cached plan, what it was built for, and whether the solver holds its
        # taps and boundary tables
        self._dev_plan = None
        self._dev_key = None
        self._dev_vmax = None
        self._dev_ready = False

    def solver(self, plan):
        key = (plan.dim, plan.shape, plan.spacing, plan.absorbing, float(plan.dt), int(plan.nt), int(self.device))
        if self._solver is None or self._solver.key != key:
            self._solver = Solver(plan.dim, plan.shape, plan.spacing, plan.absorbing, plan.dt, plan.nt,
                                  device=self.device)
            self._dev_ready = False
        return self._solver


def _np(t):
    return None if t is None else t.detach().to('cpu', torch.float32).numpy()


def _save_mode(geometry):
    return 3 if geometry.checkpoint_every is None else 4


def _configure(geometry, plan, s):
    s.set_scheme(plan.kernel_order, plan.rho_mode, plan.att_power, plan.boundary, plan.diff_source,
                 geometry.checkpoint_every or 1, _save_mode(geometry), plan.tb0, plan.tb1, 1, 0, 0)
    if plan.boundary == SPONGE:
        s.set_sponge(plan.sponge_tables)
    s.set_sources(plan.src_idx, plan.src_w)
    s.set_receivers(plan.rec_idx, plan.rec_w)


def _run_forward(geometry, plan):
    """Configure the geometry's solver for ``plan`` and run the saving forward (host path);
    returns the receiver samples (nt, nrec) and stamps the solver state with a fresh token."""
    s = geometry.solver(plan)
    _configure(geometry, plan, s)
    s.set_medium(plan.vp, plan.rho, plan.alpha)
    rec = s.forward(plan.src_data)
    geometry.state_token = object()
    return rec


# ---- device path -------------------------------------------------------------------------------

def _device_plan(geometry, vp, rho, alpha, wavelets):
    """The cached plan of the device path, rebuilt (from one host copy of the inputs) when the
    wavelets, the kind of medium or the velocity bound that picks the scheme change."""
    vmax = float(vp.detach().max())
    rho_const = None if rho is None else bool(float(rho.detach().min()) == float(rho.detach().max()))
    wav = _np(wavelets)
    key = (rho is None, rho_const, alpha is None, wav.tobytes(), wav.shape)
    if geometry._dev_plan is None or geometry._dev_key != key or vmax > geometry._dev_vmax:
        plan = build_plan(geometry.space, geometry.time, wav, _np(vp), rho=_np(rho), alpha=_np(alpha),
                          source_coordinates=geometry.source_coordinates,
                          receiver_coordinates=geometry.receiver_coordinates, grad_vp=True, **geometry.options)
        if plan.boundary == PML:
            raise ValueError('the exact adjoint does not support the CFS-PML boundary')
        geometry._dev_plan, geometry._dev_key, geometry._dev_vmax = plan, key, vmax
        geometry._dev_ready = False
        dev = vp.device
        geometry._dev_src = torch.from_numpy(np.ascontiguousarray(plan.src_data)).to(dev)
        window = np.pad(tukey(plan.tb1, 0.001), ((0, plan.nt - plan.tb1),), mode='constant')
        geometry._dev_window = torch.from_numpy(window.astype(np.float32)).to(dev)
    return geometry._dev_plan


def _run_forward_device(geometry, plan, vp, rho, alpha):
    s = geometry.solver(plan)
    if not geometry._dev_ready:
        _configure(geometry, plan, s)
        geometry._dev_ready = True
    else:
        s.set_scheme(plan.kernel_order, plan.rho_mode, plan.att_power, plan.boundary, plan.diff_source,
                     geometry.checkpoint_every or 1, _save_mode(geometry), plan.tb0, plan.tb1, 1, 0, 0)
    a_np = None
    if alpha is not None:
        a_np = (alpha.detach().to(torch.float32) * float(_db_to_neper(plan.attenuation_power))).contiguous()
    vp_ = vp.detach().to(torch.float32).contiguous()
    rho_ = None if rho is None else rho.detach().to(torch.float32).contiguous()
    if plan.rho_mode != RHO_VARIABLE:
        rho_ = None   # constant density: the div(b grad) form with no buoyancy field
    s.set_medium_dev(vp_.data_ptr(), None if rho_ is None else rho_.data_ptr(),
                     None if a_np is None else a_np.data_ptr())
    rec = torch.empty((plan.nt, plan.nrec), dtype=torch.float32, device=vp.device)
    s.forward_dev(geometry._dev_src.data_ptr(), rec.data_ptr())
    s.synchronize()
    geometry.state_token = object()
    return rec


class _IsoAcoustic(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, rho, alpha, wavelets, geometry):
        ctx.geometry = geometry
        ctx.shapes = [None if x is None else (x.dtype, x.device) for x in (vp, rho, alpha, wavelets)]
        ctx.device_path = vp.is_cuda
        if ctx.device_path:
            plan = _device_plan(geometry, vp, rho, alpha, wavelets)
            rec = _run_forward_device(geometry, plan, vp, rho, alpha)
            ctx.plan = plan
            ctx.token = geometry.state_token
            ctx.save_for_backward(vp, rho, alpha)
            return rec.t().contiguous()
        plan = build_plan(geometry.space, geometry.time, _np(wavelets), _np(vp), rho=_np(rho), alpha=_np(alpha),
                          source_coordinates=geometry.source_coordinates,
                          receiver_coordinates=geometry.receiver_coordinates, grad_vp=True, **geometry.options)
        if plan.boundary == PML:
            raise ValueError('the exact adjoint does not support the CFS-PML boundary')
        rec = _run_forward(geometry, plan)
        ctx.plan = plan
        ctx.token = geometry.state_token
        return torch.from_numpy(np.ascontiguousarray(rec.T)).to(vp.device)

    @staticmethod
    def backward(ctx, grad_traces):
        if ctx.device_path:
            return _backward_device(ctx, grad_traces)
        plan, geometry = ctx.plan, ctx.geometry
        if geometry.state_token is not ctx.token:
            # another forward ran on this geometry's solver since this one (its medium and
            # saved wavefield are in the solver now): run this forward again first
            _run_forward(geometry, plan)
            ctx.token = geometry.state_token
        s = geometry.solver(plan)
        # J = <grad_traces, traces>: the adjoint source is grad_traces itself (no window)
        res = np.ascontiguousarray(_np(grad_traces).T)
        acc = s.adjoint_exact(res, want_m=plan.kernel_order == 4, want_b=plan.rho_mode == RHO_VARIABLE)
        g = exact_gradients(plan, acc, acc['lam_src'])
        outs = []
        for name, meta in zip(('vp', 'rho', 'alpha'), ctx.shapes[:3]):
            if meta is None or not ctx.needs_input_grad[('vp', 'rho', 'alpha').index(name)]:
                outs.append(None)
                continue
            if name not in g:
                raise ValueError('no %s gradient for this scheme' % name)
            outs.append(torch.from_numpy(g[name].astype(np.float32)).to(meta[1]))
        gw = None
        if ctx.needs_input_grad[3]:
            if plan.diff_source:
                raise ValueError('wavelet gradient with diff_source is not supported')
            window = np.pad(tukey(plan.tb1, 0.001), ((0, plan.nt - plan.tb1),), mode='constant')
            gw = torch.from_numpy((g['src_samples'] * window[:, None]).T.astype(np.float32)).to(ctx.shapes[3][1])
        return outs[0], outs[1], outs[2], gw, None


def _backward_device(ctx, grad_traces):
    plan, geometry = ctx.plan, ctx.geometry
    vp, rho, alpha = ctx.saved_tensors
    if geometry.state_token is not ctx.token:
        _run_forward_device(geometry, plan, vp, rho, alpha)
        ctx.token = geometry.state_token
    s = geometry.solver(plan)
    dev = vp.device
    res = grad_traces.detach().to(torch.float32).t().contiguous()   # (nt, nrec)
    shape = plan.shape
    acc = {k: torch.empty(shape, dtype=torch.float32, device=dev) for k in ('gA', 'gB', 'gG')}
    acc['gm'] = torch.empty(shape, dtype=torch.float32, device=dev) if plan.kernel_order == 4 else None
    acc['gb'] = torch.empty(shape, dtype=torch.float32, device=dev) if plan.rho_mode == RHO_VARIABLE else None
    lam = torch.empty((plan.nt, plan.nsrc * plan.src_idx.shape[1]), dtype=torch.float32, device=dev)
    ptr = lambda t: None if t is None else t.data_ptr()
    s.adjoint_exact_dev(res.data_ptr(), ptr(acc['gA']), ptr(acc['gB']), ptr(acc['gG']), ptr(acc['gm']),
                        ptr(acc['gb']), lam.data_ptr())
    want = ctx.needs_input_grad
    g_vp = torch.zeros(shape, dtype=torch.float32, device=dev) if want[0] else None
    if want[1] and rho is not None and plan.rho_mode != RHO_VARIABLE:
        raise ValueError('no rho gradient for this scheme')
    g_rho = torch.empty(shape, dtype=torch.float32, device=dev) if want[1] and rho is not None else None
    g_alpha = torch.empty(shape, dtype=torch.float32, device=dev) if want[2] and alpha is not None else None
    g_src = None
    if want[3]:
        if plan.diff_source:
            raise ValueError('wavelet gradient with diff_source is not supported')
        g_src = torch.empty((plan.nt, plan.nsrc), dtype=torch.float32, device=dev)
    scale = float(_db_to_neper(plan.attenuation_power)) if alpha is not None else 1.0
    s.exact_gradients_dev(ptr(acc['gA']), ptr(acc['gB']), ptr(acc['gG']), ptr(acc['gm']), ptr(acc['gb']),
                          lam.data_ptr(), geometry._dev_src.data_ptr(), ptr(g_vp), ptr(g_rho), ptr(g_alpha),
                          ptr(g_src), alpha_scale=scale)
    s.synchronize()
    gw = None
    if g_src is not None:
        gw = (g_src * geometry._dev_window[:, None]).t().contiguous().to(ctx.shapes[3][1])
    return g_vp, g_rho, g_alpha, gw, None


def iso_acoustic_traces(vp, rho=None, alpha=None, wavelets=None, geometry=None):
    """Modelled traces (n_rec, nt) for extended-grid medium tensors; differentiable in
    every input that requires grad."""
    return _IsoAcoustic.apply(vp, rho, alpha, wavelets, geometry)


def residual_samples(plan, traces):
    """Windowed adjoint source as the reference builds it (for callers that want the
    as-written Tukey window applied)."""
    return adjoint_source_samples(plan, traces)
