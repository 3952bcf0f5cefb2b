"""Trace processing, misfit and gradient post-processing on the GPU.

The steps an FWI iteration runs either side of the wave solver, on device arrays (torch
tensors on the solver's GPU), through the C ABI of stride_amd/csrc/process.hip. Names,
arguments, defaults and error behaviour follow Stride's pipeline steps:

- ``CheckTraces``  stride/optimisation/pipelines/steps/check_traces.py:8-83
- ``MuteTraces``   steps/mute_traces.py:8-68
- ``FilterTraces`` steps/filter_traces.py:6-120 with the filters of stride/utils/filters.py
- ``NormPerShot`` / ``NormPerTrace`` steps/norm_traces.py:7-86
- ``ProcessTraces`` default_pipelines.py:83-142 (check, mute, filter, norm per shot)
- ``L2DistanceLoss`` stride/optimisation/loss/l2_distance.py:36-66
- ``MaskField`` / ``SmoothField`` / ``NormField`` steps/mask_field.py, smooth_field.py,
  norm_field.py, and ``ProcessGlobalGradient`` default_pipelines.py:145-175
- ``apply_prec`` stride/problem/data.py:409-446
- ``Clip`` steps/clip.py and ``GradientDescent`` optimisers/optimiser.py:126-330

As in the reference, forward steps return new arrays (the inputs are not modified) and the
adjoint of every trace step except the filter is the identity (mute_traces.py:62-68,
norm_traces.py:32-38, check_traces.py:44-53). Filter design (Butterworth sections, Kaiser FIR
taps, the cosine table) is host-side coefficient setup with scipy.signal, as in filters.py;
every pass over trace or field samples runs on the GPU. Steps the reference keeps off by
default (filter_offsets, agc, scale_per_shot/trace, time weighting, shift_traces) are not
provided.
"""

import warnings

import numpy as np
import scipy.signal

__all__ = ['CheckTraces', 'MuteTraces', 'FilterTraces', 'NormPerShot', 'NormPerTrace', 'ProcessTraces',
           'L2DistanceLoss', 'MaskField', 'SmoothField', 'NormField', 'ProcessGlobalGradient', 'apply_prec',
           'Clip', 'GradientDescent', 'traces_from_solver', 'adjoint_source_to_solver', 'cos_taps', 'crop', 'embed',
           'gradient_vp']


def _torch():
    import torch
    return torch


def _check_dev(t, what, solver=None):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError('%s must be a torch tensor on the GPU' % what)
    if t.dtype != torch.float32 or not t.is_contiguous():
        raise TypeError('%s must be a contiguous float32 tensor' % what)
    if solver is not None and t.device.index != solver.device:
        raise ValueError('%s is on cuda:%d, the solver on cuda:%d' % (what, t.device.index, solver.device))
    return t


def _enter(solver):
    """Order the solver's stream after torch's pending work on the arrays."""
    _torch().cuda.current_stream(solver.device).synchronize()


def _rows(t):
    return (t.shape[0], t.shape[-1]) if t.dim() == 2 else (1, t.shape[-1])


# ---------------------------------------------------------------------------- filters ----

def cos_taps(f_max):
    """Cosine low-pass table for a dimensionless f_max (stride/utils/filters.py:401-436):
    length 2 p + 1 with p = round(0.5 / f_max), weights 1 - cos(2 pi i / (L + 1)), i = 1..L,
    normalised to sum 1. Symmetric."""
    period = int(np.round(1 / (f_max / 0.5)))
    n = 2 * period + 1
    w = 1. - np.cos(2 * np.pi * np.arange(1, n + 1) / (n + 1))
    return w / w.sum(), period


def _fir_taps(pass_type, args, attenuation=30):
    width = 0.050
    order, beta = scipy.signal.kaiserord(attenuation, width)
    order = order // 2 * 2 + 1
    cut = [a / 0.5 for a in args]
    cut = cut[0] if len(cut) == 1 else cut
    return scipy.signal.firwin(order, cut, pass_zero=pass_type, window=('kaiser', beta), scale=True)


def _butter_sos(pass_type, args, order=8):
    cut = [a / 0.5 for a in args]
    btype = {'lowpass': 'lowpass', 'highpass': 'highpass', 'bandpass': 'band'}[pass_type]
    return scipy.signal.butter(order, cut[0] if len(cut) == 1 else cut, analog=False, btype=btype, output='sos')


class FilterTraces:
    """Causal (``zero_phase=False``) filtering of (ntr, nt) traces along time, and its adjoint
    (time-reversed filtering), as FilterTraces applies it (filter_traces.py:70-120).

    ``f_min`` / ``f_max`` in Hz; dimensionless cut-offs are f * dt (scaled by the relaxation).
    Filter type defaults to ``cos`` for a low pass and ``butterworth`` otherwise; ``fir`` is
    the Kaiser-window design. Types without a device implementation raise like the
    reference's missing-filter error."""

    def __init__(self, f_min=None, f_max=None, filter_type=None, filter_relaxation=1.0, **kwargs):
        self.f_min = f_min
        self.f_max = f_max
        self.filter_type = filter_type
        self.relaxation = filter_relaxation
        self.order = kwargs.pop('order', None)
        self.attenuation = kwargs.pop('attenuation', 30)

    def design(self, dt, **kwargs):
        """(kind, coefficients) for this step at time step dt, or None for a pass-through."""
        f_min = kwargs.pop('f_min', self.f_min)
        f_max = kwargs.pop('f_max', self.f_max)
        relaxation = kwargs.pop('filter_relaxation', self.relaxation)
        f_min_dl = relaxation * f_min * dt if f_min is not None else 0
        f_max_dl = 1 / relaxation * f_max * dt if f_max is not None else 0
        if f_min is None and f_max is not None:
            pass_type, args = 'lowpass', (f_max_dl,)
        elif f_min is not None and f_max is None:
            pass_type, args = 'highpass', (f_min_dl,)
        elif f_min is not None and f_max is not None:
            pass_type, args = 'bandpass', (f_min_dl, f_max_dl)
        else:
            return None
        filter_type = kwargs.pop('filter_type', self.filter_type or ('cos' if f_min is None else 'butterworth'))
        if filter_type == 'butterworth':
            order = self.order if self.order is not None else 8
            return 'sos', _butter_sos(pass_type, args, order)
        if filter_type == 'fir':
            return 'fir', _fir_taps(pass_type, args, self.attenuation)
        if filter_type == 'cos' and pass_type == 'lowpass':
            if self.order not in (None, 1):
                # later passes of filters.py:455-456 are centred on the padded trace and read
                # past its end: not the causal filter applied again
                raise NotImplementedError('cosine filter of order %s on the device' % self.order)
            taps, _ = cos_taps(args[0])
            return 'fir', taps
        raise Exception('Requested filter does not exist. Implemented filters are butterworth, fir & cos.')

    def _apply(self, solver, traces, dt, adjoint, **kwargs):
        _check_dev(traces, 'traces', solver)
        out = traces.clone()
        spec = self.design(dt, **kwargs)
        if spec is None or out.numel() == 0:
            return out
        _enter(solver)
        ntr, nt = _rows(out)
        kind, coef = spec
        if kind == 'sos':
            solver.traces_sosfilt(out.data_ptr(), ntr, nt, coef, adjoint=adjoint)
        else:
            # lfilter / causal convolution y[i] = sum_k b[k] x[i - k]: reversed taps, offset L - 1
            w = np.ascontiguousarray(np.asarray(coef, np.float64)[::-1])
            solver.traces_fir(out.data_ptr(), ntr, nt, w, len(w) - 1, adjoint=adjoint)
        return out

    def forward(self, solver, *traces, dt, **kwargs):
        out = tuple(self._apply(solver, t, dt, False, **kwargs) for t in traces)
        return out if len(out) > 1 else out[0]

    def adjoint(self, solver, *d_traces, dt, **kwargs):
        out = tuple(self._apply(solver, t, dt, True, **kwargs) for t in d_traces)
        return out if len(out) > 1 else out[0]


# ------------------------------------------------------------------------- trace steps ----

class CheckTraces:
    """NaN / Inf rows: raise (default), or warn and, with ``filter_incorrect``, zero them."""

    def __init__(self, raise_incorrect=True, filter_incorrect=False, **kwargs):
        self.raise_incorrect = raise_incorrect
        self.filter_incorrect = filter_incorrect

    def forward(self, solver, *traces, name='traces', shot_id=None, **kwargs):
        raise_incorrect = kwargs.pop('raise_incorrect', self.raise_incorrect)
        filter_incorrect = kwargs.pop('filter_incorrect', self.filter_incorrect)
        out = []
        for t in traces:
            _check_dev(t, 'traces', solver)
            o = t.clone()
            ntr, nt = _rows(o)
            if o.numel():
                _enter(solver)
                _, nbad = solver.traces_check(o.data_ptr(), ntr, nt, zero_bad=False)
                if nbad:
                    msg = 'Nan or inf detected in %s' % name
                    if shot_id is not None:
                        msg = '(ShotID %d) ' % shot_id + msg
                    if raise_incorrect:
                        raise RuntimeError(msg)
                    warnings.warn(msg)
                    if filter_incorrect:
                        solver.traces_check(o.data_ptr(), ntr, nt, zero_bad=True)
            out.append(o)
        return tuple(out) if len(out) > 1 else out[0]

    def adjoint(self, solver, *d_traces, **kwargs):
        return d_traces if len(d_traces) > 1 else d_traces[0]


class MuteTraces:
    """modelled x (cosine low-pass, zero phase, of the indicator observed != 0)."""

    def __init__(self, f_max=None, filter_relaxation=1, **kwargs):
        self.f_max = f_max
        self.filter_relaxation = filter_relaxation

    def forward(self, solver, modelled, observed, dt, **kwargs):
        _check_dev(modelled, 'modelled', solver)
        _check_dev(observed, 'observed', solver)
        if modelled.shape != observed.shape:
            raise ValueError('modelled %s and observed %s differ in shape' % (tuple(modelled.shape),
                                                                                tuple(observed.shape)))
        f_max = kwargs.pop('f_max', self.f_max)
        relax = kwargs.pop('filter_relaxation', self.filter_relaxation)
        f_max = f_max * dt / relax if f_max is not None else 0.5
        taps, _ = cos_taps(f_max)
        out = modelled.clone()
        if out.numel():
            _enter(solver)
            ntr, nt = _rows(out)
            solver.traces_mute(out.data_ptr(), observed.data_ptr(), ntr, nt, taps)
        return out, observed.clone()

    def adjoint(self, solver, *d_traces, **kwargs):
        return d_traces if len(d_traces) > 1 else d_traces[0]


class NormPerShot:
    """traces / (sqrt(sum over rows of sum d^2 / ntr) + 1e-31), each set by its own norm."""

    per_trace = False

    def __init__(self, **kwargs):
        pass

    def forward(self, solver, *traces, **kwargs):
        out = []
        for t in traces:
            _check_dev(t, 'traces', solver)
            o = t.clone()
            if o.numel():
                _enter(solver)
                ntr, nt = _rows(o)
                solver.traces_norm(o.data_ptr(), ntr, nt, per_trace=self.per_trace)
            out.append(o)
        return tuple(out) if len(out) > 1 else out[0]

    def adjoint(self, solver, *d_traces, **kwargs):
        return d_traces if len(d_traces) > 1 else d_traces[0]


class NormPerTrace(NormPerShot):
    """Each row / (sqrt(sum d^2) + 1e-31)."""

    per_trace = True


class ProcessTraces:
    """check_traces -> mute_traces -> filter_traces -> norm_per_shot (or norm_per_trace) on
    (modelled, observed), each step switchable by its keyword as in the reference; the adjoint
    takes d(modelled) back through the filter's adjoint (the other steps pass it through)."""

    def __init__(self, **kwargs):
        self.steps = []
        if kwargs.pop('check_traces', True):
            self.steps.append(('check_traces', CheckTraces(**kwargs)))
        if kwargs.pop('mute_traces', True):
            self.steps.append(('mute_traces', MuteTraces(**kwargs)))
        if kwargs.pop('filter_traces', True):
            self.steps.append(('filter_traces', FilterTraces(**kwargs)))
        norm_per_shot = kwargs.pop('norm_per_shot', True)
        if kwargs.pop('norm_per_trace', False):
            self.steps.append(('norm_per_trace', NormPerTrace()))
        elif norm_per_shot:
            self.steps.append(('norm_per_shot', NormPerShot()))

    def forward(self, solver, modelled, observed, dt, **kwargs):
        mod, obs = modelled, observed
        for name, step in self.steps:
            if name == 'check_traces':
                mod = step.forward(solver, mod, name='modelled', shot_id=kwargs.get('shot_id'))
                obs = step.forward(solver, obs, name='observed', shot_id=kwargs.get('shot_id'))
            elif name == 'mute_traces':
                mod, obs = step.forward(solver, mod, obs, dt)
            elif name == 'filter_traces':
                mod, obs = step.forward(solver, mod, obs, dt=dt)
            else:
                mod, obs = step.forward(solver, mod, obs)
        solver.synchronize()
        return mod, obs

    def adjoint(self, solver, d_modelled, dt):
        d = d_modelled
        for name, step in reversed(self.steps):
            if name == 'filter_traces':
                d = step.adjoint(solver, d, dt=dt)
        solver.synchronize()
        return d


class L2DistanceLoss:
    """f = 1/2 sum (modelled - observed)^2 (in double), residual = modelled - observed; the
    adjoint source for the modelled traces is the residual."""

    def __init__(self, **kwargs):
        self.residual = None
        self._solver = None

    def forward(self, solver, modelled, observed, **kwargs):
        _check_dev(modelled, 'modelled', solver)
        _check_dev(observed, 'observed', solver)
        if modelled.shape != observed.shape:
            raise ValueError('modelled %s and observed %s differ in shape' % (tuple(modelled.shape),
                                                                                tuple(observed.shape)))
        res = _torch().empty_like(modelled)
        ntr, nt = _rows(res)
        _enter(solver)
        fun = solver.traces_l2(modelled.data_ptr(), observed.data_ptr(), res.data_ptr(), ntr, nt)
        self.residual, self._solver = res, solver
        return fun, res

    def adjoint(self, d_fun=1.0, modelled_needs_grad=True, observed_needs_grad=False):
        """(+d_fun residual, -d_fun residual) as new arrays (l2_distance.py:55-66)."""
        res, s = self.residual, self._solver
        self.residual = self._solver = None
        if res is None:
            raise RuntimeError('L2DistanceLoss.adjoint called before forward')

        def scaled(a):
            out = res.clone()
            s.field_axpb(out.data_ptr(), out.numel(), a, 0.0)
            return out
        grad_mod = scaled(float(d_fun)) if modelled_needs_grad else None
        grad_obs = scaled(-float(d_fun)) if observed_needs_grad else None
        s.synchronize()
        return grad_mod, grad_obs


def traces_from_solver(solver, rec_time_major):
    """Device receiver traces (nt, nrec), as the solver writes them, to Traces rows (nrec, nt)."""
    _check_dev(rec_time_major, 'receiver traces', solver)
    nt, nrec = rec_time_major.shape
    out = _torch().empty((nrec, nt), dtype=rec_time_major.dtype, device=rec_time_major.device)
    _enter(solver)
    solver.traces_transpose(rec_time_major.data_ptr(), out.data_ptr(), nt, nrec)
    return out


def adjoint_source_to_solver(solver, residual, window):
    """Adjoint source rows (nrec, nt) to the solver's time-major layout, times the time window
    (length nt: the Tukey taper over the time bounds, devito.py:863-869)."""
    _check_dev(residual, 'adjoint source', solver)
    nrec, nt = residual.shape
    out = _torch().empty((nt, nrec), dtype=residual.dtype, device=residual.device)
    _enter(solver)
    solver.traces_transpose(residual.data_ptr(), out.data_ptr(), nrec, nt, col_scale=window)
    return out


# ------------------------------------------------------------------------- field steps ----

def _box(inner):
    return [a for a, _ in inner], [b for _, b in inner], tuple(b - a for a, b in inner)


def crop(solver, grid, inner):
    """The inner box (per axis (start, stop)) of an extended-grid device array."""
    _check_dev(grid, 'grid', solver)
    lo, hi, shape = _box(inner)
    out = _torch().empty(shape, dtype=grid.dtype, device=grid.device)
    _enter(solver)
    solver.field_crop(grid.data_ptr(), out.data_ptr(), tuple(grid.shape), lo, hi)
    solver.synchronize()
    return out


def embed(solver, box, shape, inner):
    """A zero extended-grid array of ``shape`` holding ``box`` in its inner box."""
    _check_dev(box, 'box', solver)
    lo, hi, bshape = _box(inner)
    if tuple(box.shape) != bshape:
        raise ValueError('box of shape %s, inner box %s' % (tuple(box.shape), bshape))
    out = _torch().zeros(tuple(shape), dtype=box.dtype, device=box.device)
    _enter(solver)
    solver.field_embed(box.data_ptr(), out.data_ptr(), tuple(shape), lo, hi)
    solver.synchronize()
    return out


def gradient_vp(solver, vp, inner, slowness=True):
    """get_grad_vp on the device (devito.py:1063-1095): the solver's vp gradient and
    preconditioner accumulators cut to the inner box and rescaled by the inner vp (slowness
    transform, vp's default: grad / vp, prec / vp^2). ``vp``: extended-grid device array.
    Returns (grad, prec), inner shape."""
    _check_dev(vp, 'vp', solver)
    torch = _torch()
    if tuple(vp.shape) != tuple(solver.shape):
        raise ValueError('vp of shape %s, solver grid %s' % (tuple(vp.shape), tuple(solver.shape)))
    lo, hi, shape = _box(inner)
    ext = torch.empty(tuple(solver.shape), dtype=torch.float32, device=vp.device)
    grad, prec, vpi = (torch.empty(shape, dtype=torch.float32, device=vp.device) for _ in range(3))
    _enter(solver)
    solver.field_crop(vp.data_ptr(), vpi.data_ptr(), tuple(solver.shape), lo, hi)
    for which, dst in ((0, grad), (1, prec)):
        solver.read_gradient_dev(which, ext.data_ptr())
        solver.field_crop(ext.data_ptr(), dst.data_ptr(), tuple(solver.shape), lo, hi)
    solver.field_grad_vp_finish(grad.data_ptr(), prec.data_ptr(), vpi.data_ptr(), grad.numel(), slowness)
    solver.synchronize()
    return grad, prec


def _shape3(field, dim):
    shape = tuple(field.shape[-dim:]) if dim else tuple(field.shape)
    if int(np.prod(shape)) != field.numel():
        raise ValueError('field of shape %s does not hold one %d-D grid' % (tuple(field.shape), len(shape)))
    return shape


class MaskField:
    """f x [inside the inner region] x ramp-off profile of width ``mask_rampoff`` at the edges
    of the extended grid. ``inner``: per axis (start, stop) of the inner region."""

    def __init__(self, mask_rampoff=10, **kwargs):
        self.mask_rampoff = mask_rampoff

    def forward(self, solver, field, inner, **kwargs):
        ramp = kwargs.pop('mask_rampoff', self.mask_rampoff)
        _check_dev(field, 'field', solver)
        shape = _shape3(field, len(inner))
        out = field.clone()
        _enter(solver)
        solver.field_mask(out.data_ptr(), shape, [a for a, _ in inner], [b for _, b in inner], int(ramp))
        return out


class SmoothField:
    """Gaussian smoothing (scipy.ndimage.gaussian_filter semantics, truncate 4, 'nearest'
    edges), sigma in cells, default 0.25."""

    def __init__(self, smooth_sigma=0.25, **kwargs):
        self.sigma = smooth_sigma

    def forward(self, solver, field, dim=None, **kwargs):
        _check_dev(field, 'field', solver)
        dim = dim or field.dim()
        sigma = kwargs.pop('smooth_sigma', self.sigma)
        if not np.iterable(sigma):
            sigma = (sigma,) * dim
        if all(s <= 0 for s in sigma):
            return field
        out = field.clone()
        _enter(solver)
        solver.field_gaussian(out.data_ptr(), _shape3(field, dim), sigma, mode='nearest')
        return out


class NormField:
    """field x var_corr / (max|field| + 1e-31), var_corr = transform(mid(variable)) x
    norm_guess_change / 100 (1 without a variable, or when that is 0). ``global_norm`` keeps
    the first call's max."""

    def __init__(self, global_norm=False, norm_guess_change=0.5, **kwargs):
        self.global_norm = global_norm
        self.norm_guess_change = norm_guess_change
        self.norm_value = None

    def forward(self, solver, field, variable=None, transform=None, **kwargs):
        _check_dev(field, 'field', solver)
        global_norm = kwargs.pop('global_norm', self.global_norm)
        change = kwargs.pop('norm_guess_change', self.norm_guess_change)
        _enter(solver)
        if self.norm_value is None or not global_norm:
            lo, hi, _ = solver.field_stats(field.data_ptr(), field.numel())
            self.norm_value = max(abs(lo), abs(hi)) + 1e-31
        var_corr = 1.
        if variable is not None:
            _check_dev(variable, 'variable', solver)
            vlo, vhi, _ = solver.field_stats(variable.data_ptr(), variable.numel())
            mid = (vhi + vlo) / 2.0
            if transform is not None:
                mid = transform(mid)
            var_corr = mid * change / 100
            var_corr = var_corr if np.abs(var_corr) > 0. else 1.
        out = field.clone()
        solver.field_axpb(out.data_ptr(), out.numel(), var_corr / self.norm_value, 0.0)
        solver.synchronize()
        return out


class ProcessGlobalGradient:
    """mask_field -> smooth_field -> norm_field, each switchable (mask_grad, smooth_grad,
    norm_grad)."""

    def __init__(self, **kwargs):
        self.steps = []
        if kwargs.pop('mask_grad', True):
            self.steps.append(('mask_field', MaskField(**kwargs)))
        if kwargs.pop('smooth_grad', True):
            self.steps.append(('smooth_field', SmoothField(**kwargs)))
        if kwargs.pop('norm_grad', True):
            self.steps.append(('norm_field', NormField(**kwargs)))

    def forward(self, solver, grad, inner, variable=None, transform=None):
        g = grad
        for name, step in self.steps:
            if name == 'mask_field':
                g = step.forward(solver, g, inner)
            elif name == 'smooth_field':
                g = step.forward(solver, g, dim=len(inner))
            else:
                g = step.forward(solver, g, variable=variable, transform=transform)
        solver.synchronize()
        return g


def apply_prec(solver, grad, prec, prec_scale=4.0, prec_smooth=1.0):
    """Illumination preconditioning of a gradient by its prec: prec smoothed (Gaussian, sigma
    ``prec_smooth`` cells, 'reflect' edges), scaled to mean ``prec_scale``, plus 1; the gradient
    divided by it where it is non-zero. Returns a new gradient."""
    _check_dev(grad, 'gradient', solver)
    out = grad.clone()
    if prec is None:
        return out
    _check_dev(prec, 'prec', solver)
    if prec.shape != grad.shape:
        raise ValueError('prec %s and gradient %s differ in shape' % (tuple(prec.shape), tuple(grad.shape)))
    p = prec.clone()
    _enter(solver)
    if prec_smooth is not None:
        solver.field_gaussian(p.data_ptr(), tuple(p.shape), prec_smooth, mode='reflect')
    _, _, factor = solver.field_stats(p.data_ptr(), p.numel())
    if factor > 1e-31:
        solver.field_axpb(p.data_ptr(), p.numel(), prec_scale * p.numel() / factor, 1.0)
        solver.field_div_nonzero(out.data_ptr(), p.data_ptr(), out.numel())
    solver.synchronize()
    return out


class Clip:
    """Bounds on the variable after the update (None: no bound on that side)."""

    def __init__(self, min=None, max=None, **kwargs):
        self.min = min
        self.max = max


class GradientDescent:
    """``variable -= step * direction`` in the variable's transformed space (``transform``
    'slowness': x -> 1/x, the default for vp; None: identity), then the model clip. The step is
    capped as the reference does without ``force_step``: steps above -0.2 go to
    [0.1, max_step]."""

    def __init__(self, step_size=1.0, transform='slowness', max_step=None, force_step=False, **kwargs):
        if transform not in (None, 'slowness'):
            raise ValueError('transform must be None or "slowness"')
        self.step_size = step_size
        self.transform = transform
        self.max_step = max_step
        self.force_step = force_step
        self.clip = Clip(min=kwargs.pop('min', None), max=kwargs.pop('max', None))

    def final_step(self, step_size=None):
        step = self.step_size if step_size is None else step_size
        if not self.force_step:
            max_step = self.max_step if isinstance(self.max_step, (int, float)) else np.inf
            if step > -0.2:
                step = max(0.1, min(step, max_step))
            elif max_step < np.inf and step < -max_step * 0.75:
                step = -max_step * 0.75
            elif step < -0.2:
                step = step * 0.25
        return step

    def step(self, solver, variable, direction, step_size=None):
        """Updates ``variable`` in place; returns the step taken."""
        _check_dev(variable, 'variable', solver)
        _check_dev(direction, 'direction', solver)
        if variable.shape != direction.shape:
            raise ValueError('variable %s and direction %s differ in shape' % (tuple(variable.shape),
                                                                                tuple(direction.shape)))
        step = self.final_step(step_size)
        _enter(solver)
        solver.field_descent(variable.data_ptr(), direction.data_ptr(), variable.numel(), step,
                             slowness=self.transform == 'slowness', lo=self.clip.min, hi=self.clip.max)
        solver.synchronize()
        return step
