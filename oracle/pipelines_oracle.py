"""CPU restatement of the steps either side of the solver (numpy / scipy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device kernels in
stride_amd/csrc/process.hip. The product package never imports this module.

Each function follows the reference step it names (file:line into stride v1.2). The
reference itself calls scipy for the filter design, the second-order-section filtering and the
Gaussian smoothing, so those calls are the reference algorithm, not a substitute for it.
"""

import numpy as np
import scipy.ndimage
import scipy.signal


def check_traces(d, filter_incorrect=False):
    """stride/optimisation/pipelines/steps/check_traces.py:52-83 -> (traces, bad-row mask)."""
    d = np.array(d, np.float32, copy=True)
    bad = np.any(~np.isfinite(d), axis=-1)
    if filter_incorrect:
        d[bad, :] = 0
    return d, bad


def cos_table(filter_length):
    """stride/utils/filters.py:401-411."""
    table = np.zeros((filter_length,))
    q = 0.
    for i in range(1, filter_length + 1):
        table[i - 1] = 1. - np.cos(2 * np.pi * i / (filter_length + 1))
        q += table[i - 1]
    return table / q


def lowpass_filter_cos(data, f_max, zero_phase=True, adjoint=False):
    """stride/utils/filters.py:414-465 (order 1, last axis)."""
    f_max = f_max / 0.5
    period = int(np.round(1 / f_max))
    table = cos_table(2 * period + 1)
    data = np.asarray(data, np.float64)
    if adjoint:
        data = np.flip(data, axis=-1)
    if not zero_phase:
        data = np.pad(data, [(0, 0)] * (data.ndim - 1) + [(period, 0)], mode='constant')
    out = scipy.ndimage.convolve1d(data, table, mode='constant', axis=-1)
    if not zero_phase:
        out = out[..., :out.shape[-1] - period]
    if adjoint:
        out = np.flip(out, axis=-1)
    return out


def mute_traces(modelled, observed, f_max_dimless=None):
    """stride/optimisation/pipelines/steps/mute_traces.py:26-60: modelled times the cosine
    low-passed indicator of non-zero observed samples (zero phase)."""
    mask = np.ones(observed.shape)
    mask[np.abs(observed) == 0.] = 0.
    f_max = f_max_dimless if f_max_dimless is not None else 0.5
    mask = lowpass_filter_cos(mask, f_max, zero_phase=True)
    return (mask * modelled).astype(np.float32)


def butterworth_sos(f_min=None, f_max=None, order=8):
    """Second-order sections of stride/utils/filters.py:14-175 (dimensionless frequencies)."""
    if f_min is not None and f_max is not None:
        return scipy.signal.butter(order, [f_min / 0.5, f_max / 0.5], analog=False, btype='band', output='sos')
    if f_max is not None:
        return scipy.signal.butter(order, f_max / 0.5, analog=False, btype='low', output='sos')
    return scipy.signal.butter(order, f_min / 0.5, analog=False, btype='high', output='sos')


def sosfilt(sos, data, adjoint=False):
    """stride/utils/filters.py:14-175 with zero_phase False (filter_traces.py:88)."""
    data = np.asarray(data, np.float64)
    if adjoint:
        data = np.flip(data, axis=-1)
    out = scipy.signal.sosfilt(sos, data, axis=-1)
    if adjoint:
        out = np.flip(out, axis=-1)
    return out


def fir_taps(f_min=None, f_max=None, attenuation=30):
    """Kaiser-window FIR design of stride/utils/filters.py:227-398 (dimensionless frequencies)."""
    order, beta = scipy.signal.kaiserord(attenuation, 0.050)
    order = order // 2 * 2 + 1
    if f_min is not None and f_max is not None:
        return scipy.signal.firwin(order, [f_min / 0.5, f_max / 0.5], pass_zero='bandpass', window=('kaiser', beta),
                                   scale=True)
    if f_max is not None:
        return scipy.signal.firwin(order, f_max / 0.5, pass_zero='lowpass', window=('kaiser', beta), scale=True)
    return scipy.signal.firwin(order, f_min / 0.5, pass_zero='highpass', window=('kaiser', beta), scale=True)


def fir_filter(taps, data, adjoint=False):
    """stride/utils/filters.py:227-398 with zero_phase False: scipy.signal.lfilter(taps, 1, x)."""
    data = np.asarray(data, np.float64)
    if adjoint:
        data = np.flip(data, axis=-1)
    out = scipy.signal.lfilter(taps, 1., data, axis=-1)
    if adjoint:
        out = np.flip(out, axis=-1)
    return out


def norm_per_shot(d):
    """stride/optimisation/pipelines/steps/norm_traces.py:41-68."""
    d = np.asarray(d, np.float64)
    norm = np.sqrt(np.sum(d ** 2) / d.shape[0]) + 1e-31
    return d / norm, norm


def norm_per_trace(d):
    """stride/optimisation/pipelines/steps/norm_traces.py:71-86."""
    d = np.asarray(d, np.float64)
    return d / (np.sqrt(np.sum(d ** 2, axis=-1, keepdims=True)) + 1e-31)


def l2(modelled, observed):
    """stride/optimisation/loss/l2_distance.py:36-66: the residual of float32 traces is float32;
    the sum of squares is taken here in float64."""
    r = np.asarray(modelled, np.float32) - np.asarray(observed, np.float32)
    return 0.5 * float(np.sum(r.astype(np.float64) ** 2)), r


def rampoff_mask(shape, ramp_size):
    """stride/optimisation/pipelines/steps/mask_field.py:8-40."""
    mask = np.ones(shape, dtype=np.float32)
    for dim_i in range(len(shape)):
        if 2 * ramp_size > shape[dim_i]:
            continue
        for index in range(ramp_size):
            pos = np.abs((ramp_size - index - 1) / float(ramp_size - 1))
            val = 1 - np.cos(np.pi / 2 * (1 - pos))
            ind = [slice(index, s - index + 1) for s in shape]
            ind[dim_i] = index
            mask[tuple(ind)] = val
            ind = [slice(index, s - index + 1) for s in shape]
            ind[dim_i] = shape[dim_i] - index - 1
            mask[tuple(ind)] = val
    return mask


def mask_field(f, inner, ramp=10):
    """stride/optimisation/pipelines/steps/mask_field.py:50-67 (inner: tuple of slices)."""
    mask = np.zeros(f.shape, dtype=np.float32)
    mask[inner] = 1
    mask *= rampoff_mask(mask.shape, ramp)
    return f * mask


def smooth_field(f, sigma, mode='nearest'):
    """stride/optimisation/pipelines/steps/smooth_field.py:28-47."""
    return scipy.ndimage.gaussian_filter(np.asarray(f, np.float64), sigma=sigma, mode=mode)


def norm_field(f, variable=None, norm_guess_change=0.5, transform=None):
    """stride/optimisation/pipelines/steps/norm_field.py:24-52."""
    norm_value = np.max(np.abs(f)) + 1e-31
    if variable is not None:
        mid = (np.max(variable) + np.min(variable)) / 2.0
        if transform is not None:
            mid = transform(mid)
        var_corr = mid * norm_guess_change / 100
        var_corr = var_corr if np.abs(var_corr) > 0. else 1.
    else:
        var_corr = 1.
    return f * var_corr / norm_value


def apply_prec(grad, prec, prec_scale=4.0, prec_smooth=1.0):
    """stride/problem/data.py:409-446."""
    grad = np.array(grad, np.float64)
    prec = np.array(prec, np.float64)
    prec = scipy.ndimage.gaussian_filter(prec, prec_smooth)
    factor = np.sum(np.abs(prec))
    if factor > 1e-31:
        factor = prec_scale * prec.size / factor
        prec = prec * factor + 1
        nz = np.abs(prec) > 0
        grad[nz] /= prec[nz]
    return grad


def descent(x, direction, step, slowness=True, lo=None, hi=None):
    """stride/optimisation/optimisers/optimiser.py:242-330 and steps/clip.py:28-38."""
    x = np.asarray(x, np.float64)
    v = 1 / (1 / x - step * direction) if slowness else x - step * direction
    if lo is not None or hi is not None:
        v = np.clip(v, lo, hi)
    return v
