"""Per-shot scheme selection (``IsoAcousticDevito._check_conditions``,
stride/physics/iso_acoustic/devito.py:1354-1464, ``_dt_max`` :1688-1689).

From the non-zero wavelets' -10 dB bandwidth and the velocity range it picks the
time kernel (OT2 if dt is within the OT2 stability limit, else OT4) and the
forward-wavefield undersampling factor ``clamp(int(1/(4 f_max dt)), 4, 10)``.
"""

import numpy as np

from ..utils import bandwidth

__all__ = ['Conditions', 'check_conditions', 'dt_max']


class Conditions:
    def __init__(self, **kw):
        self.__dict__.update(kw)

    def messages(self):
        return list(self._messages)


def dt_max(k, h, vp_max, dim):
    return k * h / vp_max * 1 / np.sqrt(dim)


def check_conditions(wavelets, vp_extended, space, dt, preferred_kernel=None,
                     preferred_undersampling=None):
    wavelets = np.asarray(wavelets)
    vp_min = float(np.min(vp_extended))
    vp_max = float(np.max(vp_extended))
    f_mins, f_centres, f_maxs = [], [], []
    if wavelets.ndim > 1:
        for row in wavelets:
            if np.any(row):
                lo, mid, hi = bandwidth(row, dt, cutoff=-10)
                f_mins.append(lo)
                f_centres.append(mid)
                f_maxs.append(hi)
    f_min = np.min(f_mins) if f_mins else 0.05 / dt
    f_max = np.max(f_maxs) if f_maxs else 0.05 / dt
    f_centre = np.median(f_centres) if f_centres else 0.05 / dt

    msgs = ['Estimated bandwidth for the propagated wavelet %.3f-%.3f MHz' % (f_min / 1e6, f_max / 1e6)]
    h = max(*space.spacing) if space.dim > 1 else space.spacing[0]
    wavelength = vp_min / f_max
    ppw = wavelength / h
    h_lim = wavelength / 5
    relation = 'higher than' if h > h_lim else 'below'
    msgs.append('Spatial grid spacing (%.3f mm | %.3f PPW) is %s dispersion limit (%.3f mm | %.3f PPW)'
                % (h / 1e-3, ppw, relation, h_lim / 1e-3, 5))

    limit_ot2 = dt_max(2.0 / np.pi, h, vp_max, space.dim)
    limit_ot4 = dt_max(3.6 / np.pi, h, vp_max, space.dim)
    crossing = dt * vp_max / h * 100
    if dt <= limit_ot2:
        kernel = 'OT2'
        msgs.append('Time grid spacing (%.3f μs | %d%%) is below OT2 limit (%.3f μs)'
                    % (dt / 1e-6, crossing, limit_ot2 / 1e-6))
    elif dt <= limit_ot4:
        kernel = 'OT4'
        msgs.append('Time grid spacing (%.3f μs | %d%%) is above OT2 limit (%.3f μs) and '
                    'below OT4 limit (%.3f μs)' % (dt / 1e-6, crossing, limit_ot2 / 1e-6,
                                                       limit_ot4 / 1e-6))
    else:
        kernel = 'OT4'
        msgs.append('Time grid spacing (%.3f μs | %d%%) is above OT4 limit (%.3f μs)'
                    % (dt / 1e-6, crossing, limit_ot4 / 1e-6))
    kernel = kernel if preferred_kernel is None else preferred_kernel

    undersampling = min(max(4, int((1 / (4 * f_max)) / dt)), 10)
    if preferred_undersampling is not None:
        undersampling = preferred_undersampling
    msgs.append('Selected undersampling level %d' % undersampling)
    msgs.append('Selected time stepping scheme %s' % kernel)

    return Conditions(f_min=f_min, f_centre=f_centre, f_max=f_max, ppw=ppw, h_limit=h_lim,
                      dt_max_ot2=limit_ot2, dt_max_ot4=limit_ot4, kernel=kernel,
                      undersampling=undersampling, _messages=msgs)
