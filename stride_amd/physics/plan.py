"""Host-side preparation of one shot (the numpy half of ``before_forward`` /
``before_adjoint``, stride/physics/iso_acoustic/devito.py:220-513, 689-873).

``ShotPlan`` gathers everything the device needs, in the reference's layouts:

- scheme codes (kernel, density mode, attenuation power, boundary), picked by
  ``_check_conditions`` rules (devito.py:1354-1464) unless forced;
- medium arrays on the extended grid, float32: vp; rho (variable density only);
  alpha converted from dB to nepers with ``100 (1e-6/2pi)^pw / (20 log10 e)``
  (devito.py:487-497);
- boundary description: per-axis sponge layer tables (device), the full sponge
  ``damp`` volume (checker only), CFS-PML profiles;
- source / receiver taps (linear or Kaiser-sinc), coordinates shifted by 1e-3 h
  (devito.py:307-309);
- source samples ``wavelets.T * tukey(0.001)`` (time-major ``(nt, nsrc)``),
  differentiated with ``np.gradient`` when ``diff_source`` (devito.py:500-509).
"""

import numpy as np

from .boundaries import (sponge_tables, sponge_damp, pml_profiles, boundary_code, SPONGE, PML)
from .conditions import check_conditions
from .sparse import point_taps
from ..utils import tukey

__all__ = ['ShotPlan', 'build_plan', 'adjoint_source_samples', 'ATT_NONE', 'ATT_POWER_NONE',
           'RHO_NONE', 'RHO_CONSTANT', 'RHO_VARIABLE']

RHO_NONE, RHO_CONSTANT, RHO_VARIABLE = 0, 1, 2
ATT_NONE = -2        # no attenuation field given
ATT_POWER_NONE = -1  # attenuation_power=None: p+ <- alpha p+


class ShotPlan:
    def __init__(self, **kw):
        self.__dict__.update(kw)

    @property
    def num_points(self):
        return int(np.prod(self.shape))

    @property
    def nsave(self):
        """Snapshot buffer length for the every-f-th-step save (slot t // f)."""
        return self.tb1 // self.undersampling + 1

    def config(self):
        return dict(dim=self.dim, shape=self.shape, spacing=self.spacing, absorbing=self.absorbing,
                    dt=self.dt, nt=self.nt, kernel=self.kernel_order, rho_mode=self.rho_mode,
                    att_power=self.att_power, boundary=self.boundary, diff_source=int(self.diff_source),
                    undersampling=self.undersampling, save_raw=int(self.save_raw), tb0=self.tb0,
                    tb1=self.tb1, grad_vp=int(self.grad_vp), grad_rho=int(self.grad_rho),
                    grad_src=int(self.grad_src))


def _db_to_neper(power):
    if power is None:
        return 1.
    return 100 * (1e-6 / (2 * np.pi)) ** power / (20 * np.log10(np.exp(1)))


def build_plan(space, time, wavelets, vp, rho=None, alpha=None, source_coordinates=None,
               receiver_coordinates=None, kernel=None, boundary_type='sponge_boundary_2',
               interpolation_type='linear', attenuation_power=0, save_undersampling=None,
               time_bounds=None, diff_source=False, grad_vp=False, grad_rho=False, grad_alpha=False,
               grad_src=False, **unused):
    """vp / rho / alpha: extended-grid arrays (``ScalarField.extended_data``)."""
    vp = np.ascontiguousarray(vp, dtype=np.float32)
    if vp.shape != tuple(space.extended_shape):
        raise ValueError('vp has shape %s, expected the extended shape %s'
                         % (vp.shape, tuple(space.extended_shape)))
    wavelets = np.asarray(wavelets, dtype=np.float64)
    if wavelets.ndim == 1:
        wavelets = wavelets[None]
    nt = time.num
    dt = float(time.step)

    cond = check_conditions(wavelets, vp, space, dt, preferred_kernel=kernel,
                            preferred_undersampling=save_undersampling)
    if cond.kernel not in ('OT2', 'OT4'):
        raise ValueError('Unrecognized kernel')

    if attenuation_power not in (0, 2, None):
        raise ValueError('The "attenuation_exponent" can only take values (0, 2).')
    boundary = boundary_code(boundary_type)

    if rho is not None:
        rho = np.ascontiguousarray(rho, dtype=np.float32)
        rho_mode = RHO_CONSTANT if np.isclose(np.min(rho), np.max(rho)) else RHO_VARIABLE
    else:
        rho_mode = RHO_NONE

    if alpha is not None:
        att_power = ATT_POWER_NONE if attenuation_power is None else attenuation_power
        alpha_eff = (np.asarray(alpha, dtype=np.float32) * _db_to_neper(attenuation_power)).astype(np.float32)
    else:
        att_power = ATT_NONE
        alpha_eff = None

    tb0, tb1 = time_bounds if time_bounds is not None else (0, nt)

    # source samples, time-major
    src = wavelets
    if diff_source:
        src = np.gradient(src, dt, axis=-1)
    window = np.pad(tukey(tb1, 0.001), ((0, nt - tb1),), mode='constant')
    src_data = np.ascontiguousarray((src.T * window[:, None]).astype(np.float32))

    src_idx, src_w = point_taps(space, source_coordinates, interpolation_type, smooth=True)
    rec_idx, rec_w = point_taps(space, receiver_coordinates, interpolation_type, smooth=False)

    save_raw = bool(grad_rho or grad_alpha or grad_src)

    plan = ShotPlan(
        space=space, dim=space.dim, shape=tuple(space.extended_shape), spacing=tuple(space.spacing),
        absorbing=tuple(space.absorbing), extra=tuple(space.extra), dt=dt, nt=nt,
        kernel=cond.kernel, kernel_order=2 if cond.kernel == 'OT2' else 4,
        rho_mode=rho_mode, att_power=att_power, attenuation_power=attenuation_power,
        boundary=boundary, boundary_type=boundary_type, interpolation_type=interpolation_type,
        diff_source=bool(diff_source), undersampling=int(cond.undersampling), save_raw=save_raw,
        tb0=int(tb0), tb1=int(tb1), grad_vp=bool(grad_vp), grad_rho=bool(grad_rho),
        grad_alpha=bool(grad_alpha), grad_src=bool(grad_src), conditions=cond,
        vp=vp, rho=rho if rho_mode == RHO_VARIABLE else None, rho_input=rho,
        alpha=alpha_eff,
        sponge_tables=sponge_tables(space) if boundary == SPONGE else None,
        pml=pml_profiles(space, vp, cond.f_centre) if boundary == PML else None,
        src_idx=src_idx, src_w=src_w, rec_idx=rec_idx, rec_w=rec_w,
        src_data=src_data, nsrc=src_idx.shape[0], nrec=rec_idx.shape[0],
    )
    plan._sponge_damp = None
    return plan


def sponge_volume(plan):
    """Full-grid sponge ``damp`` (checker path; the device evaluates it on the fly)."""
    if plan._sponge_damp is None:
        plan._sponge_damp = np.ascontiguousarray(
            sponge_damp(plan.space, plan.vp, plan.dt).astype(np.float32))
    return plan._sponge_damp


ShotPlan.sponge_volume = sponge_volume


def adjoint_source_samples(plan, adjoint_source):
    """Residual traces (nrec, nt) times Tukey(0.001) over the time bounds,
    time-major (devito.py:863-869)."""
    data = np.asarray(adjoint_source, dtype=np.float64)
    window = tukey(plan.tb1 - plan.tb0, 0.001)
    window = np.pad(window, ((plan.tb0, plan.nt - plan.tb1),), mode='constant')
    return np.ascontiguousarray((data.T * window[:, None]).astype(np.float32))
