"""Gradient post-processing.

``compat_finish`` applies the reference's rescale of the accumulated vp
gradient (``get_grad_vp``, stride/physics/iso_acoustic/devito.py:1062-1081):
``grad *= 1/vp, prec *= 1/vp^2`` when vp carries the slowness transform (the
default for fields named vp), else ``grad *= -2/vp^3, prec *= (2/vp)**3**2``.

``exact_gradients`` is the chain rule of the exact adjoint. The device
accumulates, per grid point, the sensitivities of J to the step coefficients
of

    u[t+1] = A u[t] + B u[t-1] + G K u[t] (+ OT4 term in m) + sources,

with A = W (2 - e)/(1 + a), B = -W (1 - a)/(1 + a), G = W dt^2 m/(1 + a),
m = vp^2 (rho var: vp^2 rho), a = dt (alpha vp + vp^2 D), e = dt^2 vp^2 D^2,
D = 7 dt vp s (sponge layer profile s), W = alpha for attenuation power None.
Their derivatives with respect to vp, rho and alpha are taken by complex step
(exact to rounding for these rational functions), so the chain rule has no
hand-derived terms.
"""

import numpy as np

from .boundaries import damping, reflection_coefficient, SPONGE
from .plan import RHO_VARIABLE, ATT_NONE, ATT_POWER_NONE

__all__ = ['compat_finish', 'step_coefficients', 'exact_gradients', 'sponge_profile']


def compat_finish(grad, prec, vp_inner, transform=None):
    is_slowness = False
    if transform is not None:
        try:
            is_slowness = bool(np.isclose(1 / transform(10), 10))
        except Exception:
            pass
    if is_slowness:
        grad = grad * (1 / vp_inner)
        prec = prec * (1 / vp_inner) ** 2
    else:
        grad = grad * (-2 / vp_inner ** 3)
        prec = prec * (2 / vp_inner) ** 3 ** 2
    return grad.astype(np.float32), prec.astype(np.float32)


def sponge_profile(plan):
    """Sine sponge layer profile without the velocity factor (the ``damp`` volume
    is 7 dt vp times this)."""
    if plan.boundary != SPONGE:
        return np.zeros(plan.shape)
    refl = reflection_coefficient(plan.space.absorbing)
    return damping(plan.space, velocity=1.0, reflection_coefficient=refl).astype(np.float64)


def step_coefficients(plan, vp, rho=None, alpha=None, profile=None):
    """A, B, G, m, b per point (any dtype incl. complex)."""
    dt = plan.dt
    v = vp
    m = v * v
    if plan.rho_mode == RHO_VARIABLE:
        m = m * rho
    a = 0 * v
    e = 0 * v
    W = 1 + 0 * v
    if plan.att_power == 0:
        a = a + dt * alpha * v
    if plan.boundary == SPONGE:
        D = 7 * dt * v * profile
        a = a + dt * v * v * D
        e = dt * dt * v * v * D * D
    if plan.att_power == ATT_POWER_NONE:
        W = alpha
    ia = 1 / (1 + a)
    out = dict(A=W * (2 - e) * ia, B=-W * (1 - a) * ia, m=m)
    out['G'] = W * dt * dt * ia * m
    out['b'] = 1 / rho if plan.rho_mode == RHO_VARIABLE else 0 * v
    return out


def _complex_step(plan, vp, rho, alpha, profile, which):
    eps = 1e-30
    args = dict(vp=vp.astype(np.complex128), rho=None if rho is None else rho.astype(np.complex128),
                alpha=None if alpha is None else alpha.astype(np.complex128))
    args[which] = args[which] + 1j * eps
    coef = step_coefficients(plan, args['vp'], args['rho'], args['alpha'], profile)
    return {k: np.imag(c) / eps for k, c in coef.items()}


def exact_gradients(plan, acc, lam_src=None):
    """acc: dict of full-grid accumulators gA, gB, gG, gm, gb. Returns dJ/dvp,
    dJ/drho, dJ/dalpha (alpha in the caller's dB units), dJ/dwavelets."""
    vp = plan.vp.astype(np.float64)
    rho = plan.rho_input.astype(np.float64) if plan.rho_input is not None else None
    alpha = plan.alpha.astype(np.float64) if plan.alpha is not None else None
    profile = sponge_profile(plan)
    keys = ('A', 'B', 'G', 'm', 'b')
    g = {k: acc.get('g' + k) for k in keys}

    def contract(d):
        total = np.zeros(plan.shape)
        for k in keys:
            if g[k] is not None:
                total += g[k].astype(np.float64) * d[k]
        return total

    out = {}
    out['vp'] = contract(_complex_step(plan, vp, rho, alpha, profile, 'vp'))
    if plan.rho_mode == RHO_VARIABLE:
        out['rho'] = contract(_complex_step(plan, vp, rho, alpha, profile, 'rho'))
    if plan.att_power != ATT_NONE:
        scale = 1. if plan.attenuation_power is None else \
            100 * (1e-6 / (2 * np.pi)) ** plan.attenuation_power / (20 * np.log10(np.exp(1)))
        out['alpha'] = contract(_complex_step(plan, vp, rho, alpha, profile, 'alpha')) * scale

    if lam_src is not None:
        # source term: c_tap = w * scale * vp[idx]
        hmax = max(plan.spacing)
        scale = 2 * plan.dt ** 2 / hmax if plan.diff_source else 2 * plan.dt / hmax
        lam = lam_src.astype(np.float64)                       # (nt, nsrc, ntap)
        src = plan.src_data.astype(np.float64)                  # (nt, nsrc)
        w = plan.src_w.astype(np.float64)                       # (nsrc, ntap)
        idx = plan.src_idx
        per_tap = np.einsum('tsq,ts->sq', lam, src) * w * scale
        flat = out['vp'].reshape(-1)
        valid = idx >= 0
        np.add.at(flat, idx[valid], per_tap[valid])
        vpt = np.where(valid, vp.reshape(-1)[np.where(valid, idx, 0)], 0.)
        coef = w * scale * vpt
        out['src_samples'] = np.einsum('tsq,sq->ts', lam, coef)    # dJ/d src_data (nt, nsrc)
    return out
