"""Shared builders for parity tests: small grids exercising every scheme option."""

import numpy as np

from stride_amd.problem import Space, Time
from stride_amd.utils import tone_burst, elliptical, ellipsoidal
from stride_amd.physics.plan import build_plan


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def small_case(dim=2, n=48, nabs=10, nt=220, dx=0.5e-3, dt=0.08e-6, rho='none', alpha=None, seed=0,
               nrec=12, **opts):
    shape = (n,) * dim
    space = Space(shape=shape, spacing=dx, extra=(nabs,) * dim, absorbing=(nabs,) * dim)
    time = Time(start=0., step=dt, num=nt)
    rng = np.random.default_rng(seed)
    ext = space.extended_shape
    vp = (1500. + 40. * rng.standard_normal(ext)).astype(np.float32)
    rho_arr = None
    if rho == 'const':
        rho_arr = np.full(ext, 1000., np.float32)
    elif rho == 'var':
        rho_arr = (1000. + 60. * rng.standard_normal(ext)).astype(np.float32)
    alpha_arr = None
    if alpha is not None:
        alpha_arr = (0.5 + 0.2 * rng.random(ext)).astype(np.float32)
        if alpha == 'none-power':
            alpha_arr = (0.9995 + 0.0004 * rng.random(ext)).astype(np.float32)
    centre = tuple(l / 2 for l in space.limit)
    r = 0.3 * min(space.limit)
    if dim == 2:
        geom = elliptical(nrec, (r, r), centre)
    else:
        geom = ellipsoidal(nrec, (r, r, r), centre, threshold=0.5)
    wav = tone_burst(0.5e6, 3, nt, dt)[None]
    kw = dict(source_coordinates=geom[:1], receiver_coordinates=geom)
    kw.update(opts)
    if alpha == 'none-power':
        kw['attenuation_power'] = None
    plan = build_plan(space, time, wav, vp, rho=rho_arr, alpha=alpha_arr, **kw)
    return plan, dict(space=space, time=time, vp=vp, rho=rho_arr, alpha=alpha_arr, wavelet=wav, kw=kw)


def solver_for(plan, save_mode=0):
    from stride_amd._native import Solver
    s = Solver(plan.dim, plan.shape, plan.spacing, plan.absorbing, plan.dt, plan.nt)
    s.set_scheme(plan.kernel_order, plan.rho_mode, plan.att_power, plan.boundary, plan.diff_source,
                 plan.undersampling, save_mode, plan.tb0, plan.tb1, plan.grad_vp, plan.grad_rho, plan.grad_src)
    s.set_medium(plan.vp, plan.rho, plan.alpha)
    if plan.boundary == 0:
        s.set_sponge(plan.sponge_tables)
    elif plan.boundary == 1:
        s.set_pml(plan.pml)
    s.set_sources(plan.src_idx, plan.src_w)
    s.set_receivers(plan.rec_idx, plan.rec_w)
    return s
