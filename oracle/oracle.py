"""ctypes front-end to the CPU oracle (acoustic_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py, always as the checker. The product package never
imports this module.
"""

import contextlib
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


class OracleCfg(ctypes.Structure):
    _fields_ = [('dim', ctypes.c_int), ('n', ctypes.c_int * 3), ('h', ctypes.c_double * 3),
                ('nabs', ctypes.c_int * 3), ('dt', ctypes.c_double), ('nt', ctypes.c_int),
                ('kernel', ctypes.c_int), ('rho_mode', ctypes.c_int), ('att_power', ctypes.c_int),
                ('boundary', ctypes.c_int), ('diff_source', ctypes.c_int), ('undersampling', ctypes.c_int),
                ('save_raw', ctypes.c_int), ('tb0', ctypes.c_int), ('tb1', ctypes.c_int),
                ('grad_vp', ctypes.c_int), ('grad_rho', ctypes.c_int), ('grad_src', ctypes.c_int)]


def build():
    subprocess.run(['make', '-s', '-C', _HERE], check=True)


_PLANTED = [0]


def _load(name):
    if name not in _LIBS:
        path = os.path.join(_HERE, '_build', 'liboracle_%s.so' % name)
        if not os.path.exists(path):
            build()
        _LIBS[name] = ctypes.CDLL(path)
    return _LIBS[name]


def _lib(precision):
    if precision == 'f64' and _PLANTED[0]:
        return _load('mut_f64')
    return _load(precision)


@contextlib.contextmanager
def planted(code):
    """FP64 oracle calls inside the block run with one planted mistake (the MUT_* codes of
    acoustic_oracle.c, from the separate mutation build). Only the oracle's own physics
    checks use this, to show that each check fails when the term it pins is wrong."""
    lib = _load('mut_f64')
    lib.oracle_set_mutation_f64(ctypes.c_int(code))
    _PLANTED[0] = code
    try:
        yield
    finally:
        lib.oracle_set_mutation_f64(ctypes.c_int(0))
        _PLANTED[0] = 0


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _cfg(plan, **override):
    d = plan.config()
    d.update(override)
    c = OracleCfg()
    c.dim = d['dim']
    for i in range(3):
        c.n[i] = d['shape'][i] if i < d['dim'] else 1
        c.h[i] = d['spacing'][i] if i < d['dim'] else 1.0
        c.nabs[i] = d['absorbing'][i] if i < d['dim'] else 0
    for key in ('dt', 'nt', 'kernel', 'rho_mode', 'att_power', 'boundary', 'diff_source', 'undersampling',
                'save_raw', 'tb0', 'tb1', 'grad_vp', 'grad_rho', 'grad_src'):
        setattr(c, key, d[key])
    return c


def _medium(plan):
    rho = _f32(plan.rho)
    alpha = _f32(plan.alpha)
    damp = plan.sponge_volume() if plan.boundary == 0 else None
    pml = _f32(plan.pml)
    pml_stride = pml.shape[-1] if pml is not None else 0
    return rho, alpha, damp, pml, pml_stride


def forward(plan, save=False, precision='f32', final_state=False, save_every=None, save_raw=None):
    """Run the forward operator. Returns (traces (nrec, nt), saved or None, final or None)."""
    lib = _lib(precision)
    over = {}
    if save_every is not None:
        over['undersampling'] = save_every
    if save_raw is not None:
        over['save_raw'] = int(save_raw)
    cfg = _cfg(plan, **over)
    rho, alpha, damp, pml, pml_stride = _medium(plan)
    rec = np.zeros((plan.nt, plan.nrec), dtype=np.float32)
    nsave = (plan.tb1 // cfg.undersampling + 1) if save else 0
    saved = np.zeros((nsave, plan.num_points), dtype=np.float32) if save else None
    final = np.zeros((2, plan.num_points), dtype=np.float32) if final_state else None
    fn = getattr(lib, 'oracle_forward_%s' % precision)
    fn.restype = ctypes.c_int
    rc = fn(ctypes.byref(cfg), _ptr(plan.vp), _ptr(rho), _ptr(alpha), _ptr(damp), _ptr(pml),
            ctypes.c_int(pml_stride),
            ctypes.c_int(plan.nsrc), ctypes.c_int(plan.src_idx.shape[1]), _ptr(plan.src_idx), _ptr(plan.src_w),
            _ptr(plan.src_data),
            ctypes.c_int(plan.nrec), ctypes.c_int(plan.rec_idx.shape[1]), _ptr(plan.rec_idx), _ptr(plan.rec_w),
            _ptr(rec), _ptr(saved), ctypes.c_int(nsave), _ptr(final))
    if rc != 0:
        raise RuntimeError('oracle_forward failed (%d)' % rc)
    shape = plan.shape
    if saved is not None:
        saved = saved.reshape((nsave,) + shape)
    if final is not None:
        final = final.reshape((2,) + shape)
    return rec.T.copy(), saved, final


def adjoint_compat(plan, residual_samples, saved, precision='f32', final_state=False):
    """As-written adjoint. residual_samples: (nt, nrec) already windowed.
    Returns dict of full-grid accumulators (grad_vp, prec_vp, grad_rho, prec_rho),
    grad_src (nsrc, nt) and optionally the final adjoint state."""
    lib = _lib(precision)
    cfg = _cfg(plan)
    rho, alpha, damp, pml, pml_stride = _medium(plan)
    n = plan.num_points
    out = {k: np.zeros(n, dtype=np.float32) for k in ('grad_vp', 'prec_vp', 'grad_rho', 'prec_rho')}
    gsrc = np.zeros((plan.nt, plan.nsrc), dtype=np.float32)
    final = np.zeros((2, n), dtype=np.float32) if final_state else None
    saved = np.ascontiguousarray(saved, dtype=np.float32).reshape(saved.shape[0], -1)
    fn = getattr(lib, 'oracle_adjoint_compat_%s' % precision)
    rc = fn(ctypes.byref(cfg), _ptr(plan.vp), _ptr(rho), _ptr(alpha), _ptr(damp), _ptr(pml),
            ctypes.c_int(pml_stride),
            ctypes.c_int(plan.nsrc), ctypes.c_int(plan.src_idx.shape[1]), _ptr(plan.src_idx), _ptr(plan.src_w),
            ctypes.c_int(plan.nrec), ctypes.c_int(plan.rec_idx.shape[1]), _ptr(plan.rec_idx), _ptr(plan.rec_w),
            _ptr(np.ascontiguousarray(residual_samples, dtype=np.float32)),
            _ptr(saved), ctypes.c_int(saved.shape[0]),
            _ptr(out['grad_vp']), _ptr(out['prec_vp']), _ptr(out['grad_rho']), _ptr(out['prec_rho']),
            _ptr(gsrc) if plan.grad_src else None, _ptr(final))
    if rc != 0:
        raise RuntimeError('oracle_adjoint_compat failed (%d)' % rc)
    res = {k: v.reshape(plan.shape) for k, v in out.items()}
    res['grad_src'] = gsrc.T.copy()
    if final is not None:
        res['final'] = final.reshape((2,) + plan.shape)
    return res


def adjoint_exact(plan, residual_samples, states, precision='f32', final_state=False):
    """Transpose of the forward map for J = sum <res, rec>. states: (nt, *shape) raw
    forward levels u_0..u_{nt-1}. Returns the coefficient accumulators and the
    adjoint field at the source taps."""
    lib = _lib(precision)
    cfg = _cfg(plan)
    rho, alpha, damp, _, _ = _medium(plan)
    n = plan.num_points
    out = {k: np.zeros(n, dtype=np.float32) for k in ('gA', 'gB', 'gG', 'gm', 'gb')}
    ntap = plan.src_idx.shape[1]
    lam = np.zeros((plan.nt, plan.nsrc * ntap), dtype=np.float32)
    final = np.zeros((2, n), dtype=np.float32) if final_state else None
    states = np.ascontiguousarray(states, dtype=np.float32).reshape(states.shape[0], -1)
    fn = getattr(lib, 'oracle_adjoint_exact_%s' % precision)
    rc = fn(ctypes.byref(cfg), _ptr(plan.vp), _ptr(rho), _ptr(alpha), _ptr(damp),
            ctypes.c_int(plan.nsrc), ctypes.c_int(ntap), _ptr(plan.src_idx),
            ctypes.c_int(plan.nrec), ctypes.c_int(plan.rec_idx.shape[1]), _ptr(plan.rec_idx), _ptr(plan.rec_w),
            _ptr(np.ascontiguousarray(residual_samples, dtype=np.float32)), _ptr(states),
            _ptr(out['gA']), _ptr(out['gB']), _ptr(out['gG']), _ptr(out['gm']), _ptr(out['gb']),
            _ptr(lam), _ptr(final))
    if rc != 0:
        raise RuntimeError('oracle_adjoint_exact: configuration not supported (%d)' % rc)
    res = {k: v.reshape(plan.shape) for k, v in out.items()}
    res['lam_src'] = lam.reshape(plan.nt, plan.nsrc, ntap)
    if final is not None:
        res['final'] = final.reshape((2,) + plan.shape)
    return res


def forward_states(plan, precision='f32'):
    """All forward levels u_0..u_{nt-1} (raw, every step)."""
    _, saved, _ = forward(plan, save=True, precision=precision, save_every=1, save_raw=True)
    # slot t holds p[t] for t = 1..tb1-2; u_{nt-1} is the last computed level
    states = np.zeros((plan.nt,) + plan.shape, dtype=np.float32)
    states[:saved.shape[0]] = saved[:plan.nt]
    _, _, final = forward(plan, save=False, precision=precision, final_state=True)
    states[plan.tb1 - 1] = final[0]
    return states


def apply_K(shape, spacing, u, rho_mode=0, rho=None, precision='f64'):
    """Spatial operator alone (SO10 Laplacian or staggered div(b grad))."""
    lib = _lib(precision)
    c = OracleCfg()
    c.dim = len(shape)
    for i in range(3):
        c.n[i] = shape[i] if i < c.dim else 1
        c.h[i] = spacing[i] if i < c.dim else 1.0
    c.rho_mode = rho_mode
    u = np.ascontiguousarray(u, dtype=np.float32)
    out = np.zeros(shape, np.float32)
    fn = getattr(lib, 'oracle_apply_K_%s' % precision)
    fn(ctypes.byref(c), _ptr(_f32(rho)), _ptr(u), _ptr(out))
    return out
