"""3D CFS-PML on the fused updates: the specialised update (no density) and the fused density
update run without a boundary term, with the auxiliary fields of every axis advanced by one launch
over the absorbing layers before the update and their term (- Cpml sum_d u1_d) applied after it
(stride/physics/boundaries/devito.py:72-193). Against the oracle and against the separate general
passes (k_pml per axis + k_update), forward and as-written adjoint.

Grids: 44^3 extended (one tile across axis 2) and 86^3 extended (two tiles, several chunks).
Bars: traces 1e-5 and gradients 1e-4 relative L2 against the FP32 oracle; fused against general
passes 1e-5 (the same terms summed in another order)."""

import os

import numpy as np
import pytest

import oracle
from helpers import small_case, solver_for, rel_l2
from stride_amd.physics.plan import adjoint_source_samples

pytestmark = pytest.mark.gpu

PML = 'complex_frequency_shift_PML_2'

CASES = [
    dict(rho='none'),
    dict(rho='none', alpha='p0'),
    dict(rho='none', alpha='p2', attenuation_power=2),
    dict(rho='var', alpha='p0'),
    dict(rho='const', alpha='none-power'),
    dict(rho='var', alpha='p2', attenuation_power=2),
]


def _label(o):
    return '-'.join('%s=%s' % kv for kv in sorted(o.items()))


def _case(n=28, nabs=8, nt=120, **opts):
    rho = opts.pop('rho')
    alpha = opts.pop('alpha', None)
    return small_case(dim=3, n=n, nabs=nabs, nt=nt, rho=rho, alpha=alpha,
                      kernel='OT2', nrec=20, boundary_type=PML, **opts)[0]


def _general(plan, save_mode=0):
    """A solver on the separate general passes: the fused density update switched off, and the
    specialised update off while it is set up (the switch is read when the scheme is prepared and
    at each step)."""
    os.environ['STRIDE_AMD_GENERAL_ONLY'] = '1'
    s = solver_for(plan, save_mode=save_mode)
    s.set_tile_density(0)
    return s


def _run(plan, general, save_mode=0, res=None, want_rho=False):
    try:
        s = _general(plan, save_mode) if general else solver_for(plan, save_mode=save_mode)
        rec = s.forward(plan.src_data)
        g = s.adjoint(res, want_vp=True, want_rho=want_rho) if res is not None else None
        s.close()
    finally:
        os.environ.pop('STRIDE_AMD_GENERAL_ONLY', None)
    return rec, g


@pytest.mark.parametrize('opts', CASES, ids=[_label(o) for o in CASES])
def test_pml3d_forward_fused_matches_general_and_oracle(gpu, opts):
    plan = _case(**dict(opts))
    assert plan.boundary == 1
    print('pml3d forward', _label(opts), plan.shape, flush=True)
    ref, _, _ = oracle.forward(plan)
    fused, _ = _run(plan, False)
    general, _ = _run(plan, True)
    e_ref, e_gen = rel_l2(fused.T, ref), rel_l2(fused, general)
    print('  vs oracle %.3e, vs general passes %.3e' % (e_ref, e_gen))
    assert np.abs(ref).max() > 0
    assert e_ref <= 1e-5
    assert e_gen <= 1e-5
    assert rel_l2(general.T, ref) <= 1e-5


@pytest.mark.parametrize('opts', [dict(rho='none', alpha='p0'), dict(rho='var', alpha='p0')],
                         ids=['vp-only', 'variable-density'])
def test_pml3d_large_grid_adjoint(gpu, opts):
    """86^3 extended: forward and as-written adjoint on the fused updates with CFS-PML (p.dt2
    saves and the fused vp gradient without density; raw saves and the vp / rho gradient pass with
    variable density) against the general passes and the oracle."""
    dens = opts['rho'] != 'none'
    plan = _case(n=70, nabs=8, nt=100, grad_vp=True, grad_rho=dens, **dict(opts))
    print('pml3d adjoint', _label(opts), plan.shape, flush=True)
    save_mode = 2 if dens else 1
    rng = np.random.default_rng(23)
    res = adjoint_source_samples(plan, rng.standard_normal((plan.nrec, plan.nt)).astype(np.float32))
    rec_f, g_f = _run(plan, False, save_mode, res, want_rho=dens)
    rec_g, g_g = _run(plan, True, save_mode, res, want_rho=dens)
    ref, saved, _ = oracle.forward(plan, save=True)
    adj = oracle.adjoint_compat(plan, res, saved)
    assert rel_l2(rec_f.T, ref) <= 1e-5
    assert rel_l2(rec_f, rec_g) <= 1e-5
    for key in ('grad_vp', 'prec_vp') + (('grad_rho', 'prec_rho') if dens else ()):
        e_ref, e_gen = rel_l2(g_f[key], adj[key]), rel_l2(g_f[key], g_g[key])
        print('  %s vs oracle %.3e, vs general passes %.3e' % (key, e_ref, e_gen))
        assert np.abs(adj[key]).max() > 0
        assert e_ref <= 1e-4, key
        assert e_gen <= 1e-5, key
