"""OT4 on the fused 3D updates (sponge boundary, attenuation off / power 0 / None): on the
specialised update (no density): a first pass forms K u and w = vp^2 K u, the update stages w, so its Laplacian is K(m K u)
of the fourth-order time stepping (stride/physics/iso_acoustic/devito.py:1466-1608), and reads u
and K u as own-cell operands; on the fused density update the same two passes with K = div(b grad)
and m = vp^2 rho. Against the oracle and against the separate general passes, forward and
as-written adjoint.

Grids: 44^3 and 51^3 extended (one tile across axis 2) and 86^3 extended (two tiles, several
chunks).
Bars: traces 1e-5 and gradients 1e-4 relative L2 against the FP32 oracle; fused against general
passes 1e-5 (the same terms summed in another order)."""

import os

import numpy as np
import pytest

import oracle
from helpers import small_case, solver_for, rel_l2
from stride_amd.physics.plan import adjoint_source_samples

pytestmark = pytest.mark.gpu

CASES = [
    dict(),
    dict(alpha='p0'),
    dict(alpha='none-power'),
    dict(n=37, nabs=7),  # partial tiles on every axis
    # the fused density update in two passes (K u and m K u; then the update staging m K u)
    dict(rho='var', alpha='p0'),
    dict(rho='const'),
]


def _label(o):
    return '-'.join('%s=%s' % kv for kv in sorted(o.items())) or 'plain'


def _case(n=28, nabs=8, nt=120, **opts):
    alpha = opts.pop('alpha', None)
    rho = opts.pop('rho', 'none')
    return small_case(dim=3, n=n, nabs=nabs, nt=nt, rho=rho, alpha=alpha, kernel='OT4', nrec=20, **opts)[0]


def _run(plan, general, save_mode=0, res=None, want_rho=False):
    try:
        if general:
            os.environ['STRIDE_AMD_GENERAL_ONLY'] = '1'
        s = solver_for(plan, save_mode=save_mode)
        if general:
            s.set_tile_density(0)  # the density schemes' fused update off as well
        rec = s.forward(plan.src_data)
        g = s.adjoint(res, want_vp=True, want_rho=want_rho) if res is not None else None
        s.close()
    finally:
        os.environ.pop('STRIDE_AMD_GENERAL_ONLY', None)
    return rec, g


@pytest.mark.parametrize('opts', CASES, ids=[_label(o) for o in CASES])
def test_ot4_3d_forward_fused_matches_general_and_oracle(gpu, opts):
    plan = _case(**dict(opts))
    assert plan.kernel_order == 4
    print('ot4 3d forward', _label(opts), plan.shape, flush=True)
    ref, _, _ = oracle.forward(plan)
    fused, _ = _run(plan, False)
    general, _ = _run(plan, True)
    e_ref, e_gen = rel_l2(fused.T, ref), rel_l2(fused, general)
    print('  vs oracle %.3e, vs general passes %.3e' % (e_ref, e_gen))
    assert np.abs(ref).max() > 0
    assert e_ref <= 1e-5
    assert e_gen <= 1e-5


@pytest.mark.parametrize('rho', ['none', 'var'], ids=['vp-only', 'variable-density'])
def test_ot4_3d_large_grid_adjoint(gpu, rho):
    """86^3 extended, attenuation power 0: the as-written adjoint against the general passes and
    the oracle (no density: p.dt2 saves and the vp gradient fused into the second pass; variable
    density: raw saves, vp and rho gradients from the gradient pass)."""
    dens = rho != 'none'
    plan = _case(n=70, nabs=8, nt=100, alpha='p0', grad_vp=True, grad_rho=dens, rho=rho)
    print('ot4 3d adjoint', rho, plan.shape, flush=True)
    rng = np.random.default_rng(29)
    res = adjoint_source_samples(plan, rng.standard_normal((plan.nrec, plan.nt)).astype(np.float32))
    save_mode = 2 if dens else 1
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
