"""GPU parity of the fused 2D update against the separate general passes it replaces (the
fused kernels switched off with a density tile of 0 rows) and against the oracle: forward
traces and the as-written adjoint's gradients, for the scheme mixes of the 2D workloads
(sponge without density; CFS-PML with variable density and power-2 attenuation, and with
constant density and power-0 attenuation). Bars: fused vs general 1e-5 (the same arithmetic),
traces 1e-5 and gradients 1e-4 against the FP32 oracle."""

import numpy as np
import pytest

import oracle
from helpers import small_case, solver_for, rel_l2
from stride_amd.physics.plan import adjoint_source_samples

pytestmark = pytest.mark.gpu

CASES = [
    dict(rho='none'),
    dict(rho='var', alpha='p2', attenuation_power=2, boundary_type='complex_frequency_shift_PML_2',
         grad_rho=True),
    dict(rho='const', alpha='p0', boundary_type='complex_frequency_shift_PML_2'),
]


def _label(o):
    return '-'.join('%s=%s' % kv for kv in sorted(o.items()))


@pytest.mark.parametrize('opts', CASES, ids=[_label(o) for o in CASES])
def test_fused_2d_matches_general_and_oracle(gpu, opts):
    opts = dict(opts)
    rho = opts.pop('rho')
    alpha = opts.pop('alpha', None)
    plan, _ = small_case(dim=2, n=100, nabs=12, nt=260, rho=rho, alpha=alpha, kernel='OT2', grad_vp=True,
                         nrec=24, **opts)
    print('fused 2D', _label(dict(opts, rho=rho)), plan.shape, flush=True)
    save_mode = 2 if plan.save_raw else 1
    rng = np.random.default_rng(19)
    res = adjoint_source_samples(plan, rng.standard_normal((plan.nrec, plan.nt)).astype(np.float32))
    runs = []
    for fused in (True, False):
        s = solver_for(plan, save_mode=save_mode)
        if not fused:
            s.set_tile_density(0)
        rec = s.forward(plan.src_data)
        g = s.adjoint(res, want_vp=True, want_rho=plan.grad_rho)
        s.close()
        runs.append((rec, g))
    ref, saved, _ = oracle.forward(plan, save=True)
    acc = oracle.adjoint_compat(plan, res, saved)
    assert rel_l2(runs[0][0], runs[1][0]) <= 1e-5
    assert rel_l2(runs[0][0].T, ref) <= 1e-5
    for key in ('grad_vp', 'prec_vp') + (('grad_rho', 'prec_rho') if plan.grad_rho else ()):
        e_gen = rel_l2(runs[0][1][key], runs[1][1][key])
        e_ref = rel_l2(runs[0][1][key], acc[key])
        print('  %s vs general %.3e, vs oracle %.3e' % (key, e_gen, e_ref))
        assert e_gen <= 1e-5, key
        assert e_ref <= 1e-4, key
