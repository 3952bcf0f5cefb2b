"""GPU parity: the HIP library against the CPU oracle on the same seeded inputs.

Bars: traces within 1e-5 relative L2 of the oracle (FP32 both sides);
gradients within 1e-4 relative L2; the FP32 GPU error against the FP64 oracle
is asserted too, so the FP32 tolerance is not hiding a scheme difference.
"""

import numpy as np
import pytest

import oracle
from helpers import small_case, solver_for, rel_l2
from stride_amd.physics.plan import adjoint_source_samples
from stride_amd.physics.gradients import exact_gradients

pytestmark = pytest.mark.gpu

FORWARD_CASES = [
    # (dim, options)
    (2, dict(kernel='OT2')),
    (2, dict(kernel='OT4')),
    (2, dict(kernel='OT2', interpolation_type='hicks')),
    (2, dict(kernel='OT2', interpolation_type='sinc')),
    (2, dict(kernel='OT2', rho='const')),
    (2, dict(kernel='OT4', rho='var')),
    (2, dict(kernel='OT2', rho='var', alpha='p0')),
    (2, dict(kernel='OT4', alpha='p0')),
    (2, dict(kernel='OT2', alpha='p2', attenuation_power=2)),
    (2, dict(kernel='OT2', alpha='none-power')),
    (2, dict(kernel='OT2', boundary_type='complex_frequency_shift_PML_2')),
    (2, dict(kernel='OT2', boundary_type='complex_frequency_shift_PML_2', interpolation_type='hicks')),
    (2, dict(kernel='OT2', rho='var', alpha='p2', attenuation_power=2,
             boundary_type='complex_frequency_shift_PML_2', interpolation_type='hicks')),
    (2, dict(kernel='OT2', diff_source=True)),
    (3, dict(kernel='OT2')),
    (3, dict(kernel='OT4', rho='var', interpolation_type='hicks')),
    (3, dict(kernel='OT2', alpha='p0', boundary_type='complex_frequency_shift_PML_2')),
    (3, dict(kernel='OT2', alpha='p2', attenuation_power=2)),
    (3, dict(kernel='OT2', alpha='p0')),
    (3, dict(kernel='OT2', alpha='none-power')),
]


def _case(dim, opts, **extra):
    opts = dict(opts)
    rho = opts.pop('rho', 'none')
    alpha = opts.pop('alpha', None)
    n = 48 if dim == 2 else 28
    nt = 220 if dim == 2 else 120
    nabs = 10 if dim == 2 else 8
    opts.update(extra)
    return small_case(dim=dim, n=n, nabs=nabs, nt=nt, rho=rho, alpha=alpha, **opts)


def _label(case):
    dim, opts = case
    return '%dd-' % dim + '-'.join('%s=%s' % kv for kv in sorted(opts.items()))


@pytest.mark.parametrize('case', FORWARD_CASES, ids=[_label(c) for c in FORWARD_CASES])
def test_forward_traces(gpu, case):
    dim, opts = case
    plan, _ = _case(dim, opts)
    print('forward', _label(case), plan.shape, plan.nt, flush=True)
    ref32, _, _ = oracle.forward(plan)
    ref64, _, _ = oracle.forward(plan, precision='f64')
    s = solver_for(plan)
    got = s.forward(plan.src_data).T
    assert np.abs(ref64).max() > 0
    err = rel_l2(got, ref32)
    err64 = rel_l2(got, ref64)
    print('  rel L2 vs f32 oracle %.3e, vs f64 %.3e (f32 oracle vs f64 %.3e)'
          % (err, err64, rel_l2(ref32, ref64)))
    assert err <= 1e-5
    assert err64 <= 1e-4


COMPAT_CASES = [
    (2, dict(kernel='OT2')),
    # time_bounds (tb0, tb1) inside the run: saves on tb0 <= t <= tb1, the adjoint from tb1-1-f
    # down to tb0+1, the residual windowed over [tb0, tb1) (devito.py:577-581, 865-866, 916-920)
    (2, dict(kernel='OT2', time_bounds=(30, 180))),
    (2, dict(kernel='OT2', rho='var', grad_rho=True, time_bounds=(25, 200))),
    (3, dict(kernel='OT2', time_bounds=(20, 100))),
    (2, dict(kernel='OT4', interpolation_type='hicks')),
    (2, dict(kernel='OT2', rho='var', grad_rho=True)),
    (2, dict(kernel='OT2', rho='var', alpha='p0', boundary_type='complex_frequency_shift_PML_2')),
    (2, dict(kernel='OT2', grad_src=True)),
    (3, dict(kernel='OT2')),
]


@pytest.mark.parametrize('case', COMPAT_CASES, ids=[_label(c) for c in COMPAT_CASES])
def test_compat_adjoint(gpu, case):
    dim, opts = case
    plan, _ = _case(dim, opts, grad_vp=True)
    print('compat adjoint', _label(case), flush=True)
    save_mode = 2 if plan.save_raw else 1
    rng = np.random.default_rng(7)
    res = rng.standard_normal((plan.nrec, plan.nt)).astype(np.float32)
    res_s = adjoint_source_samples(plan, res)
    _, saved, _ = oracle.forward(plan, save=True)
    ref = oracle.adjoint_compat(plan, res_s, saved)
    s = solver_for(plan, save_mode=save_mode)
    s.forward(plan.src_data)
    got = s.adjoint(res_s, want_vp=True, want_rho=plan.grad_rho, want_src=plan.grad_src)
    for key in ('grad_vp', 'prec_vp') + (('grad_rho', 'prec_rho') if plan.grad_rho else ()):
        err = rel_l2(got[key], ref[key])
        print('  %s rel L2 %.3e' % (key, err))
        assert np.abs(ref[key]).max() > 0
        assert err <= 1e-4, key
    if plan.grad_src:
        err = rel_l2(got['grad_src'], ref['grad_src'].T)
        print('  grad_src rel L2 %.3e' % err)
        assert err <= 1e-4


EXACT_CASES = [
    (2, dict(kernel='OT2')),
    (2, dict(kernel='OT4')),
    (2, dict(kernel='OT2', rho='var')),
    (2, dict(kernel='OT2', alpha='p0')),
    (3, dict(kernel='OT2')),
    # 3D without density: the forward, its every-step saves and the adjoint's steps on the
    # specialised update (OT4: its two marches)
    (3, dict(kernel='OT4')),
    (3, dict(kernel='OT2', alpha='p0')),
]


@pytest.mark.parametrize('case', EXACT_CASES, ids=[_label(c) for c in EXACT_CASES])
def test_exact_adjoint_matches_oracle(gpu, case):
    dim, opts = case
    plan, _ = _case(dim, opts, grad_vp=True)
    print('exact adjoint', _label(case), flush=True)
    rng = np.random.default_rng(11)
    res = rng.standard_normal((plan.nrec, plan.nt)).astype(np.float32)
    res_s = adjoint_source_samples(plan, res)
    states = oracle.forward_states(plan)
    ref = oracle.adjoint_exact(plan, res_s, states)
    s = solver_for(plan, save_mode=3)
    s.forward(plan.src_data)
    got = s.adjoint_exact(res_s, want_m=plan.kernel_order == 4, want_b=plan.rho_mode == 2)
    g_ref = exact_gradients(plan, ref, ref['lam_src'])
    g_got = exact_gradients(plan, got, got['lam_src'])
    for key in g_ref:
        err = rel_l2(g_got[key], g_ref[key])
        print('  d/d%s rel L2 %.3e' % (key, err))
        assert err <= 1e-4, key


def test_exact_adjoint_finite_difference(gpu):
    """<dJ/dvp, dv> against a central difference of J = <r, F(vp)> run on the GPU."""
    plan, info = small_case(dim=2, n=48, nabs=10, nt=240, kernel='OT2', grad_vp=True)
    rng = np.random.default_rng(3)
    r = rng.standard_normal((plan.nrec, plan.nt)).astype(np.float32)
    res_s = adjoint_source_samples(plan, r)
    s = solver_for(plan, save_mode=3)
    s.forward(plan.src_data)
    acc = s.adjoint_exact(res_s)
    grad = exact_gradients(plan, acc, acc['lam_src'])['vp']
    from stride_amd.physics.plan import build_plan

    def J(vp):
        p = build_plan(info['space'], info['time'], info['wavelet'], vp, **info['kw'])
        sv = solver_for(p)
        return float(np.sum(sv.forward(p.src_data).astype(np.float64) * res_s))

    d = np.zeros_like(info['vp'])
    d[15:50, 15:50] = rng.standard_normal((35, 35))
    eps = 4.0
    fd = (J(info['vp'] + eps * d) - J(info['vp'] - eps * d)) / (2 * eps)
    an = float(np.sum(grad * d))
    print('FD %.6e analytic %.6e' % (fd, an))
    assert abs(fd - an) <= 2e-3 * abs(fd)


@pytest.mark.parametrize('dim', [2, 3])
def test_power2_attenuation_runs_are_identical(gpu, dim):
    """The power-2 attenuation term takes the Laplacian of the level the update overwrites in
    place; that Laplacian is taken by a launch of its own, so repeated runs agree bit for bit
    (a race between the two would make them differ)."""
    plan, _ = _case(dim, dict(kernel='OT2', rho='var', alpha='p2', attenuation_power=2,
                              boundary_type='complex_frequency_shift_PML_2'))
    print('power-2 attenuation determinism %dD' % dim, flush=True)
    runs = []
    for _ in range(3):
        s = solver_for(plan)
        runs.append(s.forward(plan.src_data))
        s.close()
    for r in runs[1:]:
        np.testing.assert_array_equal(r, runs[0])
    ref, _, _ = oracle.forward(plan)
    assert rel_l2(runs[0].T, ref) < 1e-5
