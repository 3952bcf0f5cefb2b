"""Several 2D shots per launch (lockstep; one solver per shot, one launch of the fused 2D update per
time step for all shots) against each shot's one-shot run and against the oracle: traces, and the
vp / rho gradients and preconditioners of the as-written adjoint (each solver keeps its own
shot's). The shots differ in source, receivers and medium, so a shot reading another's arguments
fails. Bars: against the one-shot runs 1e-5 (the same operations in a kernel compiled apart, so
the last place of a rounding can differ and grow over the steps: ~1e-6 measured); against the FP32
oracle 1e-5 (traces) and 1e-4 (gradients)."""

import numpy as np
import pytest

import oracle
from helpers import small_case, solver_for, rel_l2
from stride_amd.physics.plan import build_plan, adjoint_source_samples

pytestmark = pytest.mark.gpu


def _shots(nshot, n, nabs, nt, **opts):
    plan0, ctx = small_case(dim=2, n=n, nabs=nabs, nt=nt, kernel='OT2', nrec=24, **opts)
    kw = dict(ctx['kw'])
    geom = kw['receiver_coordinates']
    plans = [plan0]
    for k in range(1, nshot):
        kw.update(source_coordinates=geom[5 * k:5 * k + 1], receiver_coordinates=geom[k:k + 16])
        vp = ctx['vp'] * (1. + 0.02 * k)  # each shot on a medium of its own
        plans.append(build_plan(ctx['space'], ctx['time'], ctx['wavelet'], vp.astype(np.float32), rho=ctx['rho'],
                                alpha=ctx['alpha'], **kw))
    return plans


def _residuals(plans, seed=5):
    rng = np.random.default_rng(seed)
    return [adjoint_source_samples(p, rng.standard_normal((p.nrec, p.nt)).astype(np.float32)) for p in plans]


CASES = [
    # variable density, CFS-PML, power-2 attenuation, raw saves, vp and rho gradients
    (3, 48, 10, 160, 2, dict(rho='var', alpha='p2', attenuation_power=2, grad_vp=True, grad_rho=True,
                             boundary_type='complex_frequency_shift_PML_2')),
    # no density, sponge, p.dt2 saves, vp gradient; odd sizes (partial tiles on both axes)
    (2, 61, 7, 140, 1, dict(grad_vp=True)),
    # constant density, sponge, attenuation power 0, raw saves, vp gradient
    (4, 37, 9, 120, 2, dict(rho='const', alpha='p0', grad_vp=True)),
    # time_bounds inside the run (saves on tb0 <= t <= tb1, the adjoint from tb1 - 1 - f down to
    # tb0 + 1; devito.py:577-581, 916-920), variable density with CFS-PML
    (2, 48, 10, 200, 2, dict(rho='var', grad_vp=True, grad_rho=True, time_bounds=(25, 180),
                             boundary_type='complex_frequency_shift_PML_2')),
]


@pytest.mark.parametrize('nshot,n,nabs,nt,save_mode,opts', CASES,
                         ids=['var-pml-p2', 'plain-sponge', 'const-p0', 'time-bounds'])
def test_lockstep_shots_match_one_shot_runs(gpu, nshot, n, nabs, nt, save_mode, opts):
    from stride_amd._native import shots2d_forward, shots2d_adjoint
    plans = _shots(nshot, n, nabs, nt, **dict(opts))
    want_rho = bool(opts.get('grad_rho'))
    print('lockstep 2D', nshot, plans[0].shape, nt, save_mode, sorted(opts), flush=True)
    res = _residuals(plans)
    single = []
    for p, r in zip(plans, res):
        s = solver_for(p, save_mode=save_mode)
        rec = s.forward(p.src_data)
        g = s.adjoint(r, want_vp=True, want_rho=want_rho)
        s.close()
        single.append((rec, g))
    group = [solver_for(p, save_mode=save_mode) for p in plans]
    try:
        recs = shots2d_forward(group, [p.src_data for p in plans])
        grads = shots2d_adjoint(group, res, want_vp=True, want_rho=want_rho)
        # a second run on the same solvers (buffers and the argument table reused)
        recs2 = shots2d_forward(group, [p.src_data for p in plans])
    finally:
        for s in group:
            s.close()
    keys = ('grad_vp', 'prec_vp') + (('grad_rho', 'prec_rho') if want_rho else ())
    for k in range(nshot):
        assert np.abs(single[k][0]).max() > 0
        err = rel_l2(recs[k], single[k][0])
        print('  shot %d traces vs one-shot run %.3e (max abs diff %.3e)'
              % (k, err, np.abs(recs[k] - single[k][0]).max()))
        assert err < 1e-5, k
        assert rel_l2(recs2[k], recs[k]) == 0, k  # a second run repeats the first exactly
        for key in keys:
            want = single[k][1][key]
            e = rel_l2(grads[k][key], want)
            print('  shot %d %s vs one-shot run %.3e' % (k, key, e))
            assert np.abs(want).max() > 0
            assert e < 1e-5, (k, key)


def test_lockstep_shots_match_oracle(gpu):
    from stride_amd._native import shots2d_forward, shots2d_adjoint
    plans = _shots(3, 48, 10, 160, rho='var', alpha='p0', grad_vp=True, grad_rho=True,
                   boundary_type='complex_frequency_shift_PML_2')
    res = _residuals(plans, seed=9)
    group = [solver_for(p, save_mode=2) for p in plans]
    try:
        recs = shots2d_forward(group, [p.src_data for p in plans])
        grads = shots2d_adjoint(group, res, want_vp=True, want_rho=True)
    finally:
        for s in group:
            s.close()
    for k, p in enumerate(plans):
        ref, saved, _ = oracle.forward(p, save=True)
        err = rel_l2(recs[k].T, ref)
        print('  shot %d traces rel L2 %.3e' % (k, err), flush=True)
        assert err < 1e-5, k
        adj = oracle.adjoint_compat(p, res[k], saved)
        for key in ('grad_vp', 'prec_vp', 'grad_rho', 'prec_rho'):
            e = rel_l2(grads[k][key], adj[key])
            print('  shot %d %s rel L2 %.3e' % (k, key, e))
            assert np.abs(adj[key]).max() > 0
            assert e <= 1e-4, (k, key)


def test_lockstep_shots_refuse_mixed_groups(gpu):
    from stride_amd._native import shots2d_forward, SolverError
    a = _shots(2, 48, 10, 100, grad_vp=True)
    b = _shots(1, 48, 10, 120, grad_vp=True)  # another number of steps
    group = [solver_for(a[0], save_mode=1), solver_for(b[0], save_mode=1)]
    try:
        with pytest.raises((ValueError, SolverError)):
            shots2d_forward(group, [a[0].src_data, b[0].src_data])
    finally:
        for s in group:
            s.close()
    c = small_case(dim=3, n=28, nabs=8, nt=100, kernel='OT2', grad_vp=True, nrec=20)[0]
    s3 = solver_for(c, save_mode=1)
    try:
        with pytest.raises((ValueError, SolverError)):
            shots2d_forward([s3], [c.src_data])  # 3D: not the fused 2D update
    finally:
        s3.close()
