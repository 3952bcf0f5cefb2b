"""Two density shots per pass (the fused density update reading b and the coefficients once for
both interleaved shots) against two one-shot runs and against the oracle: traces of each shot,
and the summed vp / rho gradients and preconditioners of the as-written adjoint from the raw
snapshots (stride/physics/iso_acoustic/devito.py:1115-1150, common/devito.py:541-550).

Grids: 44^3 extended (one tile across axis 2) and 86^3 extended (two tiles across axis 2, several
chunks along axis 0, partial last tiles). Bars: against the one-shot runs 1e-6 (traces) and 1e-5
(summed accumulators, the same terms summed in another order); against the FP32 oracle 1e-5
(traces) and 1e-4 (gradients)."""

import numpy as np
import pytest

import oracle
from helpers import small_case, solver_for, rel_l2
from stride_amd.physics.plan import build_plan, adjoint_source_samples

pytestmark = pytest.mark.gpu


def _two_shots(n, nabs, nt=100, rho='var', alpha='p0', **extra):
    opts = dict(attenuation_power=2) if alpha == 'p2' else {}
    opts.update(extra)
    plan0, ctx = small_case(dim=3, n=n, nabs=nabs, nt=nt, rho=rho, alpha=alpha, kernel='OT2', nrec=20,
                            grad_vp=True, grad_rho=rho == 'var', **opts)
    kw = dict(ctx['kw'])
    geom = kw['receiver_coordinates']
    # second shot: another source position and a different receiver subset
    kw.update(source_coordinates=geom[7:8], receiver_coordinates=geom[3:17])
    plan1 = build_plan(ctx['space'], ctx['time'], ctx['wavelet'], ctx['vp'], rho=ctx['rho'], alpha=ctx['alpha'],
                       **kw)
    return plan0, plan1


def _batched_solver(plan0, plan1, save_mode):
    from stride_amd._native import Solver
    s = Solver(plan0.dim, plan0.shape, plan0.spacing, plan0.absorbing, plan0.dt, plan0.nt)
    s.set_scheme(plan0.kernel_order, plan0.rho_mode, plan0.att_power, plan0.boundary, plan0.diff_source,
                 plan0.undersampling, save_mode, plan0.tb0, plan0.tb1, 1, plan0.grad_rho, 0)
    s.set_medium(plan0.vp, plan0.rho, plan0.alpha)
    s.set_sponge(plan0.sponge_tables)
    s.batch_create()
    for k, p in enumerate((plan0, plan1)):
        s.batch_set_taps(k, p.src_idx, p.src_w, p.rec_idx, p.rec_w)
    return s


def _residuals(plans, seed=7):
    rng = np.random.default_rng(seed)
    return [adjoint_source_samples(p, rng.standard_normal((p.nrec, p.nt)).astype(np.float32)) for p in plans]


@pytest.mark.parametrize('n,nabs,tile,alpha', [
    (28, 8, None, 'p0'), (70, 8, (1, 13, 4), 'p0'), (70, 8, (1, 40, 8), None), (28, 8, (1, 7, 8), 'none-power'),
])
def test_two_density_shots_match_one_shot_runs(gpu, n, nabs, tile, alpha):
    plan0, plan1 = _two_shots(n, nabs, alpha=alpha)
    print('batch density n=%d nabs=%d tile=%s alpha=%s' % (n, nabs, tile, alpha), flush=True)
    res = _residuals((plan0, plan1))
    single = []
    for p, r in zip((plan0, plan1), res):
        s = solver_for(p, save_mode=2)
        rec = s.forward(p.src_data)
        g = s.adjoint(r, want_vp=True, want_rho=True)
        s.close()
        single.append((rec, g))
    s = _batched_solver(plan0, plan1, save_mode=2)
    if tile:
        s.batch_set_tile_density(*tile)
    recs = s.batch_forward([plan0.src_data, plan1.src_data])
    g = s.batch_adjoint(res, density=True)
    s.close()
    for k in range(2):
        err = rel_l2(recs[k], single[k][0])
        print('  shot %d traces vs one-shot run %.3e' % (k, err))
        assert np.abs(single[k][0]).max() > 0
        assert err < 1e-6, k
    for key in ('grad_vp', 'prec_vp', 'grad_rho', 'prec_rho'):
        want = single[0][1][key].astype(np.float64) + single[1][1][key]
        err = rel_l2(g[key], want)
        print('  %s vs summed one-shot runs %.3e' % (key, err))
        assert np.abs(want).max() > 0
        assert err < 1e-5, key


def test_two_density_shots_time_bounds(gpu):
    """time_bounds inside the run (devito.py:577-581, 916-920): raw saves on tb0 <= t <= tb1, the
    adjoint from tb1 - 1 - f down to tb0 + 1; against two one-shot runs."""
    plan0, plan1 = _two_shots(28, 8, nt=140, time_bounds=(20, 120))
    assert (plan0.tb0, plan0.tb1) == (20, 120)
    res = _residuals((plan0, plan1), seed=3)
    single = []
    for p, r in zip((plan0, plan1), res):
        s = solver_for(p, save_mode=2)
        rec = s.forward(p.src_data)
        g = s.adjoint(r, want_vp=True, want_rho=True)
        s.close()
        single.append((rec, g))
    s = _batched_solver(plan0, plan1, save_mode=2)
    recs = s.batch_forward([plan0.src_data, plan1.src_data])
    g = s.batch_adjoint(res, density=True)
    s.close()
    for k in range(2):
        assert rel_l2(recs[k], single[k][0]) < 1e-6, k
    for key in ('grad_vp', 'prec_vp', 'grad_rho', 'prec_rho'):
        want = single[0][1][key].astype(np.float64) + single[1][1][key]
        assert np.abs(want).max() > 0
        assert rel_l2(g[key], want) < 1e-5, key


def test_two_density_shots_match_oracle(gpu):
    plan0, plan1 = _two_shots(28, 8)
    print('batch density vs oracle', plan0.shape, flush=True)
    res = _residuals((plan0, plan1), seed=11)
    s = _batched_solver(plan0, plan1, save_mode=2)
    s.batch_set_tile_density(1, 11, 4)
    recs = s.batch_forward([plan0.src_data, plan1.src_data])
    g = s.batch_adjoint(res, density=True)
    s.close()
    want = {}
    for k, p in enumerate((plan0, plan1)):
        ref, saved, _ = oracle.forward(p, save=True)
        err = rel_l2(recs[k].T, ref)
        print('  shot %d traces rel L2 %.3e' % (k, err))
        assert err < 1e-5, k
        adj = oracle.adjoint_compat(p, res[k], saved)
        for key in ('grad_vp', 'prec_vp', 'grad_rho', 'prec_rho'):
            want[key] = want.get(key, 0.) + adj[key].astype(np.float64)
    for key, ref in want.items():
        err = rel_l2(g[key], ref)
        print('  %s rel L2 %.3e' % (key, err))
        assert np.abs(ref).max() > 0
        assert err <= 1e-4, key


def test_two_density_shots_constant_density_and_errors(gpu):
    """Constant density: the two-shot forward runs (traces as one-shot runs); its adjoint (no rho
    gradient form) and power-2 attenuation are refused with an error."""
    from stride_amd._native import SolverError
    plan0, plan1 = _two_shots(28, 8, rho='const', alpha=None)
    single = []
    for p in (plan0, plan1):
        s = solver_for(p, save_mode=2)
        single.append(s.forward(p.src_data))
        s.close()
    s = _batched_solver(plan0, plan1, save_mode=2)
    recs = s.batch_forward([plan0.src_data, plan1.src_data])
    for k in range(2):
        assert rel_l2(recs[k], single[k]) < 1e-6, k
    with pytest.raises(SolverError):
        s.batch_adjoint(_residuals((plan0, plan1)), density=True)
    with pytest.raises(ValueError):
        s.batch_set_tile_density(2, 24, 8)  # rows 2: not a two-shot density tile
    s.close()
    p2, q2 = _two_shots(28, 8, alpha='p2')
    s = _batched_solver(p2, q2, save_mode=2)
    assert p2.att_power == 2
    with pytest.raises(SolverError):
        s.batch_forward([p2.src_data, q2.src_data])
    s.close()
