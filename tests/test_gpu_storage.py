"""Checkpointed raw saves (save mode 5): the forward run keeps the state (p[T], p[T-1]) of each
segment of K snapshots, and the as-written adjoint recomputes each segment's raw snapshots from it
(devito.py:366-437 stores every snapshot instead). Traces and the vp / rho / wavelet gradients must
equal those of the stored raw snapshots (save mode 2) to rounding, and the recomputation issues the
same launches as the forward run, so every segment length gives the same bits.

Cases cover the fused 2D update, the general passes (OT4), the fused 3D density update on the 3D
gradient workload's grid, segment lengths K = 1 .. 7, undersampling 1 .. 4 and time_bounds inside
the run."""

import numpy as np
import pytest

from helpers import small_case, solver_for, rel_l2
from stride_amd.physics.plan import adjoint_source_samples

pytestmark = pytest.mark.gpu

CASES = [
    # (dim, K, options)
    (2, 1, dict(kernel='OT2', rho='var', grad_rho=True)),
    (2, 3, dict(kernel='OT2', rho='var', grad_rho=True, grad_src=True)),
    (2, 7, dict(kernel='OT2', rho='var', alpha='p0', grad_rho=True, save_undersampling=1)),
    (2, 4, dict(kernel='OT2', rho='var', grad_rho=True, save_undersampling=3, time_bounds=(25, 200))),
    (2, 2, dict(kernel='OT4', rho='var', grad_rho=True, save_undersampling=2)),
    (2, 5, dict(kernel='OT2', grad_src=True)),
    (3, 2, dict(kernel='OT2', rho='var', grad_rho=True)),
    (3, 3, dict(kernel='OT2', grad_src=True, time_bounds=(20, 100))),
]


def _label(case):
    dim, K, opts = case
    return '%dd-K%d-' % (dim, K) + '-'.join('%s=%s' % kv for kv in sorted(opts.items()))


def _run(plan, res, save_mode, K=0):
    s = solver_for(plan)
    s.set_scheme(plan.kernel_order, plan.rho_mode, plan.att_power, plan.boundary, plan.diff_source,
                 plan.undersampling, save_mode, plan.tb0, plan.tb1, plan.grad_vp, plan.grad_rho, plan.grad_src,
                 checkpoint=K)
    rec = s.forward(plan.src_data)
    got = s.adjoint(res, want_vp=True, want_rho=plan.grad_rho, want_src=plan.grad_src)
    s.close()
    return rec, got


def _compare(plan, res, K):
    """Mode 5 against mode 2: equal to rounding (the 3D density update stores the snapshot inside
    the update under mode 2, a separately compiled form whose last bits can differ). Mode 5 with
    segments of K against one segment of every snapshot recomputed from the zero state: bit for
    bit, so each segment restarts from exactly the forward run's state."""
    rec2, g2 = _run(plan, res, 2)
    rec5, g5 = _run(plan, res, 5, K)
    _, g1 = _run(plan, res, 5, 1 << 20)
    print('  traces rel L2 %.3e (bitwise equal: %s)' % (rel_l2(rec5, rec2), np.array_equal(rec5, rec2)), flush=True)
    assert rel_l2(rec5, rec2) <= 1e-6
    keys = ['grad_vp', 'prec_vp'] + (['grad_rho', 'prec_rho'] if plan.grad_rho else []) + \
        (['grad_src'] if plan.grad_src else [])
    for key in keys:
        assert np.abs(g2[key]).max() > 0, key
        err = rel_l2(g5[key], g2[key])
        print('  %s rel L2 %.3e (bitwise equal: %s)' % (key, err, np.array_equal(g5[key], g2[key])), flush=True)
        assert err <= 1e-5, key
        assert np.array_equal(g5[key], g1[key]), key


@pytest.mark.parametrize('case', CASES, ids=[_label(c) for c in CASES])
def test_checkpointed_raw_saves_equal_stored(gpu, case):
    dim, K, opts = case
    opts = dict(opts)
    rho = opts.pop('rho', 'none')
    alpha = opts.pop('alpha', None)
    n, nt, nabs = (48, 220, 10) if dim == 2 else (28, 120, 8)
    plan, _ = small_case(dim=dim, n=n, nabs=nabs, nt=nt, rho=rho, alpha=alpha, grad_vp=True, **opts)
    print('checkpointed raw saves', _label(case), plan.shape, 'f', plan.undersampling, flush=True)
    rng = np.random.default_rng(11)
    res = adjoint_source_samples(plan, rng.standard_normal((plan.nrec, plan.nt)).astype(np.float32))
    _compare(plan, res, K)


def test_checkpointed_raw_saves_workload_grid(gpu):
    """The 3D gradient workload's variable vp / rho / alpha medium on its 360^3 grid (fused density
    update and gradient pass), nt 80, raw saves every 4th step, segments of 4 snapshots."""
    from stride_amd.configs import config4
    from stride_amd.physics.plan import build_plan
    from stride_amd.problem import Time
    cfg = config4(scale=1.0, nt=200, nrec=256)
    nt = 80
    time = Time(start=0., step=cfg['time'].step, num=nt)
    plan = build_plan(cfg['space'], time, cfg['wavelet'][None, :nt], cfg['vp'], rho=cfg['rho'], alpha=cfg['alpha'],
                      source_coordinates=cfg['sources'], receiver_coordinates=cfg['receivers'], kernel='OT2',
                      save_undersampling=4, grad_vp=True, grad_rho=True)
    print('checkpointed raw saves on %s, nt %d' % (plan.shape, plan.nt), flush=True)
    rng = np.random.default_rng(12)
    res = adjoint_source_samples(plan, rng.standard_normal((plan.nrec, plan.nt)).astype(np.float32))
    _compare(plan, res, 4)


def _oracle_case():
    plan, _ = small_case(dim=2, n=48, nabs=10, nt=220, rho='var', grad_vp=True, grad_rho=True)
    rng = np.random.default_rng(13)
    res = adjoint_source_samples(plan, rng.standard_normal((plan.nrec, plan.nt)).astype(np.float32))
    return plan, res


def test_snapshots_through_host_memory(gpu):
    """cache_forward (devito.py:632-668, 820-842): the stored raw snapshots go to host memory,
    the device copy is released (the adjoint refuses to run), and come back. The whole box gives
    the gradients unchanged; the interior only (margin = extra, the reference's inner_wavefield)
    gives the oracle's as-written adjoint over snapshots that are zero outside the interior."""
    import oracle
    from stride_amd._native import SolverError
    plan, res = _oracle_case()
    print('snapshots through host memory', plan.shape, flush=True)
    s = solver_for(plan, save_mode=2)
    s.forward(plan.src_data)
    ref = s.adjoint(res, want_vp=True, want_rho=True)
    s.forward(plan.src_data)
    host = s.saved_to_host()
    with pytest.raises(SolverError):
        s.adjoint(res, want_vp=True, want_rho=True)
    s.saved_from_host(host)
    got = s.adjoint(res, want_vp=True, want_rho=True)
    for key in ('grad_vp', 'grad_rho'):
        assert np.array_equal(got[key], ref[key]), key
    # interior only
    extra = plan.extra
    s.forward(plan.src_data)
    inner = s.saved_to_host(margin=extra)
    nsave = inner.size // int(np.prod([n - 2 * e for n, e in zip(plan.shape, extra)]))
    s.saved_from_host(inner, margin=extra)
    got = s.adjoint(res, want_vp=True, want_rho=True)
    s.close()
    _, saved, _ = oracle.forward(plan, save=True)
    assert saved.shape[0] == nsave
    keep = np.zeros_like(saved)
    sl = (slice(None),) + tuple(slice(e, n - e) for e, n in zip(extra, plan.shape))
    keep[sl] = saved[sl]
    want = oracle.adjoint_compat(plan, res, keep)
    full = oracle.adjoint_compat(plan, res, saved)
    for key in ('grad_vp', 'grad_rho'):
        err = rel_l2(got[key], want[key])
        print('  %s rel L2 vs the oracle %.3e (the cropped snapshots change it by %.3e)'
              % (key, err, rel_l2(want[key], full[key])), flush=True)
        assert err <= 1e-4, key


def test_wavefield_dumps(gpu):
    """dump_forward_wavefield / dump_adjoint_wavefield (devito.py:421-432, 765-778): the raw level
    of every f-th step over the extended grid, (tb1 - tb0 + f) // f + 1 slots at slot t // f. The
    forward dump equals the oracle's levels p[t]; the adjoint dump at factor 2 equals every other
    slot of the dump at factor 1, and dumping changes neither traces nor gradients."""
    import oracle
    plan, res = _oracle_case()
    s = solver_for(plan, save_mode=2)
    rec0 = s.forward(plan.src_data)
    g0 = s.adjoint(res, want_vp=True, want_rho=True)
    s.set_dump(3, 1)
    rec = s.forward(plan.src_data)
    g = s.adjoint(res, want_vp=True, want_rho=True)
    fwd, adj1 = s.dump_to_host(0), s.dump_to_host(1)
    s.set_dump(0, 2)
    s.forward(plan.src_data)
    s.adjoint(res, want_vp=True, want_rho=True)
    adj2 = s.dump_to_host(1)
    s.close()
    assert np.array_equal(rec, rec0)
    for key in ('grad_vp', 'grad_rho'):
        assert np.array_equal(g[key], g0[key])
    tb0, tb1 = plan.tb0, plan.tb1
    assert fwd.shape == ((tb1 - tb0 + 3) // 3 + 1,) + tuple(plan.shape)
    assert adj1.shape[0] == tb1 - tb0 + 2 and adj2.shape[0] == (tb1 - tb0 + 2) // 2 + 1
    states = oracle.forward_states(plan)
    for t in range(3, tb1 - 1, 3):
        err = rel_l2(fwd[t // 3], states[t])
        assert err <= 1e-5, t
    assert np.abs(adj1).max() > 0
    # the adjoint runs steps tb1-1-f .. tb0+1
    f = plan.undersampling
    for t in range(2, tb1 - 1 - f + 1, 2):
        assert np.array_equal(adj2[t // 2], adj1[t]), t
