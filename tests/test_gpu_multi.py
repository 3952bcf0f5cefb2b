"""The multi-shot schedule (sa_multi_*): shots of uneven duration and receiver count advanced
together, pairs of shots per cell, one update launch per step for every pair still running,
the whole schedule captured in one HIP graph. Each shot's traces must equal a one-shot run of
that shot alone (its own nt), graph and step-by-step launches must agree bit for bit, and on
the 3D workload grid the traces must match the oracle."""

import numpy as np
import pytest

import oracle
from helpers import small_case, solver_for, rel_l2
from stride_amd.physics.plan import build_plan
from stride_amd.problem import Time

pytestmark = pytest.mark.gpu


def _shot_plans(ctx, nts, nrecs, sources):
    geom = ctx['kw']['receiver_coordinates']
    space, dt = ctx['space'], ctx['time'].step
    plans = []
    for nt, nrec, src in zip(nts, nrecs, sources):
        plans.append(build_plan(space, Time(start=0., step=dt, num=nt), ctx['wavelet'][:, :nt], ctx['vp'],
                                source_coordinates=geom[src:src + 1], receiver_coordinates=geom[:nrec],
                                kernel='OT2', boundary_type=ctx['kw'].get('boundary_type', 'sponge_boundary_2')))
    return plans


def _multi_solver(plans, nt_max, tile=None):
    from stride_amd._native import Solver
    p0 = plans[0]
    s = Solver(p0.dim, p0.shape, p0.spacing, p0.absorbing, p0.dt, nt_max)
    s.set_scheme(p0.kernel_order, p0.rho_mode, p0.att_power, p0.boundary, p0.diff_source, p0.undersampling, 0,
                 0, nt_max, 0, 0, 0)
    s.set_medium(p0.vp)
    if p0.boundary == 0:
        s.set_sponge(p0.sponge_tables)
    s.multi_create(len(plans))
    for k, p in enumerate(plans):
        s.multi_set_shot(k, p.nt, p.src_idx, p.src_w, p.rec_idx, p.rec_w)
    if tile:
        s.batch_set_tile(*tile)
    return s


def _one_shot(p):
    s = solver_for(p, save_mode=0)
    rec = s.forward(p.src_data)
    s.close()
    return rec


@pytest.mark.parametrize('n,nabs,tile,nts', [
    (30, 6, None, (90, 61, 75, 90, 44)),
    (37, 5, (1, 12, 8), (70, 90, 33, 52)),
    (41, 7, (2, 7, 4), (88, 88, 40)),
    (30, 6, (2, 24, 8), (80,)),
])
def test_uneven_shots_match_one_shot_runs(gpu, n, nabs, tile, nts):
    nt_max = max(nts)
    _, ctx = small_case(dim=3, n=n, nabs=nabs, nt=nt_max, kernel='OT2', nrec=20)
    nrecs = [20, 14, 9, 17, 5][:len(nts)]
    plans = _shot_plans(ctx, nts, nrecs, [0, 7, 3, 12, 18][:len(nts)])
    print('multi n=%d tile=%s nts=%s' % (n, tile, nts), flush=True)
    s = _multi_solver(plans, nt_max, tile)
    graphed = s.multi_forward([p.src_data for p in plans], graph=True)
    direct = s.multi_forward([p.src_data for p in plans], graph=False)
    s.close()
    for k, p in enumerate(plans):
        assert graphed[k].shape == (p.nt, p.nrec)
        assert np.array_equal(graphed[k], direct[k]), k
        want = _one_shot(p)
        assert np.abs(want).max() > 0
        assert rel_l2(graphed[k], want) < 1e-6, k
        assert np.all(graphed[k][0] == 0) and np.all(graphed[k][-1] == 0)


def test_no_boundary_shots(gpu):
    nts = (80, 60, 71)
    _, ctx = small_case(dim=3, n=34, nabs=6, nt=80, kernel='OT2', nrec=16)
    plans = _shot_plans(ctx, nts, [16, 11, 7], [0, 5, 9])
    for p in plans:
        p.boundary = -1
    print('multi no boundary nts=%s' % (nts,), flush=True)
    s = _multi_solver(plans, 80)
    recs = s.multi_forward([p.src_data for p in plans], graph=True)
    s.close()
    for k, p in enumerate(plans):
        assert rel_l2(recs[k], _one_shot(p)) < 1e-6, k
        ref, _, _ = oracle.forward(p)
        assert rel_l2(recs[k].T, ref) < 1e-5, k


def test_workload_grid_uneven_shots(gpu):
    """Four shots of different duration and receiver count on the 360^3 grid (the hemisphere
    medium; short runs with the receivers near the sources so the traces are not zero)."""
    from stride_amd.configs import config3
    from stride_amd.utils import ellipsoidal, tone_burst
    cfg = config3(scale=1.0, nt=200)
    space, dt = cfg['space'], cfg['time'].step
    cen = np.array([l / 2 for l in space.limit])
    wav = tone_burst(400e3, 3, 200, dt)
    plans = []
    for nt, nrec, off in ((64, 64, 0.8), (40, 33, -0.5), (57, 48, 0.3), (49, 20, -0.9)):
        src = cen + off * (cfg['sources'][0] - cen)
        rec = ellipsoidal(nrec, (3e-3,) * 3, tuple(src))
        plans.append(build_plan(space, Time(start=0., step=dt, num=nt), wav[None, :nt], cfg['vp'],
                                source_coordinates=src[None], receiver_coordinates=rec, kernel='OT2'))
    print('multi 3D workload grid %s, nts %s' % (plans[0].shape, [p.nt for p in plans]), flush=True)
    s = _multi_solver(plans, 64)
    recs = s.multi_forward([p.src_data for p in plans], graph=True)
    s.close()
    for k, p in enumerate(plans):
        assert rel_l2(recs[k], _one_shot(p)) < 1e-6, k
    for k in (0, 3):
        ref, _, _ = oracle.forward(plans[k])
        assert np.linalg.norm(ref) > 0
        assert rel_l2(recs[k].T, ref) < 1e-5, k
