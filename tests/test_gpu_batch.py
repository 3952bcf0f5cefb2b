"""Two shots per pass against two one-shot runs: traces of each shot, and the summed vp
gradient / preconditioner, for awkward sizes and every two-shot tile."""

import numpy as np
import pytest

import oracle
from helpers import small_case, solver_for, rel_l2
from stride_amd.physics.plan import build_plan, adjoint_source_samples

pytestmark = pytest.mark.gpu


def _two_shots(n, nabs, nt=90, boundary='sponge_boundary_2'):
    # boundary: a reference boundary type (the sponge; the CFS-PML takes the general kernels)
    plan0, ctx = small_case(dim=3, n=n, nabs=nabs, nt=nt, kernel='OT2', grad_vp=True, nrec=20,
                            boundary_type=boundary)
    geom = ctx['kw']['receiver_coordinates']
    # second shot: another source position and a different receiver subset
    plan1 = build_plan(ctx['space'], ctx['time'], ctx['wavelet'], ctx['vp'], source_coordinates=geom[7:8],
                       receiver_coordinates=geom[3:17], kernel='OT2', grad_vp=True, boundary_type=boundary)
    return plan0, plan1


def _batched(plan0, plan1, tile, res):
    from stride_amd._native import Solver
    s = Solver(plan0.dim, plan0.shape, plan0.spacing, plan0.absorbing, plan0.dt, plan0.nt)
    s.set_scheme(plan0.kernel_order, plan0.rho_mode, plan0.att_power, plan0.boundary, plan0.diff_source,
                 plan0.undersampling, 1, plan0.tb0, plan0.tb1, 1, 0, 0)
    s.set_medium(plan0.vp)
    if plan0.boundary == 0:
        s.set_sponge(plan0.sponge_tables)
    s.batch_create()
    for k, p in enumerate((plan0, plan1)):
        s.batch_set_taps(k, p.src_idx, p.src_w, p.rec_idx, p.rec_w)
    if tile:
        s.batch_set_tile(*tile)
    recs = s.batch_forward([plan0.src_data, plan1.src_data])
    g = s.batch_adjoint(res)
    s.close()
    return recs, g


@pytest.mark.parametrize('n,nabs,tile,boundary', [
    (30, 6, (1, 24, 8), 'sponge_boundary_2'), (37, 5, (2, 12, 8), 'sponge_boundary_2'),
    (41, 7, (2, 7, 4), 'sponge_boundary_2'), (37, 5, (1, 1, 8), 'sponge_boundary_2'),
    (70, 8, (1, 36, 8), 'sponge_boundary_2'), (66, 6, (2, 24, 4), 'sponge_boundary_2'),
    (30, 6, (2, 12, 4), 'sponge_boundary_2'), (41, 7, None, 'sponge_boundary_2'),
])
def test_two_shots_match_one_shot_runs(gpu, n, nabs, tile, boundary):
    plan0, plan1 = _two_shots(n, nabs, boundary=boundary)
    print('batch n=%d nabs=%d tile=%s boundary=%s' % (n, nabs, tile, boundary), flush=True)
    rng = np.random.default_rng(7)
    res = [adjoint_source_samples(p, rng.standard_normal((p.nrec, p.nt))) for p in (plan0, plan1)]
    single = []
    for p, r in zip((plan0, plan1), res):
        s = solver_for(p, save_mode=1)
        rec = s.forward(p.src_data)
        g = s.adjoint(r, want_vp=True)
        s.close()
        single.append((rec, g))
    recs, g = _batched(plan0, plan1, tile, res)
    for k in range(2):
        assert rel_l2(recs[k], single[k][0]) < 1e-6, k
    for key in ('grad_vp', 'prec_vp'):  # raw accumulators on the extended grid
        want = single[0][1][key].astype(np.float64) + single[1][1][key]
        assert rel_l2(g[key], want) < 1e-5, key


def test_two_shot_traces_match_oracle(gpu):
    plan0, plan1 = _two_shots(37, 5)
    recs, _ = _batched(plan0, plan1, (1, 12, 8), [np.zeros((plan0.nt, plan0.nrec), np.float32),
                                                   np.zeros((plan1.nt, plan1.nrec), np.float32)])
    for k, p in enumerate((plan0, plan1)):
        ref, _, _ = oracle.forward(p, save=False)
        assert rel_l2(recs[k].T, ref) < 1e-5, k


def test_two_shot_api_errors(gpu):
    from stride_amd._native import Solver, SolverError
    plan0, plan1 = _two_shots(30, 6)
    s = solver_for(plan0, save_mode=1)
    with pytest.raises(ValueError):
        s.batch_set_tile(3, 12, 8)
    with pytest.raises(ValueError):
        s._check(s._lib.sa_batch_create(s._h, 3))
    s.batch_create()
    with pytest.raises(SolverError):
        s.batch_adjoint([np.zeros((plan0.nt, 1), np.float32)] * 2)  # no two-shot forward yet
    s.close()


def test_two_shots_anisotropic_spacing(gpu):
    """Unequal spacings take the general-weights path of the two-shot update (equal spacings
    share one weight set across the axes)."""
    from stride_amd.problem import Space, Time
    from stride_amd.utils import tone_burst, ellipsoidal
    space = Space(shape=(30, 34, 38), spacing=(0.5e-3, 0.45e-3, 0.4e-3), extra=(6, 6, 6), absorbing=(6, 6, 6))
    time = Time(start=0., step=0.06e-6, num=140)
    rng = np.random.default_rng(5)
    vp = (1500 + 30 * rng.standard_normal(space.extended_shape)).astype(np.float32)
    c = tuple(l / 2 for l in space.limit)
    geom = ellipsoidal(16, (4e-3,) * 3, c, threshold=0.5)
    wav = tone_burst(0.5e6, 3, 140, 0.06e-6)[None]
    plans = [build_plan(space, time, wav, vp, source_coordinates=geom[k:k + 1], receiver_coordinates=geom,
                        kernel='OT2', grad_vp=True) for k in (0, 9)]
    print('batch anisotropic spacing', flush=True)
    res = [adjoint_source_samples(p, rng.standard_normal((p.nrec, p.nt))) for p in plans]
    single = []
    for p, r in zip(plans, res):
        s = solver_for(p, save_mode=1)
        rec = s.forward(p.src_data)
        single.append((rec, s.adjoint(r, want_vp=True)))
        s.close()
    recs, g = _batched(plans[0], plans[1], (1, 12, 8), res)
    for k, p in enumerate(plans):
        assert rel_l2(recs[k], single[k][0]) < 1e-6, k
        ref, _, _ = oracle.forward(p, save=False)
        assert rel_l2(recs[k].T, ref) < 1e-5, k
    for key in ('grad_vp', 'prec_vp'):
        want = single[0][1][key].astype(np.float64) + single[1][1][key]
        assert rel_l2(g[key], want) < 1e-5, key


def test_two_shots_without_boundary(gpu):
    """No absorbing boundary (the zero halo): the two-shot update's plain branch on every tile,
    traces and summed gradient against one-shot runs and the oracle."""
    plan0, plan1 = _two_shots(33, 6)
    for p in (plan0, plan1):
        p.boundary = -1
    print('batch without boundary', flush=True)
    rng = np.random.default_rng(13)
    res = [adjoint_source_samples(p, rng.standard_normal((p.nrec, p.nt))) for p in (plan0, plan1)]
    single = []
    for p, r in zip((plan0, plan1), res):
        s = solver_for(p, save_mode=1)
        rec = s.forward(p.src_data)
        single.append((rec, s.adjoint(r, want_vp=True)))
        s.close()
    recs, g = _batched(plan0, plan1, (2, 12, 4), res)
    for k, p in enumerate((plan0, plan1)):
        assert rel_l2(recs[k], single[k][0]) < 1e-6, k
        ref, _, _ = oracle.forward(p, save=False)
        assert rel_l2(recs[k].T, ref) < 1e-5, k
    for key in ('grad_vp', 'prec_vp'):
        want = single[0][1][key].astype(np.float64) + single[1][1][key]
        assert rel_l2(g[key], want) < 1e-5, key
