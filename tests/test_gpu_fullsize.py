"""Parity at the workloads' full grid sizes.

- 2D forward (240^2 extended, 1000 steps, sponge, 128-element ring) through the drop-in
  operator, traces and vp gradient against the oracle;
- 2D gradient (572^2 extended, 2500 steps, CFS-PML 30, variable density, power-2 attenuation)
  traces and vp / rho gradients against the oracle;
- the 3D grid (360^3 extended, the hemisphere medium) on a short run: the fast 3D update's
  traces and vp gradient against the oracle, and a two-shot pass against two one-shot runs;
- the 3D grid with variable vp / rho / alpha on a short run: the fused density update's traces
  and the vp / rho gradients against the oracle.

Tolerances: traces 1e-5, gradients 1e-4 relative L2 (float32 kernels against the float32
oracle)."""

import numpy as np
import pytest

import oracle
from helpers import solver_for, rel_l2
from stride_amd.physics.plan import build_plan, adjoint_source_samples

pytestmark = pytest.mark.gpu


def test_config1_full_operator(gpu):
    from stride_amd import IsoAcousticDevito, ScalarField, Shot, SubProblem, Traces
    from stride_amd.configs import config1
    from stride_amd.physics.gradients import compat_finish
    cfg = config1()
    space, time = cfg['space'], cfg['time']
    print('config1 full %s nt %d' % (space.extended_shape, time.num), flush=True)
    vp = ScalarField('vp', space=space, data=cfg['vp'], needs_grad=True)
    shot = Shot(0, cfg['sources'], cfg['receivers'], wavelets=cfg['wavelet'][None].astype(np.float32))
    problem = SubProblem(shot, space, time)
    pde = IsoAcousticDevito(space=space, time=time)
    traces = pde(shot.wavelets, vp, problem=problem)
    plan = build_plan(space, time, cfg['wavelet'][None], cfg['vp'], source_coordinates=cfg['sources'],
                      receiver_coordinates=cfg['receivers'], grad_vp=True)
    ref, saved, _ = oracle.forward(plan, save=True)
    assert rel_l2(traces.data, ref) < 1e-5
    res = (0.1 * ref + 0.01 * np.abs(ref).max() * np.random.default_rng(1).standard_normal(ref.shape))
    res = res.astype(np.float32)
    out = pde.adjoint(Traces('res', data=res), shot.wavelets, vp, problem=problem)
    acc = oracle.adjoint_compat(plan, adjoint_source_samples(plan, res), saved)
    g_ref, p_ref = compat_finish(acc['grad_vp'][space.inner], acc['prec_vp'][space.inner], vp.data, vp.transform)
    assert rel_l2(out[1].data, g_ref) < 1e-4
    assert rel_l2(out[1].prec.data, p_ref) < 1e-4


def test_config2_full(gpu):
    from stride_amd.configs import config2
    cfg = config2()
    plan = build_plan(cfg['space'], cfg['time'], cfg['wavelet'][None], cfg['vp'], rho=cfg['rho'], alpha=cfg['alpha'],
                      source_coordinates=cfg['sources'][:1], receiver_coordinates=cfg['receivers'],
                      boundary_type='complex_frequency_shift_PML_2', attenuation_power=2, grad_vp=True,
                      grad_rho=True)
    print('config2 full %s nt %d' % (plan.shape, plan.nt), flush=True)
    s = solver_for(plan, save_mode=2)
    rec = s.forward(plan.src_data)
    ref, saved, _ = oracle.forward(plan, save=True)
    assert rel_l2(rec.T, ref) < 1e-5
    rng = np.random.default_rng(2)
    res = adjoint_source_samples(plan, rng.standard_normal((plan.nrec, plan.nt)) * np.abs(ref).max())
    g = s.adjoint(res, want_vp=True, want_rho=True)
    s.close()
    acc = oracle.adjoint_compat(plan, res, saved)
    for key in ('grad_vp', 'prec_vp', 'grad_rho', 'prec_rho'):
        assert rel_l2(g[key], acc[key]) < 1e-4, key


def _plan3(source_offset, nt=64):
    """The 3D grid and medium, one source inside the hemisphere and 64 receivers on a 3 mm
    sphere around it, so a short run has non-zero traces and gradient."""
    from stride_amd.configs import config3
    from stride_amd.problem import Time
    from stride_amd.utils import ellipsoidal, tone_burst
    cfg = config3(scale=1.0, nt=200)
    space = cfg['space']
    dt = cfg['time'].step
    cen = np.array([l / 2 for l in space.limit])
    src = cen + source_offset * (cfg['sources'][0] - cen)
    rec = ellipsoidal(64, (3e-3,) * 3, tuple(src))
    wav = tone_burst(400e3, 3, 200, dt)[:nt]
    return build_plan(space, Time(start=0., step=dt, num=nt), wav[None], cfg['vp'], source_coordinates=src[None],
                      receiver_coordinates=rec, kernel='OT2', save_undersampling=4, grad_vp=True)


def test_config3_grid_fast_update(gpu):
    plan = _plan3(0.8)
    print('3D full grid %s nt %d' % (plan.shape, plan.nt), flush=True)
    s = solver_for(plan, save_mode=1)
    rec = s.forward(plan.src_data)
    ref, saved, _ = oracle.forward(plan, save=True)
    assert np.linalg.norm(ref) > 0
    assert rel_l2(rec.T, ref) < 1e-5
    res = adjoint_source_samples(plan, 0.1 * ref)
    g = s.adjoint(res, want_vp=True)
    s.close()
    acc = oracle.adjoint_compat(plan, res, saved)
    for key in ('grad_vp', 'prec_vp'):
        assert np.linalg.norm(acc[key]) > 0
        assert rel_l2(g[key], acc[key]) < 1e-4, key


def test_config3_grid_two_shots(gpu):
    from stride_amd._native import Solver
    plans = (_plan3(0.8), _plan3(-0.5))
    print('3D full grid two shots %s' % (plans[0].shape,), flush=True)
    rng = np.random.default_rng(3)
    res = [adjoint_source_samples(p, rng.standard_normal((p.nrec, p.nt))) for p in plans]
    single = []
    for p, r in zip(plans, res):
        s = solver_for(p, save_mode=1)
        rec = s.forward(p.src_data)
        single.append((rec, s.adjoint(r, want_vp=True)))
        s.close()
    p0 = plans[0]
    s = Solver(p0.dim, p0.shape, p0.spacing, p0.absorbing, p0.dt, p0.nt)
    s.set_scheme(p0.kernel_order, p0.rho_mode, p0.att_power, p0.boundary, p0.diff_source, p0.undersampling, 1,
                 p0.tb0, p0.tb1, 1, 0, 0)
    s.set_medium(p0.vp)
    s.set_sponge(p0.sponge_tables)
    s.batch_create()
    for k, p in enumerate(plans):
        s.batch_set_taps(k, p.src_idx, p.src_w, p.rec_idx, p.rec_w)
    recs = s.batch_forward([p.src_data for p in plans])
    g = s.batch_adjoint(res)
    s.close()
    for k in range(2):
        assert rel_l2(recs[k], single[k][0]) < 1e-6, k
    for key in ('grad_vp', 'prec_vp'):
        want = single[0][1][key].astype(np.float64) + single[1][1][key]
        assert rel_l2(g[key], want) < 1e-5, key


def test_config4_grid_density_update(gpu):
    """The 3D grid with config 4's variable vp / rho / alpha (power 0) medium on a short run:
    the fused density update's traces and the one-pass vp / rho gradients against the oracle."""
    from stride_amd.configs import config4
    from stride_amd.problem import Time
    from stride_amd.utils import ellipsoidal, tone_burst
    nt = 48
    cfg = config4(scale=1.0, nt=200)
    space = cfg['space']
    dt = cfg['time'].step
    cen = np.array([l / 2 for l in space.limit])
    src = cen + 0.8 * (cfg['sources'][0] - cen)
    rec = ellipsoidal(64, (3e-3,) * 3, tuple(src))
    wav = tone_burst(400e3, 3, 200, dt)[:nt]
    plan = build_plan(space, Time(start=0., step=dt, num=nt), wav[None], cfg['vp'], rho=cfg['rho'],
                      alpha=cfg['alpha'], source_coordinates=src[None], receiver_coordinates=rec, kernel='OT2',
                      save_undersampling=4, grad_vp=True, grad_rho=True)
    print('3D full grid density %s nt %d' % (plan.shape, plan.nt), flush=True)
    s = solver_for(plan, save_mode=2)
    got = s.forward(plan.src_data)
    ref, saved, _ = oracle.forward(plan, save=True)
    assert np.linalg.norm(ref) > 0
    assert rel_l2(got.T, ref) < 1e-5
    res = adjoint_source_samples(plan, 0.1 * ref)
    g = s.adjoint(res, want_vp=True, want_rho=True)
    s.close()
    acc = oracle.adjoint_compat(plan, res, saved)
    for key in ('grad_vp', 'prec_vp', 'grad_rho', 'prec_rho'):
        assert np.linalg.norm(acc[key]) > 0
        assert rel_l2(g[key], acc[key]) < 1e-4, key
