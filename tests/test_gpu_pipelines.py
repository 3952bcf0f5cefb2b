"""Device trace processing, misfit and gradient post-processing (stride_amd.pipelines, HIP
kernels in csrc/process.hip) against the oracle (oracle/pipelines_oracle.py, pinned in
test_pipelines_oracle.py). Floating-point steps: relative L2 below 1e-6 (float32 results of
float64 / float32 arithmetic); data movement and selections: bit-exact."""

import numpy as np
import pytest

import pipelines_oracle as po
from helpers import rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

TOL = 1e-6


@pytest.fixture(scope='module')
def solver(gpu):
    from stride_amd._native import Solver
    s = Solver(2, (32, 32), (1e-3, 1e-3), (4, 4), 1e-7, 10)
    yield s
    s.close()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def host(t):
    return t.cpu().numpy()


def traces(rng, ntr=37, nt=513, zeros_after=None):
    d = rng.standard_normal((ntr, nt)).astype(np.float32)
    if zeros_after is not None:
        for r in range(ntr):
            d[r, zeros_after + (r * 7) % 50:] = 0
    return d


def test_check_traces(solver):
    from stride_amd.pipelines import CheckTraces
    rng = np.random.default_rng(0)
    d = traces(rng, 9, 70)
    d[2, 5] = np.nan
    d[7, 69] = -np.inf
    print('check_traces', flush=True)
    with pytest.raises(RuntimeError, match='Nan or inf'):
        CheckTraces().forward(solver, dev(d), name='observed')
    with pytest.warns(UserWarning):
        out = CheckTraces(raise_incorrect=False, filter_incorrect=True).forward(solver, dev(d))
    ref, bad = po.check_traces(d, filter_incorrect=True)
    np.testing.assert_array_equal(host(out), ref)
    clean = traces(rng, 4, 30)
    np.testing.assert_array_equal(host(CheckTraces().forward(solver, dev(clean))), clean)


@pytest.mark.parametrize('f_max', [None, 2.0e5])
def test_mute(solver, f_max):
    from stride_amd.pipelines import MuteTraces
    rng = np.random.default_rng(1)
    mod = traces(rng)
    obs = traces(rng, zeros_after=300)
    dt = 1e-7
    print('mute f_max=%s' % f_max, flush=True)
    out, obs_out = MuteTraces(f_max=f_max).forward(solver, dev(mod), dev(obs), dt)
    ref = po.mute_traces(mod, obs, None if f_max is None else f_max * dt)
    assert rel_l2(host(out), ref) < TOL
    np.testing.assert_array_equal(host(obs_out), obs)


@pytest.mark.parametrize('kind,f_min,f_max', [
    ('cos', None, 3e5), ('butterworth', None, 3e5), ('butterworth', 1e5, None), ('butterworth', 1e5, 6e5),
    ('fir', None, 3e5), ('fir', 1e5, None), ('fir', 1e5, 6e5)])
@pytest.mark.parametrize('adjoint', [False, True])
def test_filter(solver, kind, f_min, f_max, adjoint):
    from stride_amd.pipelines import FilterTraces
    rng = np.random.default_rng(2)
    d = traces(rng)
    dt = 1e-7
    print('filter %s %s %s adjoint=%s' % (kind, f_min, f_max, adjoint), flush=True)
    step = FilterTraces(f_min=f_min, f_max=f_max, filter_type=kind)
    out = (step.adjoint if adjoint else step.forward)(solver, dev(d), dt=dt)
    lo = None if f_min is None else f_min * dt
    hi = None if f_max is None else f_max * dt
    if kind == 'cos':
        ref = po.lowpass_filter_cos(d, hi, zero_phase=False, adjoint=adjoint)
    elif kind == 'butterworth':
        ref = po.sosfilt(po.butterworth_sos(lo, hi), d, adjoint=adjoint)
    else:
        ref = po.fir_filter(po.fir_taps(lo, hi), d, adjoint=adjoint)
    assert rel_l2(host(out), ref) < TOL


@pytest.mark.parametrize('order', [9, 3])
def test_sosfilt_section_counts(solver, order):
    # 9 band sections take the run-time section count kernel, 3 the compiled one
    rng = np.random.default_rng(14)
    d = traces(rng, 70, 300)
    sos = po.butterworth_sos(0.02, 0.3, order=order)
    t = dev(d)
    solver.traces_sosfilt(t.data_ptr(), 70, 300, sos, adjoint=True)
    assert rel_l2(host(t), po.sosfilt(sos, d, adjoint=True)) < TOL


def test_filter_errors(solver):
    from stride_amd.pipelines import FilterTraces
    d = dev(np.ones((2, 10)))
    with pytest.raises(Exception, match='Requested filter does not exist'):
        FilterTraces(f_min=1e5, filter_type='cos').forward(solver, d, dt=1e-7)
    np.testing.assert_array_equal(host(FilterTraces().forward(solver, d, dt=1e-7)), 1.0)  # no band: copy


@pytest.mark.parametrize('per_trace', [False, True])
def test_norm(solver, per_trace):
    from stride_amd.pipelines import NormPerShot, NormPerTrace
    rng = np.random.default_rng(3)
    d = traces(rng) * 40
    print('norm per_trace=%s' % per_trace, flush=True)
    out = (NormPerTrace() if per_trace else NormPerShot()).forward(solver, dev(d))
    ref = po.norm_per_trace(d) if per_trace else po.norm_per_shot(d)[0]
    assert rel_l2(host(out), ref) < TOL


def test_l2(solver):
    from stride_amd.pipelines import L2DistanceLoss
    rng = np.random.default_rng(4)
    mod, obs = traces(rng), traces(rng)
    loss = L2DistanceLoss()
    fun, res = loss.forward(solver, dev(mod), dev(obs))
    f_ref, r_ref = po.l2(mod, obs)
    assert fun == pytest.approx(f_ref, rel=1e-9)
    np.testing.assert_array_equal(host(res), mod - obs)  # one float32 subtraction per sample
    g_mod, g_obs = loss.adjoint(1.0, True, True)
    np.testing.assert_array_equal(host(g_mod), mod - obs)
    np.testing.assert_array_equal(host(g_obs), obs - mod)


@pytest.mark.parametrize('rows,cols', [(1, 1), (70, 130), (513, 64), (3, 1000)])
def test_transpose(solver, rows, cols):
    from stride_amd.pipelines import traces_from_solver, adjoint_source_to_solver
    rng = np.random.default_rng(5)
    a = rng.standard_normal((rows, cols)).astype(np.float32)
    np.testing.assert_array_equal(host(traces_from_solver(solver, dev(a))), a.T)
    w = rng.random(cols).astype(np.float32)
    np.testing.assert_array_equal(host(adjoint_source_to_solver(solver, dev(a), w)), (a * w).T)


def test_process_traces_chain(solver):
    from stride_amd.pipelines import ProcessTraces
    rng = np.random.default_rng(6)
    mod = traces(rng) * 3
    obs = traces(rng, zeros_after=350)
    dt = 1e-7
    pipe = ProcessTraces(f_min=1e5, f_max=6e5)
    m, o = pipe.forward(solver, dev(mod), dev(obs), dt)
    rm = po.mute_traces(mod, obs, 6e5 * dt)  # the pipeline's f_max reaches the mute step too
    rm = po.sosfilt(po.butterworth_sos(1e5 * dt, 6e5 * dt), rm)
    ro = po.sosfilt(po.butterworth_sos(1e5 * dt, 6e5 * dt), obs)
    rm, ro = po.norm_per_shot(rm)[0], po.norm_per_shot(ro)[0]
    assert rel_l2(host(m), rm) < TOL and rel_l2(host(o), ro) < TOL
    g = traces(rng)
    back = pipe.adjoint(solver, dev(g), dt)
    assert rel_l2(host(back), po.sosfilt(po.butterworth_sos(1e5 * dt, 6e5 * dt), g, adjoint=True)) < TOL


@pytest.mark.parametrize('shape,inner,ramp', [((40, 52), ((6, 34), (5, 47)), 10), ((23, 30, 41), ((3, 20), (4, 26), (5, 36)), 4),
                                              ((30, 12), ((2, 28), (1, 11)), 7),
                                              ((25, 8, 30), ((2, 23), (1, 7), (3, 27)), 5), ((21, 21), ((0, 21), (0, 21)), 10)])
def test_mask(solver, shape, inner, ramp):
    from stride_amd.pipelines import MaskField
    rng = np.random.default_rng(7)
    f = rng.standard_normal(shape).astype(np.float32)
    out = MaskField(mask_rampoff=ramp).forward(solver, dev(f), inner)
    ref = po.mask_field(f, tuple(slice(a, b) for a, b in inner), ramp)
    np.testing.assert_array_equal(host(out), ref)  # float32 profile values from double, one product


@pytest.mark.parametrize('shape,sigma,mode', [((40, 52), 0.25, 'nearest'), ((40, 52), 1.0, 'reflect'),
                                              ((17, 23, 29), 1.5, 'nearest'), ((17, 23, 29), (0.5, 0, 2.0), 'reflect'),
                                              ((9, 5), 3.0, 'reflect')])
def test_gaussian(solver, shape, sigma, mode):
    rng = np.random.default_rng(8)
    f = rng.standard_normal(shape).astype(np.float32)
    t = dev(f)
    sg = sigma if np.iterable(sigma) else (sigma,) * len(shape)
    solver.field_gaussian(t.data_ptr(), shape, sg, mode=mode)
    solver.synchronize()
    assert rel_l2(host(t), po.smooth_field(f, sigma, mode=mode)) < TOL


def test_stats_and_norm_field(solver):
    from stride_amd.pipelines import NormField
    rng = np.random.default_rng(9)
    f = rng.standard_normal(100003).astype(np.float32) * 5
    lo, hi, ab = solver.field_stats(dev(f).data_ptr(), f.size)
    assert lo == f.min() and hi == f.max()
    assert ab == pytest.approx(np.sum(np.abs(f.astype(np.float64))), rel=1e-12)
    var = (1500 + 50 * rng.random(f.size)).astype(np.float32)
    out = NormField().forward(solver, dev(f), variable=dev(var), transform=lambda v: 1 / v)
    ref = po.norm_field(f, var, transform=lambda v: 1 / v)
    assert rel_l2(host(out), ref) < TOL
    assert rel_l2(host(NormField().forward(solver, dev(f))), po.norm_field(f)) < TOL


def test_apply_prec(solver):
    from stride_amd.pipelines import apply_prec
    rng = np.random.default_rng(10)
    g = rng.standard_normal((31, 44)).astype(np.float32)
    p = (rng.random((31, 44)) ** 2).astype(np.float32)
    p[:3] = 0
    out = apply_prec(solver, dev(g), dev(p), prec_scale=4.0, prec_smooth=1.0)
    assert rel_l2(host(out), po.apply_prec(g, p, 4.0, 1.0)) < TOL
    z = apply_prec(solver, dev(g), dev(np.zeros_like(p)))
    np.testing.assert_array_equal(host(z), g)


@pytest.mark.parametrize('slowness,clip', [(True, None), (False, None), (True, (1480.0, 1530.0))])
def test_descent(solver, slowness, clip):
    from stride_amd.pipelines import GradientDescent
    rng = np.random.default_rng(11)
    x = (1500 + 20 * rng.standard_normal((20, 30))).astype(np.float32)
    d = (rng.standard_normal((20, 30)) * (1e-5 if slowness else 10)).astype(np.float32)
    kw = {} if clip is None else dict(min=clip[0], max=clip[1])
    opt = GradientDescent(step_size=0.7, transform='slowness' if slowness else None, **kw)
    xt = dev(x)
    step = opt.step(solver, xt, dev(d))
    assert step == 0.7
    lo, hi = clip if clip else (None, None)
    ref = po.descent(x, d, 0.7, slowness=slowness, lo=lo, hi=hi)
    assert rel_l2(host(xt), ref) < TOL
    assert GradientDescent(step_size=5.0, max_step=2.0).final_step() == 2.0
    assert GradientDescent(step_size=0.01).final_step() == 0.1
    assert GradientDescent(step_size=-1.0).final_step() == -0.25


def test_crop_embed_gradient_finish(solver):
    from stride_amd.pipelines import crop, embed
    from stride_amd.physics.gradients import compat_finish
    rng = np.random.default_rng(12)
    f = rng.standard_normal((19, 23, 27)).astype(np.float32)
    inner = ((2, 17), (3, 20), (4, 26))
    box = crop(solver, dev(f), inner)
    np.testing.assert_array_equal(host(box), f[2:17, 3:20, 4:26])
    big = embed(solver, box, f.shape, inner)
    ref = np.zeros_like(f)
    ref[2:17, 3:20, 4:26] = f[2:17, 3:20, 4:26]
    np.testing.assert_array_equal(host(big), ref)
    g = rng.standard_normal(500).astype(np.float32)
    p = rng.random(500).astype(np.float32)
    vp = (1500 + 30 * rng.random(500)).astype(np.float32)
    for slowness in (True, False):
        gt, pt = dev(g), dev(p)
        solver.field_grad_vp_finish(gt.data_ptr(), pt.data_ptr(), dev(vp).data_ptr(), 500, slowness)
        solver.synchronize()
        rg, rp = compat_finish(g, p, vp, transform=(lambda v: 1 / v) if slowness else None)
        assert rel_l2(host(gt), rg) < TOL and rel_l2(host(pt), rp) < TOL


def test_device_iteration_matches_host_chain(gpu):
    """One FWI iteration with everything after the forward run on the device: traces to rows,
    ProcessTraces, L2 misfit, filter adjoint, windowed adjoint source, adjoint run, vp
    gradient, apply_prec, ProcessGlobalGradient and the descent step; against the oracle
    solver and the oracle steps."""
    import oracle
    from helpers import small_case, solver_for
    from stride_amd.physics.plan import adjoint_source_samples
    from stride_amd import pipelines as pl
    from stride_amd.physics.gradients import compat_finish

    plan, info = small_case(dim=2, n=40, nabs=8, nt=200, kernel='OT2', grad_vp=True, nrec=16)
    print('device iteration 2D n=40 nt=200', flush=True)
    dt = plan.dt
    rng = np.random.default_rng(13)
    s = solver_for(plan, save_mode=1)
    src = dev(plan.src_data)
    rec = torch.zeros((plan.nt, plan.nrec), dtype=torch.float32, device='cuda')
    s.forward_dev(src.data_ptr(), rec.data_ptr())
    s.synchronize()
    mod = pl.traces_from_solver(s, rec)
    ref_rec, saved, _ = oracle.forward(plan, save=True)  # (nrec, nt)
    assert rel_l2(host(mod), ref_rec) < 1e-5
    obs_np = (ref_rec * 1.1 + 0.02 * np.abs(ref_rec).max() * rng.standard_normal(ref_rec.shape)).astype(np.float32)
    obs_np[:, :20] = 0
    pipe = pl.ProcessTraces(f_min=2e5, f_max=9e5)
    m, o = pipe.forward(s, mod, dev(obs_np), dt)
    fun, res = pl.L2DistanceLoss().forward(s, m, o)
    adj = pipe.adjoint(s, res, dt)
    window = adjoint_source_samples(plan, np.ones((1, plan.nt)))[:, 0]
    adj_tm = pl.adjoint_source_to_solver(s, adj, window)
    s.adjoint_dev(adj_tm.data_ptr(), 1)
    s.synchronize()
    inner = tuple((e, e + n) for e, n in zip(info['space'].extra, info['space'].shape))
    vp_ext = dev(plan.vp)
    grad, prec = pl.gradient_vp(s, vp_ext, inner)
    grad = pl.apply_prec(s, grad, prec)
    gfull = pl.embed(s, grad, plan.vp.shape, inner)
    vp_inner = pl.crop(s, vp_ext, inner)
    direction = pl.ProcessGlobalGradient().forward(s, gfull, inner, variable=vp_ext, transform=lambda v: 1 / v)
    d_inner = pl.crop(s, direction, inner)
    pl.GradientDescent(step_size=1.0).step(s, vp_inner, d_inner)

    # the same chain on the host from the oracle
    sb = po.butterworth_sos(2e5 * dt, 9e5 * dt)
    rm = po.norm_per_shot(po.sosfilt(sb, po.mute_traces(ref_rec, obs_np, 9e5 * dt)))[0]
    ro = po.norm_per_shot(po.sosfilt(sb, obs_np))[0]
    f_ref, r_ref = po.l2(rm, ro)
    assert fun == pytest.approx(f_ref, rel=1e-4)
    a_ref = po.sosfilt(sb, r_ref, adjoint=True)
    g_ref = oracle.adjoint_compat(plan, adjoint_source_samples(plan, a_ref), saved)
    sl = tuple(slice(a, b) for a, b in inner)
    vpi = plan.vp[sl]
    gi, pi = compat_finish(g_ref['grad_vp'][sl], g_ref['prec_vp'][sl], vpi, transform=lambda v: 1 / v)
    gi = po.apply_prec(gi, pi)
    gf = np.zeros(plan.vp.shape)
    gf[sl] = gi
    gf = po.norm_field(po.smooth_field(po.mask_field(gf, sl, 10), 0.25), plan.vp, transform=lambda v: 1 / v)
    v_ref = po.descent(vpi, gf[sl], 1.0)
    assert rel_l2(host(direction), gf) < 1e-4
    assert rel_l2(host(vp_inner) - vpi, v_ref - vpi) < 1e-4
