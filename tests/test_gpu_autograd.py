"""Exact adjoint behind torch.autograd.Function: checkpointed recomputation equals
storing every level, and autograd gradients match central finite differences of
the loss computed through the same forward."""

import numpy as np
import pytest

from helpers import small_case, solver_for, rel_l2
from stride_amd.physics.plan import adjoint_source_samples

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('dim,K', [(2, 7), (2, 16), (3, 5)])
def test_checkpointed_equals_stored(gpu, dim, K):
    n, nabs, nt = (48, 10, 200) if dim == 2 else (26, 6, 90)
    plan, _ = small_case(dim=dim, n=n, nabs=nabs, nt=nt, kernel='OT2', grad_vp=True)
    print('checkpointed exact adjoint dim=%d K=%d' % (dim, K), flush=True)
    rng = np.random.default_rng(8)
    res = adjoint_source_samples(plan, rng.standard_normal((plan.nrec, plan.nt)))
    s3 = solver_for(plan, save_mode=3)
    s3.forward(plan.src_data)
    a3 = s3.adjoint_exact(res)
    s4 = solver_for(plan, save_mode=4)
    s4.set_scheme(plan.kernel_order, plan.rho_mode, plan.att_power, plan.boundary, plan.diff_source, K, 4,
                  plan.tb0, plan.tb1, 1, 0, 0)
    rec4 = s4.forward(plan.src_data)
    a4 = s4.adjoint_exact(res)
    for key in ('gA', 'gB', 'gG'):
        assert rel_l2(a4[key], a3[key]) < 1e-6, key
    assert rel_l2(a4['lam_src'], a3['lam_src']) < 1e-6
    del rec4


@pytest.mark.parametrize('checkpoint', [None, 9])
def test_autograd_matches_finite_differences(gpu, checkpoint):
    import torch
    from stride_amd.autograd import Geometry, iso_acoustic_traces
    plan, info = small_case(dim=2, n=48, nabs=10, nt=220, kernel='OT2', rho='var', alpha='p0')
    print('autograd checkpoint=%s' % checkpoint, flush=True)
    kw = dict(info['kw'])
    geom = Geometry(info['space'], info['time'], kw.pop('source_coordinates'), kw.pop('receiver_coordinates'),
                    checkpoint_every=checkpoint, **kw)
    vp = torch.tensor(info['vp'], dtype=torch.float32, requires_grad=True)
    rho = torch.tensor(info['rho'], dtype=torch.float32, requires_grad=True)
    alpha = torch.tensor(info['alpha'], dtype=torch.float32, requires_grad=True)
    wav = torch.tensor(info['wavelet'], dtype=torch.float32, requires_grad=True)
    rng = np.random.default_rng(2)
    obs = torch.tensor(rng.standard_normal((plan.nrec, plan.nt)).astype(np.float32))

    def loss_of(vp_, rho_, alpha_, wav_):
        tr = iso_acoustic_traces(vp_, rho_, alpha_, wav_, geom)
        return 0.5 * torch.sum((tr.double() - obs.double()) ** 2)

    loss = loss_of(vp, rho, alpha, wav)
    loss.backward()
    grads = dict(vp=vp.grad.numpy(), rho=rho.grad.numpy(), alpha=alpha.grad.numpy(), wav=wav.grad.numpy())
    base = dict(vp=vp.detach(), rho=rho.detach(), alpha=alpha.detach(), wav=wav.detach())
    eps = dict(vp=4.0, rho=4.0, alpha=0.05, wav=2e-2)
    for name in ('vp', 'rho', 'alpha', 'wav'):
        d = torch.zeros_like(base[name])
        if name == 'wav':
            d[0, 10:80] = torch.tensor(rng.standard_normal(70).astype(np.float32))
        else:
            d[14:54, 14:54] = torch.tensor(rng.standard_normal((40, 40)).astype(np.float32))
        with torch.no_grad():
            args_p = dict(base)
            args_m = dict(base)
            args_p[name] = base[name] + eps[name] * d
            args_m[name] = base[name] - eps[name] * d
            fd = (loss_of(args_p['vp'], args_p['rho'], args_p['alpha'], args_p['wav']) -
                  loss_of(args_m['vp'], args_m['rho'], args_m['alpha'], args_m['wav'])).item() / (2 * eps[name])
        an = float(np.sum(grads[name].astype(np.float64) * d.numpy()))
        print('  %s: fd %.6e autograd %.6e' % (name, fd, an))
        assert abs(fd - an) <= 1e-2 * abs(fd), name


def test_backward_after_interleaved_forward(gpu):
    """A forward with another medium between a forward and its backward (what gradcheck does)
    must not change that backward's gradient: the backward runs its own forward again when the
    solver holds another run's state."""
    import torch
    from stride_amd.autograd import Geometry, iso_acoustic_traces
    plan, info = small_case(dim=2, n=40, nabs=8, nt=160, kernel='OT2')
    print('autograd interleaved forward', flush=True)
    kw = dict(info['kw'])
    geom = Geometry(info['space'], info['time'], kw.pop('source_coordinates'), kw.pop('receiver_coordinates'), **kw)
    rng = np.random.default_rng(5)
    w = torch.tensor(rng.standard_normal((plan.nrec, plan.nt)).astype(np.float32))

    def grad_of(interleave):
        vp = torch.tensor(info['vp'], dtype=torch.float32, requires_grad=True)
        loss = torch.sum(iso_acoustic_traces(vp, None, None, torch.tensor(info['wavelet']), geom) * w)
        if interleave:
            with torch.no_grad():
                other = torch.tensor(info['vp'] * 1.05, dtype=torch.float32)
                iso_acoustic_traces(other, None, None, torch.tensor(info['wavelet']), geom)
        loss.backward()
        return vp.grad.numpy().copy()

    g0 = grad_of(False)
    g1 = grad_of(True)
    assert np.abs(g0).max() > 0
    np.testing.assert_array_equal(g1, g0)


def _autograd_grads(info, geom, device, obs):
    import torch
    from stride_amd.autograd import iso_acoustic_traces
    dev = torch.device(device)
    vp = torch.tensor(info['vp'], dtype=torch.float32, device=dev, requires_grad=True)
    rho = torch.tensor(info['rho'], dtype=torch.float32, device=dev, requires_grad=True)
    alpha = torch.tensor(info['alpha'], dtype=torch.float32, device=dev, requires_grad=True)
    wav = torch.tensor(info['wavelet'], dtype=torch.float32, device=dev, requires_grad=True)
    tr = iso_acoustic_traces(vp, rho, alpha, wav, geom)
    loss = 0.5 * torch.sum((tr.double() - obs.to(dev).double()) ** 2)
    loss.backward()
    return tr.detach().cpu().numpy(), {k: v.grad.cpu().numpy() for k, v in
                                       (('vp', vp), ('rho', rho), ('alpha', alpha), ('wav', wav))}


@pytest.mark.parametrize('checkpoint', [None, 9])
def test_device_path_matches_host_path(gpu, checkpoint):
    """GPU tensors: medium set from device pointers, device accumulators, the chain rule on the
    device; the same traces and gradients as the host-tensor path."""
    import torch
    from stride_amd.autograd import Geometry
    plan, info = small_case(dim=2, n=48, nabs=10, nt=220, kernel='OT2', rho='var', alpha='p0')
    print('autograd device path checkpoint=%s' % checkpoint, flush=True)
    kw = dict(info['kw'])
    src, rec = kw.pop('source_coordinates'), kw.pop('receiver_coordinates')
    obs = torch.tensor(np.random.default_rng(12).standard_normal((plan.nrec, plan.nt)).astype(np.float32))
    tr_h, g_h = _autograd_grads(info, Geometry(info['space'], info['time'], src, rec, checkpoint_every=checkpoint,
                                               **kw), 'cpu', obs)
    geom = Geometry(info['space'], info['time'], src, rec, checkpoint_every=checkpoint, **kw)
    tr_d, g_d = _autograd_grads(info, geom, 'cuda', obs)
    assert rel_l2(tr_d, tr_h) < 1e-6
    for k in g_h:
        assert np.abs(g_h[k]).max() > 0
        assert rel_l2(g_d[k], g_h[k]) < 1e-5, k
    # the cached plan is reused on the next call with the same wavelets and kind of medium
    plan0 = geom._dev_plan
    _autograd_grads(info, geom, 'cuda', obs)
    assert geom._dev_plan is plan0


def test_device_path_workload_grid(gpu):
    """The device path on the 360^3 grid with the variable vp / rho / alpha medium (40 steps,
    checkpoints every 8): gradients equal to the host path's."""
    import torch
    from stride_amd.autograd import Geometry
    from stride_amd.configs import config4
    from stride_amd.problem import Time
    from stride_amd.utils import ellipsoidal, tone_burst
    cfg = config4(scale=1.0, nt=200)
    space, dt = cfg['space'], cfg['time'].step
    cen = np.array([l / 2 for l in space.limit])
    src = cen + 0.8 * (cfg['sources'][0] - cen)
    rec = ellipsoidal(64, (3e-3,) * 3, tuple(src))
    time = Time(start=0., step=dt, num=40)
    info = dict(vp=cfg['vp'], rho=cfg['rho'], alpha=cfg['alpha'], wavelet=tone_burst(400e3, 3, 200, dt)[:40][None])
    obs = torch.tensor(np.random.default_rng(2).standard_normal((64, 40)).astype(np.float32))
    print('autograd device path on %s' % (space.extended_shape,), flush=True)
    kw = dict(kernel='OT2')
    tr_d, g_d = _autograd_grads(info, Geometry(space, time, src[None], rec, checkpoint_every=8, **kw), 'cuda', obs)
    tr_h, g_h = _autograd_grads(info, Geometry(space, time, src[None], rec, checkpoint_every=8, **kw), 'cpu', obs)
    assert rel_l2(tr_d, tr_h) < 1e-6
    for k in g_h:
        assert np.abs(g_h[k]).max() > 0
        assert rel_l2(g_d[k], g_h[k]) < 1e-5, k
