"""The learned-preconditioner inversion on the 2D gradient grid:

- RingFWI's device-resident misfit and gradient (worker streams, per-shot residuals and
  gradients summed on the device, get_grad_vp's slowness rescale) against one-shot runs of the
  drop-in solver path summed on the host and the oracle's misfit;
- the misfit-gated loop never keeps an update that raises the misfit, whatever the network
  proposes (an untrained network here), and counts its solver passes.
"""

import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu


def test_ring_gradient_matches_shot_sum(gpu):
    from stride_amd import learned
    from stride_amd.physics.gradients import compat_finish
    from stride_amd.physics.plan import adjoint_source_samples
    from helpers import solver_for
    fwi = learned.RingFWI(nsrc=4, workers=2, nt=1200)
    true = learned.phantom(7)
    base = np.full(true.shape, 1500., np.float32)
    print('ring gradient: 4 shots, 572^2, 1200 steps', flush=True)
    fwi.observe(true, 500e3)
    f, g = fwi.gradient(base, 500e3)
    assert fwi.passes == 2
    f_only = fwi.misfit(base, 500e3)
    assert f_only == pytest.approx(f, rel=1e-6)
    # the same shots one at a time through the host-pointer path
    want = np.zeros(fwi.space.extended_shape, np.float64)
    f_want = 0.
    plan_t = fwi._plan(true, 500e3)
    plan_b = fwi._plan(base, 500e3, grad=True)
    for k in range(4):
        s = solver_for(plan_b, save_mode=1)
        s.set_sources(plan_t.src_idx[k:k + 1], plan_t.src_w[k:k + 1])
        s.set_medium(plan_t.vp)
        s.set_pml(plan_t.pml)
        obs = s.forward(np.ascontiguousarray(plan_t.src_data[:, k:k + 1]))
        s.set_medium(plan_b.vp)
        s.set_pml(plan_b.pml)
        mod = s.forward(np.ascontiguousarray(plan_b.src_data[:, k:k + 1]))
        res = mod - obs
        f_want += 0.5 * float(np.sum(res.astype(np.float64) ** 2))
        gk = s.adjoint(adjoint_source_samples(plan_b, res.T), want_vp=True)['grad_vp']
        want += gk
        s.close()
    inner = fwi.space.inner
    gw, _ = compat_finish(want[inner].astype(np.float32), want[inner].astype(np.float32), base, lambda v: 1 / v)
    assert f == pytest.approx(f_want, rel=1e-5)
    assert rel_l2(g.cpu().numpy(), gw) < 1e-5
    fwi.close()


def test_gated_inversion_never_raises_misfit(gpu):
    from stride_amd import learned
    from stride_amd.unet import UNet
    fwi = learned.RingFWI(nsrc=4, workers=2, nt=1200)
    true = learned.phantom(8)
    net = UNet(device=0, seed=5)              # untrained: its proposals are arbitrary
    print('gated inversion with an untrained network', flush=True)
    vp, hist, accepted = learned.invert_learned(fwi, net, {500e3: 1e3}, true, [500e3], iters_per_freq=3)
    f0 = fwi.misfit(np.full(true.shape, 1500., np.float32), 500e3)
    fs = [h[1] for h in hist]
    assert fs[0] <= f0 and all(b <= a for a, b in zip(fs, fs[1:]))
    assert hist[-1][0] >= 3 * 3                  # a gradient (2 passes) and a check (1) per iteration
    assert np.all(vp >= learned.VP_BOUNDS[0]) and np.all(vp <= learned.VP_BOUNDS[1])
    fwi.close()
