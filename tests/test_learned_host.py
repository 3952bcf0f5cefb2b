"""Host logic of the learned-preconditioner inversion (stride_amd.learned) without a GPU: the
solver is replaced by a quadratic misfit in slowness with its exact gradient, the network by
fixed proposals. A proposal is kept only when the misfit drops; otherwise the plain gradient
step (in slowness, backtracked) is taken; passes are counted (a gradient two, a check one)."""

import numpy as np
import pytest

from stride_amd import learned


class _Space:
    shape = (4, 4)


class _FakeFWI:
    """J(vp) = 0.5 sum (1/vp - 1/true)^2 * k; gradient in slowness."""

    def __init__(self, true, k=1e12):
        self.true = true
        self.k = k
        self.space = _Space()
        self.passes = 0

    def observe(self, true, freq):
        pass

    def _f(self, vp):
        r = 1 / vp.astype(np.float64) - 1 / self.true
        return 0.5 * self.k * float(np.sum(r ** 2)), self.k * r

    def misfit(self, vp, freq):
        self.passes += 1
        return self._f(vp)[0]

    def gradient(self, vp, freq):
        import torch
        self.passes += 2
        f, g = self._f(vp)
        return f, torch.tensor(g, dtype=torch.float32)


class _FakeNet:
    def __init__(self, proposals):
        self.proposals = list(proposals)

    def forward(self, g, scale=1.0, model=None, model_scale=1.0):
        import torch
        return torch.tensor(self.proposals.pop(0)[None] / learned.TARGET_SCALE, dtype=torch.float32)


def test_gated_updates():
    true = np.full((4, 4), 1550., np.float32)
    good = np.full((4, 4), 40., np.float32)      # toward the truth: kept
    bad = np.full((4, 4), -60., np.float32)      # away from it: rejected, plain step taken
    fwi = _FakeFWI(true)
    net = _FakeNet([good, bad, np.full((4, 4), 3., np.float32)])
    vp, hist, accepted = learned.invert_learned(fwi, net, {500e3: 1.0}, true, [500e3], iters_per_freq=3,
                                                first_step=5.0, proposal_scales=(1.0,))
    assert accepted == 2
    fs = [h[1] for h in hist]
    assert all(b < a for a, b in zip(fs, fs[1:]))
    # iteration 1: gradient + check; iteration 2: gradient + rejected check + one plain step
    # (it lowers the misfit at once); iteration 3: gradient + check
    assert [h[0] for h in hist] == [3, 7, 10]
    # the rejected proposal left the model where the plain step put it, closer to the truth
    assert np.all(np.abs(vp - true) < np.abs(1500 + 40 - true) + 1e-3)


def test_plain_step_is_in_slowness():
    vp = np.full((2, 2), 1500., np.float32)
    g = np.full((2, 2), 1e-6)
    out = learned._gradient_step(vp, g, 1.0)
    np.testing.assert_allclose(out, 1 / (1 / 1500. - 1e-6), rtol=1e-6)
    # bounds
    assert learned._gradient_step(vp, np.full((2, 2), -1.0), 1.0).max() == pytest.approx(learned.VP_BOUNDS[0])


def test_phantom_and_partial_model():
    p = learned.phantom(3)
    assert p.shape == (512, 512) and p.dtype == np.float32
    assert p.min() >= 1400 - 1e-3 and p.max() <= 1650 + 1e-3
    assert p[0, 0] == 1500. and p[256, 256] != 1500.
    rng = np.random.default_rng(0)
    m = learned.partial_model(p, rng)
    assert m.shape == p.shape and abs(float(m[0, 0]) - 1500.) < 1.0
    # between water and the phantom on average
    d = (p - 1500.).astype(np.float64)
    frac = float(np.sum((m - 1500.) * d) / np.sum(d * d))
    assert -0.01 < frac < 0.81


def test_input_scales():
    import torch
    pairs = [(torch.full((2, 2), v), None, f, None) for v, f in ((2.0, 1.0), (4.0, 1.0), (8.0, 2.0))]
    s = learned.input_scales(pairs)
    assert s[1.0] == pytest.approx(1 / 3.0) and s[2.0] == pytest.approx(1 / 8.0)
